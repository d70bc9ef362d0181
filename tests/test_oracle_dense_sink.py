"""The P_C = 0 dense sink (contract_bs_bs.jl:736-754) on the oracle: its P = 0
accumulator equals the dense mathematics, exactly on dyadic values, including
the NC = 0 scalar, and a contraction without matches produces no block."""
import numpy as np
import pytest

from oracle import oracle as O
from dense_sink_cases import cases, empty_case
from refs import labs, einsum_contract


def orc_sink(A, la, N2A, B, lb, N2B, lc, init):
    """the reference's sink restated over the oracle: P = 0 accumulator, then
    copy its single block into C, or leave C when there is none"""
    dims = tuple(dict(zip(la, A.shape), **dict(zip(lb, B.shape)))[l] for l in lc)
    acc = O.BS.new(dims, len(dims))
    O.contract(acc, labs(lc), O.BS.from_dense(A, N2A), labs(la), O.BS.from_dense(B, N2B), labs(lb))
    p = acc.parts()
    C = np.array(init, dtype=np.float64, copy=True)
    if len(p["ids"]):
        assert len(p["ids"]) == 1 and int(p["ids"][0]) == 1
        C[...] = np.asarray(p["data"][:p["blksize"]]).reshape(dims, order="F")
    return C


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_sink_equals_einsum(case):
    name, A, la, N2A, B, lb, N2B, lc = case
    ref = einsum_contract(A, la, B, lb, lc)
    got = orc_sink(A, la, N2A, B, lb, N2B, lc, np.full(ref.shape, 7.0))
    assert np.array_equal(got, ref)


def test_sink_empty_leaves_c():
    name, A, la, N2A, B, lb, N2B, lc = empty_case()
    init = np.full((2, 4), 7.0)
    got = orc_sink(A, la, N2A, B, lb, N2B, lc, init)
    assert np.array_equal(got, init)
