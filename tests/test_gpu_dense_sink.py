"""GPU: contract_bs_bs_to_dense! (contract_bs_bs.jl:736-754) through the C ABI
(sb_contract_bs_bs_to_dense) against einsum and the oracle's P = 0
accumulator: full dense outputs, the NC = 0 scalar, and a contraction without
matching blocks, which leaves C untouched."""
import numpy as np
import pytest

from dense_sink_cases import cases, empty_case
from refs import labs, einsum_contract
from test_oracle_dense_sink import orc_sink

pytestmark = pytest.mark.gpu


def dev_sink(A, la, N2A, B, lb, N2B, lc, init):
    import sparsebackends_amd as sb
    Cd = sb.DenseTensor.from_numpy(np.array(init, dtype=np.float64, order="F"))  # keeps a 0-d array 0-d
    sb.contract_bs_bs_to_dense_(Cd, labs(lc), sb.blocksparse_from_dense(A, N2A), labs(la),
                                sb.blocksparse_from_dense(B, N2B), labs(lb))
    return Cd.numpy()


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_gpu_sink(case):
    name, A, la, N2A, B, lb, N2B, lc = case
    print("case", name, flush=True)
    ref = einsum_contract(A, la, B, lb, lc)
    init = np.full(ref.shape, 7.0)
    got = dev_sink(A, la, N2A, B, lb, N2B, lc, init)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    assert np.array_equal(got, orc_sink(A, la, N2A, B, lb, N2B, lc, init))


def test_gpu_sink_empty_leaves_c():
    name, A, la, N2A, B, lb, N2B, lc = empty_case()
    init = np.arange(8, dtype=np.float64).reshape(2, 4) + 0.5
    got = dev_sink(A, la, N2A, B, lb, N2B, lc, init)
    assert np.array_equal(got, init)


def test_gpu_sink_gauss_within_fp64_bar():
    """normal values through the block GEMM: 1e-12 relative (the FP64 drop-in bar)"""
    rng = np.random.default_rng(8100)
    A = rng.standard_normal((4, 6, 5))
    B = rng.standard_normal((4, 5, 3))
    la, lb, lc = ["s", "x", "k"], ["s", "k", "y"], ["x", "y"]
    ref = einsum_contract(A, la, B, lb, lc)
    got = dev_sink(A, la, 2, B, lb, 2, lc, np.zeros(ref.shape))
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_gpu_sink_rejects_non_blocksparse():
    import sparsebackends_amd as sb
    A = np.ones((2, 3))
    Cd = sb.DenseTensor.from_numpy(np.zeros(3))
    with pytest.raises(sb.SBError) as e:
        sb.contract_bs_bs_to_dense_(Cd, labs(["x"]), sb.blocksparse_from_dense(A, 1), labs(["s", "x"]),
                                    sb.DenseTensor.from_numpy(np.ones(2)), labs(["s"]))
    assert e.value.kind == "MethodError"
