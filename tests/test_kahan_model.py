"""Pins the values of contract_prefix_outer_bd! (BS x Dense, reduced label in
A's prefix) bit for bit on inexact inputs: the oracle against
tests/kahan_model.py, a second restatement of the reference's group order,
in-group block order and compensated update written from the reference code
alone.  Plausible misreadings (another in-group order, a plain sum, a
compensation kept across groups) are shown to change the bits, so the
comparison can tell them apart."""
import numpy as np
import pytest

from oracle import oracle as O
from gen_cases import gen_single, gauss
from kahan_model import prefix_outer_bd, blocks_of
from refs import labs


def cases(n=60):
    k = 0
    for seed in range(10 * n):
        c = gen_single(np.random.default_rng(7000 + seed), "prefix_bd", gen=gauss)
        yield c
        k += 1
        if k == n:
            return


def oracle_blocks(c):
    Co = O.BS.new(c["cdims"], c["N2C"])
    O.contract(Co, labs(c["lc"]), O.BS.from_dense(c["A"], c["N2A"]), labs(c["la"]), O.Dense.new(c["B"]),
               labs(c["lb"]), 0)
    return blocks_of(Co.parts())


def test_oracle_equals_model_bitwise():
    multi = 0
    for c in cases():
        want = prefix_outer_bd(c["A"], c["la"], c["N2A"], c["B"], c["lb"], c["lc"], c["N2C"])
        got = oracle_blocks(c)
        assert set(got) == set(want)
        for key in want:
            assert np.array_equal(got[key], want[key]), key
        # cases where a C block sums more than one A block
        r = [l for l in c["la"] if l in c["lb"]][0]
        if c["A"].shape[c["la"].index(r)] > 1:
            multi += 1
    assert multi > 20


def _reversed_order(A, la, N2A, B, lb, lc, N2C):
    """a misreading: blocks of a group summed in descending r"""
    import kahan_model as K
    flipped = np.flip(A, axis=la.index([l for l in la if l in lb][0]))
    Bf = np.flip(B, axis=lb.index([l for l in la if l in lb][0]))
    return K.prefix_outer_bd(flipped, la, N2A, Bf, lb, lc, N2C)


def _plain_sum(A, la, N2A, B, lb, lc, N2C):
    """a misreading: the products summed without compensation"""
    import kahan_model as K
    PA = len(la) - N2A
    r = [l for l in la if l in lb][0]
    blocks = K._blocks(A, PA)
    pre = list(la[:PA])
    ra = pre.index(r)
    Bp = np.moveaxis(B, lb.index(r), -1)
    Bmat = Bp.reshape(-1, Bp.shape[-1], order="F")
    src = [pre.index(l) for l in lc[:len(lc) - N2C]]
    ab = list(lc[len(lc) - N2C:]) == list(la[PA:]) + [l for l in lb if l != r]
    sums = {}
    for k in sorted(blocks, key=lambda k: k[ra]):
        ck = tuple(k[j] for j in src)
        prod = np.multiply.outer(blocks[k], Bmat[:, k[ra] - 1])
        sums[ck] = prod if ck not in sums else sums[ck] + prod
    return {ck: (v.reshape(-1, order="F") if ab else v.T.reshape(-1, order="F")) for ck, v in sums.items()}


@pytest.mark.parametrize("wrong", [_reversed_order, _plain_sum])
def test_misreadings_change_bits(wrong):
    differ = 0
    for c in cases():
        want = prefix_outer_bd(c["A"], c["la"], c["N2A"], c["B"], c["lb"], c["lc"], c["N2C"])
        bad = wrong(c["A"], c["la"], c["N2A"], c["B"], c["lb"], c["lc"], c["N2C"])
        differ += any(not np.array_equal(bad[k], want[k]) for k in want)
    assert differ > 5
