"""Pins the block-creation order (C.ids) of every contraction family.

The oracle's ids are compared with tests/ids_model.py, a second restatement of
the reference's traversals written from the reference code alone (it shares
nothing with the C++ oracle), on every reference fixture and on random cases
of every family: a misreading of a traversal would have to be made twice, in
two languages, to pass.  Also: blocksparse_from_dense and permutedims give
ids 1..n (conversions.jl:47-72; storage.jl:209)."""
import numpy as np
import pytest

from oracle import oracle as O
import fixture_cases as FC
from gen_cases import gen_shared, gen_single, dyadic
from ids_model import creation_order, check_ids
from refs import labs
from sb_flags import EXACT_JOIN, SKIP_ZERO_SLICES


def _orc(kind, arr, N2, dims=None):
    if kind == "bs":
        return O.BS.from_dense(arr, N2) if arr is not None else O.BS.new(dims, N2)
    if kind == "coo":
        return O.COO.from_dense(arr) if arr is not None else O.COO.new(dims)
    return O.Dense.new(arr)


def run_and_check(ck, cdims, N2C, lc, ak, A, N2A, la, bk, B, N2B, lb, flags=0):
    Co = _orc(ck, None, N2C, cdims)
    O.contract(Co, labs(lc), _orc(ak, A, N2A), labs(la), _orc(bk, B, N2B), labs(lb), flags)
    exp = creation_order(ck, lc, N2C, ak, A, la, N2A, bk, B, lb, N2B, exact_join=bool(flags & EXACT_JOIN),
                         skip_zero_slices=bool(flags & SKIP_ZERO_SLICES))
    p = Co.parts()
    if ck == "coo":
        got = [tuple(int(x) for x in k) for k in np.asarray(p["keys"]).reshape(-1, len(lc))]
        assert got == exp
    else:
        check_ids(p, exp)
    return len(exp)


def fixture_calls():
    for c in FC.bs_dense_prefix_cases():
        yield ("bs", c["C"].shape, c["N2C"], c["lc"], "bs", c["A"], c["N2A"], c["la"], "dense", c["B"], None,
               c["lb"], 0)
    for key in ("bs_bs_prefix", "bs_bs_dense"):
        for c in FC.bs_bs_cases(key):
            yield ("bs", c["C"].shape, c["N2C"], c["lc"], "bs", c["A"], c["N2A"], c["la"], "bs", c["B"], c["N2B"],
                   c["lb"], 0)
    for flags in (0, EXACT_JOIN):
        for c in FC.bs_bs_cases("bs_bs_multi"):
            yield ("bs", c["C"].shape, c["N2C"], c["lc"], "bs", c["A"], c["N2A"], c["la"], "bs", c["B"], c["N2B"],
                   c["lb"], flags)
    for c in FC.coo_coo_cases():
        yield ("coo", c["C_shape"], None, c["lc"], "coo", c["A"], None, c["la"], "coo", c["B"], None, c["lb"], 0)
    for flags in (0, SKIP_ZERO_SLICES):
        for c in FC.coo_dense_cases():
            yield ("bs", c["C"].shape, c["N2C"], c["lc"], "coo", c["A"], None, c["la"], "dense", c["B"], None,
                   c["lb"], flags)
    for c in FC.coo_bs_cases():
        yield ("bs", c["C"].shape, c["N2C"], c["lc"], "coo", c["A"], None, c["la"], "bs", c["B"], c["N2B"], c["lb"], 0)


FAMILIES = ["prefix_bb", "dense_bb", "prefix_bd", "dense_bd", "coo_coo", "coo_bs_prefix", "coo_bs_dense", "coo_dense"]


def random_calls(nshared=40, nsingle=24):
    for b_kind in ("bs", "dense"):
        for seed in range(nshared):
            c = gen_shared(np.random.default_rng(5000 + seed), b_kind)
            for flags in ((0, EXACT_JOIN) if b_kind == "bs" else (0,)):
                yield ("bs", c["cdims"], c["N2C"], c["lc"], "bs", c["A"], c["N2A"], c["la"], b_kind, c["B"],
                       c["N2B"], c["lb"], flags)
    for fam in FAMILIES:
        for seed in range(nsingle):
            c = gen_single(np.random.default_rng(6000 + seed), fam)
            for flags in ((0, SKIP_ZERO_SLICES) if fam == "coo_dense" else (0,)):
                yield (c["ck"], c["cdims"], c["N2C"], c["lc"], c["ak"], c["A"], c["N2A"], c["la"], c["bk"], c["B"],
                       c["N2B"], c["lb"], flags)


def test_fixtures_creation_order():
    n = [run_and_check(*call) for call in fixture_calls()]
    assert sum(n) > 0 and sum(1 for x in n if x > 1) > 20  # orders with more than one block were compared


def test_random_creation_order():
    n = [run_and_check(*call) for call in random_calls()]
    assert sum(1 for x in n if x > 2) > 100


def _misreadings():
    """plausible misreadings of the traversals, each a replacement of one model function"""
    import ids_model as M
    bd, walk, bsd = M._prefix_outer_bd, M._join_walk, M._shared_bs_dense

    def bd_colmajor(A, lc, PC, rlab):  # C keys in column-major instead of lexicographic order
        return sorted(bd(A, lc, PC, rlab), key=M._cm_key)

    def walk_b_outer(A, B, posA, posB, exact):  # B blocks outside A blocks within a matched run
        return sorted(walk(A, B, posA, posB, exact), key=lambda p: (p[1], p[0]))

    def bsd_colmajor(A, lc, PC, shared):  # one block per C key in key order, not A's order
        return sorted(bsd(A, lc, PC, shared), key=M._cm_key)

    return [("_prefix_outer_bd", bd_colmajor), ("_join_walk", walk_b_outer), ("_shared_bs_dense", bsd_colmajor)]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_a_planted_misreading_fails(which, monkeypatch):
    """The comparison has teeth: each plausible misreading of a traversal, put
    into the model, disagrees with the oracle's ids on some case."""
    import ids_model as M
    name, fn = _misreadings()[which]
    monkeypatch.setattr(M, name, fn)
    fails = 0
    for call in list(fixture_calls()) + list(random_calls()):
        try:
            run_and_check(*call)
        except AssertionError:
            fails += 1
    assert fails > 0, f"misreading {name} went unnoticed"


@pytest.mark.parametrize("seed", range(6))
def test_from_dense_and_permute_ids_are_1_to_n(seed):
    rng = np.random.default_rng(7000 + seed)
    A = dyadic(rng, (3, 2, 3, 2, 2), P=3)
    t = O.BS.from_dense(A, 2)
    p = t.parts()
    assert list(p["ids"]) == list(range(1, len(p["ids"]) + 1))
    perm = list(rng.permutation(3)) + [3 + int(x) for x in rng.permutation(2)]
    q = t.permutedims([x + 1 for x in perm]).parts()
    assert list(q["ids"]) == list(range(1, len(q["ids"]) + 1))
