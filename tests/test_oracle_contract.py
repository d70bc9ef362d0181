"""Pins the oracle's contract! against the reference's own unit-test fixtures
(test/tensoralgebra/*.jl) with their closed-form dense references, exact."""
import numpy as np
import pytest

from oracle import oracle as O
from refs import labs, assert_bs_equal, einsum_contract, fixtures, dense, lab
import fixture_cases as FC
from sb_flags import EXACT_JOIN, SKIP_ZERO_SLICES

FX = fixtures()


def run_bs(A, la, N2A, B, lb, N2B, lc, N2C, Cshape, flags=0, B_dense=False):
    Ao = O.BS.from_dense(A, N2A)
    Bo = O.Dense.new(B) if B_dense else O.BS.from_dense(B, N2B)
    Co = O.BS.new(Cshape, N2C)
    O.contract(Co, labs(lc), Ao, labs(la), Bo, labs(lb), flags)
    return Co


@pytest.mark.parametrize("case", list(FC.bs_dense_prefix_cases()), ids=lambda c: c["name"])
def test_bs_dense_prefix(case):
    C = run_bs(case["A"], case["la"], case["N2A"], case["B"], case["lb"], 0, case["lc"], case["N2C"],
               case["C"].shape, B_dense=True)
    ref = O.BS.from_dense(case["C"], case["N2C"])
    assert_bs_equal(C.parts(), ref.parts())


@pytest.mark.parametrize("case", list(FC.bs_bs_cases("bs_bs_prefix")), ids=lambda c: c["name"])
def test_bs_bs_prefix(case):
    C = run_bs(case["A"], case["la"], case["N2A"], case["B"], case["lb"], case["N2B"], case["lc"], case["N2C"],
               case["C"].shape)
    assert_bs_equal(C.parts(), O.BS.from_dense(case["C"], case["N2C"]).parts())


@pytest.mark.parametrize("case", list(FC.bs_bs_cases("bs_bs_dense")), ids=lambda c: c["name"])
def test_bs_bs_dense(case):
    C = run_bs(case["A"], case["la"], case["N2A"], case["B"], case["lb"], case["N2B"], case["lc"], case["N2C"],
               case["C"].shape)
    assert_bs_equal(C.parts(), O.BS.from_dense(case["C"], case["N2C"]).parts())


@pytest.mark.parametrize("flags", [0, EXACT_JOIN])
@pytest.mark.parametrize("case", list(FC.bs_bs_cases("bs_bs_multi")), ids=lambda c: c["name"])
def test_bs_bs_multi(case, flags):
    C = run_bs(case["A"], case["la"], case["N2A"], case["B"], case["lb"], case["N2B"], case["lc"], case["N2C"],
               case["C"].shape, flags=flags)
    assert_bs_equal(C.parts(), O.BS.from_dense(case["C"], case["N2C"]).parts())


def test_bs_bs_interleaved_error():
    f = FX["bs_bs_interleaved"]
    A = dense(f["A"]["shape"], f["A"]["entries"])
    B = dense(f["B"]["shape"], f["B"]["entries"])
    with pytest.raises(O.OracleError) as e:
        run_bs(A, f["A"]["labels"], f["N2A"], B, f["B"]["labels"], f["N2B"], f["labelsC"], f["N2C"], f["C_shape"])
    assert e.value.kind == "ErrorException"


def test_multi_shared_prefix_join_order_divergence():
    """Two shared labels in the key prefix: the reference's merge-join compares
    the first shared axis first although keys are sorted with the last one most
    significant (contract_bs_bs.jl:430-442), so it can miss matches.  Pinned
    with a hand-constructed case whose dense reference is known."""
    # A(i, s1, s2 | x) with blocks at (s1,s2) = (2,1) and (1,2); B(j, s1, s2 | y) at (1,2)
    A = np.zeros((1, 2, 2, 1), order="F")
    A[0, 1, 0, 0] = 1.0
    A[0, 0, 1, 0] = 2.0
    B = np.zeros((1, 2, 2, 1), order="F")
    B[0, 0, 1, 0] = 3.0
    la, lb, lc = ["i", "s1", "s2", "x"], ["j", "s1", "s2", "y"], ["i", "j", "x", "y"]
    ref = einsum_contract(A, la, B, lb, lc)
    exact = run_bs(A, la, 1, B, lb, 1, lc, 2, ref.shape, flags=EXACT_JOIN)
    assert_bs_equal(exact.parts(), O.BS.from_dense(ref, 2).parts())
    quirk = run_bs(A, la, 1, B, lb, 1, lc, 2, ref.shape, flags=0)
    assert len(quirk.parts()["ids"]) == 0  # the reference's walk skips the (1,2) match


@pytest.mark.parametrize("case", list(FC.coo_coo_cases()), ids=lambda c: c["name"])
def test_coo_coo(case):
    Ao = O.COO.from_dense(case["A"], atol=0.0)
    Bo = O.COO.from_dense(case["B"], atol=0.0)
    Co = O.COO.new(case["C_shape"])
    O.contract(Co, labs(case["lc"]), Ao, labs(case["la"]), Bo, labs(case["lb"]))
    assert Co.parts()["dirty"]
    Co.sort_()
    ref = O.COO.from_dense(case["C"], atol=0.0).parts()
    got = Co.parts()
    assert got["keys"].tolist() == ref["keys"].tolist()
    assert got["vals"].tolist() == ref["vals"].tolist()
    assert np.array_equal(Co.to_dense(), case["C"])


@pytest.mark.parametrize("case", list(FC.coo_dense_cases()), ids=lambda c: c["name"])
def test_coo_dense(case):
    Ao = O.COO.from_dense(case["A"], atol=0.0)
    Bo = O.Dense.new(case["B"])
    Co = O.BS.new(case["C"].shape, case["N2C"])
    O.contract(Co, labs(case["lc"]), Ao, labs(case["la"]), Bo, labs(case["lb"]))
    got = Co.parts()
    ref = O.BS.from_dense(case["C"], case["N2C"]).parts()
    # coo_dense.jl compares key sets and payloads (order-free)
    gk = {tuple(k) for k in got["keys"].tolist()}
    assert gk == {tuple(k) for k in ref["keys"].tolist()}
    assert_bs_equal(got, ref)


def test_coo_dense_zero_slice_rule():
    f = FX["coo_dense"]["config3_zero_slice"]
    A = dense(f["A"]["shape"], f["A"]["entries"])
    B = dense(f["B"]["shape"], f["B"]["entries"])
    la, lb, lc = f["A"]["labels"], f["B"]["labels"], f["labelsC"]
    for flags, nblocks in ((0, 1), (SKIP_ZERO_SLICES, 0)):
        Co = O.BS.new((3, 5), 1)
        O.contract(Co, labs(lc), O.COO.from_dense(A, atol=0.0), labs(la), O.Dense.new(B), labs(lb), flags)
        p = Co.parts()
        assert len(p["ids"]) == nblocks  # reference code (flags=0) creates an all-zero block
        assert not np.any(p["data"])


@pytest.mark.parametrize("case", list(FC.coo_bs_cases()), ids=lambda c: c["name"])
def test_coo_bs(case):
    Ao = O.COO.from_dense(case["A"], atol=0.0)
    Bo = O.BS.from_dense(case["B"], case["N2B"])
    Co = O.BS.new(case["C"].shape, case["N2C"])
    O.contract(Co, labs(case["lc"]), Ao, labs(case["la"]), Bo, labs(case["lb"]))
    assert_bs_equal(Co.parts(), O.BS.from_dense(case["C"], case["N2C"]).parts())


def test_dispatch_errors():
    A = O.BS.from_dense(np.ones((2, 3), order="F"), 1)
    B = O.BS.from_dense(np.ones((2, 3), order="F"), 1)
    C = O.BS.new((3, 3), 2)
    # labels of the wrong length -> AssertionError
    with pytest.raises(O.OracleError) as e:
        O.contract(C, labs(["x", "y"]), A, labs(["r"]), B, labs(["r", "y"]))
    assert e.value.kind == "AssertionError"
    # two reduced labels in COO -> no contract_shared! method
    Ac = O.COO.from_dense(np.ones((2, 2), order="F"))
    Bc = O.COO.from_dense(np.ones((2, 2), order="F"))
    Cc = O.COO.new((1,))
    with pytest.raises(O.OracleError) as e:
        O.contract(Cc, labs(["z"]), Ac, labs(["a", "b"]), Bc, labs(["a", "b"]))
    assert e.value.kind in ("MethodError", "AssertionError")


# ---------------------------------------------------------------- random brute force
from gen_cases import gen_shared, gen_single, gauss  # noqa: E402


def _orc_case(c, flags=0):
    mk = {"bs": lambda a, n2: O.BS.from_dense(a, n2), "coo": lambda a, n2: O.COO.from_dense(a, atol=0.0),
          "dense": lambda a, n2: O.Dense.new(a)}
    Ct = O.BS.new(c["cdims"], c["N2C"]) if c["ck"] == "bs" else O.COO.new(c["cdims"])
    At = mk[c["ak"]](c["A"], c["N2A"])
    Bt = mk[c["bk"]](c["B"], c["N2B"])
    O.contract(Ct, labs(c["lc"]), At, labs(c["la"]), Bt, labs(c["lb"]), flags)
    return Ct


def _coalesced_dense(Ct, shape):
    p = Ct.parts()
    out = np.zeros(shape, order="F")
    for k, v in zip(p["keys"], p["vals"]):
        out[tuple(int(x) - 1 for x in k)] += v
    return out


@pytest.mark.parametrize("b_kind", ["bs", "dense"])
@pytest.mark.parametrize("seed", range(40))
def test_random_shared_bruteforce(b_kind, seed):
    rng = np.random.default_rng(1000 + seed)
    c = gen_shared(rng, b_kind)
    c.update(ck="bs", ak="bs", bk=b_kind)
    ref = einsum_contract(c["A"], c["la"], c["B"], c["lb"], c["lc"])
    Ct = _orc_case(c, EXACT_JOIN)
    assert np.array_equal(Ct.to_dense(), ref)
    # block list: exactly the keys the reference creates (every pair / every A block)
    Ct0 = _orc_case(c, 0)
    if b_kind == "dense":
        assert np.array_equal(Ct0.to_dense(), ref)


FAMS = ["prefix_bb", "dense_bb", "prefix_bd", "dense_bd", "coo_coo", "coo_bs_prefix", "coo_bs_dense", "coo_dense"]


@pytest.mark.parametrize("fam", FAMS)
@pytest.mark.parametrize("seed", range(20))
def test_random_single_bruteforce(fam, seed):
    rng = np.random.default_rng(2000 + seed)
    c = gen_single(rng, fam)
    ref = einsum_contract(c["A"], c["la"], c["B"], c["lb"], c["lc"])
    Ct = _orc_case(c)
    got = _coalesced_dense(Ct, c["cdims"]) if c["ck"] == "coo" else Ct.to_dense()
    assert np.array_equal(got, ref)


def test_kahan_matches_triple_loop_bound():
    """test_kernel_precision.jl:47-70 style: the Kahan path stays within a few
    ulps of an exactly rounded (math.fsum) outer-product sum"""
    import math
    rng = np.random.default_rng(42)
    for _ in range(5):
        c = gen_single(rng, "prefix_bd", gen=gauss)
        Ct = _orc_case(c)
        ref = einsum_contract(c["A"], c["la"], c["B"], c["lb"], c["lc"])
        got = Ct.to_dense()
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-14)
