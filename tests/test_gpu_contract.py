"""GPU parity: contract! through the C ABI against the oracle.

Fixtures: every label variant of the reference's unit tests.  Random cases:
every kernel family with random shapes, label orders and block sparsity,
dyadic values (so any summation order is exact) -> keys, ids and payloads
bit-exact; plus normal-distributed values within 1e-12 relative Frobenius
(the FP64 drop-in tolerance) for the GEMM families.
"""
import numpy as np
import pytest

import fixture_cases as FC
from both import run_both, compare_bs, compare_coo, dev_tensor, orc_tensor
from refs import fixtures, dense, labs, einsum_contract
from sb_flags import EXACT_JOIN, SKIP_ZERO_SLICES

pytestmark = pytest.mark.gpu
FX = fixtures()


# ---------------------------------------------------------------- reference fixtures
@pytest.mark.parametrize("case", list(FC.bs_dense_prefix_cases()), ids=lambda c: c["name"])
def test_fixture_bs_dense_prefix(case):
    d, o = run_both("bs", case["C"].shape, case["N2C"], case["lc"], "bs", case["A"], case["N2A"], case["la"], "dense",
                    case["B"], None, case["lb"])
    compare_bs(d, o)


@pytest.mark.parametrize("key", ["bs_bs_prefix", "bs_bs_dense"])
def test_fixture_bs_bs_single(key):
    for case in FC.bs_bs_cases(key):
        d, o = run_both("bs", case["C"].shape, case["N2C"], case["lc"], "bs", case["A"], case["N2A"], case["la"], "bs",
                        case["B"], case["N2B"], case["lb"])
        compare_bs(d, o)


@pytest.mark.parametrize("flags", [0, EXACT_JOIN])
def test_fixture_bs_bs_multi(flags):
    for case in FC.bs_bs_cases("bs_bs_multi"):
        d, o = run_both("bs", case["C"].shape, case["N2C"], case["lc"], "bs", case["A"], case["N2A"], case["la"], "bs",
                        case["B"], case["N2B"], case["lb"], flags)
        compare_bs(d, o)


def test_fixture_interleaved_error():
    import sparsebackends_amd as sb
    f = FX["bs_bs_interleaved"]
    A = dense(f["A"]["shape"], f["A"]["entries"])
    B = dense(f["B"]["shape"], f["B"]["entries"])
    with pytest.raises(sb.SBError) as e:
        run_both("bs", f["C_shape"], f["N2C"], f["labelsC"], "bs", A, f["N2A"], f["A"]["labels"], "bs", B, f["N2B"],
                 f["B"]["labels"])
    assert e.value.kind == "ErrorException"


def test_fixture_coo_coo():
    for case in FC.coo_coo_cases():
        d, o = run_both("coo", case["C_shape"], None, case["lc"], "coo", case["A"], None, case["la"], "coo", case["B"],
                        None, case["lb"])
        compare_coo(d, o)
        d.sort_()
        assert np.array_equal(d.to_dense(), case["C"])


@pytest.mark.parametrize("flags", [0, SKIP_ZERO_SLICES])
def test_fixture_coo_dense(flags):
    for case in FC.coo_dense_cases():
        d, o = run_both("bs", case["C"].shape, case["N2C"], case["lc"], "coo", case["A"], None, case["la"], "dense",
                        case["B"], None, case["lb"], flags)
        compare_bs(d, o)
    f = FX["coo_dense"]["config3_zero_slice"]
    A = dense(f["A"]["shape"], f["A"]["entries"])
    B = dense(f["B"]["shape"], f["B"]["entries"])
    d, o = run_both("bs", (3, 5), 1, f["labelsC"], "coo", A, None, f["A"]["labels"], "dense", B, None,
                    f["B"]["labels"], flags)
    compare_bs(d, o)
    assert d.nblocks == (0 if flags & SKIP_ZERO_SLICES else 1)


def test_fixture_coo_bs():
    for case in FC.coo_bs_cases():
        d, o = run_both("bs", case["C"].shape, case["N2C"], case["lc"], "coo", case["A"], None, case["la"], "bs",
                        case["B"], case["N2B"], case["lb"])
        compare_bs(d, o)


from gen_cases import dyadic, gauss, gen_shared, gen_single


def _run_case(c, flags=0):
    return run_both(c["ck"], c["cdims"], c["N2C"], c["lc"], c["ak"], c["A"], c["N2A"], c["la"], c["bk"], c["B"],
                    c["N2B"], c["lb"], flags)


@pytest.mark.parametrize("b_kind", ["bs", "dense"])
@pytest.mark.parametrize("seed", range(12))
def test_random_shared(b_kind, seed):
    rng = np.random.default_rng(1000 + seed)
    c = gen_shared(rng, b_kind)
    for flags in ([0, EXACT_JOIN] if b_kind == "bs" else [0]):
        d, o = run_both("bs", c["cdims"], c["N2C"], c["lc"], "bs", c["A"], c["N2A"], c["la"], b_kind, c["B"],
                        c["N2B"], c["lb"], flags)
        compare_bs(d, o)
    if b_kind == "dense" or True:
        # the exact join equals the dense mathematics
        d, _ = run_both("bs", c["cdims"], c["N2C"], c["lc"], "bs", c["A"], c["N2A"], c["la"], b_kind, c["B"],
                        c["N2B"], c["lb"], EXACT_JOIN)
        ref = einsum_contract(c["A"], c["la"], c["B"], c["lb"], c["lc"])
        assert np.array_equal(d.to_dense(), ref)


FAMILIES = ["prefix_bb", "dense_bb", "prefix_bd", "dense_bd", "coo_coo", "coo_bs_prefix", "coo_bs_dense", "coo_dense"]


@pytest.mark.parametrize("fam", FAMILIES)
@pytest.mark.parametrize("seed", range(8))
def test_random_single(fam, seed):
    rng = np.random.default_rng(2000 + seed)
    c = gen_single(rng, fam)
    d, o = _run_case(c)
    if c["ck"] == "coo":
        compare_coo(d, o)
    else:
        compare_bs(d, o)


@pytest.mark.parametrize("b_kind", ["bs", "dense"])
@pytest.mark.parametrize("seed", range(6))
def test_random_shared_gauss(b_kind, seed):
    """normal values: GEMM summation order differs from BLAS; FP64 bar 1e-12"""
    rng = np.random.default_rng(3000 + seed)
    c = gen_shared(rng, b_kind, gen=gauss)
    d, o = run_both("bs", c["cdims"], c["N2C"], c["lc"], "bs", c["A"], c["N2A"], c["la"], b_kind, c["B"], c["N2B"],
                    c["lb"])
    compare_bs(d, o, rtol=1e-12)


@pytest.mark.parametrize("fam", ["prefix_bd", "coo_dense", "coo_bs_prefix", "prefix_bb"])
def test_exact_order_kernels_bit_exact_gauss(fam):
    """Kahan / axpy / outer kernels repeat the reference's operation order:
    bit-identical to the oracle even for non-dyadic values"""
    rng = np.random.default_rng(4000)
    for _ in range(4):
        c = gen_single(rng, fam, gen=gauss)
        d, o = _run_case(c)
        compare_bs(d, o)


def test_empty_operands():
    import sparsebackends_amd as sb
    A = np.zeros((3, 4, 2), order="F")
    B = np.ones((2, 4, 3), order="F")
    d, o = run_both("bs", (3, 2, 2, 3), 3, ["i", "xa", "j", "xb"], "bs", A, 1, ["i", "r", "xa"], "dense", B, None,
                    ["j", "r", "xb"])
    compare_bs(d, o)
    assert d.nblocks == 0


def test_dispatch_errors_gpu():
    import sparsebackends_amd as sb
    A = sb.blocksparse_from_dense(np.ones((2, 3), order="F"), 1)
    B = sb.blocksparse_from_dense(np.ones((2, 3), order="F"), 1)
    C = sb.NewBlockSparseSorted.new((3, 3), 2)
    with pytest.raises(sb.SBError) as e:
        sb.contract_(C, labs(["x", "y"]), A, labs(["r"]), B, labs(["r", "y"]))
    assert e.value.kind == "AssertionError"
    Ac = sb.coo_from_dense(np.ones((2, 2), order="F"))
    Bc = sb.coo_from_dense(np.ones((2, 2), order="F"))
    Cc = sb.COOTensor.new((1,))
    with pytest.raises(sb.SBError) as e:
        sb.contract_(Cc, labs(["z"]), Ac, labs(["a", "b"]), Bc, labs(["a", "b"]))
    assert e.value.kind == "MethodError"
