"""GPU parity of AliasedBlockSparse: storage operations and every aliased
contraction kernel through the C ABI against the oracle.

Keys, template counts and alias ids bit-exact (combined templates in
first-seen order, accumulated blocks in key order on both sides); scalars
bit-exact (same sequential sums); templates bit-exact where the reference
forms them as outer products or copies, and for dyadic values everywhere,
within 1e-12 relative Frobenius for block products of normal values."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from oracle import oracle as O
from aliased_cases import aliased_parts, logical_dense, gen_single, gen_shared
from refs import labs, rel_frobenius, einsum_contract
from sb_flags import EXACT_JOIN

pytestmark = pytest.mark.gpu


def dev(p):
    return sb.AliasedBlockSparse.from_parts(p["dims"], p["N2"], p["templates"], p["ntemplates"], p["keys"],
                                            p["alias_ids"], p["scalars"])


def orc(p):
    return O.AL.from_parts(p["dims"], p["N2"], p["templates"], p["ntemplates"], p["keys"], p["alias_ids"],
                           p["scalars"])


def same(p, q, rtol=0.0):
    assert tuple(p["dims"]) == tuple(q["dims"])
    assert p["ntemplates"] == q["ntemplates"], (p["ntemplates"], q["ntemplates"])
    assert np.array_equal(p["keys"], q["keys"])
    assert np.array_equal(p["alias_ids"], q["alias_ids"])
    assert np.array_equal(p["scalars"], q["scalars"])
    if rtol == 0.0:
        assert np.array_equal(p["templates"], q["templates"])
    else:
        assert rel_frobenius(p["templates"], q["templates"]) <= rtol


def test_storage_ops():
    rng = np.random.default_rng(0)
    for t in range(6):
        p = aliased_parts(rng, [3, 4, 2], [2, 3], 10, 3)
        A, Ao = dev(p), orc(p)
        same(A.parts(), Ao.parts())
        assert np.array_equal(A.to_dense(), logical_dense(p))
        assert np.array_equal(A.to_dense(init=1.5), Ao.to_dense(init=1.5))
        bs, bo = A.to_blocksparse().parts(), Ao.to_blocksparse().parts()
        assert np.array_equal(bs["keys"], bo["keys"]) and np.array_equal(bs["ids"], bo["ids"])
        assert np.array_equal(bs["data"], bo["data"])
        assert A.compression_ratio() == Ao.compression_ratio()
        for _ in range(8):
            I = tuple(int(rng.integers(1, d + 1)) for d in p["dims"])
            assert A[I] == Ao.getindex(*I)
        perm = list(rng.permutation(3) + 1) + list(rng.permutation(2) + 4)
        same(A.permutedims(perm).parts(), Ao.permutedims(perm).parts())
    with pytest.raises(sb.SBError) as e:
        dev(p).permutedims([4, 2, 3, 1, 5])
    assert e.value.kind == "AssertionError"


def run_both(c, flags=0):
    Cd = sb.AliasedBlockSparse.new(c["cdims"], c["N2C"])
    Co = O.AL.new(c["cdims"], c["N2C"])
    if c["fam"] == "coo_dense":
        Ad, Ao = sb.coo_from_dense(c["A"], atol=0.0), O.COO.from_dense(c["A"], atol=0.0)
    else:
        Ad, Ao = dev(c["A"]), orc(c["A"])
    if isinstance(c["B"], dict):
        Bd, Bo = dev(c["B"]), orc(c["B"])
    else:
        Bd, Bo = sb.DenseTensor.from_numpy(c["B"]), O.Dense.new(c["B"])
    sb.contract_(Cd, labs(c["lc"]), Ad, labs(c["la"]), Bd, labs(c["lb"]), flags)
    O.contract(Co, labs(c["lc"]), Ao, labs(c["la"]), Bo, labs(c["lb"]), flags)
    return Cd, Co


FAMS = ["ad_prefix", "ad_dense", "aa_prefix", "aa_dense", "coo_dense"]


@pytest.mark.parametrize("fam", FAMS)
@pytest.mark.parametrize("seed", range(25))
def test_single_label(fam, seed):
    rng = np.random.default_rng(100 * seed + FAMS.index(fam))
    c = gen_single(rng, fam)
    print(f"[aliased {fam}] seed {seed}", flush=True)
    Cd, Co = run_both(c)
    same(Cd.parts(), Co.parts())
    ref = einsum_contract(c["A"] if fam == "coo_dense" else logical_dense(c["A"]), c["la"],
                          logical_dense(c["B"]) if isinstance(c["B"], dict) else c["B"], c["lb"], c["lc"])
    assert np.array_equal(Cd.to_dense(), ref)


@pytest.mark.parametrize("bk", ["dense", "al"])
@pytest.mark.parametrize("seed", range(30))
def test_shared(bk, seed):
    rng = np.random.default_rng(7000 + seed)
    c = gen_shared(rng, bk)
    print(f"[aliased shared {bk}] seed {seed}", flush=True)
    for flags in ((0, EXACT_JOIN) if bk == "al" else (0,)):
        Cd, Co = run_both(c, flags)
        same(Cd.parts(), Co.parts())


@pytest.mark.parametrize("fam", ["ad_prefix", "aa_prefix", "coo_dense", "ad_dense", "aa_dense"])
def test_normal_values(fam):
    """non-dyadic values: scalars and outer-product templates stay bit-exact;
    block-product templates (BLAS in the reference) within 1e-12"""
    rng = np.random.default_rng(55)
    gauss = lambda r, shape: r.standard_normal(shape)  # noqa: E731
    for _ in range(5):
        c = gen_single(rng, fam, vals=gauss)
        Cd, Co = run_both(c)
        p, q = Cd.parts(), Co.parts()
        exact = fam in ("ad_prefix", "aa_prefix", "coo_dense")
        same(p, q, rtol=0.0 if exact else 1e-12)


def test_plan_replay_with_new_values():
    rng = np.random.default_rng(77)
    c = gen_single(rng, "aa_prefix")
    A, B = dev(c["A"]), dev(c["B"])
    Cd = sb.AliasedBlockSparse.new(c["cdims"], c["N2C"])
    plan = sb.Plan(Cd, labs(c["lc"]), A, labs(c["la"]), B, labs(c["lb"]))
    plan.execute()
    Co = O.AL.new(c["cdims"], c["N2C"])
    O.contract(Co, labs(c["lc"]), orc(c["A"]), labs(c["la"]), orc(c["B"]), labs(c["lb"]), 0)
    same(Cd.parts(), Co.parts())
