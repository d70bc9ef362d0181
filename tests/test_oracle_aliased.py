"""Oracle: AliasedBlockSparse storage and the aliased contraction kernels.

No reference test covers them, so they are pinned by mathematics: the logical
content (scalar x template per block) against dense einsum / transposes on
dyadic values (exact), and by the aliasing rules the reference states
(contract_aliased.jl:1-25, 75-114; contract_aliased_coo_dense.jl:12-32): a
key fed by one combined template stays aliased with the summed scalar, a key
fed by two different templates becomes a concrete block (scalar 1), and the
dense-tail Aliased x Dense kernel preserves the templates one to one."""
import numpy as np
import pytest

from oracle import oracle as O
from aliased_cases import aliased_parts, logical_dense, gen_single, gen_shared, dy
from refs import labs, einsum_contract
from sb_flags import EXACT_JOIN


def al(p):
    return O.AL.from_parts(p["dims"], p["N2"], p["templates"], p["ntemplates"], p["keys"], p["alias_ids"],
                           p["scalars"])


def operand(x):
    return al(x) if isinstance(x, dict) else O.Dense.new(x)


def dense_of(x):
    return logical_dense(x) if isinstance(x, dict) else x


# ---------------------------------------------------------------- storage
def test_storage_getindex_to_dense_to_blocksparse():
    rng = np.random.default_rng(0)
    for _ in range(20):
        p = aliased_parts(rng, [3, 4], [2, 3], 7, 3)
        A = al(p)
        D = logical_dense(p)
        assert np.array_equal(A.to_dense(), D)
        assert np.array_equal(A.to_blocksparse().to_dense(), D)
        for _ in range(10):
            I = tuple(int(rng.integers(1, d + 1)) for d in p["dims"])
            assert A.getindex(*I) == D[tuple(i - 1 for i in I)]
        assert A.compression_ratio() == pytest.approx(len(p["alias_ids"]) / p["ntemplates"])
    with pytest.raises(O.OracleError) as e:
        A.getindex(4, 1, 1, 1)
    assert e.value.kind == "BoundsError"


def test_storage_init_value_and_empty():
    A = O.AL.new((3, 2), 1)
    assert A.compression_ratio() == 1.0
    assert np.array_equal(A.to_dense(init=2.5), np.full((3, 2), 2.5))
    with pytest.raises(O.OracleError) as e:
        O.AL.new((3, 2), 0)
    assert e.value.kind == "AssertionError"


@pytest.mark.parametrize("seed", range(10))
def test_permutedims_transposes_templates_only(seed):
    rng = np.random.default_rng(seed)
    p = aliased_parts(rng, [3, 2, 4], [2, 3, 2], 12, 4)
    A = al(p)
    pp = list(rng.permutation(3) + 1) + list(rng.permutation(3) + 4)
    B = A.permutedims(pp)
    assert np.array_equal(B.to_dense(), np.transpose(logical_dense(p), [x - 1 for x in pp]))
    q = B.parts()
    assert q["ntemplates"] == p["ntemplates"]  # the templates are transposed, not multiplied out
    lin = [int(np.ravel_multi_index(tuple(k - 1), q["dims"][:3], order="F")) for k in q["keys"]]
    assert lin == sorted(lin)
    with pytest.raises(O.OracleError) as e:
        A.permutedims([4, 2, 3, 1, 5, 6])
    assert e.value.kind == "AssertionError"


# ---------------------------------------------------------------- kernels
def run(c, flags=0):
    Cn = O.AL.new(c["cdims"], c["N2C"])
    A = O.COO.from_dense(c["A"], atol=0.0) if c["fam"] == "coo_dense" else operand(c["A"])
    O.contract(Cn, labs(c["lc"]), A, labs(c["la"]), operand(c["B"]), labs(c["lb"]), flags)
    return Cn


def check_structure(q):
    P = len(q["dims"]) - q["N2"]
    lin = [int(np.ravel_multi_index(tuple(k - 1), q["dims"][:P], order="F")) for k in q["keys"]] if P else [0]
    assert lin == sorted(lin) and len(set(lin)) == len(lin)
    assert all(1 <= a <= q["ntemplates"] for a in q["alias_ids"])


FAMS = ["ad_prefix", "ad_dense", "aa_prefix", "aa_dense", "coo_dense"]


@pytest.mark.parametrize("fam", FAMS)
@pytest.mark.parametrize("seed", range(25))
def test_single_label_vs_einsum(fam, seed):
    rng = np.random.default_rng(100 * seed + FAMS.index(fam))
    c = gen_single(rng, fam)
    Cn = run(c)
    ref = einsum_contract(dense_of(c["A"]), c["la"], dense_of(c["B"]), c["lb"], c["lc"])
    assert np.array_equal(Cn.to_dense(), ref)
    check_structure(Cn.parts())


@pytest.mark.parametrize("bk", ["dense", "al"])
@pytest.mark.parametrize("seed", range(30))
def test_shared_vs_einsum(bk, seed):
    rng = np.random.default_rng(7000 + seed)
    c = gen_shared(rng, bk)
    ref = einsum_contract(dense_of(c["A"]), c["la"], dense_of(c["B"]), c["lb"], c["lc"])
    Cn = run(c, EXACT_JOIN if bk == "al" else 0)
    assert np.array_equal(Cn.to_dense(), ref)
    check_structure(Cn.parts())


def test_dense_tail_preserves_templates():
    """Aliased x Dense over a tail label: one combined template per A template,
    same keys and scalars (contract_aliased.jl:234-237, 352-363)"""
    rng = np.random.default_rng(3)
    p = aliased_parts(rng, [4, 3], [2, 3], 9, 3)
    B = dy(rng, (3, 5))
    Cn = O.AL.new((4, 3, 2, 5), 2)
    O.contract(Cn, labs(["a", "b", "x", "y"]), al(p), labs(["a", "b", "x", "r"]), O.Dense.new(B), labs(["r", "y"]))
    q = Cn.parts()
    assert q["ntemplates"] == 3
    assert np.array_equal(q["keys"], p["keys"]) and np.array_equal(q["scalars"], p["scalars"])
    assert np.array_equal(q["alias_ids"], p["alias_ids"])


def test_coo_dense_aliasing_rules():
    """COO x Dense -> Aliased: a key with one entry is alpha * B[:, rv] (one template per
    distinct rv); a key with two entries is a concrete block, scalar 1
    (contract_aliased_coo_dense.jl:12-32, 96-145)"""
    B = np.arange(1.0, 7.0).reshape(2, 3, order="F")  # [y, r]: B[:, rv]
    # A [a, r]: a=1 -> rv 1 ; a=2 -> rv 1 ; a=3 -> rv 2 and rv 3
    A = np.zeros((3, 3))
    A[0, 0], A[1, 0], A[2, 1], A[2, 2] = 0.5, -2.0, 1.5, 0.25
    Cn = O.AL.new((3, 2), 1)
    O.contract(Cn, labs(["a", "y"]), O.COO.from_dense(A, atol=0.0), labs(["a", "r"]), O.Dense.new(B),
               labs(["y", "r"]))
    q = Cn.parts()
    # B[:,1] shared by keys 1 and 2; key 3 first aliases B[:,2] (a template of
    # its own), then its second entry makes it a concrete block
    assert q["ntemplates"] == 3
    assert list(q["keys"].ravel()) == [1, 2, 3]
    assert list(q["scalars"]) == [0.5, -2.0, 1.0]
    assert list(q["alias_ids"]) == [1, 1, 3]
    assert np.array_equal(q["templates"][:2], B[:, 0])
    assert np.array_equal(q["templates"][2:4], B[:, 1])
    assert np.array_equal(q["templates"][4:6], 1.5 * B[:, 1] + 0.25 * B[:, 2])
    assert np.array_equal(Cn.to_dense(), A @ B.T)


def test_prefix_outer_alias_then_demote():
    """Aliased x Aliased over a prefix label: a C key reached through one
    (tidA, tidB) combination keeps it with the summed scalar; a second,
    different combination demotes the key to a concrete block"""
    # A [a, r | x], B [r | y]: C [a | x, y]
    tA = np.array([1.0, 2.0, 3.0, 4.0])  # two templates of length 2
    tB = np.array([1.0, -1.0, 0.5, 2.0])
    # A blocks: (a=1, r=1) tid1 s=2 ; (a=1, r=2) tid1 s=3 ; (a=2, r=1) tid1 s=1 ; (a=2, r=2) tid2 s=1
    A = dict(dims=(2, 2, 2), N2=1, templates=tA, ntemplates=2,
             keys=np.array([[1, 1], [2, 1], [1, 2], [2, 2]]), alias_ids=np.array([1, 1, 1, 2]),
             scalars=np.array([2.0, 1.0, 3.0, 1.0]))
    # B blocks: r=1 tid1 s=1 ; r=2 tid1 s=0.5
    B = dict(dims=(2, 2), N2=1, templates=tB, ntemplates=2, keys=np.array([[1], [2]]),
             alias_ids=np.array([1, 1]), scalars=np.array([1.0, 0.5]))
    Cn = O.AL.new((2, 2, 2), 2)
    O.contract(Cn, labs(["a", "x", "y"]), al(A), labs(["a", "r", "x"]), al(B), labs(["r", "y"]))
    q = Cn.parts()
    assert list(q["keys"].ravel()) == [1, 2]
    # key a=1: (tid1, tid1) twice -> aliased, scalar 2*1 + 3*0.5
    assert q["scalars"][0] == 2.0 + 1.5 and q["alias_ids"][0] == 1
    # key a=2: (tid1, tid1) then (tid2, tid1) -> concrete
    assert q["scalars"][1] == 1.0 and q["ntemplates"] == 3
    ref = einsum_contract(logical_dense(A), ["a", "r", "x"], logical_dense(B), ["r", "y"], ["a", "x", "y"])
    assert np.array_equal(Cn.to_dense(), ref)


def test_method_errors():
    rng = np.random.default_rng(4)
    p = aliased_parts(rng, [3], [2], 2, 1)
    Cb = O.BS.new((3, 4), 1)
    with pytest.raises(O.OracleError) as e:
        O.contract(Cb, labs(["a", "y"]), al(p), labs(["a", "x"]), O.Dense.new(np.ones((2, 4))), labs(["x", "y"]))
    assert e.value.kind == "MethodError"
