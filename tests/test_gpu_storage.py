"""GPU parity of the storage operations against the oracle and the
reference's golden vectors."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from oracle import oracle as O
from both import compare_bs, compare_coo
from refs import fixtures, dense
from test_oracle_storage import storage_4d_dense

pytestmark = pytest.mark.gpu
FX = fixtures()


def test_from_dense_golden_and_roundtrip():
    f = FX["storage_4d"]
    D = storage_4d_dense()
    B = sb.blocksparse_from_dense(D, f["N2"])
    assert B.keys.tolist() == f["expected_keys"]
    assert np.array_equal(B.to_dense(), D)
    compare_bs(B, O.BS.from_dense(D, f["N2"]))
    f2 = FX["convertor_2d"]
    A = dense(f2["shape"], f2["entries"])
    compare_bs(sb.blocksparse_from_dense(A, 1), O.BS.from_dense(A, 1))


def test_from_dense_tolerances():
    rng = np.random.default_rng(5)
    D = rng.standard_normal((5, 4, 3)) * (rng.random((5, 4, 1)) < 0.6)
    D[0, 0, :] = 1e-9
    for atol, rtol, dz in [(0.0, 0.0, True), (1e-6, 0.0, True), (0.0, 0.5, True), (0.0, 0.0, False)]:
        compare_bs(sb.blocksparse_from_dense(D, 1, atol=atol, rtol=rtol, dropzeros=dz),
                   O.BS.from_dense(D, 1, atol=atol, rtol=rtol, dropzeros=dz))


@pytest.mark.parametrize("seed", range(6))
def test_permutedims_random(seed):
    rng = np.random.default_rng(seed)
    N, N2 = int(rng.integers(2, 6)), None
    N2 = int(rng.integers(1, N))
    shape = tuple(int(x) for x in rng.integers(1, 5, size=N))
    D = rng.standard_normal(shape) * (rng.random(shape[:N - N2] + (1,) * N2) < 0.5)
    P = N - N2
    perm = list(rng.permutation(P) + 1) + list(rng.permutation(N2) + 1 + P)
    A, Ao = sb.blocksparse_from_dense(D, N2), O.BS.from_dense(D, N2)
    Bd, Bo = A.permutedims(perm), Ao.permutedims(perm)
    compare_bs(Bd, Bo)
    assert np.array_equal(Bd.to_dense(), np.transpose(D, [p - 1 for p in perm]))
    A.permutedims_(perm)
    Ao.permutedims_(perm)
    compare_bs(A, Ao)


def test_permutedims_large_tiles():
    # tail axes >= 8 take the LDS-tiled transpose
    rng = np.random.default_rng(7)
    D = rng.standard_normal((3, 2, 40, 3, 36)) * (rng.random((3, 2, 1, 1, 1)) < 0.7)
    for perm in ([2, 1, 5, 4, 3], [1, 2, 4, 5, 3], [2, 1, 3, 5, 4]):
        compare_bs(sb.blocksparse_from_dense(D, 3).permutedims(perm), O.BS.from_dense(D, 3).permutedims(perm))


def test_permutedims_rejects_cross():
    D = np.zeros((4, 3, 2, 5), order="F")
    D[0, 1, 0, 0] = 1.0
    with pytest.raises(sb.SBError) as e:
        sb.blocksparse_from_dense(D, 2).permutedims([3, 2, 1, 4])
    assert e.value.kind == "AssertionError"


def test_dense_permute():
    rng = np.random.default_rng(8)
    D = rng.standard_normal((4, 33, 2, 17))
    T = sb.DenseTensor.from_numpy(D)
    for perm in ([4, 3, 2, 1], [2, 1, 3, 4], [1, 3, 2, 4], [3, 4, 1, 2]):
        assert np.array_equal(T.permutedims(perm).numpy(), np.transpose(D, [p - 1 for p in perm]))


def test_getindex_setindex():
    f = FX["storage_setindex"]
    A = sb.NewBlockSparseSorted.new(f["dims"], f["N2"])
    for I, v in f["sets"]:
        A[I] = v
    assert A.keys.tolist() == f["expected_keys"]
    assert A.ids.tolist() == [1, 2]
    for I, v in f["reads"]:
        assert A[I] == v
    with pytest.raises(sb.SBError) as e:
        A[(5, 1, 1, 1)]
    assert e.value.kind == "BoundsError"


def test_sort_lex_is_dense_to_dense_merged():
    D = storage_4d_dense()
    A, Ao = sb.blocksparse_from_dense(D, 2), O.BS.from_dense(D, 2)
    A.sort_()
    Ao.sort_()
    compare_bs(A, Ao)
    assert not A.is_dense()
    assert sb.blocksparse_from_dense(np.ones((2, 3, 2), order="F"), 1).is_dense()
    rng = np.random.default_rng(1)
    D2 = rng.standard_normal((2, 3, 4, 5))
    B = sb.blocksparse_from_dense(D2, 2)
    assert np.array_equal(B.to_dense(merged_axes=[[1, 3]]), O.BS.from_dense(D2, 2).to_dense(merged_axes=[[1, 3]]))
    assert np.array_equal(B.to_dense(merged_axes=[[1, 3], [2, 4]]),
                          O.BS.from_dense(D2, 2).to_dense(merged_axes=[[1, 3], [2, 4]]))


def test_fuse_axes():
    rng = np.random.default_rng(2)
    D = rng.standard_normal((3, 2, 4, 5))
    for ax in ((1, 2), (3, 4)):
        compare_bs(sb.blocksparse_from_dense(D, 2).fuse_two_axes_(*ax), O.BS.from_dense(D, 2).fuse_two_axes_(*ax))
    D3 = rng.standard_normal((2, 3, 2, 4))
    A, Ao = sb.blocksparse_from_dense(D3, 3), O.BS.from_dense(D3, 3)
    compare_bs(A.fuse_two_axes_(2, 4), Ao.fuse_two_axes_(2, 4))
    compare_bs(A, Ao)  # fuse_dense_axes! permuted A in place


def test_coo_ops():
    f = FX["coo_2d"]
    A = dense(f["shape"], f["entries"])
    S = sb.coo_from_dense(A, atol=0.0)
    assert S.keys.tolist() == f["expected_keys"] and S.vals.tolist() == f["expected_vals"]
    assert np.array_equal(S.to_dense(), A)
    f = FX["coo_permute"]
    A = dense(f["shape"], f["entries"])
    Sp = sb.coo_from_dense(A, atol=0.0).permutedims(f["perm"])
    assert Sp.keys.tolist() == f["expected_keys"] and Sp.dirty
    Sp.sort_()
    compare_coo(Sp, O.COO.from_dense(A, atol=0.0).permutedims(f["perm"]).sort_())
    f = FX["coo_setindex_dirty"]
    S = sb.COOTensor.new(f["dims"])
    for I, v in f["dirty_sets"]:
        S.setindex_(v, *I, dirty=True)
    assert S.keys.tolist() == f["expected_keys_after_sets"]
    S[tuple(f["delete"])] = 0.0
    assert S.keys.tolist() == f["expected_keys_after_delete"]
    S.sort_()
    assert S.keys.tolist() == f["expected_keys_after_sort"] and not S.dirty
    rng = np.random.default_rng(3)
    D = rng.standard_normal((3, 4, 2)) * (rng.random((3, 4, 2)) < 0.5)
    compare_coo(sb.coo_from_dense(D, atol=0.0).fuse_axes(1, 2), O.COO.from_dense(D, atol=0.0).fuse_axes(1, 2))
    assert sb.coo_from_dense(D, atol=0.0)[(1, 1, 1)] == O.COO.from_dense(D, atol=0.0).getindex(1, 1, 1)
