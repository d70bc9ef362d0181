"""Pins the oracle's storage functions against the reference's golden vectors
(test/blocksparse/*.jl, test/coo/test.jl) and against invariants."""
import numpy as np
import pytest

from oracle import oracle as O
from refs import fixtures, dense

FX = fixtures()


def storage_4d_dense():
    f = FX["storage_4d"]
    D = np.zeros(f["shape"], order="F")
    for k in range(1, 3):
        for l in range(1, 6):
            D[0, 1, k - 1, l - 1] = 10 + (k - 1) * 5 + l
            D[2, 0, k - 1, l - 1] = -((k - 1) * 5 + l)
    D[3, 2, 0, 0] = 7
    D[3, 2, 1, 4] = -9
    return D


def test_from_dense_key_order_and_roundtrip():
    f = FX["storage_4d"]
    D = storage_4d_dense()
    B = O.BS.from_dense(D, f["N2"])
    p = B.parts()
    assert p["keys"].tolist() == f["expected_keys"]
    assert p["blksize"] == 10 and p["ids"].tolist() == [1, 2, 3]
    assert np.array_equal(B.to_dense(), D)


def test_convertor_2d():
    f = FX["convertor_2d"]
    A = dense(f["shape"], f["entries"])
    B = O.BS.from_dense(A, f["N2"])
    p = B.parts()
    assert p["keys"].tolist() == f["expected_keys"] and p["blksize"] == f["expected_blksize"]
    assert len(p["data"]) == 10 and np.array_equal(B.to_dense(), A)


def test_getindex_manual_block():
    f = FX["storage_getindex_manual"]
    data = np.zeros(10)
    for lin, v in f["payload_sets"]:
        data[lin - 1] = v
    A = O.BS.from_parts(f["dims"], f["N2"], f["keys"], f["ids"], data)
    for I, v in f["reads"]:
        assert A.getindex(*I) == v
    E = O.BS.new(f["dims"], f["N2"])
    assert E.getindex(1, 1, 1, 1) == 0.0 and E.getindex(4, 3, 2, 5) == 0.0


def test_setindex_allocates_sorted():
    f = FX["storage_setindex"]
    A = O.BS.new(f["dims"], f["N2"])
    for I, v in f["sets"]:
        A.setindex(v, *I)
    p = A.parts()
    assert p["keys"].tolist() == f["expected_keys"]
    assert p["ids"].tolist() == [1, 2]
    for I, v in f["reads"]:
        assert A.getindex(*I) == v
    with pytest.raises(O.OracleError) as e:
        A.getindex(5, 1, 1, 1)
    assert e.value.kind == "BoundsError"


def test_permutedims_matches_dense():
    f = FX["storage_4d"]
    D = storage_4d_dense()
    A = O.BS.from_dense(D, f["N2"])
    for perm in f["perms"]:
        B = A.permutedims(perm)
        assert np.array_equal(B.to_dense(), np.transpose(D, [p - 1 for p in perm]))
        p = B.parts()
        assert p["ids"].tolist() == list(range(1, len(p["ids"]) + 1))
    with pytest.raises(O.OracleError) as e:
        A.permutedims(f["bad_perm"])
    assert e.value.kind == "AssertionError"


def test_permutedims_inplace_equals_copy():
    D = storage_4d_dense()
    for perm in FX["storage_4d"]["perms"]:
        A = O.BS.from_dense(D, 2)
        B = A.permutedims(perm).parts()
        A.permutedims_(perm)
        q = A.parts()
        assert q["keys"].tolist() == B["keys"].tolist() and np.array_equal(q["data"], B["data"])


def test_sort_lex_and_is_dense():
    D = storage_4d_dense()
    A = O.BS.from_dense(D, 2)
    A.sort_()
    p = A.parts()
    assert p["keys"].tolist() == [[1, 2], [3, 1], [4, 3]]
    assert p["ids"].tolist() == [2, 1, 3]
    assert not A.is_dense()
    full = O.BS.from_dense(np.ones((2, 3, 2), order="F"), 1)
    assert full.is_dense()


def test_to_dense_merged_axes():
    rng = np.random.default_rng(1)
    D = rng.standard_normal((2, 3, 4, 5))
    A = O.BS.from_dense(D, 2)
    M = A.to_dense(merged_axes=[[1, 3]])
    # keep axis 1 (minor), fold axis 3 (major): index i1 + (i3-1)*d1
    ref = np.transpose(D, (0, 2, 1, 3)).reshape((8, 3, 5), order="F")
    assert np.array_equal(M, ref)


def test_fuse_axes():
    rng = np.random.default_rng(2)
    D = rng.standard_normal((3, 2, 4, 5))
    A = O.BS.from_dense(D, 2)
    F = A.fuse_two_axes_(1, 2)
    ref = D.reshape((6, 4, 5), order="F")
    assert np.array_equal(F.to_dense(), ref)
    A = O.BS.from_dense(D, 2)
    G = A.fuse_two_axes_(3, 4)
    assert np.array_equal(G.to_dense(), D.reshape((3, 2, 20), order="F"))
    D3 = rng.standard_normal((2, 3, 2, 4))
    A = O.BS.from_dense(D3, 3)
    H = A.fuse_two_axes_(2, 4)  # non-adjacent tail axes: moves 4 after 2
    ref = np.transpose(D3, (0, 1, 3, 2)).reshape((2, 12, 2), order="F")
    assert np.array_equal(H.to_dense(), ref)
    with pytest.raises(O.OracleError) as e:
        O.BS.from_dense(D, 2).fuse_two_axes_(2, 3)
    assert e.value.kind == "ArgumentError"


def test_coo_golden_orders():
    f = FX["coo_2d"]
    A = dense(f["shape"], f["entries"])
    S = O.COO.from_dense(A, atol=0.0)
    p = S.parts()
    assert p["keys"].tolist() == f["expected_keys"] and p["vals"].tolist() == f["expected_vals"]
    assert np.array_equal(S.to_dense(), A)
    f = FX["coo_4d"]
    S = O.COO.from_dense(dense(f["shape"], f["entries"]), atol=0.0)
    assert S.parts()["keys"].tolist() == f["expected_keys"]


def test_coo_setindex_dirty():
    f = FX["coo_setindex_dirty"]
    S = O.COO.new(f["dims"])
    for I, v in f["dirty_sets"]:
        S.setindex(v, *I, dirty=True)
    p = S.parts()
    assert p["keys"].tolist() == f["expected_keys_after_sets"] and p["dirty"]
    S.setindex(0.0, *f["delete"])
    assert S.parts()["keys"].tolist() == f["expected_keys_after_delete"]
    S.sort_()
    p = S.parts()
    assert p["keys"].tolist() == f["expected_keys_after_sort"] and not p["dirty"]


def test_coo_permute():
    f = FX["coo_permute"]
    A = dense(f["shape"], f["entries"])
    S = O.COO.from_dense(A, atol=0.0)
    Sp = S.permutedims(f["perm"])
    assert Sp.parts()["keys"].tolist() == f["expected_keys"]
    assert np.array_equal(Sp.to_dense(), np.transpose(A, [p - 1 for p in f["perm"]]))


def test_coo_fuse_axes():
    rng = np.random.default_rng(3)
    D = rng.standard_normal((3, 4, 2)) * (rng.random((3, 4, 2)) < 0.5)
    S = O.COO.from_dense(D, atol=0.0)
    F = S.fuse_axes(1, 2)
    assert np.array_equal(F.to_dense(), D.reshape((12, 2), order="F"))
    k = F.parts()["keys"]
    lin = k[:, 0] + 12 * (k[:, 1] - 1)
    assert np.all(np.diff(lin) > 0)
