"""The oracle's blocksparse_svd (ops_factorize.jl:50-148) pinned on the CPU:
a literal case worked by hand from the reference's loop order, exact
reconstruction without truncation, orthonormal U within each L group, the
per-group singular values of numpy's SVD, and the truncation rules."""
import numpy as np
import pytest

from oracle import factorize_oracle as F
from oracle import oracle as O


def literal():
    # L=1: blocks (1,1) = diag(3, 0), (1,3) = [[0,0],[0,4]]  -> sigma 4, 3
    # L=3: block (3,2) = diag(2, 1)                          -> sigma 2, 1
    blocks = {(1, 1): np.array([[3.0, 0.0], [0.0, 0.0]]), (3, 2): np.array([[2.0, 0.0], [0.0, 1.0]]),
              (1, 3): np.array([[0.0, 0.0], [0.0, 4.0]])}
    keys = [(1, 1), (3, 2), (1, 3)]  # column-major order
    data = np.concatenate([blocks[k].reshape(-1, order="F") for k in keys])
    return dict(dims=(3, 4, 2, 2), N2=2, keys=np.array(keys), ids=np.array([1, 2, 3]), data=data)


def test_literal_keys_ids_and_truncation():
    p = literal()
    U, SV, svs, eigs, terr = F.blocksparse_svd(p)
    assert svs.tolist() == [4.0, 3.0, 2.0, 1.0] and terr == 0.0
    assert U["keys"].tolist() == [[1, 1], [1, 2], [3, 3], [3, 4]] and U["ids"].tolist() == [1, 2, 3, 4]
    # SV blocks created nb by nb, R ascending: (1,1) (1,3) (2,1) (2,3) (3,2) (4,2); keys sorted R-major
    assert SV["keys"].tolist() == [[1, 1], [2, 1], [3, 2], [4, 2], [1, 3], [2, 3]]
    assert SV["ids"].tolist() == [1, 3, 5, 6, 2, 4]
    assert U["dims"] == (3, 4, 2) and SV["dims"] == (4, 4, 2)
    np.testing.assert_allclose(F.reconstruct(U, SV), O.BS.from_parts(**p).to_dense(), atol=1e-15)
    # maxdim 3: one of L=3's values dropped
    U, SV, svs, _, terr = F.blocksparse_svd(p, maxdim=3)
    assert svs.tolist() == [4.0, 3.0, 2.0] and terr == 1.0 and U["keys"].tolist() == [[1, 1], [1, 2], [3, 3]]
    # cutoff 0.6: 2 and 1 are below 0.6 * 4
    _, _, svs, _, terr = F.blocksparse_svd(p, cutoff=0.6)
    assert svs.tolist() == [4.0, 3.0] and terr == 5.0
    # mindim 3 stops the cutoff loop
    _, _, svs, _, terr = F.blocksparse_svd(p, cutoff=0.6, mindim=3)
    assert svs.tolist() == [4.0, 3.0, 2.0] and terr == 1.0


def random_parts(rng, nL, nR, d1, d2, fill=0.4):
    D = rng.standard_normal((nL, nR, d1, d2))
    D[rng.random((nL, nR)) > fill] = 0.0
    return O.BS.from_dense(D, 2).parts(), D


@pytest.mark.parametrize("shape", [(6, 7, 2, 2), (5, 9, 4, 4), (8, 3, 4, 2), (4, 12, 2, 4)])
def test_random_reconstruction_and_isometry(shape):
    rng = np.random.default_rng(sum(shape))
    p, D = random_parts(rng, *shape)
    for ortho in ("left", "right"):
        U, SV, svs, eigs, terr = F.blocksparse_svd(p, ortho=ortho)
        assert terr == 0.0
        np.testing.assert_allclose(F.reconstruct(U, SV), D, atol=1e-13 * np.abs(D).max())
    U, SV, svs, _, _ = F.blocksparse_svd(p, ortho="left")
    d1 = shape[2]
    by_L = {}
    for (L, nb), i in zip(U["keys"], U["ids"]):
        by_L.setdefault(L, []).append(U["data"][(i - 1) * d1:i * d1])
    for L, cols in by_L.items():
        Q = np.array(cols).T
        np.testing.assert_allclose(Q.T @ Q, np.eye(Q.shape[1]), atol=1e-13)
        # the group's singular values are numpy's for its block row
        row = D[L - 1].transpose(1, 0, 2).reshape(d1, -1)  # [s1, (R, s2)]
        ref = np.linalg.svd(row, compute_uv=False)  # zero columns of absent blocks change nothing
        mine = sorted((svs[nb - 1] for (LL, nb) in U["keys"] if LL == L), reverse=True)
        np.testing.assert_allclose(mine, ref[:len(mine)], rtol=1e-13, atol=1e-13 * ref[0])


def test_truncation_global_order():
    rng = np.random.default_rng(5)
    p, D = random_parts(rng, 10, 10, 4, 2)
    _, _, full, _, _ = F.blocksparse_svd(p)
    s = np.sort(full)[::-1]
    for m in (1, 5, 12):
        _, _, svs, _, terr = F.blocksparse_svd(p, maxdim=m)
        assert sorted(svs, reverse=True) == sorted(s[:m], reverse=True)
        assert abs(terr - np.sum(s[m:] ** 2)) <= 1e-12 * np.sum(s ** 2)
