"""Pins the oracle's blocksparse_left_qr / left_factorize
(ops_factorize.jl:8-43, tensor_wrappers.jl:954-970): reconstruction of the
dense (l, s) x r matrix, orthonormal Q, LAPACK's sign convention on a case
worked by hand, and the key/id rule of Q (one block per l value, ascending)."""
import numpy as np
import pytest

from oracle import factorize_oracle as F
from oracle import oracle as O


def rand_parts(rng, nl, nr, ds, fill=0.4):
    A = rng.standard_normal((nl, nr, ds)) * (rng.random((nl, nr, 1)) < fill)
    A[0, 0, 0] += 1.0  # at least one block
    return O.BS.from_dense(np.asfortranarray(A), 1).parts(), A


def test_hand_worked_case():
    # one r value, blocks at l = 1 and l = 3 with d_s = 1: M = [3; 4]
    parts = dict(dims=(3, 2, 1), N2=1, keys=np.array([[1, 2], [3, 2]]), ids=np.array([1, 2]),
                 data=np.array([3.0, 4.0]))
    q, R, lv, rv = F.blocksparse_left_qr(parts)
    # Householder with beta = -sign(alpha) |x|: R = -5, Q = -[0.6; 0.8]
    assert R.shape == (1, 1) and R[0, 0] == pytest.approx(-5.0, abs=1e-15)
    np.testing.assert_allclose(q["data"], [-0.6, -0.8], atol=1e-15)
    assert list(lv) == [1, 3] and list(rv) == [2]
    assert q["keys"].ravel().tolist() == [1, 3] and q["ids"].tolist() == [1, 2] and q["dims"] == (3, 1, 1)
    _, full = F.left_factorize(parts)
    np.testing.assert_allclose(full, [[0.0, -5.0]], atol=1e-15)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("shape", [(6, 5, 2), (3, 9, 2), (8, 8, 1), (4, 4, 3)])
def test_reconstruction_and_orthonormality(seed, shape):
    rng = np.random.default_rng(100 + seed)
    parts, A = rand_parts(rng, *shape)
    q, R, lv, rv = F.blocksparse_left_qr(parts)
    ds, k = shape[2], R.shape[0]
    Q = np.concatenate([q["data"][i * ds * k:(i + 1) * ds * k].reshape(ds, k, order="F") for i in range(len(lv))])
    np.testing.assert_allclose(Q.T @ Q, np.eye(k), atol=1e-13)
    M = np.concatenate([A[l - 1][rv - 1].T for l in lv])  # rows (l, s), columns r
    np.testing.assert_allclose(Q @ R, M, atol=1e-13 * max(1.0, np.abs(M).max()))
    assert np.allclose(np.tril(R, -1), 0.0)
    _, full = F.left_factorize(parts)
    assert full.shape == (k, shape[1])
    np.testing.assert_allclose(full[:, rv - 1], R, atol=0)
    assert not np.any(np.delete(full, rv - 1, axis=1))
