"""GPU blocksparse_svd (ops_factorize.jl:50-148) through the C ABI against
the oracle: keys, ids and kept counts bit-exact; singular values within
1e-13 of the largest; the truncated product U*SV within 1e-12 relative
Frobenius; each U block equal to the oracle's up to the sign of its singular
vector (the gauge the SVD leaves free; LAPACK and the one-sided Jacobi pick
signs differently)."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from oracle import factorize_oracle as F
from oracle import oracle as O
from test_oracle_factorize import literal, random_parts

pytestmark = pytest.mark.gpu


def dev(p):
    return sb.NewBlockSparseSorted.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"])


def compare(p, ortho, **kw):
    U, SV, svs, spec = sb.blocksparse_svd(dev(p), ortho=ortho, **kw)
    Uo, SVo, so, eo, to = F.blocksparse_svd(p, ortho=ortho, **kw)
    u, s = U.parts(), SV.parts()
    for a, b in ((u, Uo), (s, SVo)):
        assert tuple(a["dims"]) == tuple(b["dims"])
        assert np.array_equal(a["keys"], b["keys"]) and np.array_equal(a["ids"], b["ids"])
    scale = so.max() if so.size else 1.0
    assert np.abs(svs - so).max() <= 1e-13 * scale
    assert abs(spec.truncerr - to) <= 1e-12 * max(1.0, to) * scale ** 2
    np.testing.assert_allclose(spec.eigs, svs ** 2)
    R, Ro = F.reconstruct(u, s), F.reconstruct(Uo, SVo)
    assert np.linalg.norm(R - Ro) <= 1e-12 * max(np.linalg.norm(Ro), 1e-300)
    d1 = p["dims"][2]
    for i in range(len(u["ids"])):
        a = u["data"][(u["ids"][i] - 1) * d1:u["ids"][i] * d1]
        b = Uo["data"][(Uo["ids"][i] - 1) * d1:Uo["ids"][i] * d1]
        sgn = 1.0 if a @ b >= 0 else -1.0
        assert np.abs(a - sgn * b).max() <= 1e-11 * max(1.0, np.abs(b).max())
    return svs


def test_literal():
    p = literal()
    for ortho in ("left", "right"):
        svs = compare(p, ortho)
        assert svs.tolist() == [4.0, 3.0, 2.0, 1.0]
    compare(p, "left", maxdim=3)
    compare(p, "left", cutoff=0.6)
    compare(p, "right", cutoff=0.6, mindim=3)


@pytest.mark.parametrize("shape", [(6, 7, 2, 2), (5, 9, 4, 4), (8, 3, 4, 2), (4, 12, 2, 4), (6, 1, 4, 2),
                                   (64, 64, 4, 4), (40, 300, 2, 2)])
@pytest.mark.parametrize("ortho", ["left", "right"])
def test_random(shape, ortho):
    rng = np.random.default_rng(sum(shape) + (ortho == "right"))
    p, _ = random_parts(rng, *shape)
    print(f"[svd] {shape} {ortho}", flush=True)
    compare(p, ortho)
    full = F.blocksparse_svd(p)[2]
    m = max(1, len(full) // 2)
    compare(p, ortho, maxdim=m)
    compare(p, ortho, cutoff=0.3, mindim=2)


def test_errors():
    p = literal()
    A = sb.NewBlockSparseSorted.from_parts((3, 4, 2, 2), 1, np.array([[1, 1, 1]]), np.array([1]), np.zeros(2))
    with pytest.raises(sb.SBError) as e:
        sb.blocksparse_svd(A)
    assert e.value.kind == "MethodError"
    empty = sb.NewBlockSparseSorted.new((3, 4, 2, 2), 2)
    with pytest.raises(sb.SBError) as e:
        sb.blocksparse_svd(empty)  # nothing kept: the factors' constructor rejects a zero dimension
    assert e.value.kind == "AssertionError"
    with pytest.raises(ValueError):
        sb.blocksparse_svd(dev(p), ortho="up")


def test_rank_deficient_right_isometry():
    """cutoff = 0 keeps the zero singular values of groups of lower rank: with
    ortho right the SV rows of each group (V^T) are still orthonormal, as
    LAPACK's are, and the factorisation still matches the oracle"""
    rng = np.random.default_rng(3)
    nL, nR, d1, d2 = 5, 6, 4, 3
    D = rng.standard_normal((nL, nR, d1, d2))
    D[rng.random((nL, nR)) > 0.5] = 0.0
    D[:, :, 1, :] = D[:, :, 0, :]  # two equal rows in every group: rank <= d1 - 1
    D[0] = 0.0
    D[0, 0, 0, 0] = 1.0  # a group of rank 1
    p = O.BS.from_dense(D, 2).parts()
    svs = compare(p, "right")
    assert np.sum(svs <= 1e-13 * svs.max()) >= nL  # null triplets kept
    U, SV, svs, spec = sb.blocksparse_svd(dev(p), ortho="right")
    u, s = U.parts(), SV.parts()
    k = len(svs)
    rows = np.zeros((k, nR * d2))
    for (nb, R), i in zip(s["keys"], s["ids"]):
        rows[nb - 1, (R - 1) * d2:R * d2] = s["data"][(i - 1) * d2:i * d2]
    group = {}
    for (L, nb) in u["keys"]:
        group.setdefault(L, []).append(nb - 1)
    for L, nbs in group.items():
        V = rows[nbs]
        np.testing.assert_allclose(V @ V.T, np.eye(len(nbs)), atol=1e-12)
