"""Sector SVD of a two-site psi [Ql, Qr | s1, s2, l, r]: the oracle against
the mathematics (exact reconstruction over the full bonds, isometric
factors, singular values equal to numpy's SVD of the whole (s1, l) x (s2, r)
matrix, truncation rules), and the GPU against the oracle (keys, ids and
kept counts bit-exact, singular values to 1e-13 of the largest, the
reconstruction to 1e-12, isometry to 1e-12)."""
import numpy as np
import pytest

from sparsebackends_amd import heff as H
from sparsebackends_amd import dmrg as D
from oracle import factorize_oracle as F
from heff_dense import expand


def problems():
    yield "heisenberg", H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(24, qmax=3, width=4.0),
                                      H.heisenberg_bond(20, qmax=3, width=4.0), seed=3)
    yield "hubbard", H.HeffProblem(H.hubbard_model(1.0, 4.0), H.hubbard_bond(40, n0=0, sigma=1.2),
                                   H.hubbard_bond(36, n0=2, sigma=1.2), seed=5)
    # fused sectors up to ~120 x 120: the rows in LDS, the rotations in a global workspace
    yield "heisenberg-mid", H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(240, qmax=3, width=4.0),
                                          H.heisenberg_bond(230, qmax=3, width=4.0), seed=9)
    # fused sectors up to ~200 x 200: beyond the LDS of a workgroup, factorised in a global workspace
    yield "heisenberg-big", H.HeffProblem(H.heisenberg_model(1.0), H.heisenberg_bond(400, qmax=3, width=4.0),
                                          H.heisenberg_bond(380, qmax=3, width=4.0), seed=7)


def new_bond(charges, kept):
    live = [i for i, k in enumerate(kept) if k > 0]
    return H.Bond([charges[i] for i in live], [int(kept[i]) for i in live])


def dense_factors(prob, A, B, charges, kept):
    Bm = new_bond(charges, kept)
    Ad, _ = expand(A, [prob.Bl, Bm], [None, 0, 1])  # [s1, l, m]
    Bd, _ = expand(B, [Bm, prob.Br], [None, 0, 1])  # [s2, m, r]
    return Ad, Bd


@pytest.mark.parametrize("name,prob", list(problems()), ids=["heisenberg", "hubbard", "heisenberg-mid", "heisenberg-big"])
def test_oracle_sector_svd(name, prob):
    rowsec, colsec, charges = D.fused_sectors(prob.m, prob.Bl, prob.Br)
    psi, _ = expand(prob.psi_parts, [prob.Bl, prob.Br], [None, None, 0, 1])  # [s1, s2, l, r]
    d = prob.m.d
    full = psi.transpose(0, 2, 1, 3).reshape(d * prob.Bl.dim, d * prob.Br.dim)
    ref = np.linalg.svd(full, compute_uv=False)
    for ortho in ("left", "right"):
        A, B, svs, kept, terr = F.svd_sectors(prob.psi_parts, rowsec, colsec, prob.Bl.dims, prob.Br.dims,
                                              len(charges), ortho=ortho)
        assert terr == 0.0
        np.testing.assert_allclose(np.sort(svs)[::-1], ref[:len(svs)], rtol=0, atol=1e-12 * ref[0])
        assert np.all(ref[len(svs):] <= 1e-12 * ref[0])
        Ad, Bd = dense_factors(prob, A, B, charges, kept)
        np.testing.assert_allclose(np.einsum("alm,bmr->ablr", Ad, Bd), psi, atol=1e-12 * np.abs(psi).max())
        iso = Ad.reshape(-1, Ad.shape[2]) if ortho == "left" else Bd.transpose(1, 0, 2).reshape(Bd.shape[1], -1).T
        np.testing.assert_allclose(iso.T @ iso, np.eye(iso.shape[1]), atol=1e-12)
    # truncation: the kept values are the largest of the whole spectrum
    m = len(ref) // 3
    A, B, svs, kept, terr = F.svd_sectors(prob.psi_parts, rowsec, colsec, prob.Bl.dims, prob.Br.dims, len(charges),
                                          maxdim=m)
    np.testing.assert_allclose(np.sort(svs)[::-1], ref[:m], rtol=1e-12)
    assert abs(terr - np.sum(ref[m:] ** 2)) <= 1e-10 * np.sum(ref ** 2)


@pytest.mark.gpu
@pytest.mark.parametrize("name,prob", list(problems()), ids=["heisenberg", "hubbard", "heisenberg-mid", "heisenberg-big"])
@pytest.mark.parametrize("ortho", ["left", "right"])
@pytest.mark.parametrize("trunc", [{}, {"maxdim": 17}, {"cutoff": 0.05, "mindim": 3}])
def test_gpu_sector_svd(name, prob, ortho, trunc):
    from sparsebackends_amd import factorize as Fz
    rowsec, colsec, charges = D.fused_sectors(prob.m, prob.Bl, prob.Br)
    print(f"[sector svd] {name} {ortho} {trunc}", flush=True)
    psi = D.upload(prob.psi_parts)
    A, B, svs, kept, spec = Fz.svd_sectors(psi, rowsec, colsec, prob.Bl.dims, prob.Br.dims, len(charges),
                                           ortho=ortho, **trunc)
    Ao, Bo, so, ko, to = F.svd_sectors(prob.psi_parts, rowsec, colsec, prob.Bl.dims, prob.Br.dims, len(charges),
                                       ortho=ortho, **trunc)
    assert np.array_equal(kept, ko)
    a, b = A.parts(), B.parts()
    for p, q in ((a, Ao), (b, Bo)):
        assert tuple(p["dims"]) == tuple(q["dims"])
        assert np.array_equal(p["keys"], q["keys"]) and np.array_equal(p["ids"], q["ids"])
    assert np.abs(svs - so).max() <= 1e-13 * so.max()
    assert abs(spec.truncerr - to) <= 1e-12 * so.max() ** 2 * max(1, len(so))
    Ad, Bd = dense_factors(prob, a, b, charges, kept)
    Aod, Bod = dense_factors(prob, Ao, Bo, charges, kept)
    R, Ro = np.einsum("alm,bmr->ablr", Ad, Bd), np.einsum("alm,bmr->ablr", Aod, Bod)
    assert np.linalg.norm(R - Ro) <= 1e-12 * np.linalg.norm(Ro)
    iso = Ad.reshape(-1, Ad.shape[2]) if ortho == "left" else Bd.transpose(1, 0, 2).reshape(Bd.shape[1], -1).T
    np.testing.assert_allclose(iso.T @ iso, np.eye(iso.shape[1]), atol=1e-12)


def rank_deficient(prob, keep_every=3):
    """psi with most blocks zeroed: sectors of lower rank than their size, and
    sectors whose whole matrix is zero (every singular value 0)"""
    import copy
    p = copy.copy(prob)
    pp = dict(prob.psi_parts)
    data = pp["data"].copy()
    bs = len(data) // len(pp["ids"])
    for i, bid in enumerate(pp["ids"]):
        if i % keep_every:
            data[(bid - 1) * bs:bid * bs] = 0.0
    pp["data"] = data
    p.psi_parts = pp
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("name,prob", list(problems())[:3], ids=["heisenberg", "hubbard", "heisenberg-mid"])
@pytest.mark.parametrize("ortho", ["left", "right"])
def test_gpu_sector_svd_rank_deficient_isometry(name, prob, ortho):
    """cutoff = 0 keeps the triplets of zero singular value: the isometric
    factor must still be an isometry (LAPACK's is), and A B = psi"""
    from sparsebackends_amd import factorize as Fz
    p = rank_deficient(prob)
    rowsec, colsec, charges = D.fused_sectors(p.m, p.Bl, p.Br)
    print(f"[sector svd rank-deficient] {name} {ortho}", flush=True)
    A, B, svs, kept, spec = Fz.svd_sectors(D.upload(p.psi_parts), rowsec, colsec, p.Bl.dims, p.Br.dims,
                                           len(charges), ortho=ortho)
    Ao, Bo, so, ko, to = F.svd_sectors(p.psi_parts, rowsec, colsec, p.Bl.dims, p.Br.dims, len(charges), ortho=ortho)
    assert np.array_equal(kept, ko)
    assert np.sum(so <= 1e-13 * so.max()) > 0  # the case has null triplets to keep
    assert np.abs(svs - so).max() <= 1e-13 * so.max()
    a, b = A.parts(), B.parts()
    for q, r in ((a, Ao), (b, Bo)):
        assert np.array_equal(q["keys"], r["keys"]) and np.array_equal(q["ids"], r["ids"])
    Ad, Bd = dense_factors(p, a, b, charges, kept)
    psi, _ = expand(p.psi_parts, [p.Bl, p.Br], [None, None, 0, 1])
    np.testing.assert_allclose(np.einsum("alm,bmr->ablr", Ad, Bd), psi, atol=1e-12 * np.abs(psi).max())
    iso = Ad.reshape(-1, Ad.shape[2]) if ortho == "left" else Bd.transpose(1, 0, 2).reshape(Bd.shape[1], -1).T
    np.testing.assert_allclose(iso.T @ iso, np.eye(iso.shape[1]), atol=1e-12)
