"""GPU: blocksparse_left_qr / left_factorize (ops_factorize.jl:8-43,
tensor_wrappers.jl:954-970) through the C ABI against the oracle (numpy's
LAPACK QR): Q's block list bit for bit, Q and R to rounding with the same
signs, reconstruction and orthonormality; the sector QR of a sector-format
site tensor: orthonormal factors per right sector, Q R = A, and the left
orthonormalisation of a random MPS."""
import numpy as np
import pytest

import sparsebackends_amd as sb
from sparsebackends_amd import dmrg as D
from sparsebackends_amd import heff as H
from oracle import factorize_oracle as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rand_dense(rng, nl, nr, ds, fill=0.4):
    A = rng.standard_normal((nl, nr, ds)) * (rng.random((nl, nr, 1)) < fill)
    A[0, 0, 0] += 1.0
    return np.asfortranarray(A)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("shape", [(6, 5, 2), (3, 9, 2), (40, 30, 2), (128, 128, 2), (8, 8, 1)])
def test_left_qr_vs_oracle(seed, shape):
    rng = np.random.default_rng(200 + seed)
    A = rand_dense(rng, *shape)
    parts = O.BS.from_dense(A, 1).parts()
    q, R, lv, rv = F.blocksparse_left_qr(parts)
    Qd, Rd, lvd, rvd = sb.blocksparse_left_qr(sb.blocksparse_from_dense(A, 1))
    p = Qd.parts()
    assert tuple(p["dims"]) == tuple(q["dims"])
    assert np.array_equal(p["keys"].ravel(), q["keys"].ravel()) and np.array_equal(p["ids"], q["ids"])
    assert np.array_equal(lvd, lv) and np.array_equal(rvd, rv)
    scale = max(1.0, np.abs(R).max())
    np.testing.assert_allclose(Rd, R, atol=1e-13 * scale)
    np.testing.assert_allclose(p["data"], q["data"], atol=1e-13)
    _, full = F.left_factorize(parts)
    _, fulld = sb.left_factorize(sb.blocksparse_from_dense(A, 1))
    np.testing.assert_allclose(fulld, full, atol=1e-13 * scale)


def test_left_qr_hand_case_and_errors():
    parts = dict(dims=(3, 2, 1), N2=1, keys=np.array([[1, 2], [3, 2]]), ids=np.array([1, 2]),
                 data=np.array([3.0, 4.0]))
    A = sb.NewBlockSparseSorted.from_parts(parts["dims"], 1, parts["keys"], parts["ids"], parts["data"])
    Q, R, lv, rv = sb.blocksparse_left_qr(A)
    assert R[0, 0] == pytest.approx(-5.0, abs=1e-14)
    np.testing.assert_allclose(Q.parts()["data"], [-0.6, -0.8], atol=1e-15)
    with pytest.raises(sb.SBError) as e:
        sb.blocksparse_left_qr(sb.NewBlockSparseSorted.new((3, 2, 2), 1))
    assert e.value.kind == "ArgumentError"
    with pytest.raises(sb.SBError) as e:
        sb.blocksparse_left_qr(sb.blocksparse_from_dense(np.ones((2, 2, 2, 2)), 2))
    assert e.value.kind == "MethodError"


def sector_matrix(parts, rowsec, dl, dr, ir):
    """the rows (il, s, l) of right sector ir as the library orders them"""
    nl, nr, d, Dl, Dr = parts["dims"]
    bs = d * Dl * Dr
    rows = []
    for k, i in sorted(((tuple(int(x) for x in k), int(i)) for k, i in zip(parts["keys"], parts["ids"])),
                       key=lambda t: t[0][0]):
        il, r = k[0] - 1, k[1] - 1
        if r != ir:
            continue
        blk = parts["data"][(i - 1) * bs:i * bs].reshape((d, Dl, Dr), order="F")
        for s in range(d):
            if rowsec[il, s] == ir:
                for l in range(dl[il]):
                    rows.append(blk[s, l, :dr[ir]])
    return np.array(rows)


@pytest.mark.parametrize("chi", [16, 64])
def test_sector_qr(chi):
    m = H.heisenberg_model(1.0)
    bonds = D.config_bonds(12, chi=chi)
    A = D.random_mps(m, bonds, seed=3)
    j = 6
    Bl, Br = bonds[j], bonds[j + 1]
    rowsec = np.full((Bl.nsec, 2), -1, dtype=np.int64)
    for il in range(Bl.nsec):
        for s in range(2):
            rowsec[il, s] = Br.index.get(H.qadd(Bl.charges[il], m.charges[s]), -1)
    Q, R, kept = sb.qr_sectors(A[j], rowsec, Bl.dims, Br.dims)
    pa, pq, pr = A[j].parts(), Q.parts(), R.parts()
    assert np.array_equal(pq["keys"], pa["keys"]) and np.array_equal(pq["ids"], pa["ids"])
    Dr = pa["dims"][4]
    rblocks = {int(k[0]) - 1: pr["data"][(int(i) - 1) * Dr * Dr:int(i) * Dr * Dr].reshape(Dr, Dr, order="F")
               for k, i in zip(pr["keys"], pr["ids"])}
    for ir in range(Br.nsec):
        M = sector_matrix(pa, rowsec, Bl.dims, Br.dims, ir)
        Qm = sector_matrix(pq, rowsec, Bl.dims, [kept[ir]] * Br.nsec, ir)
        k = kept[ir]
        assert k == min(M.shape)
        np.testing.assert_allclose(Qm.T @ Qm, np.eye(k), atol=1e-13)
        Rm = rblocks[ir][:k, :Br.dims[ir]]
        np.testing.assert_allclose(Qm @ Rm, M, atol=1e-12 * np.abs(M).max())
        ref_r = np.linalg.qr(M, mode="r")  # LAPACK: the same R, signs included
        np.testing.assert_allclose(Rm, ref_r, atol=1e-12 * np.abs(M).max())


def test_left_orthonormal_mps():
    """after the sweep of QRs every site is a left isometry and the state is normalised"""
    m = H.heisenberg_model(1.0)
    bonds = D.config_bonds(10, chi=32)
    A, nb = D.left_orthonormalize(m, D.random_mps(m, bonds, seed=0), bonds)
    for j, Aj in enumerate(A):
        Bl, Br = nb[j], nb[j + 1]
        p = Aj.parts()
        nl, nr, d, Dl, Dr = p["dims"]
        G = np.zeros((Br.dim, Br.dim))
        off = np.concatenate([[0], np.cumsum(Br.dims)])
        bs = d * Dl * Dr
        for k, i in zip(p["keys"], p["ids"]):
            il, ir = int(k[0]) - 1, int(k[1]) - 1
            blk = p["data"][(int(i) - 1) * bs:int(i) * bs].reshape((d, Dl, Dr), order="F")[:, :Bl.dims[il], :Br.dims[ir]]
            X = blk.reshape(-1, Br.dims[ir], order="F")
            G[off[ir]:off[ir + 1], off[ir]:off[ir + 1]] += X.T @ X
        np.testing.assert_allclose(G, np.eye(Br.dim), atol=1e-12)
