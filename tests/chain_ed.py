"""Exact diagonalisation of short open chains, built from spin and fermion
operators (not from the MPOs), and the operator of a HeffProblem on the
oracle as a matvec over psi payload vectors."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from oracle.heff_oracle import HeffOracle


def _site_op(op, i, N, d, fill=None):
    """op on site i (0-based, site 0 most significant), identity (or fill) on
    the sites before it and identity after"""
    mats = []
    for k in range(N):
        if k == i:
            mats.append(sp.csr_matrix(op))
        elif k < i and fill is not None:
            mats.append(sp.csr_matrix(fill))
        else:
            mats.append(sp.identity(d, format="csr"))
    out = mats[0]
    for m in mats[1:]:
        out = sp.kron(out, m, format="csr")
    return out


def spin_chain(N, couplings):
    """sum over (range r, J, Jz) of J/2 (S+_i S-_{i+r} + h.c.) + Jz Sz_i Sz_{i+r};
    basis per site [up, dn]; charge 2Sz"""
    Sz = np.diag([0.5, -0.5])
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    H = sp.csr_matrix((2 ** N, 2 ** N))
    for r, J, Jz in couplings:
        for i in range(N - r):
            H = H + 0.5 * J * (_site_op(Sp, i, N, 2) @ _site_op(Sp.T, i + r, N, 2) +
                               _site_op(Sp.T, i, N, 2) @ _site_op(Sp, i + r, N, 2))
            H = H + Jz * _site_op(Sz, i, N, 2) @ _site_op(Sz, i + r, N, 2)
    q = np.zeros(2 ** N, dtype=np.int64)
    for idx in range(2 ** N):
        bits = np.unravel_index(idx, (2,) * N)
        q[idx] = sum(1 if b == 0 else -1 for b in bits)
    return H, [(int(v),) for v in q]


def hubbard_chain(N, t, U):
    """-t sum_{i,s} (c+_{i s} c_{i+1 s} + h.c.) + U sum_i n_{i up} n_{i dn};
    site basis [0, up, dn, updn] with |updn> = c+_up c+_dn |0>, modes ordered
    (1up, 1dn, 2up, 2dn, ...) for the Jordan-Wigner strings; charge (n, 2Sz)"""
    d = 4
    cu = np.zeros((d, d))
    cd = np.zeros((d, d))
    cu[0, 1] = 1.0
    cu[2, 3] = 1.0
    cd[0, 2] = 1.0
    cd[1, 3] = -1.0
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    c = {}
    for i in range(N):
        c[(i, "u")] = _site_op(cu, i, N, d, fill=F)
        c[(i, "d")] = _site_op(cd, i, N, d, fill=F)
    H = sp.csr_matrix((d ** N, d ** N))
    for i in range(N - 1):
        for s in "ud":
            hop = c[(i, s)].T @ c[(i + 1, s)]
            H = H - t * (hop + hop.T)
    for i in range(N):
        H = H + U * (c[(i, "u")].T @ c[(i, "u")]) @ (c[(i, "d")].T @ c[(i, "d")])
    charges = [(0, 0), (1, 1), (1, -1), (2, 0)]
    q = []
    for idx in range(d ** N):
        digs = np.unravel_index(idx, (d,) * N)
        q.append((sum(charges[x][0] for x in digs), sum(charges[x][1] for x in digs)))
    return H.tocsr(), q


def ground_energy(H, q, total):
    sel = np.array([i for i, c in enumerate(q) if tuple(c) == tuple(total)])
    Hs = H[sel][:, sel].toarray()
    return float(np.linalg.eigvalsh(Hs)[0]), len(sel)


class OracleHeff:
    """y = H_eff x over psi payload vectors (x in the layout of psi's host
    parts), through the oracle's four contract! calls; output blocks whose key
    psi lacks are dropped (they are zero)"""

    def __init__(self, prob):
        self.prob = prob
        self.o = HeffOracle(prob)
        p = prob.psi_parts
        self.bs = int(np.prod(p["dims"][2:]))
        self.index = {tuple(map(int, k)): int(i) for k, i in zip(p["keys"], p["ids"])}

    def __call__(self, x):
        p = self.prob.psi_parts
        self.o.psi = O.BS.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], np.asarray(x, dtype=np.float64))
        q = self.o.run().parts()
        y = np.zeros(len(p["ids"]) * self.bs)
        for k, i in zip(q["keys"], q["ids"]):
            j = self.index.get(tuple(map(int, k)))
            if j is None:  # outside psi's block space (structurally zero for a charge-conserving H_eff)
                assert not np.any(q["data"][(i - 1) * self.bs:i * self.bs])
                continue
            y[(j - 1) * self.bs:j * self.bs] = q["data"][(i - 1) * self.bs:i * self.bs]
        return y


def lowest_sparse(H, q, total):
    """ground energy of a larger sector through ARPACK (checker only)"""
    sel = np.array([i for i, c in enumerate(q) if tuple(c) == tuple(total)])
    Hs = H[sel][:, sel]
    return float(spla.eigsh(Hs, k=1, which="SA", tol=1e-14)[0][0])
