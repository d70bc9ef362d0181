"""Two-site DMRG on symmetry-sector block-sparse tensors, driven through the
library's C ABI: H_eff*psi as four contract! plans, the Lanczos eigensolver,
the sector SVD and the environment updates makeL!/makeR! as three contract!
calls each (ITensorMPS's L = L * psi[j] * H[j] * dag(prime(psi[j])), the
order the reference's tests decompose at test_product_dmrg_state.jl:313-378).

Tensors use the sector format of heff.py: every bond index is a sector axis
in the key prefix plus an in-sector axis in the dense tail, padded to the
bond's largest sector; bond charges accumulate from the left
(q(right bond) = q(left bond) + q(s)).

    site A_j  [Ql, Qr | s, l, r]
    L_j       [Ql', Ql | w, l', l]      (left environment of site j+1)
    R_j       [Qr', Qr | w, r', r]      (right environment of site j-1)
"""
from __future__ import annotations

import numpy as np

from .storage import NewBlockSparseSorted
from .tensoralgebra import Plan, contract_


def lab(name, plev=0):
    """a contraction label (id, plev) from a short name (at most 8 bytes)"""
    return (int.from_bytes(("" + name).encode()[:8].ljust(8, b"\0"), "little") | (1 << 63), plev)


# labels of the environment updates
_A = [lab("Ql"), lab("Qr"), lab("s"), lab("l"), lab("r")]
_Ap = [lab("Ql", 1), lab("Qr", 1), lab("s", 1), lab("l", 1), lab("r", 1)]  # the bra copy (primed)


def make_L(L, A, W, flags=0):
    """L_new[Qr', Qr | wr, r', r] = sum L[Ql', Ql | wl, l', l] A[Ql, Qr | s, l, r]
    W[wl, s', s, wr] A[Ql', Qr' | s', l', r'] as three contract! calls:
    X1 = L*A, X2 = X1*W, L_new = X2*A' (makeL!'s order).  Returns (L_new, the
    three plans, the intermediates)."""
    lL = [lab("Ql", 1), lab("Ql"), lab("wl"), lab("l", 1), lab("l")]
    lW = [lab("wl"), lab("s", 1), lab("s"), lab("wr")]
    dL, dA, dW = L.dims, A.dims, W.dims
    l1 = [lab("Ql", 1), lab("Qr"), lab("wl"), lab("l", 1), lab("s"), lab("r")]
    X1 = NewBlockSparseSorted.new((dL[0], dA[1], dL[2], dL[3], dA[2], dA[4]), 4, device=A.device)
    p1 = Plan(X1, l1, L, lL, A, _A, flags)
    l2 = [lab("Ql", 1), lab("Qr"), lab("l", 1), lab("r"), lab("s", 1), lab("wr")]
    X2 = NewBlockSparseSorted.new((dL[0], dA[1], dL[3], dA[4], dW[1], dW[3]), 4, device=A.device)
    p2 = Plan(X2, l2, X1, l1, W, lW, flags)
    # the bra index comes from the second operand and sits between the first
    # operand's w and ket index in the environment layout: contract to
    # [r', wr, r] (second operand's group first) and permute the tail
    l3 = [lab("Qr", 1), lab("Qr"), lab("r", 1), lab("wr"), lab("r")]
    X3 = NewBlockSparseSorted.new((dA[1], dA[1], dA[4], dW[3], dA[4]), 3, device=A.device)
    p3 = Plan(X3, l3, X2, l2, A, _Ap, flags)
    return X3.permutedims([1, 2, 4, 3, 5]), [p1, p2, p3], (X1, X2, X3)


def make_R(R, B, W, flags=0):
    """R_new[Ql', Ql | wl, l', l] = sum R[Qr', Qr | wr, r', r] B[Ql, Qr | s, l, r]
    W[wl, s', s, wr] B[Ql', Qr' | s', l', r'] (makeR!'s order: R*B, *W, *B')."""
    lR = [lab("Qr", 1), lab("Qr"), lab("wr"), lab("r", 1), lab("r")]
    lW = [lab("wl"), lab("s", 1), lab("s"), lab("wr")]
    dR, dB, dW = R.dims, B.dims, W.dims
    l1 = [lab("Qr", 1), lab("Ql"), lab("wr"), lab("r", 1), lab("s"), lab("l")]
    X1 = NewBlockSparseSorted.new((dR[0], dB[0], dR[2], dR[3], dB[2], dB[3]), 4, device=B.device)
    p1 = Plan(X1, l1, R, lR, B, _A, flags)
    l2 = [lab("Qr", 1), lab("Ql"), lab("r", 1), lab("l"), lab("s", 1), lab("wl")]
    X2 = NewBlockSparseSorted.new((dR[0], dB[0], dR[3], dB[3], dW[1], dW[0]), 4, device=B.device)
    p2 = Plan(X2, l2, X1, l1, W, lW, flags)
    l3 = [lab("Ql", 1), lab("Ql"), lab("l", 1), lab("wl"), lab("l")]
    X3 = NewBlockSparseSorted.new((dB[0], dB[0], dB[3], dW[0], dB[3]), 3, device=B.device)
    p3 = Plan(X3, l3, X2, l2, B, _Ap, flags)
    return X3.permutedims([1, 2, 4, 3, 5]), [p1, p2, p3], (X1, X2, X3)


# ---------------------------------------------------------------- host-side parts


def site_parts(model, Bl, Br, rng=None, fill=None):
    """A [Ql, Qr | s, l, r] host parts: blocks where q(r) = q(l) + q(s) for some s;
    values random (rng) or from fill(ql_index, qr_index, s) -> (dl, dr) array,
    zero on the charge-forbidden s and in the padding"""
    d = model.d
    cl, cr, cs = np.array(Bl.charges), np.array(Br.charges), np.array(model.charges)
    ok = np.all(cl[:, None, None, :] + cs[None, None, :, :] == cr[None, :, None, :], axis=-1)  # [il, ir, s]
    keys = [(int(il) + 1, int(ir) + 1) for ir in range(Br.nsec) for il in range(Bl.nsec) if ok[il, ir].any()]
    tdims = (d, Bl.dmax, Br.dmax)
    bs = int(np.prod(tdims))
    data = np.zeros(len(keys) * bs)
    for i, (a, b) in enumerate(keys):
        blk = np.zeros(tdims, order="F")
        dl, dr = Bl.dims[a - 1], Br.dims[b - 1]
        for s_ in range(d):
            if ok[a - 1, b - 1, s_]:
                blk[s_, :dl, :dr] = fill(a, b, s_) if fill is not None else rng.standard_normal((dl, dr))
        data[i * bs:(i + 1) * bs] = blk.reshape(-1, order="F")
    return dict(dims=(Bl.nsec, Br.nsec) + tdims, N2=3, keys=np.array(keys, dtype=np.int64).reshape(-1, 2),
                ids=np.arange(1, len(keys) + 1), data=data)


def env_parts(model, B, rng):
    """random environment [Q', Q | w, x', x] on the blocks the charge rule allows
    (q(x') - q(x) = -Qlink(w)), zero elsewhere"""
    from .heff import HeffProblem
    c, cw = np.array(B.charges), np.array(model.Qlink)
    m = np.all(c[:, None, None, :] - c[None, :, None, :] == -cw[None, None, :, :], axis=-1)
    blocks = {(int(ip) + 1, int(i) + 1): m[ip, i].astype(np.float64) for ip, i in zip(*np.nonzero(m.any(axis=2)))}
    w = model.w
    return HeffProblem._parts([B.nsec, B.nsec], (w, B.dmax, B.dmax), blocks,
                              lambda k: (w, B.dims[k[0] - 1], B.dims[k[1] - 1]),
                              lambda k, blk: blk * blocks[k][:, None, None], rng)


def mpo_parts(model):
    return dict(dims=model.W.shape, N2=4, keys=np.zeros((1, 0), dtype=np.int64), ids=np.array([1]),
                data=np.asfortranarray(model.W).reshape(-1, order="F"))


def upload(p, device=0):
    return NewBlockSparseSorted.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"], device=device)


def fused_sectors(model, Bl, Br):
    """sector tables of a two-site psi [Ql, Qr | s1, s2, l, r] for its SVD:
    rows (Ql, s1) carry q(Ql) + q(s1), columns (Qr, s2) carry q(Qr) - q(s2);
    the sectors are the charges present on both sides, ascending.  Returns
    (rowsec (nl, d), colsec (nr, d), charges)."""
    from .heff import qadd, qsub
    d = model.d
    rq = {(il, s): qadd(Bl.charges[il], model.charges[s]) for il in range(Bl.nsec) for s in range(d)}
    cq = {(ir, s): qsub(Br.charges[ir], model.charges[s]) for ir in range(Br.nsec) for s in range(d)}
    charges = sorted(set(rq.values()) & set(cq.values()))
    idx = {q: i for i, q in enumerate(charges)}
    rowsec = np.full((Bl.nsec, d), -1, dtype=np.int64)
    colsec = np.full((Br.nsec, d), -1, dtype=np.int64)
    for (il, s), q in rq.items():
        rowsec[il, s] = idx.get(q, -1)
    for (ir, s), q in cq.items():
        colsec[ir, s] = idx.get(q, -1)
    return rowsec, colsec, charges


# ---------------------------------------------------------------- two-site DMRG

_PSI_AB = [lab("Ql"), lab("Qr"), lab("s1"), lab("l"), lab("s2"), lab("r")]


def two_site(A, B):
    """psi [Ql, Qr | s1, s2, l, r] = A [Ql, Qm | s1, l, m] * B [Qm, Qr | s2, m, r]
    (contracted to [s1, l, s2, r], each operand's group contiguous, then the
    tail permuted)"""
    lA = [lab("Ql"), lab("Qm"), lab("s1"), lab("l"), lab("m")]
    lB = [lab("Qm"), lab("Qr"), lab("s2"), lab("m"), lab("r")]
    nl, _, d1, Dl, _ = A.dims
    _, nr, d2, _, Dr = B.dims
    X = NewBlockSparseSorted.new((nl, nr, d1, Dl, d2, Dr), 4, device=A.device)
    contract_(X, _PSI_AB, A, lA, B, lB)
    return X.permutedims([1, 2, 3, 5, 4, 6])


def config_bonds(N, chi=256, qmax=8, width=32.0):
    """The bond layout of the plain configuration (S=1/2 chain, q = 2 Sz, total
    Sz = 0): bond k (k sites to its left) holds the sectors q in {-qmax..qmax}
    that the left block reaches with the right block still able to cancel it
    (q = k mod 2, |q| <= min(k, N - k)), with dims proportional to
    exp(-q^2 / width) summing (largest remainder) to min(chi, 2^k, 2^(N-k)),
    each capped by the states the two blocks have in that sector
    (binomial counts); bonds 0 and N are the trivial sector of dim 1."""
    from math import comb
    from .heff import Bond, largest_remainder
    out = []
    for k in range(N + 1):
        qs = [q for q in range(-qmax, qmax + 1) if (q - k) % 2 == 0 and abs(q) <= min(k, N - k)]
        cap = min(chi, 2 ** min(k, 60), 2 ** min(N - k, 60))
        dims = largest_remainder([np.exp(-q * q / width) for q in qs], cap) if qs else []
        dims = [min(int(dm), comb(k, (k + q) // 2), comb(N - k, (N - k - q) // 2)) for q, dm in zip(qs, dims)]
        out.append(Bond([(q,) for q in qs], dims))
    return out


def random_mps(model, bonds, seed=0, device=0):
    """site tensors A_j [Ql, Qr | s, l, r] on the bonds' sectors with i.i.d.
    N(0, 1) entries (numpy PCG64, the given seed) on the charge-allowed ones"""
    rng = np.random.default_rng(seed)
    return [upload(site_parts(model, bonds[j], bonds[j + 1], rng=rng), device) for j in range(len(bonds) - 1)]


def left_orthonormalize(model, A, bonds):
    """Left-orthonormalise an MPS by QR, sector by sector (sb_bs_qr_sectors):
    A_j = Q_j R_j, A_{j+1} <- R_j A_{j+1}, from the first site to the last;
    the last site's R (the state's norm) is dropped, so the state ends
    normalised.  Returns (the new site tensors, the new bonds)."""
    from .factorize import qr_sectors
    from .heff import Bond, qadd
    A = list(A)
    bonds = list(bonds)
    d = model.d
    lR = [lab("Qm"), lab("Qr"), lab("m"), lab("r")]
    lA = [lab("Qr"), lab("Qr2"), lab("s"), lab("r"), lab("r2")]
    lC = [lab("Qm"), lab("Qr2"), lab("m"), lab("s"), lab("r2")]
    for j in range(len(A)):
        Bl, Br = bonds[j], bonds[j + 1]
        rowsec = np.full((Bl.nsec, d), -1, dtype=np.int64)
        for il in range(Bl.nsec):
            for s_ in range(d):
                q = qadd(Bl.charges[il], model.charges[s_])
                rowsec[il, s_] = Br.index.get(q, -1)
        Q, R, kept = qr_sectors(A[j], rowsec, Bl.dims, Br.dims)
        if np.any(kept == 0):
            raise ValueError(f"bond {j + 1}: a sector without rows (unreachable charge)")
        A[j] = Q
        bonds[j + 1] = Bond(Br.charges, [int(x) for x in kept])
        if j + 1 < len(A):
            nx = A[j + 1]
            dims = nx.dims
            C_ = NewBlockSparseSorted.new((dims[0], dims[1], dims[3], d, dims[4]), 3, device=nx.device)
            contract_(C_, lC, R, lR, nx, lA)  # (R A)[Qm, Qr2 | m, s, r2]
            A[j + 1] = C_.permutedims([1, 2, 4, 3, 5])
    return A, bonds


class DMRG:
    """Two-site DMRG of an open chain with a uniform site model, in a fixed
    total-charge sector, from a product state (or from given site tensors,
    DMRG.from_mps).

    Tensors, environments and bonds stay on the device between steps; the
    host keeps the sector structure (charges and extents of every bond)."""

    def __init__(self, model, states, device=0, flags=0, eigsolver=None, environments=True):
        """eigsolver(plans, psi, krylovdim=, maxiter=, tol=) -> (vals, vecs, info)
        replaces krylov.eigsolve (a serving loop's co-batching hook)."""
        from .heff import Bond
        from .krylov import eigsolve
        self.m, self.N, self.device, self.flags = model, len(states), device, flags
        self.eigsolver = eigsolver if eigsolver is not None else eigsolve
        d, w = model.d, model.w
        zero = tuple(0 for _ in model.charges[0])
        # bonds 0..N: bond j sits between sites j and j+1; charges accumulate from the left
        charges = [zero]
        for s_ in states:
            charges.append(tuple(a + b for a, b in zip(charges[-1], model.charges[s_])))
        self.total = charges[-1]
        self.bonds = [Bond([q], [1]) for q in charges]
        self.A = []
        for j, s_ in enumerate(states):
            blk = np.zeros((d, 1, 1))
            blk[s_, 0, 0] = 1.0
            p = dict(dims=(1, 1, d, 1, 1), N2=3, keys=np.array([[1, 1]]), ids=np.array([1]),
                     data=blk.reshape(-1, order="F"))
            self.A.append(upload(p, device))
        self.W = upload(mpo_parts(model), device)
        # boundary environments: L0 selects the start link state, R_{N+1} the done state
        e = np.zeros((w, 1, 1))
        e[w - 1, 0, 0] = 1.0
        self.L = {0: upload(dict(dims=(1, 1, w, 1, 1), N2=3, keys=np.array([[1, 1]]), ids=np.array([1]),
                                 data=e.reshape(-1, order="F")), device)}
        e = np.zeros((w, 1, 1))
        e[0, 0, 0] = 1.0
        self.R = {self.N + 1: upload(dict(dims=(1, 1, w, 1, 1), N2=3, keys=np.array([[1, 1]]), ids=np.array([1]),
                                          data=e.reshape(-1, order="F")), device)}
        if environments:
            for j in range(self.N, 2, -1):  # R_j from R_{j+1} and site j (1-based)
                self.R[j] = make_R(self.R[j + 1], self.A[j - 1], self.W, flags)[0]
        self.energies = []
        self.truncerr = 0.0
        # phase timing of the two-site steps: a dict {phase: seconds} turns it
        # on (each phase then ends with a device synchronisation)
        self.phase_times = None

    @classmethod
    def from_mps(cls, model, A, bonds, device=0, flags=0, eigsolver=None):
        """DMRG from a left-orthonormal MPS (site tensors A, bonds): the right
        environments are built from it, as from a product state"""
        self = cls(model, [0] * len(A), device=device, flags=flags, eigsolver=eigsolver, environments=False)
        self.A = list(A)
        self.bonds = list(bonds)
        self.total = bonds[-1].charges[0]
        for j in range(self.N, 2, -1):
            self.R[j] = make_R(self.R[j + 1], self.A[j - 1], self.W, flags)[0]
        return self

    def _phase(self, name, t0):
        """with phase_times set (a dict), the device is synchronised and the
        seconds since t0 are added to phase_times[name]; returns the new t0"""
        if self.phase_times is None:
            return t0
        import time
        from .storage import synchronize
        synchronize(self.device)
        t = time.perf_counter()
        self.phase_times[name] = self.phase_times.get(name, 0.0) + (t - t0)
        return t

    def step(self, b, direction, maxdim, cutoff, krylovdim, maxiter, tol, mindim=1):
        """optimise sites b, b+1 (1-based): eigsolve on H_eff, sector SVD, environment update"""
        import time
        from .heff import Bond, heff_plans
        from .factorize import svd_sectors
        t0 = 0.0
        if self.phase_times is not None:
            from .storage import synchronize
            synchronize(self.device)
            t0 = time.perf_counter()
        psi = two_site(self.A[b - 1], self.A[b])
        t0 = self._phase("two_site", t0)
        plans, Hp = heff_plans(psi, self.L[b - 1], self.W, self.W, self.R[b + 2], self.flags)
        t0 = self._phase("heff_plans", t0)
        vals, _, info = self.eigsolver(plans, psi, krylovdim=krylovdim, maxiter=maxiter, tol=tol)
        t0 = self._phase("eigsolve", t0)
        Bl, Br = self.bonds[b - 1], self.bonds[b + 1]
        rowsec, colsec, charges = fused_sectors(self.m, Bl, Br)
        A, B, svs, kept, spec = svd_sectors(psi, rowsec, colsec, Bl.dims, Br.dims, len(charges),
                                            ortho="left" if direction > 0 else "right", maxdim=maxdim,
                                            mindim=mindim, cutoff=cutoff)
        t0 = self._phase("svd", t0)
        live = [i for i in range(len(charges)) if kept[i] > 0]
        self.bonds[b] = Bond([charges[i] for i in live], [int(kept[i]) for i in live])
        self.A[b - 1], self.A[b] = A, B
        # the environment this step used on the side it leaves behind is rebuilt
        # before its next use (by the steps of the opposite direction), so it is
        # freed now: only one side's environments are resident at a time
        if direction > 0:
            self.L[b] = make_L(self.L[b - 1], A, self.W, self.flags)[0]
            if b + 2 <= self.N:
                self.R.pop(b + 2, None)
        else:
            self.R[b + 1] = make_R(self.R[b + 2], B, self.W, self.flags)[0]
            if b - 1 >= 1:
                self.L.pop(b - 1, None)
        self._phase("environment", t0)
        self.truncerr = max(self.truncerr, spec.truncerr)
        return vals[0], info

    def sweep(self, maxdim=None, cutoff=0.0, krylovdim=3, maxiter=1, tol=1e-14):
        """one left-to-right and one right-to-left half sweep; returns the last energy"""
        self.truncerr = 0.0
        e = None
        for b in range(1, self.N):
            e, _ = self.step(b, +1, maxdim, cutoff, krylovdim, maxiter, tol)
        for b in range(self.N - 1, 0, -1):
            e, _ = self.step(b, -1, maxdim, cutoff, krylovdim, maxiter, tol)
        self.energies.append(e)
        return e

    def maxlinkdim(self):
        return max(bd.dim for bd in self.bonds)
