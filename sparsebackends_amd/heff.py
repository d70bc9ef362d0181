"""Two-site H_eff*psi of DMRG on symmetry-sector tensors, expressed as four
reference contract! calls on NewBlockSparseSorted tensors.

Sector structure (U(1) or U(1)xU(1) charges) is mapped onto the reference's
prefix-sparse format: every bond index becomes a sector axis (in the key
prefix) plus an in-sector axis (in the dense tail, padded to the largest
sector of the bond); physical and MPO-link indices stay in the dense tail.
The MPO site tensor is one dense block (P = 0).

The contraction order is psi*L*W1*W2*R (the order ITensorMPS uses for
product(ProjMPO, psi)); each step is contract_shared! BS x BS
(tensoralgebra/contract_bs_bs.jl:572-730) with at most one shared key axis.
Within each dense tail the small physical / MPO axes come first and the
padded in-sector axes last, so the nonzero part of a block is a few long
contiguous runs.

    psi  [Ql, Qr | s1, s2, l, r]
    L    [Ql', Ql | wl, l', l]          R [Qr', Qr | wr, r', r]
    W1   [ | wl, s1', s1, wm]           W2 [ | wm, s2', s2, wr]
    T1 = psi*L   [Ql', Qr | s1, s2, r, wl, l']
    T2 = T1*W1   [Ql', Qr | s1', wm, s2, r, l']
    T3 = T2*W2   [Ql', Qr | s2', wr, s1', r, l']
    Hpsi = T3*R  [Ql', Qr' | s1', s2', l', r']
"""
from __future__ import annotations

import numpy as np

from .storage import NewBlockSparseSorted
from .tensoralgebra import HeffPlan, Plan, contract_

# ------------------------------------------------------------------ sectors


def largest_remainder(weights, total):
    """Integer dims proportional to weights summing to total (largest remainder)."""
    w = np.asarray(weights, dtype=np.float64)
    raw = w / w.sum() * total
    base = np.floor(raw).astype(np.int64)
    rem = int(total - base.sum())
    order = np.argsort(-(raw - base), kind="stable")
    base[order[:rem]] += 1
    return base


class Bond:
    """A bond index split into charge sectors: charges[i] (tuple), dims[i]."""

    def __init__(self, charges, dims):
        keep = [i for i, d in enumerate(dims) if d > 0]
        self.charges = [tuple(charges[i]) for i in keep]
        self.dims = [int(dims[i]) for i in keep]
        self.index = {q: i for i, q in enumerate(self.charges)}

    @property
    def nsec(self):
        return len(self.charges)

    @property
    def dmax(self):
        return max(self.dims)

    @property
    def dim(self):
        return sum(self.dims)


def qadd(a, b):
    return tuple(x + y for x, y in zip(a, b))


def qsub(a, b):
    return tuple(x - y for x, y in zip(a, b))


# ------------------------------------------------------------------ models


class SiteModel:
    """Local space: charges of the d states; MPO W[wl, s', s, wr] (dense) and
    the charge Qlink[w] of every MPO link state, with W nonzero only where
    q(s') - q(s) = Qlink[wl] - Qlink[wr]."""

    def __init__(self, name, charges, W, Qlink, link_conj):
        self.name = name
        self.charges = [tuple(c) for c in charges]
        self.W = np.asarray(W, dtype=np.float64)
        self.Qlink = [tuple(q) for q in Qlink]
        self.d = len(charges)
        self.w = self.W.shape[0]
        # link involution with W[a, :, :, b].T == W[conj[a], :, :, conj[b]]: the
        # Hamiltonian is symmetric, and environments with L[conj[w]] = L[w].T keep
        # H_eff symmetric
        self.link_conj = list(link_conj)


def heisenberg_model(J=1.0, Jz=None):
    """S=1/2 XXZ chain MPO, w=5: index 0 = done, 4 = start; charge q = 2Sz."""
    Jz = J if Jz is None else Jz
    I = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])  # |up><dn|, state 0 = up
    Sm = Sp.T
    W = np.zeros((5, 2, 2, 5))
    W[0, :, :, 0] = I
    W[1, :, :, 0] = Sp
    W[2, :, :, 0] = Sm
    W[3, :, :, 0] = Sz
    W[4, :, :, 1] = 0.5 * J * Sm
    W[4, :, :, 2] = 0.5 * J * Sp
    W[4, :, :, 3] = Jz * Sz
    W[4, :, :, 4] = I
    return SiteModel("heisenberg", [(1,), (-1,)], W, [(0,), (2,), (-2,), (0,), (0,)], [0, 2, 1, 3, 4])


def j1j2_model(J1=1.0, J2=0.0):
    """S=1/2 J1-J2 chain MPO (w=8): nearest and next-nearest Heisenberg couplings."""
    I = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sm = Sp.T
    W = np.zeros((8, 2, 2, 8))
    # 0 done, 7 start; 1..3 carry (S+, S-, Sz) one site; 4..6 carry them two sites
    W[0, :, :, 0] = I
    W[1, :, :, 0] = Sp
    W[2, :, :, 0] = Sm
    W[3, :, :, 0] = Sz
    W[4, :, :, 1] = I
    W[5, :, :, 2] = I
    W[6, :, :, 3] = I
    W[7, :, :, 1] = 0.5 * J1 * Sm
    W[7, :, :, 2] = 0.5 * J1 * Sp
    W[7, :, :, 3] = J1 * Sz
    W[7, :, :, 4] = 0.5 * J2 * Sm
    W[7, :, :, 5] = 0.5 * J2 * Sp
    W[7, :, :, 6] = J2 * Sz
    W[7, :, :, 7] = I
    return SiteModel("j1j2", [(1,), (-1,)], W, [(0,), (2,), (-2,), (0,), (2,), (-2,), (0,), (0,)],
                     [0, 2, 1, 3, 5, 4, 6, 7])


def hubbard_model(t=1.0, U=4.0):
    """Hubbard chain MPO (w=6) with Jordan-Wigner strings; states |0>,|up>,|dn>,|updn>
    with charges (n, 2Sz); index 0 = done, 5 = start."""
    d = 4
    # c_up, c_dn on the site basis [0, up, dn, updn] with up ordered before dn
    cu = np.zeros((d, d))
    cd = np.zeros((d, d))
    cu[0, 1] = 1.0   # c_up |up> = |0>
    cu[2, 3] = 1.0   # c_up |updn> = |dn>
    cd[0, 2] = 1.0   # c_dn |dn> = |0>
    cd[1, 3] = -1.0  # c_dn |updn> = -|up>  (passes c_up^dag)
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    I = np.eye(d)
    nu = cu.T @ cu
    nd = cd.T @ cd
    W = np.zeros((6, d, d, 6))
    W[0, :, :, 0] = I
    W[5, :, :, 5] = I
    W[5, :, :, 0] = U * nu @ nd
    # c^dag_{i s} c_{i+1 s} and h.c., strings on the left operator
    W[5, :, :, 1] = -t * cu.T @ F
    W[5, :, :, 2] = -t * cd.T @ F
    W[5, :, :, 3] = t * cu @ F
    W[5, :, :, 4] = t * cd @ F
    W[1, :, :, 0] = cu
    W[2, :, :, 0] = cd
    W[3, :, :, 0] = cu.T
    W[4, :, :, 0] = cd.T
    charges = [(0, 0), (1, 1), (1, -1), (2, 0)]
    Qlink = [(0, 0), (-1, -1), (-1, 1), (1, 1), (1, -1), (0, 0)]
    return SiteModel("hubbard", charges, W, Qlink, [0, 3, 4, 1, 2, 5])


# ------------------------------------------------------------------ sector layouts of the configs


def heisenberg_bond(chi=256, qmax=8, width=32.0):
    """q in {-qmax..qmax}, dims proportional to exp(-q^2/width), summing to chi."""
    qs = list(range(-qmax, qmax + 1))
    dims = largest_remainder([np.exp(-q * q / width) for q in qs], chi)
    return Bond([(q,) for q in qs], dims)


def hubbard_bond(chi=2048, n0=0, sigma=7.0, span=3.0):
    """(n, 2Sz) sectors around (n0, 0) with dims from a 2D Gaussian of width sigma."""
    r = int(np.ceil(span * sigma))
    qs, w = [], []
    for n in range(n0 - r, n0 + r + 1):
        for s in range(-r, r + 1):
            if (n - s) % 2:
                continue  # 2Sz has the parity of n
            qs.append((n, s))
            w.append(np.exp(-((n - n0) ** 2 + s ** 2) / (2 * sigma * sigma)))
    dims = largest_remainder(w, chi)
    return Bond(qs, dims)


# ------------------------------------------------------------------ problem


def _lab(name):
    return (int.from_bytes(("h:" + name).encode()[:8].ljust(8, b"\0"), "little"), 0)


# labels of the four H_eff*psi contractions (layouts in the module docstring)
L_PSI = [_lab("Ql"), _lab("Qr"), _lab("s1"), _lab("s2"), _lab("l"), _lab("r")]
L_L = [_lab("Qlp"), _lab("Ql"), _lab("wl"), _lab("lp"), _lab("l")]
L_W1 = [_lab("wl"), _lab("s1p"), _lab("s1"), _lab("wm")]
L_W2 = [_lab("wm"), _lab("s2p"), _lab("s2"), _lab("wr")]
L_R = [_lab("Qrp"), _lab("Qr"), _lab("wr"), _lab("rp"), _lab("r")]
L_T1 = [_lab("Qlp"), _lab("Qr"), _lab("s1"), _lab("s2"), _lab("r"), _lab("wl"), _lab("lp")]
L_T2 = [_lab("Qlp"), _lab("Qr"), _lab("s1p"), _lab("wm"), _lab("s2"), _lab("r"), _lab("lp")]
L_T3 = [_lab("Qlp"), _lab("Qr"), _lab("s2p"), _lab("wr"), _lab("s1p"), _lab("r"), _lab("lp")]
L_OUT = [_lab("Qlp"), _lab("Qrp"), _lab("s1p"), _lab("s2p"), _lab("lp"), _lab("rp")]


def heff_plans(psi, L, W1, W2, R, flags=0):
    """the four H_eff*psi contractions of psi [Ql, Qr | s1, s2, l, r] with
    L [Ql', Ql | wl, l', l], W [wl, s', s, wr] and R [Qr', Qr | wr, r', r] as
    plans with fresh outputs; returns (plans, Hpsi)"""
    nl, nr, d1, d2, Dl, Dr = psi.dims
    w, wr = W1.dims[0], W2.dims[3]
    wm = W1.dims[3]
    dev = psi.device
    T1 = NewBlockSparseSorted.new((nl, nr, d1, d2, Dr, w, Dl), 5, device=dev)
    T2 = NewBlockSparseSorted.new((nl, nr, d1, wm, d2, Dr, Dl), 5, device=dev)
    T3 = NewBlockSparseSorted.new((nl, nr, d2, wr, d1, Dr, Dl), 5, device=dev)
    Hp = NewBlockSparseSorted.new((nl, nr, d1, d2, Dl, Dr), 4, device=dev)
    plans = [Plan(T1, L_T1, psi, L_PSI, L, L_L, flags), Plan(T2, L_T2, T1, L_T1, W1, L_W1, flags),
             Plan(T3, L_T3, T2, L_T2, W2, L_W2, flags), Plan(Hp, L_OUT, T3, L_T3, R, L_R, flags)]
    return plans, Hp


class HeffProblem:
    """A two-site H_eff*psi at one bond: left bond sectors Bl, right Br,
    site model m.  Random psi / L / R of the implied block structure (random
    values on the charge-allowed entries, zero elsewhere and in the padding).

    hermitian: environments symmetrised so that H_eff is symmetric
    (L[w] <- (L[w] + L[conj w]^T) / 2 with the model's link involution, same for R).
    envs: (Lfun, Rfun) giving exact environment blocks instead of random ones:
    Lfun(ip, i) -> array [w, Bl.dims[ip-1], Bl.dims[i-1]] (sector numbers 1-based)."""

    def __init__(self, model, Bl, Br, seed=0, device=0, hermitian=False, envs=None):
        self.m, self.Bl, self.Br = model, Bl, Br
        self.device = device
        rng = np.random.default_rng(seed)
        d, w = model.d, model.w
        Dl, Dr = Bl.dmax, Br.dmax
        self.Dl, self.Dr = Dl, Dr
        qs, Ql = model.charges, model.Qlink
        # psi [Ql, Qr | s1, s2, l, r] : q(r) = q(l) + q(s1) + q(s2)
        cl, cr, cs = np.array(Bl.charges), np.array(Br.charges), np.array(qs)
        allowed = np.all(cl[:, None, None, None, :] + cs[None, None, :, None, :] + cs[None, None, None, :, :]
                         == cr[None, :, None, None, :], axis=-1)  # [il, ir, s1, s2]
        psi = {(int(il) + 1, int(ir) + 1): allowed[il, ir].astype(np.float64)
               for il, ir in zip(*np.nonzero(allowed.any(axis=(2, 3))))}
        self.psi_parts = self._parts([Bl.nsec, Br.nsec], (d, d, Dl, Dr), psi,
                                     lambda k: (d, d, Bl.dims[k[0] - 1], Br.dims[k[1] - 1]),
                                     lambda k, blk: blk * psi[k][:, :, None, None], rng)

        # environments: X[w, x', x] nonzero where q(x') - q(x) = -Qlink(w)
        def env(B):
            c, cw = np.array(B.charges), np.array(Ql)
            m = np.all(c[:, None, None, :] - c[None, :, None, :] == -cw[None, None, :, :], axis=-1)  # [ip, i, w]
            return {(int(ip) + 1, int(i) + 1): m[ip, i].astype(np.float64) for ip, i in zip(*np.nonzero(m.any(axis=2)))}
        Lb, Rb = env(Bl), env(Br)
        self.L_parts = self._parts([Bl.nsec, Bl.nsec], (w, Dl, Dl), Lb,
                                   lambda k: (w, Bl.dims[k[0] - 1], Bl.dims[k[1] - 1]),
                                   lambda k, blk: blk * Lb[k][:, None, None], rng)
        self.R_parts = self._parts([Br.nsec, Br.nsec], (w, Dr, Dr), Rb,
                                   lambda k: (w, Br.dims[k[0] - 1], Br.dims[k[1] - 1]),
                                   lambda k, blk: blk * Rb[k][:, None, None], rng)
        if envs is not None:
            self._set_env(self.L_parts, (w, Dl, Dl), envs[0])
            self._set_env(self.R_parts, (w, Dr, Dr), envs[1])
        if hermitian:
            self._hermitize(self.L_parts, (w, Dl, Dl), model.link_conj)
            self._hermitize(self.R_parts, (w, Dr, Dr), model.link_conj)
        # W [ | wl, s', s, wr] : one dense block
        self.W_parts = dict(dims=(w, d, d, w), N2=4, keys=np.zeros((1, 0), dtype=np.int64),
                            ids=np.array([1]), data=np.asfortranarray(model.W).reshape(-1, order="F"))
        (self.l_psi, self.l_L, self.l_W1, self.l_W2, self.l_R, self.l_T1, self.l_T2, self.l_T3,
         self.l_out) = (L_PSI, L_L, L_W1, L_W2, L_R, L_T1, L_T2, L_T3, L_OUT)
        nl, nr = Bl.nsec, Br.nsec
        self.dims_T1 = (nl, nr, d, d, Dr, w, Dl)
        self.dims_T2 = (nl, nr, d, w, d, Dr, Dl)
        self.dims_T3 = (nl, nr, d, w, d, Dr, Dl)
        self.dims_out = (nl, nr, d, d, Dl, Dr)

    @staticmethod
    def _block(parts, i, tdims):
        bs = int(np.prod(tdims))
        return parts["data"][i * bs:(i + 1) * bs].reshape(tdims, order="F")  # a view

    def _set_env(self, parts, tdims, fun):
        for i, (a, b) in enumerate(parts["keys"]):
            blk = self._block(parts, i, tdims)
            v = np.asarray(fun(int(a), int(b)), dtype=np.float64)
            blk[...] = 0.0
            blk[:, :v.shape[1], :v.shape[2]] = v

    def _hermitize(self, parts, tdims, conj):
        index = {tuple(map(int, k)): i for i, k in enumerate(parts["keys"])}
        old = parts["data"].copy()
        src = dict(parts, data=old)
        for (a, b), i in index.items():
            j = index[(b, a)]  # present: the charge rule maps to itself under the involution
            A, B = self._block(src, i, tdims), self._block(src, j, tdims)
            self._block(parts, i, tdims)[...] = 0.5 * (A + B[conj].swapaxes(1, 2))

    @staticmethod
    def _parts(pdims, tdims, blocks, ext, mask, rng):
        """random payload on the actual extents ext(key), masked; padding zero"""
        keys = sorted(blocks, key=lambda k: tuple(reversed(k)))  # column-major order of the prefix
        bs = int(np.prod(tdims))
        data = np.zeros(len(keys) * bs)
        for i, k in enumerate(keys):
            e = ext(k)
            blk = np.zeros(tdims, order="F")
            sub = np.zeros(tdims)
            sl = tuple(slice(0, x) for x in e)
            sub[sl] = rng.standard_normal(e)
            blk[...] = mask(k, sub)
            data[i * bs:(i + 1) * bs] = blk.reshape(-1, order="F")
        return dict(dims=tuple(pdims) + tuple(tdims), N2=len(tdims),
                    keys=np.array(keys, dtype=np.int64).reshape(len(keys), len(pdims)),
                    ids=np.arange(1, len(keys) + 1), data=data)

    def upload(self):
        mk = lambda p: NewBlockSparseSorted.from_parts(p["dims"], p["N2"], p["keys"], p["ids"], p["data"],  # noqa
                                                       device=self.device)
        self.psi, self.L, self.R = mk(self.psi_parts), mk(self.L_parts), mk(self.R_parts)
        self.W1 = mk(self.W_parts)
        self.W2 = mk(self.W_parts)
        return self

    def outputs(self):
        new = lambda dims, N2: NewBlockSparseSorted.new(dims, N2, device=self.device)  # noqa
        self.T1 = new(self.dims_T1, 5)
        self.T2 = new(self.dims_T2, 5)
        self.T3 = new(self.dims_T3, 5)
        self.Hpsi = new(self.dims_out, 4)
        return self

    def contract_eager(self, flags=0):
        """Hpsi through four contract! calls (block structure recomputed each call)."""
        contract_(self.T1, self.l_T1, self.psi, self.l_psi, self.L, self.l_L, flags)
        contract_(self.T2, self.l_T2, self.T1, self.l_T1, self.W1, self.l_W1, flags)
        contract_(self.T3, self.l_T3, self.T2, self.l_T2, self.W2, self.l_W2, flags)
        contract_(self.Hpsi, self.l_out, self.T3, self.l_T3, self.R, self.l_R, flags)
        return self.Hpsi

    def make_plans(self, flags=0):
        """The four contractions as plans (structure fixed, values recomputed per execute)."""
        self.plans = [Plan(self.T1, self.l_T1, self.psi, self.l_psi, self.L, self.l_L, flags)]
        self.plans.append(Plan(self.T2, self.l_T2, self.T1, self.l_T1, self.W1, self.l_W1, flags))
        self.plans.append(Plan(self.T3, self.l_T3, self.T2, self.l_T2, self.W2, self.l_W2, flags))
        self.plans.append(Plan(self.Hpsi, self.l_out, self.T3, self.l_T3, self.R, self.l_R, flags))
        return self.plans

    def make_fused_plan(self, flags=0):
        """H_eff*psi as ONE plan with one fused kernel (HeffPlan), writing
        self.Hpsi_fused: the chain's block list, the chain's values"""
        self.Hpsi_fused = NewBlockSparseSorted.new(self.dims_out, 4, device=self.device)
        self.fused = HeffPlan(self.Hpsi_fused, self.psi, self.L, self.W1, self.W2, self.R, flags)
        return self.fused

    def execute(self):
        for p in self.plans:
            p.execute()

    # ---------------------------------------------------------- accounting
    def algorithmic_flops(self):
        """sum of 2*M*K*N over the matched block pairs of the four steps, with
        the actual sector extents in place of the padded ones"""
        d, w = self.m.d, self.m.w
        dl, dr = self.Bl.dims, self.Br.dims
        psi = [tuple(k) for k in self.psi_parts["keys"]]
        L_by = {}
        for k in self.L_parts["keys"]:
            L_by.setdefault(int(k[1]), []).append(int(k[0]))
        R_by = {}
        for k in self.R_parts["keys"]:
            R_by.setdefault(int(k[1]), []).append(int(k[0]))
        f = [0.0, 0.0, 0.0, 0.0]
        T1 = set()
        for (il, ir) in psi:
            for ilp in L_by.get(il, []):
                f[0] += 2.0 * (d * d * dr[ir - 1]) * dl[il - 1] * (dl[ilp - 1] * w)
                T1.add((ilp, ir))
        for (ilp, ir) in T1:
            f[1] += 2.0 * (d * dr[ir - 1] * dl[ilp - 1]) * (d * w) * (d * w)
            f[2] += 2.0 * (dr[ir - 1] * dl[ilp - 1] * d) * (d * w) * (d * w)
            for irp in R_by.get(ir, []):
                f[3] += 2.0 * (dl[ilp - 1] * d * d) * (dr[ir - 1] * w) * dr[irp - 1]
        return dict(total=sum(f), steps=f)

    def useful_flops(self):
        """algorithmic_flops with the two MPO steps counted only over the MPO's
        nonzero entries: the chain multiplies every T1 block by the whole
        dense (d w) x (d w) MPO block, most of whose entries are zero"""
        f = self.algorithmic_flops()["steps"]
        frac = float(np.count_nonzero(self.m.W)) / float(self.m.W.size)
        steps = [f[0], f[1] * frac, f[2] * frac, f[3]]
        return dict(total=sum(steps), steps=steps, mpo_nonzero_fraction=frac)

    def padded_flops(self):
        """the same sum with the padded extents the kernels execute"""
        d, w, Dl, Dr = self.m.d, self.m.w, self.Dl, self.Dr
        psi = [tuple(k) for k in self.psi_parts["keys"]]
        L_by = {}
        for k in self.L_parts["keys"]:
            L_by.setdefault(int(k[1]), []).append(int(k[0]))
        R_by = {}
        for k in self.R_parts["keys"]:
            R_by.setdefault(int(k[1]), []).append(int(k[0]))
        n1 = sum(len(L_by.get(il, [])) for (il, ir) in psi)
        T1 = {(ilp, ir) for (il, ir) in psi for ilp in L_by.get(il, [])}
        n4 = sum(len(R_by.get(ir, [])) for (ilp, ir) in T1)
        f = [2.0 * (d * d * Dr) * Dl * (Dl * w) * n1, 2.0 * (d * Dr * Dl) * (d * w) ** 2 * len(T1),
             2.0 * (Dr * Dl * d) * (d * w) ** 2 * len(T1), 2.0 * (Dl * d * d) * (Dr * w) * Dr * n4]
        return dict(total=sum(f), steps=f)

    def algorithmic_bytes(self):
        """float64 bytes each step must move at the least: its two operands read
        once and its result written once, every block counted with the actual
        sector extents (no padding)"""
        d, w = self.m.d, self.m.w
        dl, dr = self.Bl.dims, self.Br.dims
        psi = [tuple(map(int, k)) for k in self.psi_parts["keys"]]
        Lk = [tuple(map(int, k)) for k in self.L_parts["keys"]]
        Rk = [tuple(map(int, k)) for k in self.R_parts["keys"]]
        L_by, R_by = {}, {}
        for (a, b) in Lk:
            L_by.setdefault(b, []).append(a)
        for (a, b) in Rk:
            R_by.setdefault(b, []).append(a)
        T1 = {(a, ir) for (il, ir) in psi for a in L_by.get(il, [])}
        out = {(a, c) for (a, ir) in T1 for c in R_by.get(ir, [])}
        b_psi = sum(d * d * dl[a - 1] * dr[b - 1] for (a, b) in psi)
        b_L = sum(w * dl[a - 1] * dl[b - 1] for (a, b) in Lk)
        b_R = sum(w * dr[a - 1] * dr[b - 1] for (a, b) in Rk)
        b_T = sum(d * d * w * dl[a - 1] * dr[b - 1] for (a, b) in T1)
        b_out = sum(d * d * dl[a - 1] * dr[c - 1] for (a, c) in out)
        b_W = w * d * d * w
        steps = [8.0 * (b_psi + b_L + b_T), 8.0 * (b_T + b_W + b_T), 8.0 * (b_T + b_W + b_T), 8.0 * (b_T + b_R + b_out)]
        return dict(total=sum(steps), steps=steps)


def heisenberg_chi256_problem(seed=0, device=0, hermitian=False):
    """Heisenberg S=1/2, J=1, chi=256 in Sz sectors q in {-8..8}, dims ~ exp(-q^2/32)."""
    B = heisenberg_bond(256)
    return HeffProblem(heisenberg_model(1.0), B, B, seed=seed, device=device, hermitian=hermitian)


def hubbard_chi2048_problem(seed=0, sigma=7.0, device=0, hermitian=False):
    """Hubbard t=1, U=4, chi=2048 in (n, 2Sz) sectors with a 2D Gaussian law of width sigma."""
    Bl = hubbard_bond(2048, n0=0, sigma=sigma)
    Br = hubbard_bond(2048, n0=2, sigma=sigma)
    return HeffProblem(hubbard_model(1.0, 4.0), Bl, Br, seed=seed, device=device, hermitian=hermitian)


def j1j2_batch_problems(count=64, chi_lo=64, chi_hi=1024, seed0=0, device=0, hermitian=True):
    """Independent J1-J2 H_eff*psi problems at the centre bond of a long chain:
    problem i (seed seed0 + i) draws J2 uniform in [0, 0.5] and chi log-uniform
    in [chi_lo, chi_hi] from numpy's PCG64 with that seed; both bonds hold
    chi in Sz sectors q in {-8..8} with dims ~ exp(-q^2/32).  Returns a list of
    (HeffProblem, J2, chi)."""
    out = []
    for i in range(count):
        rng = np.random.default_rng(seed0 + i)
        J2 = float(rng.uniform(0.0, 0.5))
        chi = int(round(float(np.exp(rng.uniform(np.log(chi_lo), np.log(chi_hi))))))
        B = heisenberg_bond(chi)
        out.append((HeffProblem(j1j2_model(1.0, J2), B, B, seed=seed0 + i, device=device, hermitian=hermitian), J2,
                    chi))
    return out


# ------------------------------------------------------------------ exact chains


def product_bond(model, nsites, total=None):
    """The product basis of nsites sites (site 1 most significant, the order of
    np.kron) grouped into charge sectors.  total=None: a left block, bond charge
    = the block's charge; else a right block, bond charge = total - the block's
    charge.  Returns (Bond, sector of each state, position in its sector)."""
    d = model.d
    n = d ** nsites
    q = []
    for idx in range(n):
        digits = np.unravel_index(idx, (d,) * nsites) if nsites else ()
        c = tuple(0 for _ in model.charges[0])
        for s_ in digits:
            c = qadd(c, model.charges[s_])
        q.append(c if total is None else qsub(tuple(total), c))
    charges = sorted(set(q))
    sec = {c: i for i, c in enumerate(charges)}
    pos, counts = [], [0] * len(charges)
    for c in q:
        pos.append(counts[sec[c]])
        counts[sec[c]] += 1
    return Bond(charges, counts), [sec[c] for c in q], pos


def block_mpo_left(model, nsites):
    """M[w] = sum W[start, :, :, a1] x W[a1, :, :, a2] x ... x W[., :, :, w]
    over the product basis of the block (start = last link state)"""
    W = model.W
    M = [W[model.w - 1, :, :, a] for a in range(model.w)]
    for _ in range(nsites - 1):
        M = [sum(np.kron(M[a], W[a, :, :, b]) for a in range(model.w)) for b in range(model.w)]
    return M


def block_mpo_right(model, nsites):
    """N[w] = sum W[w, :, :, a] x ... x W[., :, :, done] (done = link state 0)"""
    W = model.W
    N = [W[a, :, :, 0] for a in range(model.w)]
    for _ in range(nsites - 1):
        N = [sum(np.kron(W[b, :, :, a], N[a]) for a in range(model.w)) for b in range(model.w)]
    return N


def chain_problem(model, nl, nr, total, seed=0, device=0):
    """H_eff of sites nl+1, nl+2 of an (nl+2+nr)-site open chain whose left and
    right bonds are the exact product bases of the outer sites: H_eff is then
    the whole chain's Hamiltonian in the sector of total charge `total`."""
    Bl, secl, posl = product_bond(model, nl)
    Br, secr, posr = product_bond(model, nr, total)
    Ml, Nr = block_mpo_left(model, nl), block_mpo_right(model, nr)
    st_l = [[i for i in range(len(secl)) if secl[i] == k] for k in range(Bl.nsec)]
    st_r = [[i for i in range(len(secr)) if secr[i] == k] for k in range(Br.nsec)]

    def Lfun(a, b):
        ra, rb = st_l[a - 1], st_l[b - 1]
        return np.stack([Ml[x][np.ix_(ra, rb)] for x in range(model.w)])

    def Rfun(a, b):
        ra, rb = st_r[a - 1], st_r[b - 1]
        return np.stack([Nr[x][np.ix_(ra, rb)] for x in range(model.w)])

    return HeffProblem(model, Bl, Br, seed=seed, device=device, envs=(Lfun, Rfun))
