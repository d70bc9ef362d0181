"""The element-sparse (COO) workload: two-site DMRG pieces of a PXP
(Rydberg-blockade) chain, local dimension 2 (|0>, |1>), in a bond basis with
no symmetry blocks, where the blockade n_i n_{i+1} = 1 is enforced by masking:
every bond basis state carries the occupation of the site next to the bond,
and psi [l, s1, s2, r] is zero wherever two neighbouring sites are occupied
(l's boundary site with s1, s1 with s2, s2 with r's boundary site).  Values
are N(0, 1) on the allowed entries (numpy PCG64, the given seed); the left
environment [l', w, l] is dense; the MPO is the Rydberg chain's
H = sum_i X_i + V sum_i n_i n_{i+1} with bond dimension 3.

The contraction measured is the reference's sparse-psi step psi (COO) x L
(dense) -> block-sparse [s1, s2, r | l', w] (contract_coo_dense.jl:48-125),
followed by the MPO step [s1, s2, r | l', w] x W [w, s1', s1, w'] (BS x Dense
contract_shared!, contract_bs_dense.jl:411-619)."""
from __future__ import annotations

import numpy as np


def _lab(name):
    return (int.from_bytes(("x:" + name).encode()[:8].ljust(8, b"\0"), "little"), 0)


L_PSI = [_lab("l"), _lab("s1"), _lab("s2"), _lab("r")]
L_ENV = [_lab("lp"), _lab("w"), _lab("l")]
L_C1 = [_lab("s1"), _lab("s2"), _lab("r"), _lab("lp"), _lab("w")]
L_W = [_lab("w"), _lab("s1p"), _lab("s1"), _lab("wr")]
L_C2 = [_lab("s2"), _lab("r"), _lab("lp"), _lab("s1p"), _lab("wr")]


def rydberg_mpo(V=1.0):
    """W [wl, s', s, wr] for H = sum_i X_i + V sum_i n_i n_{i+1} (w = 3; the
    left boundary selects link state 3, the right boundary link state 1)"""
    I = np.eye(2)
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    n = np.diag([0.0, 1.0])
    W = np.zeros((3, 2, 2, 3))
    W[0, :, :, 0] = I
    W[1, :, :, 0] = n
    W[2, :, :, 0] = X
    W[2, :, :, 1] = V * n
    W[2, :, :, 2] = I
    return W


def problem(chi=128, seed=0, p_occupied=0.9):
    """Host arrays of the workload: psi (chi, 2, 2, chi) with the blockade
    mask, L (chi, 3, chi), W (3, 2, 2, 3) and the fill fraction of psi.
    A boundary site is occupied with probability p_occupied; the expected fill
    is (3 - 2 p) / 4, so the default 0.9 gives the ~30 % of the workload."""
    rng = np.random.default_rng(seed)
    bl = rng.random(chi) < p_occupied  # occupation of the site left of the two sites
    br = rng.random(chi) < p_occupied  # occupation of the site right of them
    s = np.array([False, True])
    ok = (~(bl[:, None, None, None] & s[None, :, None, None]) & ~(s[None, :, None, None] & s[None, None, :, None])
          & ~(s[None, None, :, None] & br[None, None, None, :]))
    psi = np.where(ok, rng.standard_normal((chi, 2, 2, chi)), 0.0)
    L = rng.standard_normal((chi, 3, chi))
    return dict(psi=np.asfortranarray(psi), L=np.asfortranarray(L), W=np.asfortranarray(rydberg_mpo()),
                fill=float(ok.mean()), chi=chi)


def algorithmic_bytes(nnz, chi, nblocks, w=3):
    """psi (COO: four int32 coordinates + one float64 per entry) read once, L
    read once, the block-sparse result written once"""
    return nnz * (4 * 4 + 8) + chi * w * chi * 8 + nblocks * chi * w * 8
