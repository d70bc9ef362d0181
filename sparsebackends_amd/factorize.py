"""Factorisations on block-sparse storage: blocksparse_left_qr and
left_factorize (SparseBackends/src/ops_factorize.jl:8-43,
tensor_wrappers.jl:954-970), the two-site SVD with truncation
blocksparse_svd (ops_factorize.jl:50-148), and their sector-format
counterparts, through the library's sb_bs_left_qr / sb_bs_svd family."""
from __future__ import annotations

import ctypes as C
from collections import namedtuple

import numpy as np

from ._abi import check, load
from .storage import DenseTensor, NewBlockSparseSorted

# ITensors.Spectrum(eigs, truncerr): eigs = kept singular values squared
Spectrum = namedtuple("Spectrum", ["eigs", "truncerr"])


def blocksparse_left_qr(A):
    """A: NewBlockSparseSorted with prefix (l, r) and dense s.  Returns
    (Q, R, l_vals, r_vals) like the reference: Q keyed by l with dense
    (s, new bond), R the k x n_r matrix (numpy) over the ascending nonzero r
    values, l_vals / r_vals the sorted distinct l and r values (1-based)."""
    Q, R = C.c_void_p(), C.c_void_p()
    check(load().sb_bs_left_qr(A.ctx, A.handle, C.byref(Q), C.byref(R)))
    keys = A.parts()["keys"]
    l_vals = np.unique(keys[:, 0]).astype(np.int64)
    r_vals = np.unique(keys[:, 1]).astype(np.int64)
    return NewBlockSparseSorted(Q, A.device), DenseTensor(R, A.device).numpy(), l_vals, r_vals


def left_factorize(A):
    """left_factorize below its ITensor layer: (Q, R_full) with R_full the
    k x d_r array (numpy) whose column r holds R's column of r value r"""
    Q, R = C.c_void_p(), C.c_void_p()
    check(load().sb_bs_left_factorize(A.ctx, A.handle, C.byref(Q), C.byref(R)))
    return NewBlockSparseSorted(Q, A.device), DenseTensor(R, A.device).numpy()


def qr_sectors(A, rowsec, dl, dr):
    """Sector QR of a site tensor A [Ql, Qr | s, l, r] (sb_bs_qr_sectors):
    rowsec[il, s] = the right sector of (Ql, s) or -1 (shape (nl, d)); dl, dr
    the sector extents.  Returns (Q with A's block list, R [Qr, Qr | m, r],
    kept: the new extent of every right sector)."""
    i64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64))  # noqa: E731
    rs = i64(np.asarray(rowsec).reshape(-1, order="F"))
    dl_, dr_ = i64(dl), i64(dr)
    kept = np.zeros(max(1, len(dr_)), dtype=np.int64)
    Q, R = C.c_void_p(), C.c_void_p()
    p64 = C.POINTER(C.c_int64)
    check(load().sb_bs_qr_sectors(A.ctx, A.handle, rs.ctypes.data_as(p64), dl_.ctypes.data_as(p64),
                                  dr_.ctypes.data_as(p64), C.byref(Q), C.byref(R), kept.ctypes.data_as(p64)))
    return NewBlockSparseSorted(Q, A.device), NewBlockSparseSorted(R, A.device), kept[:len(dr_)].copy()


def blocksparse_svd(A, ortho="left", maxdim=None, mindim=1, cutoff=0.0):
    """A: NewBlockSparseSorted with prefix (L, R) and dense tail (s1, s2).
    Returns (U, SV, svs_kept, spec) like the reference: U keys (L, nb) with
    tail s1, SV keys (nb, R) with tail s2; the singular value sits in SV for
    ortho "left" and in U for ortho "right"."""
    if ortho not in ("left", "right"):
        raise ValueError("ortho must be 'left' or 'right'")
    dims = A.dims
    cap = int(dims[2]) * max(1, A.nblocks)
    svs = np.zeros(cap)
    U, SV = C.c_void_p(), C.c_void_p()
    k, terr = C.c_int64(), C.c_double()
    check(load().sb_bs_svd(A.ctx, A.handle, 1 if ortho == "right" else 0, -1 if maxdim is None else int(maxdim),
                           int(mindim), float(cutoff), C.byref(U), C.byref(SV),
                           svs.ctypes.data_as(C.POINTER(C.c_double)), cap, C.byref(k), C.byref(terr)))
    s = svs[:k.value].copy()
    return (NewBlockSparseSorted(U, A.device), NewBlockSparseSorted(SV, A.device), s, Spectrum(s ** 2, terr.value))


def svd_sectors(psi, rowsec, colsec, dl, dr, nsec, ortho="left", maxdim=None, mindim=1, cutoff=0.0):
    """Block-diagonal SVD of psi [Ql, Qr | s1, s2, l, r] by fused charge sector
    (sb_bs_svd_sectors).  rowsec[il, s1] / colsec[ir, s2]: sector number or -1
    (arrays of shape (nl, d1) / (nr, d2)); dl, dr: actual bond sector extents.
    Returns (A [Ql, Qm | s1, l, m], B [Qm, Qr | s2, m, r], svs, kept per sector, spec)."""
    if ortho not in ("left", "right"):
        raise ValueError("ortho must be 'left' or 'right'")
    i64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int64))  # noqa: E731
    rs = i64(np.asarray(rowsec).reshape(-1, order="F"))
    cs = i64(np.asarray(colsec).reshape(-1, order="F"))
    dl_, dr_ = i64(dl), i64(dr)
    kept = np.zeros(max(1, nsec), dtype=np.int64)
    cap = int(sum(int(x) for x in dl_)) * int(psi.dims[2]) + 1
    svs = np.zeros(cap)
    A, B = C.c_void_p(), C.c_void_p()
    k, terr = C.c_int64(), C.c_double()
    p64 = C.POINTER(C.c_int64)
    check(load().sb_bs_svd_sectors(psi.ctx, psi.handle, int(nsec), rs.ctypes.data_as(p64), cs.ctypes.data_as(p64),
                                   dl_.ctypes.data_as(p64), dr_.ctypes.data_as(p64), 1 if ortho == "right" else 0,
                                   -1 if maxdim is None else int(maxdim), int(mindim), float(cutoff), C.byref(A),
                                   C.byref(B), kept.ctypes.data_as(p64), svs.ctypes.data_as(C.POINTER(C.c_double)),
                                   cap, C.byref(k), C.byref(terr)))
    s = svs[:k.value].copy()
    return (NewBlockSparseSorted(A, psi.device), NewBlockSparseSorted(B, psi.device), s, kept[:nsec].copy(),
            Spectrum(s ** 2, terr.value))
