"""Two-site SVD with truncation on block-sparse storage:
blocksparse_svd (SparseBackends/src/ops_factorize.jl:50-148) through the
library's sb_bs_svd."""
from __future__ import annotations

import ctypes as C
from collections import namedtuple

import numpy as np

from ._abi import check, load
from .storage import NewBlockSparseSorted

# ITensors.Spectrum(eigs, truncerr): eigs = kept singular values squared
Spectrum = namedtuple("Spectrum", ["eigs", "truncerr"])


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
