"""eigsolve for two-site DMRG: the lowest eigenpair of H_eff given as a chain
of contraction plans, solved by the library's FP64 Lanczos.

The call shape follows KrylovKit's eigsolve(A, x0, howmany, which; ishermitian,
krylovdim, maxiter, tol) as ITensorMPS-based DMRG uses it on H_eff
(test_product_dmrg_state.jl:221-226 with krylovdim=3, maxiter=1, tol=1e-14;
test_product.jl:235-238 with krylovdim=8, tol=1e-12).  Defaults are KrylovKit's
(krylovdim=30, maxiter=100, tol=1e-12).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import check, load


@dataclass
class ConvergenceInfo:
    """converged: 1 if the lowest Ritz pair met tol; normres: its residual norm;
    numiter: Krylov factorisations used; numops: operator applications;
    ritz: Ritz values of the last factorisation, ascending."""
    converged: int
    normres: float
    numiter: int
    numops: int
    ritz: np.ndarray = field(repr=False)


def eigsolve(plans, x0, howmany=1, which="SR", *, y=None, ishermitian=True, krylovdim=30, maxiter=100, tol=1e-12):
    """Lowest eigenpair of the Hermitian operator y = plans(x0).

    plans: Plan objects applied in order; the first reads x0, the last writes
    y (by default the last plan's output).  x0 (a NewBlockSparseSorted) is the
    start vector and is overwritten with the normalised Ritz vector.
    Returns (vals, vecs, info) like KrylovKit: vals = [lowest Ritz value],
    vecs = [x0], info = ConvergenceInfo."""
    if howmany != 1 or which != "SR" or not ishermitian:
        raise ValueError("eigsolve supports howmany=1, which='SR', ishermitian=True (the DMRG use)")
    plans = list(plans)
    if y is None:
        y = plans[-1]._ops[0]
    k = max(int(krylovdim), 2)
    arr = (C.c_void_p * len(plans))(*[p._h.value for p in plans])
    val, nres = C.c_double(), C.c_double()
    ritz = np.zeros(k, dtype=np.float64)
    nritz, numiter, conv = C.c_int(), C.c_int(), C.c_int()
    numops = C.c_int64()
    check(load().sb_eigsolve_lanczos(x0.ctx, C.cast(arr, C.POINTER(C.c_void_p)), len(plans), x0.handle, y.handle, k,
                                     int(maxiter), float(tol), C.byref(val),
                                     ritz.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nritz), C.byref(nres),
                                     C.byref(numops), C.byref(numiter), C.byref(conv)))
    info = ConvergenceInfo(conv.value, nres.value, numiter.value, numops.value, ritz[:nritz.value].copy())
    return [val.value], [x0], info


def eigsolve_refined(plans_lo, plans_hi, x_lo, x_hi, *, krylovdim=30, maxiter=100, tol_lo=1e-7, tol=1e-12,
                     polish_krylovdim=8, polish_maxiter=10):
    """Mixed-precision solve: Lanczos on the low-precision operator plans_lo
    (FP32 or bf16x3 block products) down to tol_lo, then FP64 Lanczos on
    plans_hi from that Ritz vector down to tol.  The energy of the second
    stage is a Rayleigh-Ritz value of the FP64 operator, so its error is
    quadratic in the first stage's eigenvector error.  x_lo and x_hi are the
    vectors the two chains read (same block structure); x_hi ends with the
    Ritz vector.  Returns (vals, vecs, info_lo, info_hi)."""
    _, _, info_lo = eigsolve(plans_lo, x_lo, krylovdim=krylovdim, maxiter=maxiter, tol=tol_lo)
    x_hi.axpby_(1.0, x_lo, 0.0)
    vals, vecs, info_hi = eigsolve(plans_hi, x_hi, krylovdim=polish_krylovdim, maxiter=polish_maxiter, tol=tol)
    return vals, vecs, info_lo, info_hi


def eigsolve_batch(batch, xs, *, ys=None, krylovdim=30, maxiter=100, tol=1e-12):
    """eigsolve for every chain of a PlanBatch at once: chain c is the
    operator from xs[c] (its first plan reads it) to ys[c] (by default its
    last plan's output).  The solves run in lock step, one batched operator
    application per iteration for the chains still running; each result and
    Ritz vector (written into xs[c]) equals what eigsolve gives for that
    chain alone, bit for bit.  Returns (vals, infos) with one entry per chain
    (infos without Ritz values)."""
    n = batch.nchains
    xs = list(xs)
    if len(xs) != n:
        raise ValueError("one start vector per chain of the batch")
    if ys is None:
        ys = [ch[-1]._ops[0] for ch in batch._chains]
    k = max(int(krylovdim), 2)
    ax = (C.c_void_p * n)(*[x.handle.value if hasattr(x.handle, "value") else x.handle for x in xs])
    ay = (C.c_void_p * n)(*[y.handle.value if hasattr(y.handle, "value") else y.handle for y in ys])
    vals, nres = np.zeros(n), np.zeros(n)
    numops = np.zeros(n, dtype=np.int64)
    numiter, conv = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    f64 = C.POINTER(C.c_double)
    i32 = C.POINTER(C.c_int)
    check(load().sb_eigsolve_lanczos_batch(batch.ctx, batch._h, C.cast(ax, C.POINTER(C.c_void_p)),
                                           C.cast(ay, C.POINTER(C.c_void_p)), n, k, int(maxiter), float(tol),
                                           vals.ctypes.data_as(f64), nres.ctypes.data_as(f64),
                                           numops.ctypes.data_as(C.POINTER(C.c_int64)), numiter.ctypes.data_as(i32),
                                           conv.ctypes.data_as(i32)))
    infos = [ConvergenceInfo(int(conv[c]), float(nres[c]), int(numiter[c]), int(numops[c]), np.zeros(0))
             for c in range(n)]
    return [float(v) for v in vals], infos


def eigsolve_multi(chains, xs, *, ys=None, ctx=None, krylovdim=30, maxiter=100, tol=1e-12):
    """eigsolve_batch over chains whose plans and vectors may belong to the
    contexts of other host threads (one context per thread, same device): the
    solves run in lock step on the stream of ctx (default: the calling
    thread's context), which the other contexts borrow for the duration of the
    call (sb_eigsolve_lanczos_multi).  chains[c] is the plan list of problem c
    (every chain the same length), xs[c] its start vector (overwritten with the
    Ritz vector).  The threads owning those contexts must not use them during
    the call.  Results equal eigsolve per chain, bit for bit.
    Returns (vals, infos)."""
    from .storage import context
    chains = [list(ch) for ch in chains]
    n = len(chains)
    if n == 0:
        return [], []
    ns = len(chains[0])
    if ns == 0 or any(len(ch) != ns for ch in chains):
        raise ValueError("every chain must have the same, nonzero number of plans")
    xs = list(xs)
    if len(xs) != n:
        raise ValueError("one start vector per chain")
    if ys is None:
        ys = [ch[-1]._ops[0] for ch in chains]
    if ctx is None:
        ctx = context(xs[0].device)
    k = max(int(krylovdim), 2)
    hv = lambda t: t.handle.value if hasattr(t.handle, "value") else t.handle  # noqa: E731
    ap = (C.c_void_p * (n * ns))(*[p._h.value for ch in chains for p in ch])
    ax = (C.c_void_p * n)(*[hv(x) for x in xs])
    ay = (C.c_void_p * n)(*[hv(y) for y in ys])
    vals, nres = np.zeros(n), np.zeros(n)
    numops = np.zeros(n, dtype=np.int64)
    numiter, conv = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    f64 = C.POINTER(C.c_double)
    i32 = C.POINTER(C.c_int)
    pp = C.POINTER(C.c_void_p)
    check(load().sb_eigsolve_lanczos_multi(ctx, C.cast(ap, pp), n, ns, C.cast(ax, pp), C.cast(ay, pp), k,
                                           int(maxiter), float(tol), vals.ctypes.data_as(f64),
                                           nres.ctypes.data_as(f64), numops.ctypes.data_as(C.POINTER(C.c_int64)),
                                           numiter.ctypes.data_as(i32), conv.ctypes.data_as(i32)))
    infos = [ConvergenceInfo(int(conv[c]), float(nres[c]), int(numiter[c]), int(numops[c]), np.zeros(0))
             for c in range(n)]
    return [float(v) for v in vals], infos
