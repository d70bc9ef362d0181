"""CPU restatement of the eigensolver DMRG runs on H_eff, and of the Krylov
vector operations on block-sparse storage.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline): nothing in the product
imports this module.

eigsolve: KrylovKit 0.10.2's eigsolve(A, x0, 1, :SR; ishermitian=true,
krylovdim, maxiter, tol) is what the reference's DMRG calls
(test_product_dmrg_state.jl:221-226, test_product.jl:235-238); KrylovKit is
not in the reference's sources, so its published method is restated: Lanczos
with full reorthogonalisation, Rayleigh-Ritz on the projected matrix after
every expansion, the residual norm |beta * u[m-1]| as the convergence test,
and thick restarts that keep the lowest (3*krylovdim)/5 Ritz vectors (the
keep rule KrylovKit uses when nothing has converged) plus the residual
direction.  Orthogonalisation is classical Gram-Schmidt applied twice; the
projected eigenproblem uses numpy's eigh.

Pinned by (tests/test_oracle_lanczos.py): the lowest eigenvalue of random
symmetric matrices against numpy.linalg.eigh; the first factorisation's Ritz
values against Rayleigh-Ritz on the Krylov space built by QR of
[x0, A x0, ..., A^(k-1) x0]; an eigenvector start (one application, exact
value); dimension below krylovdim (exact after breakdown); the application
count of a restart; and exact-chain H_eff ground energies against exact
diagonalisation of the chain Hamiltonians built from spin and fermion
operators.

Vector operations (SparseBackends/src/tensor_wrappers.jl:792-923): on host
parts dicts (dims, N2, keys, ids, data).
"""
from __future__ import annotations

import numpy as np


def eigsolve(matvec, x0, krylovdim=30, maxiter=100, tol=1e-12):
    """Lowest eigenpair of the symmetric operator matvec; returns
    (value, vector, info) with info = ritz, normres, numops, numiter, converged."""
    kdim = int(krylovdim)
    if kdim < 2:
        raise ValueError("krylovdim must be at least 2")
    x0 = np.asarray(x0, dtype=np.float64)
    n = x0.size
    V = np.zeros((kdim + 1, n))
    b0 = np.linalg.norm(x0)
    if not b0 > 0:
        raise ValueError("initial vector has zero norm")
    V[0] = x0 / b0
    ld = kdim + 1
    T = np.zeros((ld, ld))
    j, cycles, numops = 0, 1, 0
    converged = False
    while True:
        w = np.asarray(matvec(V[j]), dtype=np.float64).copy()
        numops += 1
        nAv = np.linalg.norm(w)
        B = V[:j + 1]
        h1 = B @ w
        w -= B.T @ h1
        h2 = B @ w
        w -= B.T @ h2
        beta = np.linalg.norm(w)
        h = h1 + h2
        T[:j + 1, j] = h
        T[j, :j + 1] = h
        m = j + 1
        theta, U = np.linalg.eigh(T[:m, :m])
        breakdown = beta <= 1e-13 * max(nAv, 1e-300)
        resid = 0.0 if breakdown else beta * abs(U[m - 1, 0])
        if not breakdown:
            V[j + 1] = w / beta
        if resid <= tol or breakdown:
            converged = True
            break
        if m < kdim:
            j += 1
            continue
        if cycles == maxiter:
            break
        keep = max(1, min(kdim - 1, (3 * kdim) // 5))
        Vk = U[:, :keep].T @ V[:m]
        V[keep] = V[m]
        V[:keep] = Vk
        T[:] = 0.0
        for i in range(keep):
            T[i, i] = theta[i]
            T[keep, i] = T[i, keep] = beta * U[m - 1, i]
        j = keep
        cycles += 1
    x = U[:, 0] @ V[:m]
    x /= np.linalg.norm(x)
    return theta[0], x, dict(ritz=theta, normres=resid, numops=numops, numiter=cycles, converged=converged)


# ---------------------------------------------------------------- vector operations on host parts

def _lin(keys, dims, P):
    lin = np.zeros(len(keys), dtype=np.int64)
    s = 1
    for d in range(P):
        lin += (np.asarray(keys)[:, d] - 1) * s
        s *= dims[d]
    return lin


def add(pa, pb):
    """wA + wB (tensor_wrappers.jl:826-840): copy of A; each block of B, in B's
    key order, added into the block of equal key, created by _ensure_block!
    (storage.jl:73-83: appended payload, next id, key inserted sorted)."""
    dims, N2 = tuple(pa["dims"]), pa["N2"]
    P = len(dims) - N2
    bs = int(np.prod(dims[P:]))
    keys = [tuple(k) for k in np.asarray(pa["keys"]).reshape(-1, P)]
    ids = list(np.asarray(pa["ids"], dtype=np.int64))
    data = list(np.asarray(pa["data"], dtype=np.float64))
    lin = list(_lin(np.asarray(pa["keys"]).reshape(-1, P), dims, P)) if P else [0] * len(keys)
    for kb, ib in zip(np.asarray(pb["keys"]).reshape(-1, P), pb["ids"]):
        kb = tuple(kb)
        lb = int(_lin(np.array([kb]).reshape(1, P), dims, P)[0]) if P else 0
        pos = int(np.searchsorted(lin, lb))
        if pos < len(lin) and lin[pos] == lb:
            idc = ids[pos]
        else:
            idc = len(data) // bs + 1
            data.extend([0.0] * bs)
            lin.insert(pos, lb)
            keys.insert(pos, kb)
            ids.insert(pos, idc)
        src = np.asarray(pb["data"])[(ib - 1) * bs:ib * bs]
        for t in range(bs):
            data[(idc - 1) * bs + t] += src[t]
    return dict(dims=dims, N2=N2, keys=np.array(keys, dtype=np.int64).reshape(-1, P),
                ids=np.array(ids, dtype=np.int64), data=np.array(data))


def dot(pa, pb):
    """sum over blocks of equal key of <A block, B block>"""
    dims, N2 = tuple(pa["dims"]), pa["N2"]
    P = len(dims) - N2
    bs = int(np.prod(dims[P:]))
    kb = {tuple(k): int(i) for k, i in zip(np.asarray(pb["keys"]).reshape(-1, P), pb["ids"])}
    s = 0.0
    for k, i in zip(np.asarray(pa["keys"]).reshape(-1, P), pa["ids"]):
        j = kb.get(tuple(k))
        if j is not None:
            s += float(np.dot(pa["data"][(i - 1) * bs:i * bs], pb["data"][(j - 1) * bs:j * bs]))
    return s
