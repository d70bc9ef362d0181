"""Block-sparse operands whose blocks carry zero padding: every block holds
values only inside a random box (a prefix range along each tail axis), the
layout of symmetry sectors stored in blocks padded to the largest sector.
Used to exercise the box-aware block products against the oracle."""
import itertools

import numpy as np


def dyadic_vals(rng, shape):
    v = rng.integers(-8, 9, size=shape).astype(np.float64) / 4.0
    v[v == 0] = 0.25
    return v


def gauss_vals(rng, shape):
    return rng.standard_normal(shape)


def padded_parts(rng, pdims, tdims, nblocks, vals=dyadic_vals, empty_frac=0.1, full_frac=0.2, shuffle_ids=True):
    """host parts (dims, N2, keys, ids, data) of a NewBlockSparseSorted with
    random keys and, per block, values inside a random box, zero elsewhere"""
    P = len(pdims)
    space = list(itertools.product(*[range(1, d + 1) for d in pdims]))
    pick = rng.choice(len(space), size=min(nblocks, len(space)), replace=False)
    keys = sorted((space[i] for i in pick), key=lambda k: tuple(reversed(k)))  # column-major order
    n = len(keys)
    bs = int(np.prod(tdims))
    ids = (rng.permutation(n) + 1) if shuffle_ids else np.arange(1, n + 1)
    data = np.zeros(n * bs)
    for i in range(n):
        u = rng.random()
        if u < empty_frac:
            ext = [int(rng.integers(0, t + 1)) for t in tdims]
            ext[int(rng.integers(0, len(tdims)))] = 0  # an all-zero stored block
        elif u < empty_frac + full_frac:
            ext = list(tdims)
        else:
            ext = [int(rng.integers(1, t + 1)) for t in tdims]
        blk = np.zeros(tdims, order="F")
        if all(e > 0 for e in ext):
            blk[tuple(slice(0, e) for e in ext)] = vals(rng, ext)
        data[(ids[i] - 1) * bs:ids[i] * bs] = blk.reshape(-1, order="F")
    return dict(dims=tuple(pdims) + tuple(tdims), N2=len(tdims),
                keys=np.array(keys, dtype=np.int64).reshape(n, P), ids=np.asarray(ids, dtype=np.int64), data=data)


def _shuffle(rng, x):
    x = list(x)
    rng.shuffle(x)
    return x


def gen_shared_labels(rng, big=False):
    """labels and dims of a BS x BS contract_shared! (>= 2 reduced labels):
    shared prefix sp, shared tail sd, A-only prefix/tail, B-only prefix/tail"""
    while True:
        nsp, nsd = int(rng.integers(0, 3)), int(rng.integers(0, 3))
        if nsp + nsd >= 2 and nsd >= 1:
            break
    kpa, kda = int(rng.integers(0, 3)), int(rng.integers(1, 3))
    kpb, kdb = int(rng.integers(0, 2)), int(rng.integers(1, 3))
    nm = lambda p, k: [f"{p}{i}" for i in range(k)]  # noqa: E731
    sp, sd, kpA, kdA, kpB, kdB = nm("s", nsp), nm("k", nsd), nm("a", kpa), nm("x", kda), nm("b", kpb), nm("y", kdb)
    hi_t = 12 if big else 8
    dims = {}
    for n in sp + kpA + kpB:
        dims[n] = int(rng.integers(1, 5))
    for n in sd + kdA + kdB:
        dims[n] = int(rng.integers(2, hi_t + 1))
    la = _shuffle(rng, sp + kpA) + _shuffle(rng, sd + kdA)
    lb = _shuffle(rng, sp + kpB) + _shuffle(rng, sd + kdB)
    cp = _shuffle(rng, kpA + kpB)
    ga, gb = _shuffle(rng, kdA), _shuffle(rng, kdB)
    lc = cp + (ga + gb if rng.random() < 0.5 else gb + ga)
    return dict(la=la, PA=nsp + kpa, lb=lb, PB=nsp + kpb, lc=lc, PC=len(cp), dims=dims)


def gen_padded_case(rng, big=False, vals=dyadic_vals, density=0.6):
    g = gen_shared_labels(rng, big)
    d = g["dims"]

    def parts(labels, P):
        pd = [d[n] for n in labels[:P]]
        td = [d[n] for n in labels[P:]]
        nspace = int(np.prod(pd)) if pd else 1
        return padded_parts(rng, pd, td, max(1, int(round(density * nspace))), vals=vals)

    A = parts(g["la"], g["PA"])
    B = parts(g["lb"], g["PB"])
    cdims = tuple(d[n] for n in g["lc"])
    return dict(A=A, la=g["la"], B=B, lb=g["lb"], lc=g["lc"], cdims=cdims, N2C=len(g["lc"]) - g["PC"])
