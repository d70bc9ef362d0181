"""Random AliasedBlockSparse operands and contraction cases (host parts).

An aliased tensor is given by dims, N2, templates (ntemplates x blksize),
keys (sorted by the column-major prefix lin), alias ids and scalars; its
logical content is scalars[i] * templates[alias_ids[i]] in block i."""
import itertools

import numpy as np


def dy(rng, shape):
    v = rng.integers(-6, 7, size=shape).astype(np.float64) / 4.0
    v[v == 0] = 0.5
    return v


def aliased_parts(rng, pdims, tdims, nblocks, ntemplates, vals=dy):
    P = len(pdims)
    space = list(itertools.product(*[range(1, d + 1) for d in pdims]))
    pick = rng.choice(len(space), size=min(nblocks, len(space)), replace=False)
    keys = sorted((space[i] for i in pick), key=lambda k: tuple(reversed(k)))
    n = len(keys)
    bs = int(np.prod(tdims)) if tdims else 1
    tm = vals(rng, (ntemplates, bs)).reshape(-1)
    ids = rng.integers(1, ntemplates + 1, size=n)
    sc = vals(rng, n)
    return dict(dims=tuple(pdims) + tuple(tdims), N2=len(tdims), templates=tm, ntemplates=ntemplates,
                keys=np.array(keys, dtype=np.int64).reshape(n, P), alias_ids=ids, scalars=sc)


def logical_dense(p):
    """the logical dense array of aliased host parts"""
    dims = p["dims"]
    N2 = p["N2"]
    P = len(dims) - N2
    out = np.zeros(dims)
    bs = int(np.prod(dims[P:]))
    for k, a, s in zip(p["keys"], p["alias_ids"], p["scalars"]):
        blk = s * p["templates"][(a - 1) * bs:a * bs].reshape(dims[P:], order="F")
        out[tuple(int(x) - 1 for x in k)] = blk
    return out


def _names(prefix, n):
    return [f"{prefix}{i}" for i in range(n)]


def _shuffle(rng, x):
    x = list(x)
    rng.shuffle(x)
    return x


def gen_single(rng, fam, vals=dy):
    """single reduced label r; fam in ad_prefix, ad_dense, aa_prefix, aa_dense, coo_dense"""
    r = "r"
    R = int(rng.integers(1, 4))
    if fam in ("ad_prefix", "ad_dense"):
        kpa = int(rng.integers(1, 3)) if fam == "ad_dense" else int(rng.integers(0, 3))
        kda = int(rng.integers(1, 3))
        kb = int(rng.integers(1, 3))
        pA, dA, dB = _names("a", kpa), _names("x", kda), _names("y", kb)
        dims = {n: int(rng.integers(1, 4)) for n in pA + dA + dB}
        dims[r] = R
        if fam == "ad_prefix":
            prefA = _shuffle(rng, pA + [r])
            tailA = _shuffle(rng, dA)
        else:
            prefA = _shuffle(rng, pA)
            tailA = _shuffle(rng, dA + [r])
        la = prefA + tailA
        lb = _shuffle(rng, dB + [r])
        Ad = [n for n in tailA if n != r]
        Bwo = [n for n in lb if n != r]
        if fam == "ad_prefix":
            Bwo = [n for n in _permuted_without(lb, r)]
        cp = _shuffle(rng, pA) if fam == "ad_prefix" else _shuffle(rng, pA)
        tail = Ad + Bwo if rng.random() < 0.5 else Bwo + Ad
        lc = cp + tail
        A = aliased_parts(rng, [dims[n] for n in prefA], [dims[n] for n in tailA],
                          max(1, int(0.7 * np.prod([dims[n] for n in prefA]))), int(rng.integers(1, 4)), vals)
        B = vals(rng, tuple(dims[n] for n in lb))
        return dict(fam=fam, A=A, la=la, B=B, lb=lb, lc=lc, cdims=tuple(dims[n] for n in lc), N2C=len(tail))
    if fam in ("aa_prefix", "aa_dense"):
        kpa, kpb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        kda, kdb = int(rng.integers(1, 3)), int(rng.integers(1, 3))
        if fam == "aa_dense" and kpa + kpb == 0:
            kpa = 1
        pA, pB, dA, dB = _names("a", kpa), _names("b", kpb), _names("x", kda), _names("y", kdb)
        dims = {n: int(rng.integers(1, 4)) for n in pA + pB + dA + dB}
        dims[r] = R
        if fam == "aa_prefix":
            prefA, tailA = _shuffle(rng, pA + [r]), _shuffle(rng, dA)
            prefB, tailB = _shuffle(rng, pB + [r]), _shuffle(rng, dB)
            Ad, Bd = tailA, tailB
        else:
            prefA, tailA = _shuffle(rng, pA), _shuffle(rng, dA + [r])
            prefB, tailB = _shuffle(rng, pB), _shuffle(rng, dB + [r])
            Ad, Bd = [n for n in tailA if n != r], [n for n in tailB if n != r]
        la, lb = prefA + tailA, prefB + tailB
        cp = _shuffle(rng, pA + pB)
        tail = Ad + Bd if rng.random() < 0.5 else Bd + Ad
        lc = cp + tail
        A = aliased_parts(rng, [dims[n] for n in prefA], [dims[n] for n in tailA],
                          max(1, int(0.7 * np.prod([dims[n] for n in prefA]))), int(rng.integers(1, 4)), vals)
        B = aliased_parts(rng, [dims[n] for n in prefB], [dims[n] for n in tailB],
                          max(1, int(0.7 * np.prod([dims[n] for n in prefB]))), int(rng.integers(1, 4)), vals)
        return dict(fam=fam, A=A, la=la, B=B, lb=lb, lc=lc, cdims=tuple(dims[n] for n in lc), N2C=len(tail))
    if fam == "coo_dense":
        kp, kb = int(rng.integers(1, 3)), int(rng.integers(1, 3))
        pA, dB = _names("a", kp), _names("y", kb)
        dims = {n: int(rng.integers(1, 4)) for n in pA + dB}
        dims[r] = R
        la = _shuffle(rng, pA + [r])
        lb = _shuffle(rng, dB + [r])
        cp = _shuffle(rng, pA)
        lc = cp + [n for n in lb if n != r]
        Ad = vals(rng, tuple(dims[n] for n in la))
        Ad[rng.random(Ad.shape) < 0.5] = 0.0
        B = vals(rng, tuple(dims[n] for n in lb))
        return dict(fam=fam, A=Ad, la=la, B=B, lb=lb, lc=lc, cdims=tuple(dims[n] for n in lc), N2C=len(lc) - len(cp))
    raise ValueError(fam)


def _permuted_without(lb, r):
    return [n for n in lb if n != r]


def gen_shared(rng, b_kind, vals=dy):
    """contract_shared! with >= 2 reduced labels: Aliased A x (Dense | Aliased) B"""
    while True:
        nsp, nsd = int(rng.integers(0, 3)), int(rng.integers(0, 3))
        if nsp + nsd >= 2:
            break
    kpa, kda, kdb = int(rng.integers(1, 3)), int(rng.integers(0, 3)), int(rng.integers(1, 3))
    kpb = int(rng.integers(0, 2)) if b_kind == "al" else 0
    sp, sd = _names("s", nsp), _names("k", nsd)
    kpA, kdA, kdB, kpB = _names("a", kpa), _names("x", kda), _names("y", kdb), _names("b", kpb)
    dims = {n: int(rng.integers(1, 4)) for n in sp + sd + kpA + kdA + kdB + kpB}
    prefA, tailA = _shuffle(rng, sp + kpA), _shuffle(rng, sd + kdA)
    if not tailA:  # an aliased tensor needs a dense tail
        kdA = ["x0"]
        dims["x0"] = int(rng.integers(1, 4))
        tailA = _shuffle(rng, sd + kdA)
    la = prefA + tailA
    if b_kind == "al":
        prefB, tailB = _shuffle(rng, sp + kpB), _shuffle(rng, sd + kdB)
        lb = prefB + tailB
    else:
        lb = _shuffle(rng, sp + sd + kdB)
    cp = _shuffle(rng, kpA + kpB)
    ga, gb = _shuffle(rng, kdA), _shuffle(rng, kdB)
    lc = cp + (ga + gb if rng.random() < 0.5 else gb + ga)
    A = aliased_parts(rng, [dims[n] for n in prefA], [dims[n] for n in tailA],
                      max(1, int(0.7 * np.prod([dims[n] for n in prefA]))), int(rng.integers(1, 4)), vals)
    if b_kind == "al":
        B = aliased_parts(rng, [dims[n] for n in prefB], [dims[n] for n in tailB],
                          max(1, int(0.7 * np.prod([dims[n] for n in prefB]) if prefB else 1)),
                          int(rng.integers(1, 4)), vals)
    else:
        B = vals(rng, tuple(dims[n] for n in lb))
    return dict(fam="shared_" + b_kind, A=A, la=la, B=B, lb=lb, lc=lc, cdims=tuple(dims[n] for n in lc),
                N2C=len(lc) - len(cp))
