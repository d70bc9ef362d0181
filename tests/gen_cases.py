"""Random contraction cases for every kernel family (shared by the oracle
CPU tests and the GPU parity tests)."""
import numpy as np

def dyadic(rng, shape, fill=0.5, P=0):
    """dyadic values; whole prefix blocks zeroed at random (block sparsity)"""
    a = rng.integers(-8, 9, size=shape).astype(np.float64) / 4.0
    a *= rng.random(shape) < 0.8
    if P > 0:
        mask = rng.random(shape[:P]) < fill
        a *= mask.reshape(mask.shape + (1,) * (len(shape) - P))
    return np.asfortranarray(a)


def gauss(rng, shape, fill=0.5, P=0):
    a = rng.standard_normal(shape)
    if P > 0:
        mask = rng.random(shape[:P]) < fill
        a *= mask.reshape(mask.shape + (1,) * (len(shape) - P))
    return np.asfortranarray(a)


def _names(prefix, n):
    return [f"{prefix}{i}" for i in range(n)]


def _shuffle(rng, x):
    x = list(x)
    rng.shuffle(x)
    return x


def gen_shared(rng, b_kind, gen=dyadic):
    """contract_shared! case: BS A x (BS | Dense) B, all shared labels reduced"""
    nsp = int(rng.integers(0, 3))
    nsd = int(rng.integers(0, 3))
    if nsp + nsd < 2:
        nsd = 2 - nsp  # >= 2 shared labels (else the single-label kernels run)
    kpa = int(rng.integers(1, 3))
    kda = int(rng.integers(0, 3))
    kdb = int(rng.integers(1, 3))
    kpb = int(rng.integers(0, 2)) if b_kind == "bs" else 0
    sp, sd = _names("s", nsp), _names("k", nsd)
    kpA, kdA, kdB, kpB = _names("a", kpa), _names("x", kda), _names("y", kdb), _names("b", kpb)
    dims = {n: int(rng.integers(1, 4)) for n in sp + sd + kpA + kdA + kdB + kpB}
    la = _shuffle(rng, sp + kpA) + _shuffle(rng, sd + kdA)
    PA = nsp + kpa
    if b_kind == "bs":
        lb = _shuffle(rng, sp + kpB) + _shuffle(rng, sd + kdB)
        PB = nsp + kpb
    else:
        lb = _shuffle(rng, sp + sd + kdB)
        PB = 0
    cp = _shuffle(rng, kpA + kpB)
    ga, gb = _shuffle(rng, kdA), _shuffle(rng, kdB)
    lc = cp + (ga + gb if rng.random() < 0.5 else gb + ga)
    A = gen(rng, [dims[n] for n in la], P=PA)
    B = gen(rng, [dims[n] for n in lb], P=PB)
    cdims = [dims[n] for n in lc]
    return dict(A=A, la=la, N2A=len(la) - PA, B=B, lb=lb, N2B=len(lb) - PB, lc=lc, cdims=cdims,
                N2C=len(lc) - len(cp))


def gen_single(rng, fam, gen=dyadic):
    """single-label kernel families (exactly one shared label r)"""
    R = int(rng.integers(1, 5))
    d = lambda: int(rng.integers(1, 4))  # noqa: E731
    if fam in ("prefix_bb", "dense_bb"):
        pa, ta, pb, tb = _names("a", int(rng.integers(1, 3))), _names("x", int(rng.integers(1, 3))), \
            _names("b", int(rng.integers(1, 3))), _names("y", int(rng.integers(1, 3)))
        if fam == "prefix_bb":
            la = _shuffle(rng, pa + ["r"]) + _shuffle(rng, ta)
            lb = _shuffle(rng, pb + ["r"]) + _shuffle(rng, tb)
            PA, PB = len(pa) + 1, len(pb) + 1
            gtail = (ta, tb)
        else:
            la = _shuffle(rng, pa) + _shuffle(rng, ta + ["r"])
            lb = _shuffle(rng, pb) + _shuffle(rng, tb + ["r"])
            PA, PB = len(pa), len(pb)
            gtail = ([l for l in la[PA:] if l != "r"], [l for l in lb[PB:] if l != "r"])
        dims = {n: d() for n in pa + ta + pb + tb}
        dims["r"] = R
        cp = _shuffle(rng, pa + pb)
        if fam == "prefix_bb":
            ta_o, tb_o = [l for l in la[PA:]], [l for l in lb[PB:]]
        else:
            ta_o, tb_o = gtail
        lc = cp + (ta_o + tb_o if rng.random() < 0.5 else tb_o + ta_o)
        A = gen(rng, [dims[n] for n in la], P=PA)
        B = gen(rng, [dims[n] for n in lb], P=PB)
        return dict(ck="bs", ak="bs", bk="bs", A=A, la=la, N2A=len(la) - PA, B=B, lb=lb, N2B=len(lb) - PB, lc=lc,
                    cdims=[dims[n] for n in lc], N2C=len(lc) - len(cp))
    if fam in ("prefix_bd", "dense_bd"):
        pa, ta, tb = _names("a", int(rng.integers(1, 3))), _names("x", int(rng.integers(1, 3))), \
            _names("y", int(rng.integers(1, 3)))
        dims = {n: d() for n in pa + ta + tb}
        dims["r"] = R
        if fam == "prefix_bd":
            la = _shuffle(rng, pa + ["r"]) + _shuffle(rng, ta)
            PA = len(pa) + 1
            ta_o = la[PA:]
        else:
            la = _shuffle(rng, pa) + _shuffle(rng, ta + ["r"])
            PA = len(pa)
            ta_o = [l for l in la[PA:] if l != "r"]
        lb = _shuffle(rng, tb + ["r"])
        tb_o = [l for l in lb if l != "r"]
        cp = _shuffle(rng, pa)
        lc = cp + (ta_o + tb_o if rng.random() < 0.5 else tb_o + ta_o)
        A = gen(rng, [dims[n] for n in la], P=PA)
        B = gen(rng, [dims[n] for n in lb])
        return dict(ck="bs", ak="bs", bk="dense", A=A, la=la, N2A=len(la) - PA, B=B, lb=lb, N2B=None, lc=lc,
                    cdims=[dims[n] for n in lc], N2C=len(lc) - len(cp))
    if fam == "coo_coo":
        oa, ob = _names("a", int(rng.integers(1, 3))), _names("b", int(rng.integers(1, 3)))
        dims = {n: d() for n in oa + ob}
        dims["r"] = R
        la, lb = _shuffle(rng, oa + ["r"]), _shuffle(rng, ob + ["r"])
        lc = _shuffle(rng, oa + ob)
        A = gen(rng, [dims[n] for n in la])
        B = gen(rng, [dims[n] for n in lb])
        return dict(ck="coo", ak="coo", bk="coo", A=A, la=la, N2A=None, B=B, lb=lb, N2B=None, lc=lc,
                    cdims=[dims[n] for n in lc], N2C=None)
    if fam in ("coo_bs_prefix", "coo_bs_dense"):
        oa, pb, tb = _names("a", int(rng.integers(1, 3))), _names("b", int(rng.integers(1, 3))), \
            _names("y", int(rng.integers(1, 3)))
        dims = {n: d() for n in oa + pb + tb}
        dims["r"] = R
        la = _shuffle(rng, oa + ["r"])
        if fam == "coo_bs_prefix":
            lb = _shuffle(rng, pb + ["r"]) + _shuffle(rng, tb)
            PB = len(pb) + 1
            ctail = lb[PB:]
        else:
            lb = _shuffle(rng, pb) + _shuffle(rng, tb + ["r"])
            PB = len(pb)
            ctail = [l for l in lb[PB:] if l != "r"]
        cp = _shuffle(rng, oa + pb)
        lc = cp + ctail
        A = gen(rng, [dims[n] for n in la])
        B = gen(rng, [dims[n] for n in lb], P=PB)
        return dict(ck="bs", ak="coo", bk="bs", A=A, la=la, N2A=None, B=B, lb=lb, N2B=len(lb) - PB, lc=lc,
                    cdims=[dims[n] for n in lc], N2C=len(ctail))
    if fam == "coo_dense":
        oa, tb = _names("a", int(rng.integers(1, 3))), _names("y", int(rng.integers(1, 3)))
        dims = {n: d() for n in oa + tb}
        dims["r"] = R
        la, lb = _shuffle(rng, oa + ["r"]), _shuffle(rng, tb + ["r"])
        lc = _shuffle(rng, oa) + _shuffle(rng, tb)
        A = gen(rng, [dims[n] for n in la])
        B = gen(rng, [dims[n] for n in lb])
        return dict(ck="bs", ak="coo", bk="dense", A=A, la=la, N2A=None, B=B, lb=lb, N2B=None, lc=lc,
                    cdims=[dims[n] for n in lc], N2C=len(tb))
    raise ValueError(fam)


