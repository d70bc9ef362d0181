"""Contraction cases built from the reference's unit-test fixtures, shared by
the oracle tests and the GPU parity tests.  Each case yields the operands as
dense arrays plus labels, and the brute-force dense reference of the result."""
import itertools

import numpy as np

from refs import fixtures, dense, permute_named, einsum_contract, labs

FX = fixtures()


def _canon(spec):
    return dense(spec["shape"], spec["entries"]), spec["labels"]


def bs_dense_prefix_cases():
    f = FX["bs_dense_prefix"]
    A0, la0 = _canon(f["A"])
    B0, lb0 = _canon(f["B"])
    for la, lb, lc in itertools.product(f["labelsA_variants"], f["labelsB_variants"], f["labelsC_variants"]):
        A = permute_named(A0, la0, la)
        B = permute_named(B0, lb0, lb)
        Cd = einsum_contract(A, la, B, lb, lc)
        yield dict(name=f"{la}-{lb}-{lc}", A=A, la=la, N2A=f["N2A"], B=B, lb=lb, lc=lc, N2C=f["N2C"], C=Cd)


def bs_bs_cases(key):
    f = FX[key]
    A0, la0 = _canon(f["A"])
    B0, lb0 = _canon(f["B"])
    for la, lb, lc in itertools.product(f["labelsA_variants"], f["labelsB_variants"], f["labelsC_variants"]):
        A = permute_named(A0, la0, la)
        B = permute_named(B0, lb0, lb)
        Cd = einsum_contract(A, la, B, lb, lc)
        yield dict(name=f"{la}-{lb}-{lc}", A=A, la=la, N2A=f["N2A"], B=B, lb=lb, N2B=f["N2B"], lc=lc,
                   N2C=f["N2C"], C=Cd)


def coo_coo_cases():
    for i, c in enumerate(FX["coo_coo"]["cases"]):
        A, la = _canon(c["A"])
        B, lb = _canon(c["B"])
        yield dict(name=f"coo{i}", A=A, la=la, B=B, lb=lb, lc=c["labelsC"], C_shape=c["C_shape"],
                   C=einsum_contract(A, la, B, lb, c["labelsC"]))


def coo_dense_cases():
    f = FX["coo_dense"]
    for cfg in ("config1", "config2"):
        g = f[cfg]
        A0, la0 = _canon(g["A"])
        B0, lb0 = _canon(g["B"])
        for la, lb, lc in itertools.product(g["labelsA_variants"], g["labelsB_variants"], g["labelsC_variants"]):
            A = permute_named(A0, la0, la)
            B = permute_named(B0, lb0, lb)
            yield dict(name=f"{cfg}-{la}-{lb}-{lc}", A=A, la=la, B=B, lb=lb, lc=lc, N2C=len(lb) - 1,
                       C=einsum_contract(A, la, B, lb, lc))


def coo_bs_cases():
    for i, c in enumerate(FX["coo_bs"]["cases"]):
        A, la = _canon(c["A"])
        B, lb = _canon(c["B"])
        yield dict(name=f"coobs{i}", A=A, la=la, B=B, lb=lb, N2B=c["B"]["N2"], lc=c["labelsC"],
                   N2C=c["N2C"], C=einsum_contract(A, la, B, lb, c["labelsC"]))


def rand_sparse(rng, shape, fill=0.4, dyadic=True):
    """Random array whose values are small dyadic rationals (exact products/sums)."""
    a = rng.integers(-8, 9, size=shape).astype(np.float64) / 4.0 if dyadic else rng.standard_normal(shape)
    mask = rng.random(shape) < fill
    return np.asfortranarray(a * mask)
