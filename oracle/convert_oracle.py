"""CPU restatement of rekey / retypeval (numpy).

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker; the product never
imports it.

rekey(A, K2; check) (SparseBackends/src/blocksparse/conversions.jl:327-342,
coo/conversions.jl:41-53, aliased/conversions.jl:76-93): with check, first
every prefix dimension (block-sparse and aliased; not COO) must be <=
typemax(K2), then every key coordinate must lie in [typemin(K2),
typemax(K2)], key by key in stored order and axis by axis within a key;
the first failure raises ErrorException.  retypeval(A::COO, T2)
(coo/conversions.jl:65-82): values in stored order; an integer T2 needs an
integral value, then T2(v) must not throw (inside T2's range), then
Float64(T2(v)) == v (NaN never passes).

Pinned by tests/test_oracle_convert.py against numpy's integer ranges
(np.iinfo), numpy's float32/float16 round trips and hand-worked cases.
"""
from __future__ import annotations

import numpy as np


def rekey_failure(dims, P, keys, K2, kind):
    """None, or the reference's error message for the first failing check.
    kind: "bs", "aliased" (prefix dimensions checked) or "coo"."""
    info = np.iinfo(K2)
    name = np.dtype(K2).name.replace("int", "Int").replace("uInt", "UInt")
    if kind in ("bs", "aliased"):
        for i in range(P):
            if dims[i] > info.max:
                return f"Dimension {dims[i]} at prefix axis {i + 1} cannot be indexed by {name} (max={info.max})"
    # `for key in A.keys, v in key`: key by key in stored order, axis by axis
    for v in np.asarray(keys, dtype=np.int64).reshape(-1):
        if v < info.min or v > info.max:
            what = "Coordinate" if kind == "coo" else "Key coordinate"
            return f"{what} {int(v)} is out of range for {name} ({info.min}..{info.max})"
    return None


def retypeval_failure(vals, T2):
    """None, or (index of the first value that does not survive, reason):
    reason 1 non-integer (integer T2), 2 outside T2's range, 3 round trip lossy"""
    v = np.asarray(vals, dtype=np.float64)
    t = np.dtype(T2)
    for i, x in enumerate(v):
        if t.kind == "f":
            with np.errstate(over="ignore", invalid="ignore"):
                c = np.array([x]).astype(t)[0]
            if not (np.float64(c) == x):
                return i, 3
        else:
            if not (np.isfinite(x) and x == np.floor(x)):
                return i, 1
            info = np.iinfo(t)
            if not (x >= float(info.min) and x < float(info.max) + 1.0):
                return i, 2
    return None
