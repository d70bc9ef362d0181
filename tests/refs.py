"""Test helpers: fixture loading, brute-force dense references, comparisons.

The dense references here are the closed forms the reference's own tests use
(permute + reshape + one matmul, e.g. test/tensoralgebra/bs_bs.jl:91-208),
written as numpy einsum over named labels.
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixtures():
    with open(os.path.join(GOLDEN, "reference_fixtures.json")) as f:
        return json.load(f)


def lab(name, plev=0):
    """A label (id, plev) for a reference Symbol name; ids are stable."""
    return (int.from_bytes(name.encode()[:8].ljust(8, b"\0"), "little"), plev)


def labs(names):
    return [lab(n) for n in names]


def dense(shape, entries):
    a = np.zeros(shape, dtype=np.float64, order="F")
    for idx, v in entries:
        a[tuple(i - 1 for i in idx)] = v
    return a


def permute_named(arr, names, target):
    return np.transpose(arr, [names.index(t) for t in target])


def einsum_contract(A, namesA, B, namesB, namesC):
    """C[namesC] = sum over shared names of A[namesA] * B[namesB]."""
    letters = {}
    for n in list(namesA) + list(namesB) + list(namesC):
        if n not in letters:
            letters[n] = chr(ord("a") + len(letters))
    sa = "".join(letters[n] for n in namesA)
    sb = "".join(letters[n] for n in namesB)
    sc = "".join(letters[n] for n in namesC)
    return np.einsum(f"{sa},{sb}->{sc}", A, B, optimize=False)


def blocks_by_key(p):
    """key tuple -> payload (the reference tests' _blocks_by_key, bs_bs.jl:8-17)."""
    bs = p["blksize"]
    out = {}
    for i, k in enumerate(p["keys"]):
        idv = int(p["ids"][i])
        out[tuple(int(x) for x in k)] = np.asarray(p["data"][(idv - 1) * bs: idv * bs])
    return out


def assert_bs_equal(p, q, exact=True, rtol=0.0, atol=0.0, check_ids=False):
    """_assert_blocksparse_equal (bs_bs.jl:19-29): dims, blksize, key order,
    key -> payload.  check_ids additionally compares the slot ids and the raw
    payload array (bit-exact block lists)."""
    assert tuple(p["dims"]) == tuple(q["dims"])
    assert p["blksize"] == q["blksize"]
    assert np.array_equal(np.asarray(p["keys"]).reshape(len(p["ids"]), -1),
                          np.asarray(q["keys"]).reshape(len(q["ids"]), -1)), (p["keys"], q["keys"])
    da, db = blocks_by_key(p), blocks_by_key(q)
    assert da.keys() == db.keys()
    for k in da:
        if exact:
            assert np.array_equal(da[k], db[k]), (k, da[k], db[k])
        else:
            np.testing.assert_allclose(da[k], db[k], rtol=rtol, atol=atol)
    if check_ids:
        assert np.array_equal(np.asarray(p["ids"]), np.asarray(q["ids"]))


def rel_frobenius(x, y):
    x = np.asarray(x, dtype=np.float64).ravel()
    y = np.asarray(y, dtype=np.float64).ravel()
    n = np.linalg.norm(y)
    return np.linalg.norm(x - y) / (n if n > 0 else 1.0)
