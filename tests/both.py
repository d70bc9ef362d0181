"""Build the same operands in the oracle and on the device, run both, compare.

The device side goes through the product's C ABI (sparsebackends_amd); the
oracle side through oracle/oracle.py.  Comparisons: dims, blksize, key order
and ids bit-exact; payloads exact or within a stated relative tolerance.
"""
import numpy as np

import sparsebackends_amd as sb
from oracle import oracle as O
from refs import labs, rel_frobenius


def dev_tensor(kind, arr=None, N2=None, dims=None, atol=0.0):
    if kind == "bs":
        return sb.blocksparse_from_dense(arr, N2) if arr is not None else sb.NewBlockSparseSorted.new(dims, N2)
    if kind == "coo":
        return sb.coo_from_dense(arr, atol=atol) if arr is not None else sb.COOTensor.new(dims)
    if kind == "dense":
        return sb.DenseTensor.from_numpy(arr)
    raise ValueError(kind)


def orc_tensor(kind, arr=None, N2=None, dims=None, atol=0.0):
    if kind == "bs":
        return O.BS.from_dense(arr, N2) if arr is not None else O.BS.new(dims, N2)
    if kind == "coo":
        return O.COO.from_dense(arr, atol=atol) if arr is not None else O.COO.new(dims)
    if kind == "dense":
        return O.Dense.new(arr)
    raise ValueError(kind)


def run_both(ckind, cdims, N2C, lc, akind, A, N2A, la, bkind, B, N2B, lb, flags=0):
    """Returns (device C, oracle C) after contract! on both."""
    Cd = dev_tensor(ckind, dims=cdims, N2=N2C)
    Co = orc_tensor(ckind, dims=cdims, N2=N2C)
    Ad, Ao = dev_tensor(akind, A, N2A), orc_tensor(akind, A, N2A)
    Bd, Bo = dev_tensor(bkind, B, N2B), orc_tensor(bkind, B, N2B)
    sb.contract_(Cd, labs(lc), Ad, labs(la), Bd, labs(lb), flags)
    O.contract(Co, labs(lc), Ao, labs(la), Bo, labs(lb), flags)
    return Cd, Co


def compare_bs(dev, orc, rtol=0.0, check_ids=True):
    p, q = dev.parts(), orc.parts()
    assert tuple(p["dims"]) == tuple(q["dims"])
    assert p["blksize"] == q["blksize"]
    assert p["keys"].shape == q["keys"].shape, (p["keys"].shape, q["keys"].shape)
    assert np.array_equal(p["keys"], q["keys"]), "block keys differ"
    if check_ids:
        assert np.array_equal(p["ids"], q["ids"]), "block ids differ"
    if rtol == 0.0:
        assert np.array_equal(p["data"], q["data"]), "payload differs"
    else:
        assert rel_frobenius(p["data"], q["data"]) <= rtol, rel_frobenius(p["data"], q["data"])


def compare_coo(dev, orc, rtol=0.0):
    p, q = dev.parts(), orc.parts()
    assert tuple(p["dims"]) == tuple(q["dims"])
    assert p["dirty"] == q["dirty"]
    assert np.array_equal(p["keys"], q["keys"]), "COO keys differ"
    if rtol == 0.0:
        assert np.array_equal(p["vals"], q["vals"])
    else:
        assert rel_frobenius(p["vals"], q["vals"]) <= rtol
