"""Python front of the oracle (oracle/sb_oracle.cpp).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker; the product never imports it.

The oracle restates the reference SparseBackends (Julia) algorithms on the
CPU.  Each C++ function cites the reference file:line it follows.  How each
part is pinned (tests/test_oracle_*.py):

  pinned by the reference's own golden vectors / fixtures
    blocksparse_from_dense, to_dense, permutedims, getindex/setindex!
        (test/blocksparse/test_storage.jl:56,75,97-121; test_convertor.jl:13-41)
    COO from/to dense, setindex! (dirty), permutedims, sort!
        (test/coo/test.jl:24-25,41,57,61-63,79)
    contract! BS x Dense, prefix r (all 12 variants, test/tensoralgebra/bs_dense.jl:116-195)
    contract! BS x BS prefix r / dense r (test/tensoralgebra/bs_bs.jl:210-388)
    contract! BS x BS multi-label (test/tensoralgebra/bs_bs_multi_labels.jl:172-301)
    contract! COO x COO (test/tensoralgebra/coo.jl:8-77)
    contract! COO x BS (test/tensoralgebra/coo_bs.jl:21-251)
    contract! COO x Dense (test/tensoralgebra/coo_dense.jl:69-189)
  pinned by mathematics (brute-force dense einsum on random inputs,
  invariants: permute/to_dense commutation, round trips, closed forms)
    every contraction family, every permute, fusion of axes
  AliasedBlockSparse storage and kernels (no reference test covers them):
    logical content pinned by mathematics (to_dense of inputs and outputs
    against a dense einsum; to_blocksparse / getindex / permutedims against
    to_dense); the aliasing rules checked structurally (a key fed by one
    combined template stays aliased with the summed scalar, two different
    templates give a concrete block with scalar 1, template counts)
  parity unpinned (only an import of the Julia reference could pin them, and
  Julia is absent): the `ids` allocation order of each kernel, the tie order of
  Julia's stable sortperm in contract_prefix_outer_bd!, the Kahan summation
  order there, and the template ids of accumulated aliased blocks (Julia Dict
  order; here column-major key order).  They are restated by reading the
  reference.
"""
from __future__ import annotations

import ctypes as C
import glob
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SB_OK = 0
ERROR_NAMES = {1: "AssertionError", 2: "ErrorException", 3: "ArgumentError", 4: "KeyError",
               5: "MethodError", 6: "BoundsError", 7: "DeviceError", 8: "OutOfMemory",
               9: "InvalidHandle", 10: "LimitExceeded"}


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code
        self.kind = ERROR_NAMES.get(code, str(code))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libsb_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = C.CDLL(path)
        _LIB.orc_last_error.restype = C.c_char_p
    return _LIB


def openblas_path():
    """Path of the OpenBLAS 0.3.29 that numpy ships (the reference's BLAS pin)."""
    import numpy
    d = os.path.join(os.path.dirname(os.path.dirname(numpy.__file__)), "numpy.libs")
    c = sorted(glob.glob(os.path.join(d, "libscipy_openblas64_*.so")))
    return c[0] if c else None


def use_blas(on=True, threads=1):
    path = openblas_path() if on else ""
    _chk(lib().orc_set_blas((path or "").encode()))
    if on and path:
        _chk(lib().orc_blas_threads(C.c_int64(threads)))
    return path


def _chk(rc):
    if rc != SB_OK:
        raise OracleError(rc, lib().orc_last_error().decode())


def _i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def labels_arr(labels):
    """labels: sequence of (id, plev) pairs -> uint64 array (2*n)."""
    a = np.zeros(2 * len(labels), dtype=np.uint64)
    for i, (lid, plev) in enumerate(labels):
        a[2 * i] = lid
        a[2 * i + 1] = plev
    return a


class Handle:
    def __init__(self, h):
        self.h = C.c_void_p(h) if not isinstance(h, C.c_void_p) else h

    def __del__(self):
        try:
            if self.h and _LIB is not None:
                _LIB.orc_free(self.h)
        except Exception:
            pass


class BS(Handle):
    """Oracle NewBlockSparseSorted."""

    @staticmethod
    def new(dims, N2):
        out = C.c_void_p()
        d = _i64(dims)
        _chk(lib().orc_bs_new(C.c_int(len(dims)), C.c_int(N2), _p(d, C.c_int64), C.byref(out)))
        return BS(out)

    @staticmethod
    def from_parts(dims, N2, keys, ids, data):
        d, k, i, x = _i64(dims), _i64(keys).reshape(-1), _i64(ids), _f64(data).reshape(-1)
        out = C.c_void_p()
        _chk(lib().orc_bs_from_parts(C.c_int(len(dims)), C.c_int(N2), _p(d, C.c_int64), C.c_int64(len(i)),
                                     _p(k, C.c_int64), _p(i, C.c_int64), _p(x, C.c_double), C.byref(out)))
        return BS(out)

    @staticmethod
    def from_dense(arr, N2, atol=0.0, rtol=0.0, dropzeros=True):
        D = Dense.new(arr)
        out = C.c_void_p()
        _chk(lib().orc_bs_from_dense(D.h, C.c_int(N2), C.c_double(atol), C.c_double(rtol),
                                     C.c_int(1 if dropzeros else 0), C.byref(out)))
        return BS(out)

    def info(self):
        N, N2, nb, bs = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
        dims = np.zeros(16, dtype=np.int64)
        _chk(lib().orc_bs_info(self.h, C.byref(N), C.byref(N2), _p(dims, C.c_int64), C.byref(nb), C.byref(bs)))
        return tuple(int(x) for x in dims[:N.value]), N2.value, nb.value, bs.value

    def parts(self):
        dims, N2, nb, bs = self.info()
        P = len(dims) - N2
        keys = np.zeros((nb, P), dtype=np.int64)
        ids = np.zeros(nb, dtype=np.int64)
        data = np.zeros(nb * bs, dtype=np.float64)
        _chk(lib().orc_bs_get(self.h, _p(keys, C.c_int64), _p(ids, C.c_int64), _p(data, C.c_double)))
        return dict(dims=dims, N2=N2, blksize=bs, keys=keys, ids=ids, data=data)

    def getindex(self, *I):
        v = C.c_double()
        a = _i64(I)
        _chk(lib().orc_bs_getindex(self.h, _p(a, C.c_int64), C.byref(v)))
        return v.value

    def setindex(self, v, *I):
        a = _i64(I)
        _chk(lib().orc_bs_setindex(self.h, C.c_double(v), _p(a, C.c_int64)))

    def permutedims(self, perm):
        p = _i64(perm)
        out = C.c_void_p()
        _chk(lib().orc_bs_permutedims(self.h, _p(p, C.c_int64), C.byref(out)))
        return BS(out)

    def permutedims_(self, perm):
        p = _i64(perm)
        _chk(lib().orc_bs_permutedims_inplace(self.h, _p(p, C.c_int64)))
        return self

    def sort_(self):
        _chk(lib().orc_bs_sort_lex(self.h))
        return self

    def is_dense(self):
        o = C.c_int()
        _chk(lib().orc_bs_is_dense(self.h, C.byref(o)))
        return bool(o.value)

    def to_dense(self, init=0.0, merged_axes=None):
        g = _i64(merged_axes if merged_axes else np.zeros((0, 2))).reshape(-1)
        out = C.c_void_p()
        _chk(lib().orc_bs_to_dense(self.h, C.c_double(init), C.c_int(len(g) // 2), _p(g, C.c_int64), C.byref(out)))
        return Dense(out).array()

    def fuse_two_axes_(self, ax1, ax2):
        out = C.c_void_p()
        _chk(lib().orc_bs_fuse_two_axes(self.h, C.c_int64(ax1), C.c_int64(ax2), C.byref(out)))
        return BS(out)


class AL(Handle):
    """Oracle AliasedBlockSparse: block i = scalars[i] * templates[alias_ids[i]]."""

    @staticmethod
    def new(dims, N2):
        out = C.c_void_p()
        d = _i64(dims)
        _chk(lib().orc_al_new(C.c_int(len(dims)), C.c_int(N2), _p(d, C.c_int64), C.byref(out)))
        return AL(out)

    @staticmethod
    def from_parts(dims, N2, templates, ntemplates, keys, alias_ids, scalars):
        d, t = _i64(dims), _f64(templates).reshape(-1)
        k, a, sc = _i64(keys).reshape(-1), _i64(alias_ids), _f64(scalars)
        out = C.c_void_p()
        _chk(lib().orc_al_from_parts(C.c_int(len(dims)), C.c_int(N2), _p(d, C.c_int64), C.c_int64(ntemplates),
                                     _p(t, C.c_double), C.c_int64(len(a)), _p(k, C.c_int64), _p(a, C.c_int64),
                                     _p(sc, C.c_double), C.byref(out)))
        return AL(out)

    def info(self):
        N, N2, nb, bs, nt = C.c_int(), C.c_int(), C.c_int64(), C.c_int64(), C.c_int64()
        dims = np.zeros(16, dtype=np.int64)
        _chk(lib().orc_al_info(self.h, C.byref(N), C.byref(N2), _p(dims, C.c_int64), C.byref(nb), C.byref(bs),
                               C.byref(nt)))
        return tuple(int(x) for x in dims[:N.value]), N2.value, nb.value, bs.value, nt.value

    def parts(self):
        dims, N2, nb, bs, nt = self.info()
        P = len(dims) - N2
        keys = np.zeros((nb, P), dtype=np.int64)
        ids = np.zeros(nb, dtype=np.int64)
        sc = np.zeros(nb, dtype=np.float64)
        tm = np.zeros(nt * bs, dtype=np.float64)
        _chk(lib().orc_al_get(self.h, _p(keys, C.c_int64), _p(ids, C.c_int64), _p(sc, C.c_double),
                              _p(tm, C.c_double)))
        return dict(dims=dims, N2=N2, blksize=bs, ntemplates=nt, keys=keys, alias_ids=ids, scalars=sc, templates=tm)

    def getindex(self, *I):
        v = C.c_double()
        a = _i64(I)
        _chk(lib().orc_al_getindex(self.h, _p(a, C.c_int64), C.byref(v)))
        return v.value

    def permutedims(self, perm):
        p = _i64(perm)
        out = C.c_void_p()
        _chk(lib().orc_al_permutedims(self.h, _p(p, C.c_int64), C.byref(out)))
        return AL(out)

    def to_blocksparse(self):
        out = C.c_void_p()
        _chk(lib().orc_al_to_blocksparse(self.h, C.byref(out)))
        return BS(out)

    def to_dense(self, init=0.0):
        out = C.c_void_p()
        _chk(lib().orc_al_to_dense(self.h, C.c_double(init), C.byref(out)))
        return Dense(out).array()

    def compression_ratio(self):
        v = C.c_double()
        _chk(lib().orc_al_compression_ratio(self.h, C.byref(v)))
        return v.value


class COO(Handle):
    """Oracle COOTensor."""

    @staticmethod
    def new(dims):
        d = _i64(dims)
        out = C.c_void_p()
        _chk(lib().orc_coo_new(C.c_int(len(dims)), _p(d, C.c_int64), C.byref(out)))
        return COO(out)

    @staticmethod
    def from_parts(dims, keys, vals, dirty=False):
        d, k, v = _i64(dims), _i64(keys).reshape(-1), _f64(vals)
        out = C.c_void_p()
        _chk(lib().orc_coo_from_parts(C.c_int(len(dims)), _p(d, C.c_int64), C.c_int64(len(v)), _p(k, C.c_int64),
                                      _p(v, C.c_double), C.c_int(1 if dirty else 0), C.byref(out)))
        return COO(out)

    @staticmethod
    def from_dense(arr, atol=1e-12, rtol=0.0):
        D = Dense.new(arr)
        out = C.c_void_p()
        _chk(lib().orc_coo_from_dense(D.h, C.c_double(atol), C.c_double(rtol), C.byref(out)))
        return COO(out)

    def parts(self):
        N, nnz, dirty = C.c_int(), C.c_int64(), C.c_int()
        dims = np.zeros(16, dtype=np.int64)
        _chk(lib().orc_coo_info(self.h, C.byref(N), _p(dims, C.c_int64), C.byref(nnz), C.byref(dirty)))
        keys = np.zeros((nnz.value, N.value), dtype=np.int64)
        vals = np.zeros(nnz.value, dtype=np.float64)
        _chk(lib().orc_coo_get(self.h, _p(keys, C.c_int64), _p(vals, C.c_double)))
        return dict(dims=tuple(int(x) for x in dims[:N.value]), keys=keys, vals=vals, dirty=bool(dirty.value))

    def getindex(self, *I):
        v = C.c_double()
        a = _i64(I)
        _chk(lib().orc_coo_getindex(self.h, _p(a, C.c_int64), C.byref(v)))
        return v.value

    def setindex(self, v, *I, dirty=False):
        a = _i64(I)
        _chk(lib().orc_coo_setindex(self.h, C.c_double(v), _p(a, C.c_int64), C.c_int(1 if dirty else 0)))

    def permutedims(self, perm):
        p = _i64(perm)
        out = C.c_void_p()
        _chk(lib().orc_coo_permutedims(self.h, _p(p, C.c_int64), C.byref(out)))
        return COO(out)

    def permutedims_(self, perm):
        p = _i64(perm)
        _chk(lib().orc_coo_permutedims_inplace(self.h, _p(p, C.c_int64)))
        return self

    def sort_(self):
        _chk(lib().orc_coo_sort(self.h))
        return self

    def to_dense(self):
        out = C.c_void_p()
        _chk(lib().orc_coo_to_dense(self.h, C.byref(out)))
        return Dense(out).array()

    def fuse_axes(self, ax1, ax2):
        out = C.c_void_p()
        _chk(lib().orc_coo_fuse_axes(self.h, C.c_int64(ax1), C.c_int64(ax2), C.byref(out)))
        return COO(out)


class Dense(Handle):
    @staticmethod
    def new(arr):
        a = np.asarray(arr, dtype=np.float64)
        flat = _f64(a.reshape(-1, order="F"))
        d = _i64(a.shape)
        out = C.c_void_p()
        _chk(lib().orc_dense_new(C.c_int(a.ndim), _p(d, C.c_int64), _p(flat, C.c_double), C.byref(out)))
        return Dense(out)

    def array(self):
        N = C.c_int()
        dims = np.zeros(16, dtype=np.int64)
        _chk(lib().orc_dense_info(self.h, C.byref(N), _p(dims, C.c_int64)))
        shape = tuple(int(x) for x in dims[:N.value])
        flat = np.zeros(int(np.prod(shape)) if shape else 1, dtype=np.float64)
        _chk(lib().orc_dense_get(self.h, _p(flat, C.c_double)))
        return flat.reshape(shape, order="F")


def contract(Ct, labelsC, At, labelsA, Bt, labelsB, flags=0):
    """contract!(C, labelsC, A, labelsA, B, labelsB) (contract.jl:37-104); C is overwritten."""
    lc, la, lb = labels_arr(labelsC), labels_arr(labelsA), labels_arr(labelsB)
    _chk(lib().orc_contract(Ct.h, _p(lc, C.c_uint64), C.c_int(len(labelsC)), At.h, _p(la, C.c_uint64),
                            C.c_int(len(labelsA)), Bt.h, _p(lb, C.c_uint64), C.c_int(len(labelsB)),
                            C.c_uint32(flags)))
    return Ct
