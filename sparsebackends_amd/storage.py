"""Device-resident storage types mirroring the reference's
NewBlockSparseSorted (SparseBackends/src/blocksparse/storage.jl:3-9),
COOTensor (coo/storage.jl:1-6) and plain dense arrays.

Field names, 1-based keys/ids and column-major payloads are the reference's.
Reading .keys/.ids/.data copies them to the host (numpy).
"""
from __future__ import annotations

import ctypes as C
import threading
import weakref

import numpy as np

from ._abi import load, check

_local = threading.local()  # per thread: .own {device: _Context}, .bound {device: context}
_registry = weakref.WeakValueDictionary()  # handle value -> its _Context
_pending = []  # handles of contexts no longer referenced, destroyed at the next context() call
_pending_lock = threading.Lock()


class _Context:
    """A library context (HIP device + stream + device-memory cache) owned by
    Python: the thread that made it holds it in thread-local storage, and every
    tensor and plan made on it holds it too, so it lives until the thread has
    ended and the last of them is gone.  Then it is destroyed (sb_ctx_destroy)
    at the next context() call of any thread: not in __del__ itself, because
    a garbage-collection pass may finalise the context before tensors of the
    same pass whose device memory its cache still owns."""

    __slots__ = ("h", "device", "owned", "__weakref__")

    def __init__(self, device, handle=None):
        if handle is None:
            h = C.c_void_p()
            check(load().sb_ctx_create(device, C.byref(h)))
            self.h, self.owned = h, True
        else:  # a context made outside this module: used, never destroyed here
            self.h, self.owned = handle, False
        self.device = device
        if self.owned:
            _registry[self.h.value] = self

    def __del__(self):
        try:
            if self.owned and self.h:
                with _pending_lock:
                    _pending.append(self.h)
        except Exception:
            pass


def _drain():
    with _pending_lock:
        hs = list(_pending)
        _pending.clear()
    for h in hs:
        load().sb_ctx_destroy(h)


def _context_obj(device=0):
    if _pending:
        _drain()
    b = getattr(_local, "bound", None)
    if b is not None and device in b:
        return b[device]
    own = getattr(_local, "own", None)
    if own is None:
        own = _local.own = {}
    c = own.get(device)
    if c is None:
        c = own[device] = _Context(device)
    return c


def context(device=0):
    """The context handle (HIP device + stream) this thread uses on a device:
    the one bound by bind_context, else the thread's own default context,
    made at its first use.  A thread's default context is its own (a thread
    started later never continues it) and is destroyed once the thread has
    ended and no tensor or plan made on it is left."""
    return _context_obj(device).h


def bind_context(handle, device=0):
    """Make the calling thread use the context handle on device (None: back to
    its own default).  Threads bound to one context share its stream and its
    device memory cache, so tensors and plans made by one of them can be used
    by the others (a tensor belongs to the context that made it); their
    library calls are ordered by the stream, and the binding keeps the context
    alive.  Binding turns the context's graph replays of eigensolve operators
    off (a capture on the shared stream would take in the other threads'
    launches); PlanGraph must not be used on a context that several threads
    use.  Timing measurements on a shared context (the autotuner's candidate
    runs, profile()) take in the other threads' work on the same stream."""
    b = getattr(_local, "bound", None)
    if b is None:
        b = _local.bound = {}
    if handle is None:
        b.pop(device, None)
        return
    h = handle if isinstance(handle, C.c_void_p) else C.c_void_p(handle)
    obj = _registry.get(h.value)
    if obj is None:
        obj = _Context(device, h)
    # a graph capture on the shared stream would take in the other threads'
    # launches: the shared context's eigensolves run without graphs
    check(load().sb_ctx_graphs(obj.h, 0))
    b[device] = obj


def synchronize(device=0):
    check(load().sb_ctx_synchronize(context(device)))


def memory(device=0):
    """(bytes in use, bytes cached) of this thread's context's device memory cache"""
    a, b = C.c_int64(), C.c_int64()
    check(load().sb_ctx_memory(context(device), C.byref(a), C.byref(b)))
    return int(a.value), int(b.value)


def trim(device=0):
    """return the device memory this thread's context keeps cached for reuse
    (blocks of live tensors stay); synchronises the device"""
    check(load().sb_ctx_trim(context(device)))


def profile(enable=True, device=0, products_only=False):
    """Time every numeric plan step (products_only: only the block products,
    grouped_gemm) with HIP events on the library's stream."""
    check(load().sb_ctx_profile(context(device), (2 if products_only else 1) if enable else 0))


def set_svd_sweeps(sweeps=40, device=0):
    """the sweep cap of the one-sided Jacobi SVDs (default 40); a factorisation
    still rotating after it raises SBError of kind "LAPACKException", as the
    reference's LAPACK SVD does when it fails to converge"""
    check(load().sb_ctx_set_svd_sweeps(context(device), int(sweeps)))


def profile_reset(device=0):
    check(load().sb_ctx_profile_reset(context(device)))


def profile_read(device=0):
    """{kernel name: dict(count, total_ms, min_ms, max_ms)} (synchronises)."""
    need = C.c_int64()
    check(load().sb_ctx_profile_read(context(device), None, 0, C.byref(need)))
    buf = C.create_string_buffer(int(need.value) + 1)
    check(load().sb_ctx_profile_read(context(device), buf, len(buf), C.byref(need)))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, tot, mn, mx = line.split()
        out[name] = dict(count=int(cnt), total_ms=float(tot), min_ms=float(mn), max_ms=float(mx))
    return out


def _i64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


_status = check  # the status check, for functions whose keyword argument is named check

# declared key and payload types (include/sb_status.h SB_KEY_*, SB_VAL_*): the
# device holds int32 keys and float64 payloads; parts() returns the declared ones
KEY_TYPES = [np.int64, np.int32, np.int16, np.int8, np.uint64, np.uint32, np.uint16, np.uint8]
VAL_TYPES = [np.float64, np.float32, np.float16, np.int64, np.int32, np.int16, np.int8]


class _Tensor:
    kind = 0

    def __init__(self, handle, device=0):
        self._h = handle if isinstance(handle, C.c_void_p) else C.c_void_p(handle)
        self.device = device
        self._owner = _context_obj(device)  # the context whose cache holds the device memory

    @property
    def handle(self):
        return self._h

    @property
    def ctx(self):
        return context(self.device)

    def __del__(self):
        try:
            if self._h:
                load().sb_tensor_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def permutedims(self, perm):
        """Base.permutedims (new tensor); perm is 1-based."""
        out = C.c_void_p()
        p = _i64(perm)
        check(load().sb_permutedims(self.ctx, self._h, _ptr(p, C.c_int64), C.byref(out)))
        return type(self)(out, self.device)

    def permutedims_(self, perm):
        """permutedims! (in place)."""
        p = _i64(perm)
        check(load().sb_permutedims_inplace(self.ctx, self._h, _ptr(p, C.c_int64)))
        return self

    def types(self):
        """(key dtype, value dtype) the tensor declares (rekey / retypeval change them)"""
        k, v = C.c_int(), C.c_int()
        check(load().sb_tensor_types(self._h, C.byref(k), C.byref(v)))
        return KEY_TYPES[k.value], VAL_TYPES[v.value]

    def _declared(self, d):
        kt, vt = self.types()
        if "keys" in d:
            d["keys"] = d["keys"].astype(kt)
        for f in ("data", "vals", "templates", "scalars"):
            if f in d and vt is not np.float64:
                d[f] = d[f].astype(vt)
        return d


class DenseTensor(_Tensor):
    """A dense column-major float64 array on the device."""
    kind = 3

    @staticmethod
    def from_numpy(arr, device=0):
        a = np.asarray(arr, dtype=np.float64)
        flat = np.ascontiguousarray(a.reshape(-1, order="F"))
        d = _i64(a.shape)
        out = C.c_void_p()
        check(load().sb_dense_from_host(context(device), a.ndim, _ptr(d, C.c_int64), _ptr(flat, C.c_double),
                                        C.byref(out)))
        return DenseTensor(out, device)

    @staticmethod
    def from_device_ptr(ptr, shape, borrow=True, device=0):
        """Wrap (borrow) or copy caller device memory, e.g. torch's tensor.data_ptr()."""
        d = _i64(shape)
        out = C.c_void_p()
        check(load().sb_dense_from_device(context(device), len(shape), _ptr(d, C.c_int64), C.c_void_p(ptr),
                                          1 if borrow else 0, C.byref(out)))
        return DenseTensor(out, device)

    @property
    def dims(self):
        N = C.c_int()
        d = np.zeros(16, dtype=np.int64)
        check(load().sb_dense_info(self._h, C.byref(N), _ptr(d, C.c_int64)))
        return tuple(int(x) for x in d[:N.value])

    def device_ptr(self):
        p = C.c_void_p()
        check(load().sb_dense_device_ptr(self._h, C.byref(p)))
        return p.value

    def numpy(self):
        shape = self.dims
        flat = np.zeros(int(np.prod(shape)) if shape else 1, dtype=np.float64)
        check(load().sb_dense_to_host(self.ctx, self._h, _ptr(flat, C.c_double)))
        return flat.reshape(shape, order="F")


class NewBlockSparseSorted(_Tensor):
    """Prefix-sparse block tensor: the first P = N - N2 axes are sparse keys
    sorted by their column-major linear index; the last N2 axes form one dense
    block per key (blocksparse/storage.jl:3-13)."""
    kind = 1

    @staticmethod
    def new(dims, N2, device=0):
        """NewBlockSparseSorted{T,N,N2}(dims): no blocks (storage.jl:86-93)."""
        d = _i64(dims)
        out = C.c_void_p()
        check(load().sb_bs_new(context(device), len(dims), N2, _ptr(d, C.c_int64), C.byref(out)))
        return NewBlockSparseSorted(out, device)

    @staticmethod
    def from_parts(dims, N2, keys, ids, data, device=0):
        d = _i64(dims)
        k = _i64(keys).reshape(-1)
        i = _i64(ids)
        x = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
        out = C.c_void_p()
        check(load().sb_bs_from_host(context(device), len(dims), N2, _ptr(d, C.c_int64), len(i), _ptr(k, C.c_int64),
                                     _ptr(i, C.c_int64), _ptr(x, C.c_double), C.byref(out)))
        return NewBlockSparseSorted(out, device)

    def info(self):
        N, N2, nb, bs = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
        d = np.zeros(16, dtype=np.int64)
        check(load().sb_bs_info(self._h, C.byref(N), C.byref(N2), _ptr(d, C.c_int64), C.byref(nb), C.byref(bs)))
        return tuple(int(x) for x in d[:N.value]), N2.value, nb.value, bs.value

    @property
    def dims(self):
        return self.info()[0]

    @property
    def N2(self):
        return self.info()[1]

    @property
    def nblocks(self):
        return self.info()[2]

    @property
    def blksize(self):
        return self.info()[3]

    def parts(self):
        """dims, N2, blksize, keys (nblocks x P), ids, data copied to the host."""
        dims, N2, nb, bs = self.info()
        P = len(dims) - N2
        keys = np.zeros((nb, P), dtype=np.int64)
        ids = np.zeros(nb, dtype=np.int64)
        data = np.zeros(nb * bs, dtype=np.float64)
        check(load().sb_bs_to_host(self.ctx, self._h, _ptr(keys, C.c_int64), _ptr(ids, C.c_int64),
                                   _ptr(data, C.c_double)))
        return self._declared(dict(dims=dims, N2=N2, blksize=bs, keys=keys, ids=ids, data=data))

    @property
    def keys(self):
        return self.parts()["keys"]

    @property
    def ids(self):
        return self.parts()["ids"]

    @property
    def data(self):
        return self.parts()["data"]

    def device_ptrs(self):
        k, i, d = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(load().sb_bs_device_ptrs(self._h, C.byref(k), C.byref(i), C.byref(d)))
        return k.value, i.value, d.value

    def set_data_(self, data):
        """A.data .= data (the payload field, storage.jl:3-9): keys and ids stay."""
        x = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
        if x.size != self.nblocks * self.blksize:
            raise ValueError("payload size must be nblocks * blksize")
        check(load().sb_bs_set_payload(self.ctx, self._h, x.ctypes.data, 0))
        return self

    def set_data_device_(self, ptr):
        """the payload copied from device memory at ptr (nblocks * blksize float64)."""
        check(load().sb_bs_set_payload(self.ctx, self._h, ptr, 1))
        return self

    def __getitem__(self, I):
        a = _i64(I)
        v = C.c_double()
        check(load().sb_bs_getindex(self.ctx, self._h, _ptr(a, C.c_int64), C.byref(v)))
        return v.value

    def __setitem__(self, I, v):
        a = _i64(I)
        check(load().sb_bs_setindex(self.ctx, self._h, float(v), _ptr(a, C.c_int64)))

    # ---- Krylov vector operations (SparseBackends/src/tensor_wrappers.jl:792-923)
    def norm(self):
        """norm(data) (tensor_wrappers.jl:793)."""
        v = C.c_double()
        check(load().sb_bs_norm(self.ctx, self._h, C.byref(v)))
        return v.value

    def norm_noalloc(self, out_device_ptr):
        """norm(data) written to out_device_ptr[0] (device memory) on the
        context's stream, with no device allocation and no synchronisation
        (sb_bs_norm_noalloc; for graph capture and allocation-free loops)."""
        check(load().sb_bs_norm_noalloc(self.ctx, self._h, C.c_void_p(out_device_ptr)))

    def copy(self):
        """copy: keys, ids and data (tensor_wrappers.jl:817-823)."""
        out = C.c_void_p()
        check(load().sb_bs_copy(self.ctx, self._h, C.byref(out)))
        return NewBlockSparseSorted(out, self.device)

    def similar(self):
        """_zero_similar: same keys and ids, zero data (tensor_wrappers.jl:898-904)."""
        out = C.c_void_p()
        check(load().sb_bs_similar(self.ctx, self._h, C.byref(out)))
        return NewBlockSparseSorted(out, self.device)

    def __rmul__(self, alpha):
        """alpha * A: a new tensor (tensor_wrappers.jl:801-807)."""
        out = C.c_void_p()
        check(load().sb_bs_scale(self.ctx, self._h, float(alpha), C.byref(out)))
        return NewBlockSparseSorted(out, self.device)

    def fill_(self, v):
        """fill!(data, v) (tensor_wrappers.jl:850-853)."""
        check(load().sb_bs_fill(self.ctx, self._h, float(v)))
        return self

    def axpby_(self, alpha, A, beta):
        """self.data = alpha*A.data + beta*self.data position by position
        (_apply_elementwise!, tensor_wrappers.jl:856-862)."""
        check(load().sb_bs_axpby(self.ctx, float(alpha), A.handle, float(beta), self._h))
        return self

    def __add__(self, other):
        """block-wise addition over the union of keys (tensor_wrappers.jl:826-840)."""
        out = C.c_void_p()
        check(load().sb_bs_add(self.ctx, self._h, other.handle, C.byref(out)))
        return NewBlockSparseSorted(out, self.device)

    def restrict(self, other):
        """other's blocks in this tensor's block structure (zero where other
        has no block of a key): the projection onto this vector's block space"""
        out = C.c_void_p()
        check(load().sb_bs_restrict(self.ctx, self._h, other.handle, C.byref(out)))
        return NewBlockSparseSorted(out, self.device)

    def dot(self, other):
        """inner product over blocks of equal key."""
        v = C.c_double()
        check(load().sb_bs_dot(self.ctx, self._h, other.handle, C.byref(v)))
        return v.value

    def sort_(self):
        """sort! (storage.jl:403-410): lexicographic keys, ids follow, data stays."""
        check(load().sb_bs_sort_lex(self.ctx, self._h))
        return self

    def is_dense(self):
        o = C.c_int()
        check(load().sb_bs_is_dense(self._h, C.byref(o)))
        return bool(o.value)

    def to_dense(self, init=0.0, merged_axes=None):
        """to_dense(B; init, merged_axes) (blocksparse/conversions.jl:83-187) -> numpy."""
        return self.to_dense_tensor(init, merged_axes).numpy()

    def to_dense_tensor(self, init=0.0, merged_axes=None):
        g = _i64(merged_axes if merged_axes else np.zeros((0, 2))).reshape(-1)
        out = C.c_void_p()
        check(load().sb_bs_to_dense(self.ctx, self._h, float(init), len(g) // 2, _ptr(g, C.c_int64), C.byref(out)))
        return DenseTensor(out, self.device)

    def fuse_two_axes_(self, ax1, ax2):
        """fuse_two_axes! (conversions.jl:299-314)."""
        out = C.c_void_p()
        check(load().sb_bs_fuse_two_axes(self.ctx, self._h, ax1, ax2, C.byref(out)))
        return NewBlockSparseSorted(out, self.device)


class AliasedBlockSparse(_Tensor):
    """Block-sparse tensor whose block i is scalars[i] * templates[alias_ids[i]]
    (aliased/storage.jl:46-59); keys sorted by the column-major prefix lin."""
    kind = 4

    @staticmethod
    def new(dims, N2, device=0):
        """AliasedBlockSparse{T,N,N2}(dims): no blocks (aliased/storage.jl:122-132)."""
        d = _i64(dims)
        out = C.c_void_p()
        check(load().sb_al_new(context(device), len(dims), N2, _ptr(d, C.c_int64), C.byref(out)))
        return AliasedBlockSparse(out, device)

    @staticmethod
    def from_parts(dims, N2, templates, ntemplates, keys, alias_ids, scalars, device=0):
        d = _i64(dims)
        t = np.ascontiguousarray(np.asarray(templates, dtype=np.float64).reshape(-1))
        k = _i64(keys).reshape(-1)
        a = _i64(alias_ids)
        sc = np.ascontiguousarray(np.asarray(scalars, dtype=np.float64).reshape(-1))
        out = C.c_void_p()
        check(load().sb_al_from_host(context(device), len(dims), N2, _ptr(d, C.c_int64), int(ntemplates),
                                     _ptr(t, C.c_double), len(a), _ptr(k, C.c_int64), _ptr(a, C.c_int64),
                                     _ptr(sc, C.c_double), C.byref(out)))
        return AliasedBlockSparse(out, device)

    def info(self):
        N, N2 = C.c_int(), C.c_int()
        nb, bs, nt = C.c_int64(), C.c_int64(), C.c_int64()
        d = np.zeros(16, dtype=np.int64)
        check(load().sb_al_info(self._h, C.byref(N), C.byref(N2), _ptr(d, C.c_int64), C.byref(nb), C.byref(bs),
                                C.byref(nt)))
        return tuple(int(x) for x in d[:N.value]), N2.value, nb.value, bs.value, nt.value

    def parts(self):
        dims, N2, nb, bs, nt = self.info()
        P = len(dims) - N2
        keys = np.zeros((nb, P), dtype=np.int64)
        ids = np.zeros(nb, dtype=np.int64)
        sc = np.zeros(nb, dtype=np.float64)
        tm = np.zeros(nt * bs, dtype=np.float64)
        check(load().sb_al_to_host(self.ctx, self._h, _ptr(keys, C.c_int64), _ptr(ids, C.c_int64),
                                   _ptr(sc, C.c_double), _ptr(tm, C.c_double)))
        return self._declared(dict(dims=dims, N2=N2, blksize=bs, ntemplates=nt, keys=keys, alias_ids=ids,
                                   scalars=sc, templates=tm))

    def __getitem__(self, I):
        a = _i64(I)
        v = C.c_double()
        check(load().sb_al_getindex(self.ctx, self._h, _ptr(a, C.c_int64), C.byref(v)))
        return v.value

    def compression_ratio(self):
        """n_blocks / n_templates (aliased/storage.jl:163-164)."""
        v = C.c_double()
        check(load().sb_al_compression_ratio(self._h, C.byref(v)))
        return v.value

    def to_blocksparse(self):
        """to_blocksparse (aliased/conversions.jl:18-32)."""
        out = C.c_void_p()
        check(load().sb_al_to_blocksparse(self.ctx, self._h, C.byref(out)))
        return NewBlockSparseSorted(out, self.device)

    def to_dense(self, init=0.0):
        """to_dense(A; init) (aliased/conversions.jl:45-64) -> numpy."""
        out = C.c_void_p()
        check(load().sb_al_to_dense(self.ctx, self._h, float(init), C.byref(out)))
        return DenseTensor(out, self.device).numpy()


class COOTensor(_Tensor):
    """Element-sparse tensor: keys (nnz x N), vals, dirty flag (coo/storage.jl:1-6)."""
    kind = 2

    @staticmethod
    def new(dims, device=0):
        d = _i64(dims)
        out = C.c_void_p()
        check(load().sb_coo_new(context(device), len(dims), _ptr(d, C.c_int64), C.byref(out)))
        return COOTensor(out, device)

    @staticmethod
    def from_parts(dims, keys, vals, dirty=False, device=0):
        d = _i64(dims)
        k = _i64(keys).reshape(-1)
        v = np.ascontiguousarray(np.asarray(vals, dtype=np.float64))
        out = C.c_void_p()
        check(load().sb_coo_from_host(context(device), len(dims), _ptr(d, C.c_int64), len(v), _ptr(k, C.c_int64),
                                      _ptr(v, C.c_double), 1 if dirty else 0, C.byref(out)))
        return COOTensor(out, device)

    def parts(self):
        N, nnz, dirty = C.c_int(), C.c_int64(), C.c_int()
        d = np.zeros(16, dtype=np.int64)
        check(load().sb_coo_info(self._h, C.byref(N), _ptr(d, C.c_int64), C.byref(nnz), C.byref(dirty)))
        keys = np.zeros((nnz.value, N.value), dtype=np.int64)
        vals = np.zeros(nnz.value, dtype=np.float64)
        check(load().sb_coo_to_host(self.ctx, self._h, _ptr(keys, C.c_int64), _ptr(vals, C.c_double)))
        return self._declared(dict(dims=tuple(int(x) for x in d[:N.value]), keys=keys, vals=vals,
                                   dirty=bool(dirty.value)))

    @property
    def dims(self):
        return self.parts()["dims"]

    @property
    def keys(self):
        return self.parts()["keys"]

    @property
    def vals(self):
        return self.parts()["vals"]

    @property
    def dirty(self):
        return self.parts()["dirty"]

    def __getitem__(self, I):
        a = _i64(I)
        v = C.c_double()
        check(load().sb_coo_getindex(self.ctx, self._h, _ptr(a, C.c_int64), C.byref(v)))
        return v.value

    def setindex_(self, v, *I, dirty=False):
        """setindex!(A, v, I...; dirty) (coo/storage.jl:84-134)."""
        a = _i64(I)
        check(load().sb_coo_setindex(self.ctx, self._h, float(v), _ptr(a, C.c_int64), 1 if dirty else 0))
        return self

    def __setitem__(self, I, v):
        self.setindex_(v, *I)

    def sort_(self):
        check(load().sb_coo_sort(self.ctx, self._h))
        return self

    def to_dense(self):
        out = C.c_void_p()
        check(load().sb_coo_to_dense(self.ctx, self._h, C.byref(out)))
        return DenseTensor(out, self.device).numpy()

    def fuse_axes(self, ax1, ax2):
        """fuse_axes_coo (coo/conversions.jl:174-223)."""
        out = C.c_void_p()
        check(load().sb_coo_fuse_axes(self.ctx, self._h, ax1, ax2, C.byref(out)))
        return COOTensor(out, self.device)


def rekey(A, K2, check=True):
    """rekey(A, K2; check=true) (blocksparse/conversions.jl:327-342, coo/conversions.jl:41-53,
    aliased/conversions.jl:76-93): a copy whose keys have the integer type K2 (a numpy
    integer type); with check, an ErrorException (SBError) when a prefix dimension or a
    key coordinate does not fit K2."""
    code = [np.dtype(t) for t in KEY_TYPES].index(np.dtype(K2))
    out = C.c_void_p()
    _status(load().sb_rekey(A.ctx, A.handle, code, 1 if check else 0, C.byref(out)))
    return type(A)(out, A.device)


def retypeval(A, T2):
    """retypeval(A::COOTensor, T2) (coo/conversions.jl:65-82): a copy whose values have
    the type T2 (a numpy type); ErrorException (SBError) when a value does not survive
    the round trip Float64(T2(v)) == v."""
    code = [np.dtype(t) for t in VAL_TYPES].index(np.dtype(T2))
    out = C.c_void_p()
    check(load().sb_coo_retypeval(A.ctx, A.handle, code, C.byref(out)))
    return COOTensor(out, A.device)


def blocksparse_from_dense(arr, N2, atol=0.0, rtol=0.0, dropzeros=True, device=0):
    """blocksparse_from_dense(A, Val(N2); atol, rtol, dropzeros) (blocksparse/conversions.jl:21-75)."""
    D = arr if isinstance(arr, DenseTensor) else DenseTensor.from_numpy(arr, device)
    out = C.c_void_p()
    check(load().sb_bs_from_dense(D.ctx, D.handle, N2, float(atol), float(rtol), 1 if dropzeros else 0,
                                  C.byref(out)))
    return NewBlockSparseSorted(out, D.device)


def coo_from_dense(arr, atol=1e-12, rtol=0.0, device=0):
    """coo_from_dense(A; atol=1e-12, rtol=0) (coo/conversions.jl:13-26)."""
    D = arr if isinstance(arr, DenseTensor) else DenseTensor.from_numpy(arr, device)
    out = C.c_void_p()
    check(load().sb_coo_from_dense(D.ctx, D.handle, float(atol), float(rtol), C.byref(out)))
    return COOTensor(out, D.device)


def to_dense(t, **kw):
    """to_dense for any storage kind -> numpy."""
    if isinstance(t, DenseTensor):
        return t.numpy()
    return t.to_dense(**kw)
