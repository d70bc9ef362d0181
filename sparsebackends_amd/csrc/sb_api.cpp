// The C ABI (include/sparsebackends_amd.h): handle management, host/device
// transfer, and the thin wrappers that map C++ exceptions to status codes.
#include <cstring>
#include <mutex>
#include <set>

#include "../../include/sparsebackends_amd.h"
#include "sb_batch.h"
#include "sb_convert.h"
#include "sb_factorize.h"
#include "sb_graph.h"
#include "sb_krylov.h"

using namespace sb;

static thread_local std::string g_err;

#define SB_TRY(...)                   \
  try {                               \
    __VA_ARGS__;                      \
    return SB_OK;                     \
  } catch (const sb::Error& e) {      \
    g_err = e.msg;                    \
    return e.code;                    \
  } catch (const std::bad_alloc&) {   \
    g_err = "host out of memory";     \
    return SB_E_NOMEM;                \
  } catch (const std::exception& e) { \
    g_err = e.what();                 \
    return SB_E_ERROR;                \
  }

static Ctx* C_(sb_ctx_t c) {
  if (!c) fail(SB_E_HANDLE, "null context");
  Ctx* ctx = reinterpret_cast<Ctx*>(c);
  HIP_CHECK(hipSetDevice(ctx->device));
  return ctx;
}
static Tensor* T_(sb_tensor_t t, int kind = 0) { return as_tensor(reinterpret_cast<void*>(t), kind); }
static sb_tensor_t H_(TP&& t) { return reinterpret_cast<sb_tensor_t>(t.release()); }
static std::vector<i64> vec(const int64_t* p, int n) { return std::vector<i64>(p, p + n); }
static Labels labs(const uint64_t* p, int n) {
  if (n < 0) fail(SB_E_ARGUMENT, "negative label count");
  Labels l(n);
  for (int i = 0; i < n; ++i) l[i] = Label{p[2 * i], p[2 * i + 1]};
  return l;
}

static void up_i32(Ctx* ctx, DBuf<i32>& dst, const int64_t* src, i64 n) {
  std::vector<i32> h(n);
  for (i64 i = 0; i < n; ++i) {
    if (src[i] < INT32_MIN || src[i] > INT32_MAX) fail(SB_E_LIMIT, "key or id outside int32");
    h[i] = (i32)src[i];
  }
  dst.alloc(ctx, n ? n : 1);
  if (n) HIP_CHECK(hipMemcpyAsync(dst.get(), h.data(), n * sizeof(i32), hipMemcpyHostToDevice, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
}
static void down_i32(Ctx* ctx, int64_t* dst, const i32* src, i64 n) {
  std::vector<i32> h(n);
  if (n) HIP_CHECK(hipMemcpyAsync(h.data(), src, n * sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (i64 i = 0; i < n; ++i) dst[i] = h[i];
}

extern "C" {

const char* sb_last_error(void) { return g_err.c_str(); }
int sb_version(void) { return 10000; }

int sb_ctx_create(int device, sb_ctx_t* out) {
  SB_TRY({
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(SB_E_ARGUMENT, "no such HIP device");
    Ctx* ctx = new Ctx(device);  // makes the device current
    try {
      noalloc_setup(device);
    } catch (...) {
      delete ctx;
      throw;
    }
    *out = reinterpret_cast<sb_ctx_t>(ctx);
  })
}
int sb_ctx_destroy(sb_ctx_t c) { SB_TRY(delete C_(c)) }
int sb_ctx_set_stream(sb_ctx_t c, void* s) {
  SB_TRY({
    Ctx* ctx = C_(c);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) HIP_CHECK(hipStreamDestroy(ctx->stream));
    if (s) {
      ctx->stream = static_cast<hipStream_t>(s);
      ctx->own_stream = false;
    } else {
      HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
      ctx->own_stream = true;
    }
  })
}
int sb_ctx_synchronize(sb_ctx_t c) { SB_TRY(HIP_CHECK(hipStreamSynchronize(C_(c)->stream))) }
int sb_ctx_memory(sb_ctx_t c, int64_t* in_use, int64_t* cached) {
  SB_TRY({
    Ctx* ctx = C_(c);
    *in_use = (int64_t)ctx->cache.bytes_in_use();
    *cached = (int64_t)ctx->cache.bytes_cached();
  })
}
int sb_ctx_trim(sb_ctx_t c) {
  SB_TRY({
    Ctx* ctx = C_(c);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->cache.trim();
  })
}

int sb_ctx_profile(sb_ctx_t c, int enable) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!enable) ctx->prof_collect();
    if (enable < 0 || enable > 2) fail(SB_E_ARGUMENT, "ArgumentError: profile level must be 0, 1 or 2");
    ctx->profile = enable;
  })
}
int sb_ctx_profile_reset(sb_ctx_t c) {
  SB_TRY({
    Ctx* ctx = C_(c);
    ctx->prof_collect();
    ctx->prof.clear();
  })
}
int sb_ctx_profile_read(sb_ctx_t c, char* buf, int64_t cap, int64_t* needed) {
  SB_TRY({
    Ctx* ctx = C_(c);
    ctx->prof_collect();
    std::string s;
    char line[256];
    for (auto& kv : ctx->prof) {
      snprintf(line, sizeof line, "%s %ld %.6f %.6f %.6f\n", kv.first.c_str(), kv.second.count, kv.second.ms,
               kv.second.min_ms, kv.second.max_ms);
      s += line;
    }
    if (needed) *needed = (int64_t)s.size() + 1;
    if (buf && cap > 0) {
      size_t n = std::min((size_t)(cap - 1), s.size());
      std::memcpy(buf, s.data(), n);
      buf[n] = 0;
    }
  })
}

int sb_tensor_free(sb_tensor_t t) { SB_TRY(delete T_(t)) }
int sb_ctx_set_svd_sweeps(sb_ctx_t c, int sweeps) {
  SB_TRY({
    if (sweeps < 1 || sweeps > 10000) fail(SB_E_ARGUMENT, "ArgumentError: the SVD sweep cap must be in 1..10000");
    C_(c)->svd_sweeps = sweeps;
  })
}
int sb_tensor_kind(sb_tensor_t t, int* kind) { SB_TRY(*kind = T_(t)->kind) }
int sb_tensor_types(sb_tensor_t t, int* key_type, int* val_type) {
  SB_TRY({
    if (key_type) *key_type = T_(t)->key_type;
    if (val_type) *val_type = T_(t)->val_type;
  })
}
int sb_rekey(sb_ctx_t c, sb_tensor_t A, int key_type, int check, sb_tensor_t* out) {
  SB_TRY(*out = H_(tensor_rekey(C_(c), *T_(A), key_type, check != 0)))
}
int sb_coo_retypeval(sb_ctx_t c, sb_tensor_t A, int val_type, sb_tensor_t* out) {
  SB_TRY(*out = H_(coo_retypeval(C_(c), *T_(A, COO), val_type)))
}

// ---------------------------------------------------------------- block-sparse
int sb_bs_new(sb_ctx_t c, int N, int N2, const int64_t* dims, sb_tensor_t* out) {
  SB_TRY(*out = H_(bs_new(C_(c), vec(dims, N), N2)))
}
int sb_bs_from_host(sb_ctx_t c, int N, int N2, const int64_t* dims, int64_t nblocks, const int64_t* keys,
                    const int64_t* ids, const double* data, sb_tensor_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    TP t = bs_make(ctx, vec(dims, N), N2);
    int P = N - N2;
    t->n = nblocks;
    up_i32(ctx, t->keys, keys, nblocks * P);
    up_i32(ctx, t->ids, ids, nblocks);
    for (i64 i = 0; i < nblocks; ++i)
      if (ids[i] < 1 || ids[i] > nblocks) fail(SB_E_BOUNDS, "block id outside 1:nblocks");
    // keys sorted by column-major lin and unique? (else the tensor is marked
    // unsorted and permutes re-sort it, as the reference's sortperm would)
    bool sorted = true;
    for (i64 i = 1; i < nblocks && sorted; ++i) {
      for (int d = P - 1; d >= 0; --d) {
        i64 a = keys[(i - 1) * P + d], b = keys[i * P + d];
        if (a < b) break;
        if (a > b || d == 0) {
          sorted = false;
          break;
        }
      }
    }
    t->dirty = !sorted;
    t->data.alloc(ctx, nblocks * t->blksize);
    if (nblocks * t->blksize)
      HIP_CHECK(hipMemcpyAsync(t->data.get(), data, nblocks * t->blksize * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = H_(std::move(t));
  })
}
int sb_bs_info(sb_tensor_t h, int* N, int* N2, int64_t* dims, int64_t* nblocks, int64_t* blksize) {
  SB_TRY({
    Tensor* t = T_(h, BS);
    *N = t->N();
    *N2 = t->N2;
    if (dims)
      for (int i = 0; i < t->N(); ++i) dims[i] = t->dims[i];
    *nblocks = t->n;
    *blksize = t->blksize;
  })
}
int sb_bs_to_host(sb_ctx_t c, sb_tensor_t h, int64_t* keys, int64_t* ids, double* data) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* t = T_(h, BS);
    if (keys) down_i32(ctx, keys, t->keys.get(), t->n * t->P());
    if (ids) down_i32(ctx, ids, t->ids.get(), t->n);
    if (data && t->n * t->blksize) {
      HIP_CHECK(hipMemcpyAsync(data, t->data.get(), t->n * t->blksize * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
  })
}
int sb_bs_device_ptrs(sb_tensor_t h, const int32_t** keys, const int32_t** ids, double** data) {
  SB_TRY({
    Tensor* t = T_(h, BS);
    if (keys) *keys = t->keys.get();
    if (ids) *ids = t->ids.get();
    if (data) {
      *data = t->data.get();
      t->external = true;  // writes through the pointer are invisible to the block boxes
    }
  })
}
int sb_bs_set_payload(sb_ctx_t c, sb_tensor_t h, const double* data, int from_device) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* t = T_(h, BS);
    i64 cnt = t->n * t->blksize;
    if (cnt && !data) fail(SB_E_ARGUMENT, "null payload");
    t->touch();
    if (cnt)
      HIP_CHECK(hipMemcpyAsync(t->data.get(), data, cnt * sizeof(double),
                               from_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    if (!from_device) HIP_CHECK(hipStreamSynchronize(ctx->stream));
  })
}
int sb_bs_getindex(sb_ctx_t c, sb_tensor_t h, const int64_t* I, double* v) {
  SB_TRY(*v = bs_getindex(C_(c), *T_(h, BS), I))
}
int sb_bs_setindex(sb_ctx_t c, sb_tensor_t h, double v, const int64_t* I) {
  SB_TRY({
    Tensor* t = T_(h, BS);
    t->touch();
    bs_setindex(C_(c), *t, v, I);
  })
}
int sb_bs_sort_lex(sb_ctx_t c, sb_tensor_t h) {
  SB_TRY({
    Tensor* t = T_(h, BS);
    t->touch();
    bs_sort_lex(C_(c), *t);
  })
}
int sb_bs_is_dense(sb_tensor_t h, int* out) { SB_TRY(*out = bs_is_dense(*T_(h, BS)) ? 1 : 0) }
int sb_bs_from_dense(sb_ctx_t c, sb_tensor_t d, int N2, double atol, double rtol, int dropzeros, sb_tensor_t* out) {
  SB_TRY(*out = H_(bs_from_dense(C_(c), *T_(d, DENSE), N2, atol, rtol, dropzeros != 0)))
}
int sb_bs_to_dense(sb_ctx_t c, sb_tensor_t h, double init, int ngroups, const int64_t* groups, sb_tensor_t* out) {
  SB_TRY({
    std::vector<std::pair<i64, i64>> g;
    for (int i = 0; i < ngroups; ++i) g.push_back({groups[2 * i], groups[2 * i + 1]});
    *out = H_(bs_to_dense(C_(c), *T_(h, BS), init, g));
  })
}
int sb_bs_fuse_two_axes(sb_ctx_t c, sb_tensor_t h, int64_t ax1, int64_t ax2, sb_tensor_t* out) {
  SB_TRY(*out = H_(bs_fuse_two_axes(C_(c), *T_(h, BS), ax1, ax2)))
}

// ---------------------------------------------------------------- AliasedBlockSparse
int sb_al_new(sb_ctx_t c, int N, int N2, const int64_t* dims, sb_tensor_t* out) {
  SB_TRY(*out = H_(al_new(C_(c), vec(dims, N), N2)))
}
int sb_al_from_host(sb_ctx_t c, int N, int N2, const int64_t* dims, int64_t ntemplates, const double* templates,
                    int64_t nblocks, const int64_t* keys, const int64_t* alias_ids, const double* scalars,
                    sb_tensor_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    TP t = al_new(ctx, vec(dims, N), N2);
    int P = N - N2;
    if (ntemplates < 0 || nblocks < 0) fail(SB_E_ARGUMENT, "negative count");
    for (i64 i = 0; i < nblocks; ++i)
      if (alias_ids[i] < 1 || alias_ids[i] > ntemplates) fail(SB_E_BOUNDS, "alias id outside 1:ntemplates");
    t->n = nblocks;
    t->ntemplates = ntemplates;
    up_i32(ctx, t->keys, keys, nblocks * P);
    up_i32(ctx, t->ids, alias_ids, nblocks);
    t->scalars.alloc(ctx, nblocks ? nblocks : 1);
    t->data.alloc(ctx, ntemplates * t->blksize);
    if (nblocks)
      HIP_CHECK(hipMemcpyAsync(t->scalars.get(), scalars, nblocks * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
    if (ntemplates * t->blksize)
      HIP_CHECK(hipMemcpyAsync(t->data.get(), templates, ntemplates * t->blksize * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = H_(std::move(t));
  })
}
int sb_al_info(sb_tensor_t h, int* N, int* N2, int64_t* dims, int64_t* nblocks, int64_t* blksize,
               int64_t* ntemplates) {
  SB_TRY({
    Tensor* t = T_(h, ALIASED);
    *N = t->N();
    *N2 = t->N2;
    std::copy(t->dims.begin(), t->dims.end(), dims);
    *nblocks = t->n;
    *blksize = t->blksize;
    *ntemplates = t->ntemplates;
  })
}
int sb_al_to_host(sb_ctx_t c, sb_tensor_t h, int64_t* keys, int64_t* alias_ids, double* scalars, double* templates) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* t = T_(h, ALIASED);
    if (keys) down_i32(ctx, keys, t->keys.get(), t->n * t->P());
    if (alias_ids) down_i32(ctx, alias_ids, t->ids.get(), t->n);
    if (scalars && t->n)
      HIP_CHECK(hipMemcpyAsync(scalars, t->scalars.get(), t->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (templates && t->ntemplates * t->blksize)
      HIP_CHECK(hipMemcpyAsync(templates, t->data.get(), t->ntemplates * t->blksize * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  })
}
int sb_al_getindex(sb_ctx_t c, sb_tensor_t h, const int64_t* I, double* v) {
  SB_TRY(*v = al_getindex(C_(c), *T_(h, ALIASED), I))
}
int sb_al_compression_ratio(sb_tensor_t h, double* out) {
  SB_TRY({
    Tensor* t = T_(h, ALIASED);
    *out = t->ntemplates == 0 ? 1.0 : (double)t->n / (double)t->ntemplates;
  })
}
int sb_al_to_blocksparse(sb_ctx_t c, sb_tensor_t h, sb_tensor_t* out) {
  SB_TRY(*out = H_(al_to_blocksparse(C_(c), *T_(h, ALIASED))))
}
int sb_al_to_dense(sb_ctx_t c, sb_tensor_t h, double init, sb_tensor_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    TP b = al_to_blocksparse(ctx, *T_(h, ALIASED));
    *out = H_(bs_to_dense(ctx, *b, init, {}));
  })
}

// ---------------------------------------------------------------- COO
int sb_coo_new(sb_ctx_t c, int N, const int64_t* dims, sb_tensor_t* out) { SB_TRY(*out = H_(coo_new(C_(c), vec(dims, N)))) }
int sb_coo_from_host(sb_ctx_t c, int N, const int64_t* dims, int64_t nnz, const int64_t* keys, const double* vals,
                     int dirty, sb_tensor_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    TP t = coo_new(ctx, vec(dims, N));
    t->n = nnz;
    up_i32(ctx, t->keys, keys, nnz * N);
    t->data.alloc(ctx, nnz);
    if (nnz) HIP_CHECK(hipMemcpyAsync(t->data.get(), vals, nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    t->dirty = dirty != 0;
    *out = H_(std::move(t));
  })
}
int sb_coo_info(sb_tensor_t h, int* N, int64_t* dims, int64_t* nnz, int* dirty) {
  SB_TRY({
    Tensor* t = T_(h, COO);
    *N = t->N();
    if (dims)
      for (int i = 0; i < t->N(); ++i) dims[i] = t->dims[i];
    *nnz = t->n;
    *dirty = t->dirty ? 1 : 0;
  })
}
int sb_coo_to_host(sb_ctx_t c, sb_tensor_t h, int64_t* keys, double* vals) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* t = T_(h, COO);
    if (keys) down_i32(ctx, keys, t->keys.get(), t->n * t->N());
    if (vals && t->n) {
      HIP_CHECK(hipMemcpyAsync(vals, t->data.get(), t->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
  })
}
int sb_coo_device_ptrs(sb_tensor_t h, const int32_t** keys, double** vals) {
  SB_TRY({
    Tensor* t = T_(h, COO);
    if (keys) *keys = t->keys.get();
    if (vals) *vals = t->data.get();
  })
}
int sb_coo_getindex(sb_ctx_t c, sb_tensor_t h, const int64_t* I, double* v) {
  SB_TRY(*v = coo_getindex(C_(c), *T_(h, COO), I))
}
int sb_coo_setindex(sb_ctx_t c, sb_tensor_t h, double v, const int64_t* I, int dirty) {
  SB_TRY(coo_setindex(C_(c), *T_(h, COO), v, I, dirty != 0))
}
int sb_coo_sort(sb_ctx_t c, sb_tensor_t h) { SB_TRY(coo_sort(C_(c), *T_(h, COO))) }
int sb_coo_from_dense(sb_ctx_t c, sb_tensor_t d, double atol, double rtol, sb_tensor_t* out) {
  SB_TRY(*out = H_(coo_from_dense(C_(c), *T_(d, DENSE), atol, rtol)))
}
int sb_coo_to_dense(sb_ctx_t c, sb_tensor_t h, sb_tensor_t* out) { SB_TRY(*out = H_(coo_to_dense(C_(c), *T_(h, COO)))) }
int sb_coo_fuse_axes(sb_ctx_t c, sb_tensor_t h, int64_t ax1, int64_t ax2, sb_tensor_t* out) {
  SB_TRY(*out = H_(coo_fuse_axes(C_(c), *T_(h, COO), ax1, ax2)))
}

// ---------------------------------------------------------------- dense
int sb_dense_from_host(sb_ctx_t c, int N, const int64_t* dims, const double* data, sb_tensor_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    TP t = dense_new(ctx, vec(dims, N));
    if (t->numel())
      HIP_CHECK(hipMemcpyAsync(t->data.get(), data, t->numel() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = H_(std::move(t));
  })
}
int sb_dense_from_device(sb_ctx_t c, int N, const int64_t* dims, double* dptr, int borrow, sb_tensor_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    TP t;
    if (borrow) {
      t.reset(new Tensor);
      t->ctx = ctx;
      t->kind = DENSE;
      t->dims = vec(dims, N);
      t->N2 = N;
      t->blksize = t->numel();
      t->n = 1;
      t->borrowed = dptr;
    } else {
      t = dense_new(ctx, vec(dims, N));
      if (t->numel())
        HIP_CHECK(hipMemcpyAsync(t->data.get(), dptr, t->numel() * sizeof(double), hipMemcpyDeviceToDevice,
                                 ctx->stream));
    }
    *out = H_(std::move(t));
  })
}
int sb_dense_info(sb_tensor_t h, int* N, int64_t* dims) {
  SB_TRY({
    Tensor* t = T_(h, DENSE);
    *N = t->N();
    if (dims)
      for (int i = 0; i < t->N(); ++i) dims[i] = t->dims[i];
  })
}
int sb_dense_to_host(sb_ctx_t c, sb_tensor_t h, double* data) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* t = T_(h, DENSE);
    if (t->numel())
      HIP_CHECK(hipMemcpyAsync(data, t->dptr(), t->numel() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  })
}
int sb_dense_device_ptr(sb_tensor_t h, double** data) { SB_TRY(*data = T_(h, DENSE)->dptr()) }

// ---------------------------------------------------------------- permutes
int sb_permutedims(sb_ctx_t c, sb_tensor_t h, const int64_t* perm, sb_tensor_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* t = T_(h);
    std::vector<i64> p = vec(perm, t->N());
    if (t->kind == BS)
      *out = H_(bs_permute(ctx, *t, p));
    else if (t->kind == COO)
      *out = H_(coo_permute(ctx, *t, p));
    else if (t->kind == ALIASED)
      *out = H_(al_permute(ctx, *t, p));
    else if (t->kind == DENSE) {
      std::vector<i64> s(p);
      std::sort(s.begin(), s.end());
      for (int i = 0; i < t->N(); ++i)
        if (s[i] != i + 1) fail(SB_E_ARGUMENT, "perm must be a permutation of 1:N");
      *out = H_(dense_permute(ctx, *t, p));
    } else
      fail(SB_E_METHOD, "permutedims: unsupported storage kind");
  })
}
int sb_permutedims_inplace(sb_ctx_t c, sb_tensor_t h, const int64_t* perm) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* t = T_(h);
    std::vector<i64> p = vec(perm, t->N());
    if (t->kind == BS) {
      t->touch();
      bs_permute_inplace(ctx, *t, p);
    }
    else if (t->kind == COO) {
      TP r = coo_permute(ctx, *t, p);
      t->dims = r->dims;
      t->keys = std::move(r->keys);
      t->dirty = true;
    } else
      fail(SB_E_METHOD, "permutedims!: unsupported storage kind");
  })
}

// ---------------------------------------------------------------- contraction
int sb_contract(sb_ctx_t c, sb_tensor_t C, const uint64_t* labelsC, int nC, sb_tensor_t A, const uint64_t* labelsA,
                int nA, sb_tensor_t B, const uint64_t* labelsB, int nB, uint32_t flags) {
  SB_TRY({
    Ctx* ctx = C_(c);
    auto pl = contract_plan(ctx, *T_(C), labs(labelsC, nC), *T_(A), labs(labelsA, nA), *T_(B), labs(labelsB, nB),
                            flags);
    pl->run_pending();
    // the plan's scratch returns to the cache; stream order keeps reuse safe
  })
}

int sb_contract_bs_bs_to_dense(sb_ctx_t c, sb_tensor_t Cd, const uint64_t* labelsC, int nC, sb_tensor_t A,
                               const uint64_t* labelsA, int nA, sb_tensor_t B, const uint64_t* labelsB, int nB,
                               uint32_t flags) {
  SB_TRY({
    Ctx* ctx = C_(c);
    Tensor* D = T_(Cd, DENSE);
    Tensor* a = T_(A);
    Tensor* b = T_(B);
    if (a->kind != BS || b->kind != BS) fail(SB_E_METHOD, "MethodError: contract_bs_bs_to_dense! needs BS x BS");
    int NC = D->N();
    TP Cbs = bs_new(ctx, D->dims, NC);
    auto pl = contract_plan(ctx, *Cbs, labs(labelsC, nC), *a, labs(labelsA, nA), *b, labs(labelsB, nB), flags);
    pl->run_pending();
    if (Cbs->n > 0) {
      HIP_CHECK(hipMemcpyAsync(D->dptr(), Cbs->data.get(), D->numel() * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream));
    }
  })
}

int sb_plan_create(sb_ctx_t c, sb_tensor_t C, const uint64_t* labelsC, int nC, sb_tensor_t A,
                   const uint64_t* labelsA, int nA, sb_tensor_t B, const uint64_t* labelsB, int nB, uint32_t flags,
                   sb_plan_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    auto pl = contract_plan(ctx, *T_(C), labs(labelsC, nC), *T_(A), labs(labelsA, nA), *T_(B), labs(labelsB, nB),
                            flags);
    pl->run_pending();
    *out = reinterpret_cast<sb_plan_t>(pl.release());
  })
}
int sb_heff_plan_create(sb_ctx_t c, sb_tensor_t Hpsi, sb_tensor_t psi, sb_tensor_t L, sb_tensor_t W1, sb_tensor_t W2,
                        sb_tensor_t R, uint32_t flags, sb_plan_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!out) fail(SB_E_ARGUMENT, "ArgumentError: null output handle");
    auto pl = heff_plan(ctx, *T_(Hpsi), *T_(psi), *T_(L), *T_(W1), *T_(W2), *T_(R), flags);
    pl->run_pending();
    *out = reinterpret_cast<sb_plan_t>(pl.release());
  })
}
int sb_plan_execute(sb_ctx_t c, sb_plan_t p) {
  SB_TRY({
    C_(c);
    if (!p) fail(SB_E_HANDLE, "null plan");
    reinterpret_cast<Plan*>(p)->run();
  })
}
int sb_plan_stats(sb_plan_t p, double* flops, double* gemm_bytes, double* perm_bytes, int64_t* npairs,
                  int64_t* nblocks_out) {
  SB_TRY({
    if (!p) fail(SB_E_HANDLE, "null plan");
    Plan* pl = reinterpret_cast<Plan*>(p);
    if (flops) *flops = pl->flops;
    if (gemm_bytes) *gemm_bytes = pl->gemm_bytes;
    if (perm_bytes) *perm_bytes = pl->perm_bytes;
    if (npairs) *npairs = pl->npairs;
    if (nblocks_out) *nblocks_out = pl->nblocks_out;
  })
}
int sb_plan_box_config(sb_plan_t p, int* tile, int* wg_per_cu, int* candidates, double* best_ms) {
  SB_TRY({
    if (!p) fail(SB_E_HANDLE, "null plan");
    Plan* pl = reinterpret_cast<Plan*>(p);
    if (tile) *tile = pl->box ? pl->box->g.cfg : -1;
    if (wg_per_cu) *wg_per_cu = pl->box ? pl->box->g.wgcu : 0;
    if (candidates) *candidates = pl->tune.candidates;
    if (best_ms) *best_ms = pl->tune.ms;
  })
}
int sb_autotune_reset(void) {
  SB_TRY({ autotune_cache_clear(); })
}
int sb_plan_operand_bytes(sb_plan_t p, double* a_bytes, double* b_bytes, double* c_bytes) {
  SB_TRY({
    if (!p) fail(SB_E_HANDLE, "null plan");
    Plan* pl = reinterpret_cast<Plan*>(p);
    if (a_bytes) *a_bytes = pl->a_bytes;
    if (b_bytes) *b_bytes = pl->b_bytes;
    if (c_bytes) *c_bytes = pl->c_bytes;
  })
}
int sb_plan_mfma_flops(sb_plan_t p, double* out) {
  SB_TRY({
    if (!p || !out) fail(SB_E_HANDLE, "null plan or output");
    *out = reinterpret_cast<Plan*>(p)->mfma_flops;
  })
}
int sb_plan_free(sb_plan_t p) { SB_TRY(delete reinterpret_cast<Plan*>(p)) }

int sb_graph_create(sb_ctx_t c, const sb_plan_t* plans, int nplans, sb_graph_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!plans || !out) fail(SB_E_ARGUMENT, "ArgumentError: null argument");
    if (nplans < 1) fail(SB_E_ARGUMENT, "ArgumentError: a plan chain needs at least one plan");
    std::vector<const Plan*> v(nplans);
    for (int i = 0; i < nplans; ++i) v[i] = reinterpret_cast<const Plan*>(plans[i]);
    *out = reinterpret_cast<sb_graph_t>(new ChainGraph(ctx, v));
  })
}
int sb_batch_graph_create(sb_ctx_t c, sb_batch_t b, sb_graph_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!b || !out) fail(SB_E_ARGUMENT, "ArgumentError: null argument");
    Batch* bt = reinterpret_cast<Batch*>(b);
    if (bt->ctx != ctx) fail(SB_E_ARGUMENT, "ArgumentError: the batch belongs to another context");
    *out = reinterpret_cast<sb_graph_t>(batch_graph(ctx, *bt).release());
  })
}
int sb_graph_launch(sb_ctx_t c, sb_graph_t g) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!g) fail(SB_E_HANDLE, "null graph");
    reinterpret_cast<ChainGraph*>(g)->launch();
    (void)ctx;
  })
}
int sb_graph_stats(sb_graph_t g, int64_t* eager, int64_t* captures, int64_t* replays, int64_t* nodes) {
  SB_TRY({
    if (!g) fail(SB_E_HANDLE, "null graph");
    ChainGraph* cg = reinterpret_cast<ChainGraph*>(g);
    if (eager) *eager = cg->eager;
    if (captures) *captures = cg->captures;
    if (replays) *replays = cg->replays;
    if (nodes) *nodes = (int64_t)cg->nodes();
  })
}
int sb_graph_free(sb_graph_t g) {
  SB_TRY({ delete reinterpret_cast<ChainGraph*>(g); })
}
int sb_ctx_graphs(sb_ctx_t c, int mode) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (mode < -1 || mode > 1) fail(SB_E_ARGUMENT, "ArgumentError: graph mode must be -1, 0 or 1");
    ctx->graphs = mode;
  })
}

int sb_batch_create(sb_ctx_t c, const sb_plan_t* plans, int nchains, int nstages, sb_batch_t* out) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!plans || !out) fail(SB_E_ARGUMENT, "ArgumentError: null argument");
    if (nchains < 1 || nstages < 1) fail(SB_E_ARGUMENT, "ArgumentError: a batch needs at least one chain and one stage");
    std::vector<Plan*> v((size_t)nchains * nstages);
    for (size_t i = 0; i < v.size(); ++i) v[i] = reinterpret_cast<Plan*>(plans[i]);
    *out = reinterpret_cast<sb_batch_t>(make_batch(ctx, v, nchains, nstages).release());
  })
}
int sb_batch_execute(sb_ctx_t c, sb_batch_t b, const int* active, int nactive) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!b) fail(SB_E_HANDLE, "null batch");
    Batch* bt = reinterpret_cast<Batch*>(b);
    if (bt->ctx != ctx) fail(SB_E_ARGUMENT, "ArgumentError: the batch belongs to another context");
    if (nactive < 0 || (nactive > 0 && !active)) fail(SB_E_ARGUMENT, "ArgumentError: active chains");
    bt->run(std::vector<int>(active, active + (active ? nactive : 0)));
  })
}
int sb_batch_stats(sb_batch_t b, double* flops, int* fused_stages) {
  SB_TRY({
    if (!b) fail(SB_E_HANDLE, "null batch");
    Batch* bt = reinterpret_cast<Batch*>(b);
    if (flops) *flops = bt->flops;
    if (fused_stages) {
      int f = 0;
      for (auto& s : bt->stages) f += s.fused;
      *fused_stages = f;
    }
  })
}
int sb_batch_autotune(sb_ctx_t c, sb_batch_t b) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!b) fail(SB_E_HANDLE, "null batch");
    Batch* bt = reinterpret_cast<Batch*>(b);
    if (bt->ctx != ctx) fail(SB_E_ARGUMENT, "ArgumentError: the batch belongs to another context");
    for (int s = 0; s < bt->nstages; ++s)
      if (bt->stages[s].fused) bt->tune_stage(s);
  })
}
int sb_batch_stage_tile(sb_batch_t b, int stage, int* tile, double* tuned_ms) {
  SB_TRY({
    if (!b) fail(SB_E_HANDLE, "null batch");
    Batch* bt = reinterpret_cast<Batch*>(b);
    if (stage < 0 || stage >= bt->nstages) fail(SB_E_BOUNDS, "BoundsError: stage out of range");
    const BatchStage& st = bt->stages[stage];
    if (tile) *tile = st.fused ? st.cfg : -1;
    if (tuned_ms) *tuned_ms = st.tuned_ms;
  })
}
int sb_batch_free(sb_batch_t b) { SB_TRY(delete reinterpret_cast<Batch*>(b)) }

int sb_bs_norm(sb_ctx_t c, sb_tensor_t A, double* out) { SB_TRY(*out = bs_norm(C_(c), *T_(A, BS))) }
int sb_bs_copy(sb_ctx_t c, sb_tensor_t A, sb_tensor_t* out) { SB_TRY(*out = H_(bs_copy(C_(c), *T_(A, BS)))) }
int sb_bs_similar(sb_ctx_t c, sb_tensor_t A, sb_tensor_t* out) { SB_TRY(*out = H_(bs_similar(C_(c), *T_(A, BS)))) }
int sb_bs_scale(sb_ctx_t c, sb_tensor_t A, double alpha, sb_tensor_t* out) {
  SB_TRY(*out = H_(bs_scale(C_(c), *T_(A, BS), alpha)))
}
int sb_bs_fill(sb_ctx_t c, sb_tensor_t A, double v) { SB_TRY(bs_fill(C_(c), *T_(A, BS), v)) }
int sb_bs_axpby(sb_ctx_t c, double alpha, sb_tensor_t A, double beta, sb_tensor_t R) {
  SB_TRY(bs_axpby(C_(c), alpha, *T_(A, BS), beta, *T_(R, BS)))
}
int sb_bs_add(sb_ctx_t c, sb_tensor_t A, sb_tensor_t B, sb_tensor_t* out) {
  SB_TRY(*out = H_(bs_add(C_(c), *T_(A, BS), *T_(B, BS))))
}
int sb_bs_dot(sb_ctx_t c, sb_tensor_t A, sb_tensor_t B, double* out) {
  SB_TRY(*out = bs_dot(C_(c), *T_(A, BS), *T_(B, BS)))
}
int sb_bs_restrict(sb_ctx_t c, sb_tensor_t A, sb_tensor_t B, sb_tensor_t* out) {
  SB_TRY(*out = H_(bs_restrict(C_(c), *T_(A, BS), *T_(B, BS))))
}
int sb_bs_norm_noalloc(sb_ctx_t c, sb_tensor_t A, double* out_device) {
  SB_TRY(bs_norm_noalloc(C_(c), *T_(A), out_device))
}

int sb_eigsolve_lanczos(sb_ctx_t c, const sb_plan_t* plans, int nplans, sb_tensor_t x, sb_tensor_t y, int krylovdim,
                        int maxiter, double tol, double* value, double* ritz, int* nritz, double* normres,
                        int64_t* numops, int* numiter, int* converged) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (nplans < 1 || !plans) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve needs at least one plan");
    std::vector<const Plan*> pl(nplans);
    for (int i = 0; i < nplans; ++i) pl[i] = reinterpret_cast<const Plan*>(plans[i]);
    EigResult r = eigsolve_lanczos(ctx, pl, *T_(x, BS), *T_(y, BS), krylovdim, maxiter, tol);
    if (value) *value = r.value;
    if (ritz)
      for (size_t i = 0; i < r.ritz.size(); ++i) ritz[i] = r.ritz[i];
    if (nritz) *nritz = (int)r.ritz.size();
    if (normres) *normres = r.normres;
    if (numops) *numops = r.numops;
    if (numiter) *numiter = r.numiter;
    if (converged) *converged = r.converged ? 1 : 0;
  })
}

int sb_eigsolve_lanczos_batch(sb_ctx_t c, sb_batch_t b, const sb_tensor_t* xs, const sb_tensor_t* ys, int nchains,
                              int krylovdim, int maxiter, double tol, double* values, double* normres,
                              int64_t* numops, int* numiter, int* converged) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!b) fail(SB_E_HANDLE, "null batch");
    Batch* bt = reinterpret_cast<Batch*>(b);
    if (bt->ctx != ctx) fail(SB_E_ARGUMENT, "ArgumentError: the batch belongs to another context");
    if (!xs || !ys || nchains != bt->nchains)
      fail(SB_E_ARGUMENT, "ArgumentError: one vector x and one output y per chain of the batch");
    std::vector<Tensor*> vx(nchains), vy(nchains);
    for (int i = 0; i < nchains; ++i) {
      vx[i] = T_(xs[i], BS);
      vy[i] = T_(ys[i], BS);
    }
    auto r = eigsolve_lanczos_batch(ctx, *bt, vx, vy, krylovdim, maxiter, tol);
    for (int i = 0; i < nchains; ++i) {
      if (values) values[i] = r[i].value;
      if (normres) normres[i] = r[i].normres;
      if (numops) numops[i] = r[i].numops;
      if (numiter) numiter[i] = r[i].numiter;
      if (converged) converged[i] = r[i].converged ? 1 : 0;
    }
  })
}

int sb_eigsolve_lanczos_multi(sb_ctx_t c, const sb_plan_t* plans, int nchains, int nstages, const sb_tensor_t* xs,
                              const sb_tensor_t* ys, int krylovdim, int maxiter, double tol, double* values,
                              double* normres, int64_t* numops, int* numiter, int* converged) {
  SB_TRY({
    Ctx* ctx = C_(c);
    if (!plans || !xs || !ys) fail(SB_E_ARGUMENT, "ArgumentError: null argument");
    if (nchains < 1 || nstages < 1) fail(SB_E_ARGUMENT, "ArgumentError: a batch needs at least one chain and one stage");
    std::vector<Plan*> v((size_t)nchains * nstages);
    std::vector<Ctx*> guests;
    for (size_t i = 0; i < v.size(); ++i) {
      v[i] = reinterpret_cast<Plan*>(plans[i]);
      if (!v[i]) fail(SB_E_HANDLE, "null plan");
      guests.push_back(v[i]->ctx);
    }
    std::vector<Tensor*> vx(nchains), vy(nchains);
    for (int i = 0; i < nchains; ++i) {
      vx[i] = T_(xs[i], BS);
      vy[i] = T_(ys[i], BS);
      guests.push_back(vx[i]->ctx);
      guests.push_back(vy[i]->ctx);
    }
    std::vector<EigResult> r;
    {
      StreamLoan loan(ctx, guests);
      auto bt = make_batch(ctx, v, nchains, nstages);
      r = eigsolve_lanczos_batch(ctx, *bt, vx, vy, krylovdim, maxiter, tol);
    }
    for (int i = 0; i < nchains; ++i) {
      if (values) values[i] = r[i].value;
      if (normres) normres[i] = r[i].normres;
      if (numops) numops[i] = r[i].numops;
      if (numiter) numiter[i] = r[i].numiter;
      if (converged) converged[i] = r[i].converged ? 1 : 0;
    }
  })
}

int sb_bs_svd(sb_ctx_t c, sb_tensor_t A, int ortho_right, int64_t maxdim, int64_t mindim, double cutoff,
              sb_tensor_t* U, sb_tensor_t* SV, double* svs, int64_t svs_cap, int64_t* newdim, double* truncerr) {
  SB_TRY({
    if (!U || !SV) fail(SB_E_ARGUMENT, "ArgumentError: null output handle");
    SvdResult r = bs_svd(C_(c), *T_(A, BS), ortho_right != 0, maxdim, mindim, cutoff);
    if (svs)
      for (int64_t i = 0; i < std::min<int64_t>(svs_cap, (int64_t)r.svs.size()); ++i) svs[i] = r.svs[i];
    if (newdim) *newdim = (int64_t)r.svs.size();
    if (truncerr) *truncerr = r.truncerr;
    *U = H_(std::move(r.U));
    *SV = H_(std::move(r.SV));
  })
}

int sb_bs_left_qr(sb_ctx_t c, sb_tensor_t A, sb_tensor_t* Q, sb_tensor_t* R) {
  SB_TRY({
    if (!Q || !R) fail(SB_E_ARGUMENT, "ArgumentError: null output handle");
    LeftQrResult r = bs_left_qr(C_(c), *T_(A));
    *Q = H_(std::move(r.Q));
    *R = H_(std::move(r.R));
  })
}
int sb_bs_left_factorize(sb_ctx_t c, sb_tensor_t A, sb_tensor_t* Q, sb_tensor_t* R) {
  SB_TRY({
    if (!Q || !R) fail(SB_E_ARGUMENT, "ArgumentError: null output handle");
    LeftQrResult r = bs_left_factorize(C_(c), *T_(A));
    *Q = H_(std::move(r.Q));
    *R = H_(std::move(r.R));
  })
}
int sb_bs_qr_sectors(sb_ctx_t c, sb_tensor_t A, const int64_t* rowsec, const int64_t* dl, const int64_t* dr,
                     sb_tensor_t* Q, sb_tensor_t* R, int64_t* kept) {
  SB_TRY({
    if (!Q || !R) fail(SB_E_ARGUMENT, "ArgumentError: null output handle");
    if (!rowsec || !dl || !dr) fail(SB_E_ARGUMENT, "ArgumentError: null sector table");
    Tensor* t = T_(A);
    SectQrResult r = bs_qr_sectors(C_(c), *t, rowsec, dl, dr);
    if (kept)
      for (size_t i = 0; i < r.kept.size(); ++i) kept[i] = r.kept[i];
    *Q = H_(std::move(r.Q));
    *R = H_(std::move(r.R));
  })
}

int sb_bs_svd_sectors(sb_ctx_t c, sb_tensor_t psi, int64_t nsec, const int64_t* rowsec, const int64_t* colsec,
                      const int64_t* dl, const int64_t* dr, int ortho_right, int64_t maxdim, int64_t mindim,
                      double cutoff, sb_tensor_t* A, sb_tensor_t* B, int64_t* kept, double* svs, int64_t svs_cap,
                      int64_t* newdim, double* truncerr) {
  SB_TRY({
    if (!A || !B) fail(SB_E_ARGUMENT, "ArgumentError: null output handle");
    Tensor* t = T_(psi, BS);
    if (t->N() != 6) fail(SB_E_METHOD, "MethodError: the sector SVD needs psi [Ql, Qr | s1, s2, l, r]");
    if (!rowsec || !colsec || !dl || !dr) fail(SB_E_ARGUMENT, "ArgumentError: null sector table");
    SectSvdResult r = bs_svd_sectors(C_(c), *t, nsec, rowsec, colsec, dl, dr, ortho_right != 0, maxdim, mindim, cutoff);
    if (kept)
      for (int64_t i = 0; i < nsec; ++i) kept[i] = r.kept[i];
    if (svs)
      for (int64_t i = 0; i < std::min<int64_t>(svs_cap, (int64_t)r.svs.size()); ++i) svs[i] = r.svs[i];
    if (newdim) *newdim = (int64_t)r.svs.size();
    if (truncerr) *truncerr = r.truncerr;
    *A = H_(std::move(r.A));
    *B = H_(std::move(r.B));
  })
}

}  // extern "C"
