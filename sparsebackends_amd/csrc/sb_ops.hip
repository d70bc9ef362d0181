// Storage operations of the three storage types on the device:
// NewBlockSparseSorted (SparseBackends/src/blocksparse/storage.jl,
// conversions.jl) and COOTensor (coo/storage.jl, coo/conversions.jl).
#include <cstring>

#include <algorithm>
#include <numeric>

#include <rocprim/rocprim.hpp>

#include "sb_ops.h"

namespace sb {

static inline unsigned nb256(i64 n) { return (unsigned)((n + 255) / 256); }
#define KLAUNCH(kern, n, ...)                                                          \
  do {                                                                                 \
    if ((n) > 0) {                                                                     \
      hipLaunchKernelGGL(kern, dim3(nb256(n)), dim3(256), 0, ctx->stream, __VA_ARGS__); \
      HIP_CHECK(hipGetLastError());                                                    \
    }                                                                                  \
  } while (0)

static void check_dims(const std::vector<i64>& dims) {
  for (auto d : dims) {
    SB_ASSERT(d >= 1, "dims .>= 1");
    if (d >= (1ll << 31)) fail(SB_E_LIMIT, "dimension exceeds 2^31-1");
  }
  if ((int)dims.size() > 16) fail(SB_E_LIMIT, "rank above 16");
}

// the inner constructor: any 0 <= N2 <= N (blocksparse_from_dense and
// permutedims build tensors this way, conversions.jl:74, storage.jl:263)
TP bs_make(Ctx* ctx, const std::vector<i64>& dims, int N2) {
  int N = (int)dims.size();
  SB_ASSERT(0 <= N2 && N2 <= N, "0 <= N2 <= N");
  check_dims(dims);
  if (N - N2 > MAXR || N2 > MAXR) fail(SB_E_LIMIT, "prefix or dense-tail rank above 8");
  TP t(new Tensor);
  t->ctx = ctx;
  t->kind = BS;
  t->dims = dims;
  t->N2 = N2;
  t->blksize = prod(dims, N - N2, N);
  t->n = 0;
  return t;
}

// NewBlockSparseSorted{T,N,N2}(dims) (storage.jl:86-93)
TP bs_new(Ctx* ctx, const std::vector<i64>& dims, int N2) {
  int N = (int)dims.size();
  SB_ASSERT((N == 0 && N2 == 0) || (0 < N2 && N2 <= N), "(N == 0 && N2 == 0) || (0 < N2 <= N)");
  return bs_make(ctx, dims, N2);
}

// AliasedBlockSparse{T,N,N2}(dims) (aliased/storage.jl:122-132)
TP al_new(Ctx* ctx, const std::vector<i64>& dims, int N2) {
  int N = (int)dims.size();
  SB_ASSERT(0 < N2 && N2 <= N, "0 < N2 <= N");
  for (auto d : dims) SB_ASSERT(d >= 1, "all(dims .>= 1)");
  if (N - N2 > MAXR || N2 > MAXR) fail(SB_E_LIMIT, "prefix or dense-tail rank above 8");
  TP t(new Tensor);
  t->ctx = ctx;
  t->kind = ALIASED;
  t->dims = dims;
  t->N2 = N2;
  t->blksize = prod(dims, N - N2, N);
  t->n = 0;
  t->ntemplates = 0;
  return t;
}

TP dense_new(Ctx* ctx, const std::vector<i64>& dims) {
  check_dims(dims);
  TP t(new Tensor);
  t->ctx = ctx;
  t->kind = DENSE;
  t->dims = dims;
  t->N2 = (int)dims.size();
  t->blksize = t->numel();
  t->n = 1;
  t->data.alloc(ctx, t->numel());
  return t;
}

TP coo_new(Ctx* ctx, const std::vector<i64>& dims) {
  check_dims(dims);
  if ((int)dims.size() > MAXR) fail(SB_E_LIMIT, "COO rank above 8");
  TP t(new Tensor);
  t->ctx = ctx;
  t->kind = COO;
  t->dims = dims;
  t->N2 = 0;
  t->n = 0;
  return t;
}

DBuf<i64> block_offsets(Ctx* ctx, const Tensor& A) {
  DBuf<i64> off(ctx, A.n ? A.n : 1);
  k_block_offsets(ctx, A.ids.get(), A.n, A.blksize, off.get());
  return off;
}

// ---------------------------------------------------------------- permutes
void check_no_cross_perm(const std::vector<i64>& perm, int N, int P) {
  SB_ASSERT((int)perm.size() == N, "length(perm) == N");
  std::vector<i64> s(perm);
  std::sort(s.begin(), s.end());
  for (int i = 0; i < N; ++i) SB_ASSERT(s[i] == i + 1, "perm must be a permutation of 1:N");
  for (int i = 0; i < P; ++i)
    SB_ASSERT(perm[i] <= P, "No-cross reorder violated: output prefix axis uses old block axis");
  for (int i = P; i < N; ++i)
    SB_ASSERT(perm[i] > P, "No-cross reorder violated: output block axis uses old prefix axis");
}

TP bs_permute(Ctx* ctx, const Tensor& A, const std::vector<i64>& perm) {
  int N = A.N(), P = A.P(), N2 = A.N2;
  check_no_cross_perm(perm, N, P);
  std::vector<i64> dimsB(N);
  for (int j = 0; j < N; ++j) dimsB[j] = A.dims[perm[j] - 1];
  TP B = bs_make(ctx, dimsB, N2);
  i64 n = A.n;
  B->n = n;
  if (n == 0) return B;
  std::vector<int> sel(P), pblk(N2);
  bool prefix_id = true;
  for (int j = 0; j < P; ++j) {
    sel[j] = (int)perm[j];
    prefix_id = prefix_id && perm[j] == j + 1;
  }
  for (int i = 0; i < N2; ++i) pblk[i] = (int)(perm[P + i] - P - 1);
  B->keys.alloc(ctx, (size_t)n * (P ? P : 1));
  B->ids.alloc(ctx, n);
  B->data.alloc(ctx, (size_t)n * B->blksize);
  k_iota(ctx, B->ids.get(), n, 1);
  DBuf<i64> src(ctx, n);
  if (P > 0) {
    DBuf<i32> nk(ctx, (size_t)n * P), idx(ctx, n);
    DBuf<u64> lin(ctx, n);
    k_remap_keys(ctx, A.keys.get(), n, P, sel.data(), P, dimsB.data(), nk.get(), lin.get(), idx.get());
    // stable sortperm by the new column-major lin (storage.jl:207)
    u64 maxlin = (u64)prod(dimsB, 0, P) - 1;
    if (!(prefix_id && A.dirty == false)) sort_pairs(ctx, lin.get(), idx.get(), n, bits_for(maxlin));
    k_gather_rows(ctx, nk.get(), P, idx.get(), n, B->keys.get());
    k_src_offsets(ctx, A.ids.get(), idx.get(), n, A.blksize, src.get());
  } else {
    k_block_offsets(ctx, A.ids.get(), n, A.blksize, src.get());
  }
  std::vector<i64> tdims(A.dims.begin() + P, A.dims.end());
  permute_blocks(ctx, A.data.get(), src.get(), n, N2, tdims.data(), pblk.data(), B->data.get());
  return B;
}

void bs_permute_inplace(Ctx* ctx, Tensor& A, const std::vector<i64>& perm) {
  int N = A.N(), P = A.P();
  check_no_cross_perm(perm, N, P);
  // packed ids required (storage.jl:281)
  if (A.n > 0) {
    std::vector<i32> ids(A.n);
    HIP_CHECK(hipMemcpyAsync(ids.data(), A.ids.get(), A.n * sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (i64 i = 0; i < A.n; ++i) SB_ASSERT(ids[i] == i + 1, "permutedims! requires packed ids (ids[i]==i)");
  }
  bool prefix_id = true;
  for (int i = 0; i < P; ++i) prefix_id = prefix_id && perm[i] == i + 1;
  TP B = bs_permute(ctx, A, perm);
  if (prefix_id) {  // dense-tail-only: keys untouched (storage.jl:337-341)
    A.dims = B->dims;
    A.data = std::move(B->data);
    return;
  }
  A.dims = B->dims;
  A.keys = std::move(B->keys);
  A.ids = std::move(B->ids);
  A.data = std::move(B->data);
}

TP dense_permute(Ctx* ctx, const Tensor& Bt, const std::vector<i64>& perm) {
  int N = Bt.N();
  std::vector<i64> nd(N);
  std::vector<int> p0(N);
  for (int j = 0; j < N; ++j) {
    nd[j] = Bt.dims[perm[j] - 1];
    p0[j] = (int)perm[j] - 1;
  }
  TP R = dense_new(ctx, nd);
  if (N > MAXR) fail(SB_E_LIMIT, "dense permute of rank above 8");
  permute_blocks(ctx, Bt.dptr(), nullptr, 1, N, Bt.dims.data(), p0.data(), R->data.get());
  return R;
}

// ---------------------------------------------------------------- COO
__global__ void coo_remap_kernel(const i32* in, i64 n, int N, IArr perm, i32* out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * N) return;
  i64 i = t / N;
  int j = (int)(t - i * N);
  out[t] = in[i * N + perm.v[j] - 1];
}

static void coo_check_perm(const std::vector<i64>& perm, int N) {
  SB_ASSERT((int)perm.size() == N, "length(perm) == N");
  std::vector<bool> seen(N + 1, false);
  for (int j = 0; j < N; ++j) {
    i64 p = perm[j];
    if (!(1 <= p && p <= N)) fail(SB_E_ARGUMENT, "perm out of range");
    if (seen[p]) fail(SB_E_ARGUMENT, "perm has duplicates");
    seen[p] = true;
  }
}

TP coo_permute(Ctx* ctx, const Tensor& A, const std::vector<i64>& perm) {
  int N = A.N();
  coo_check_perm(perm, N);
  std::vector<i64> nd(N);
  IArr pa{};
  for (int j = 0; j < N; ++j) {
    nd[j] = A.dims[perm[j] - 1];
    pa.v[j] = perm[j];
  }
  TP B = coo_new(ctx, nd);
  B->n = A.n;
  B->dirty = true;
  if (A.n == 0) return B;
  B->keys.alloc(ctx, (size_t)A.n * N);
  B->data.alloc(ctx, A.n);
  KLAUNCH(coo_remap_kernel, A.n * N, A.keys.get(), A.n, N, pa, B->keys.get());
  HIP_CHECK(hipMemcpyAsync(B->data.get(), A.data.get(), A.n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return B;
}

__global__ void coo_col_kernel(const i32* keys, i64 n, int N, int col, unsigned* out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (unsigned)(keys[i * N + col] - 1);
}
__global__ void gather_f64_kernel(const double* in, const i32* idx, i64 n, double* out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

// stable sort of COO entries by the column-major order of their coordinates
// (ColMajorOrder, coo/storage.jl:34-46): one 64-bit radix sort on the lin
// when prod(dims) fits, else LSD passes axis by axis
static void coo_sort_perm(Ctx* ctx, const Tensor& A, DBuf<i32>& idx) {
  int N = A.N();
  i64 n = A.n;
  idx.alloc(ctx, n);
  k_iota(ctx, idx.get(), n, 0);
  long double tot = 1;
  for (auto d : A.dims) tot *= (long double)d;
  if (tot < 9.0e18L) {
    std::vector<int> cols(N);
    std::iota(cols.begin(), cols.end(), 1);
    DBuf<u64> lin(ctx, n);
    k_cols_lin(ctx, A.keys.get(), n, N, cols.data(), N, A.dims.data(), lin.get());
    sort_pairs(ctx, lin.get(), idx.get(), n, bits_for((u64)(tot - 1)));
    return;
  }
  DBuf<unsigned> kc(ctx, n);
  DBuf<i32> rows(ctx, (size_t)n * N);
  for (int d = 0; d < N; ++d) {
    // current order's column d
    k_gather_rows(ctx, A.keys.get(), N, idx.get(), n, rows.get());
    KLAUNCH(coo_col_kernel, n, rows.get(), n, N, d, kc.get());
    sort_pairs32(ctx, kc.get(), idx.get(), n, bits_for((u64)A.dims[d]));
  }
}

void coo_sort(Ctx* ctx, Tensor& A) {
  if (!A.dirty) return;
  int N = A.N();
  i64 n = A.n;
  if (n > 1) {
    DBuf<i32> idx;
    coo_sort_perm(ctx, A, idx);
    DBuf<i32> k2(ctx, (size_t)n * N);
    DBuf<double> v2(ctx, n);
    k_gather_rows(ctx, A.keys.get(), N, idx.get(), n, k2.get());
    KLAUNCH(gather_f64_kernel, n, A.data.get(), idx.get(), n, v2.get());
    A.keys = std::move(k2);
    A.data = std::move(v2);
  }
  A.dirty = false;
}

TP coo_permute_sorted(Ctx* ctx, const Tensor& A, const std::vector<i64>& perm, DBuf<i32>& idx) {
  TP B = coo_permute(ctx, A, perm);
  int N = B->N();
  i64 n = B->n;
  if (n == 0) {
    B->dirty = false;
    return B;
  }
  coo_sort_perm(ctx, *B, idx);
  DBuf<i32> k2(ctx, (size_t)n * N);
  k_gather_rows(ctx, B->keys.get(), N, idx.get(), n, k2.get());
  B->keys = std::move(k2);
  B->dirty = false;
  return B;
}

// ---------------------------------------------------------------- BS sort! (lexicographic)
void bs_sort_lex(Ctx* ctx, Tensor& A) {
  int P = A.P();
  i64 n = A.n;
  if (n <= 1 || P == 0) return;
  // lexicographic = column-major lin of the reversed key
  std::vector<int> cols(P);
  std::vector<i64> cd(P);
  for (int j = 0; j < P; ++j) {
    cols[j] = P - j;
    cd[j] = A.dims[P - 1 - j];
  }
  DBuf<u64> lin(ctx, n);
  DBuf<i32> idx(ctx, n);
  k_iota(ctx, idx.get(), n, 0);
  k_cols_lin(ctx, A.keys.get(), n, P, cols.data(), P, cd.data(), lin.get());
  sort_pairs(ctx, lin.get(), idx.get(), n, bits_for((u64)prod(A.dims, 0, P)));
  DBuf<i32> k2(ctx, (size_t)n * P), id2(ctx, n);
  k_gather_rows(ctx, A.keys.get(), P, idx.get(), n, k2.get());
  k_gather_rows(ctx, A.ids.get(), 1, idx.get(), n, id2.get());
  A.keys = std::move(k2);
  A.ids = std::move(id2);
  A.dirty = true;  // keys no longer in column-major order
}

bool bs_is_dense(const Tensor& A) {
  int P = A.P();
  if (P == 0) {
    SB_ASSERT(A.n <= 1, "P=0 means no prefix, so at most one block (the whole tensor) is allowed");
    return true;
  }
  return A.n == prod(A.dims, 0, P);
}

// ---------------------------------------------------------------- from/to dense
__global__ void block_max_kernel(const double* D, i64 np, i64 bs, double* mx) {
  i64 pl = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (pl >= np) return;
  double m = 0.0;
  for (i64 s = 0; s < bs; ++s) {
    double a = fabs(D[pl + np * s]);
    if (a > m) m = a;
  }
  mx[pl] = m;
}
__global__ void keep_kernel(const double* mx, i64 np, double tol, int dropzeros, unsigned char* keep, i32* iota) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  keep[i] = (!dropzeros || mx[i] > tol) ? 1 : 0;
  iota[i] = (i32)i;
}
__global__ void decode_keys_kernel(const i32* pl, i64 n, int P, IArr pd, i32* keys) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  i64 r = pl[i];
  for (int d = 0; d < P; ++d) {
    keys[i * P + d] = (i32)(r % pd.v[d]) + 1;
    r /= pd.v[d];
  }
}
__global__ void gather_dense_blocks_kernel(const double* D, const i32* pl, i64 n, i64 np, i64 bs, double* out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * bs) return;
  i64 o = t / bs, s = t - o * bs;
  out[t] = D[pl[o] + np * s];
}

TP bs_from_dense(Ctx* ctx, const Tensor& D, int N2, double atol, double rtol, bool dropzeros) {
  int N = D.N();
  SB_ASSERT(0 <= N2 && N2 <= N, "0 <= N2 <= N");
  int P = N - N2;
  TP B = bs_make(ctx, D.dims, N2);
  i64 np = prod(D.dims, 0, P), bs = B->blksize;
  double gmax = 0.0;
  if (rtol != 0) {
    DBuf<double> r(ctx, 1);
    size_t tb = 0;
    i64 tot = D.numel();
    HIP_CHECK(rocprim::reduce(nullptr, tb, D.dptr(), r.get(), 0.0, (size_t)tot,
                              [] __device__(double a, double b) { return fmax(fabs(a), fabs(b)); }, ctx->stream));
    DBuf<unsigned char> tmp(ctx, tb ? tb : 1);
    HIP_CHECK(rocprim::reduce(tmp.get(), tb, D.dptr(), r.get(), 0.0, (size_t)tot,
                              [] __device__(double a, double b) { return fmax(fabs(a), fabs(b)); }, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(&gmax, r.get(), sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  double tol = std::max(atol, rtol * gmax);
  DBuf<double> mx(ctx, np);
  DBuf<unsigned char> keep(ctx, np);
  DBuf<i32> pl(ctx, np), none;
  KLAUNCH(block_max_kernel, np, D.dptr(), np, bs, mx.get());
  KLAUNCH(keep_kernel, np, mx.get(), np, tol, dropzeros ? 1 : 0, keep.get(), pl.get());
  i64 n = compact_pairs(ctx, keep.get(), pl, none, np);
  B->n = n;
  if (n == 0) return B;
  IArr pd{};
  for (int d = 0; d < P; ++d) pd.v[d] = D.dims[d];
  B->keys.alloc(ctx, (size_t)n * (P ? P : 1));
  B->ids.alloc(ctx, n);
  B->data.alloc(ctx, (size_t)n * bs);
  if (P > 0) KLAUNCH(decode_keys_kernel, n, pl.get(), n, P, pd, B->keys.get());
  k_iota(ctx, B->ids.get(), n, 1);
  KLAUNCH(gather_dense_blocks_kernel, n * bs, D.dptr(), pl.get(), n, np, bs, B->data.get());
  return B;
}

struct Strides16 {
  i64 v[16];
};
__global__ void bs_scatter_kernel(const i32* keys, const i32* ids, i64 n, int P, int N2, i64 bs, IArr tdims,
                                  Strides16 ostr, const double* data, double* out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * bs) return;
  i64 i = t / bs, s = t - i * bs;
  i64 off = 0;
  for (int d = 0; d < P; ++d) off += (i64)(keys[i * P + d] - 1) * ostr.v[d];
  i64 r = s;
  for (int d = 0; d < N2; ++d) {
    i64 c = r % tdims.v[d];
    r /= tdims.v[d];
    off += c * ostr.v[P + d];
  }
  out[off] = data[(i64)(ids[i] - 1) * bs + s];
}

// to_dense with merged axes (conversions.jl:83-187): the merge plan is a
// reshape, so the output offset is linear in the original coordinates; the
// strides are derived on the host by replaying the plan.
TP bs_to_dense(Ctx* ctx, const Tensor& B, double init, const std::vector<std::pair<i64, i64>>& groups) {
  int N = B.N(), P = B.P();
  if (N > 16) fail(SB_E_LIMIT, "rank above 16");
  std::vector<std::pair<i64, i64>> plan;
  std::vector<i64> out_dims = B.dims;
  if (!groups.empty()) {
    std::vector<bool> used(N + 1, false);
    for (auto g : groups) {
      i64 a = g.first, b = g.second;
      if (!(1 <= a && a <= N) || !(1 <= b && b <= N)) fail(SB_E_ERROR, "Axis out of range in merge group");
      if (a == b) fail(SB_E_ERROR, "Duplicate axes in merge group");
      i64 keep = std::min(a, b), other = std::max(a, b);
      if (used[keep] || used[other]) fail(SB_E_ERROR, "Axis appears in multiple merge groups");
      used[keep] = used[other] = true;
      plan.push_back({keep, other});
    }
    std::stable_sort(plan.begin(), plan.end(), [](auto& x, auto& y) { return x.second > y.second; });
    std::vector<i64> dw = B.dims;
    for (auto [keep, other] : plan) {
      dw[keep - 1] *= dw[other - 1];
      dw.erase(dw.begin() + (other - 1));
    }
    out_dims = dw;
  }
  // replay: each current axis is a list of (original axis, multiplier)
  std::vector<std::vector<std::pair<int, i64>>> ax(N);
  std::vector<i64> dimw = B.dims;
  for (int d = 0; d < N; ++d) ax[d] = {{d, 1}};
  for (auto [keep, other] : plan) {
    for (auto& e : ax[other - 1]) ax[keep - 1].push_back({e.first, e.second * dimw[keep - 1]});
    dimw[keep - 1] *= dimw[other - 1];
    ax.erase(ax.begin() + (other - 1));
    dimw.erase(dimw.begin() + (other - 1));
  }
  Strides16 os{};
  i64 st = 1;
  for (size_t j = 0; j < ax.size(); ++j) {
    for (auto& e : ax[j]) os.v[e.first] += e.second * st;
    st *= out_dims[j];
  }
  TP D = dense_new(ctx, out_dims);
  k_fill_f64(ctx, D->data.get(), D->numel(), init);
  IArr td{};
  for (int d = 0; d < B.N2; ++d) td.v[d] = B.dims[P + d];
  KLAUNCH(bs_scatter_kernel, B.n * B.blksize, B.keys.get(), B.ids.get(), B.n, P, B.N2, B.blksize, td, os,
          B.data.get(), D->data.get());
  return D;
}

__global__ void coo_keep_kernel(const double* D, i64 n, double atol, double rtol, unsigned char* keep, i32* iota) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = D[i];
  keep[i] = fabs(v) > atol + rtol * fabs(v) ? 1 : 0;
  iota[i] = (i32)i;
}

TP coo_from_dense(Ctx* ctx, const Tensor& D, double atol, double rtol) {
  int N = D.N();
  TP A = coo_new(ctx, D.dims);
  i64 tot = D.numel();
  if (tot >= (1ll << 31)) fail(SB_E_LIMIT, "coo_from_dense of more than 2^31 elements");
  DBuf<unsigned char> keep(ctx, tot);
  DBuf<i32> lin(ctx, tot), none;
  KLAUNCH(coo_keep_kernel, tot, D.dptr(), tot, atol, rtol, keep.get(), lin.get());
  i64 n = compact_pairs(ctx, keep.get(), lin, none, tot);
  A->n = n;
  if (n == 0) return A;
  IArr pd{};
  for (int d = 0; d < N; ++d) pd.v[d] = D.dims[d];
  A->keys.alloc(ctx, (size_t)n * N);
  A->data.alloc(ctx, n);
  KLAUNCH(decode_keys_kernel, n, lin.get(), n, N, pd, A->keys.get());
  KLAUNCH(gather_f64_kernel, n, D.dptr(), lin.get(), n, A->data.get());
  return A;
}

__global__ void coo_winner_kernel(const i32* keys, i64 n, int N, IArr st, int* win) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  i64 off = 0;
  for (int d = 0; d < N; ++d) off += (i64)(keys[i * N + d] - 1) * st.v[d];
  atomicMax(&win[off], (int)i);
}
__global__ void coo_write_kernel(const i32* keys, const double* vals, i64 n, int N, IArr st, const int* win,
                                 double* out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  i64 off = 0;
  for (int d = 0; d < N; ++d) off += (i64)(keys[i * N + d] - 1) * st.v[d];
  if (win[off] == (int)i) out[off] = vals[i];
}
__global__ void fill_i32_kernel(int* p, i64 n, int v) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// to_dense(::COOTensor) (coo/conversions.jl:4-10): for duplicate coordinates
// the last entry wins
TP coo_to_dense(Ctx* ctx, const Tensor& A) {
  int N = A.N();
  TP D = dense_new(ctx, A.dims);
  i64 tot = D->numel();
  k_fill_f64(ctx, D->data.get(), tot, 0.0);
  if (A.n == 0) return D;
  IArr st{};
  i64 s = 1;
  for (int d = 0; d < N; ++d) {
    st.v[d] = s;
    s *= A.dims[d];
  }
  DBuf<int> win(ctx, tot);
  KLAUNCH(fill_i32_kernel, tot, win.get(), tot, -1);
  KLAUNCH(coo_winner_kernel, A.n, A.keys.get(), A.n, N, st, win.get());
  KLAUNCH(coo_write_kernel, A.n, A.keys.get(), A.data.get(), A.n, N, st, win.get(), D->data.get());
  return D;
}

// ---------------------------------------------------------------- axis fusion
__global__ void fuse_keys_kernel(const i32* in, i64 n, int P, int a1, int a2, i64 d1, i32* out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const i32* k = in + i * P;
  i32 fused = (i32)((k[a2 - 1] - 1) * d1 + k[a1 - 1]);
  for (int j = 1; j <= P - 1; ++j) out[i * (P - 1) + j - 1] = j < a2 ? (j == a1 ? fused : k[j - 1]) : k[j];
}

// fuse_two_axes! (conversions.jl:299-314) with fuse_sparse_axes (:228-263,
// keys not re-sorted) and fuse_dense_axes! (:265-297, mutates A)
TP bs_fuse_two_axes(Ctx* ctx, Tensor& A, i64 ax1, i64 ax2) {
  int N = A.N(), P = A.P(), N2 = A.N2;
  if (!(1 <= ax1 && ax1 <= N && 1 <= ax2 && ax2 <= N)) fail(SB_E_ARGUMENT, "axes must be in 1:N");
  if (ax1 == ax2) fail(SB_E_ARGUMENT, "axes must be distinct");
  bool s1 = ax1 <= P, s2 = ax2 <= P;
  i64 a1 = std::min(ax1, ax2), a2 = std::max(ax1, ax2);
  if (s1 && s2) {
    i64 d1 = A.dims[a1 - 1], d2 = A.dims[a2 - 1];
    std::vector<i64> nd;
    for (int i = 1; i <= N - 1; ++i) nd.push_back(i < a2 ? (i == a1 ? d1 * d2 : A.dims[i - 1]) : A.dims[i]);
    TP B = bs_make(ctx, nd, N2);
    B->n = A.n;
    B->dirty = true;  // keys are not re-sorted
    if (A.n) {
      B->keys.alloc(ctx, (size_t)A.n * (P - 1 ? P - 1 : 1));
      B->ids.alloc(ctx, A.n);
      B->data.alloc(ctx, (size_t)A.n * A.blksize);
      if (P > 1) KLAUNCH(fuse_keys_kernel, A.n, A.keys.get(), A.n, P, (int)a1, (int)a2, d1, B->keys.get());
      HIP_CHECK(hipMemcpyAsync(B->ids.get(), A.ids.get(), A.n * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
      HIP_CHECK(hipMemcpyAsync(B->data.get(), A.data.get(), A.n * A.blksize * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    return B;
  }
  if (s1 || s2) fail(SB_E_ARGUMENT, "Cannot fuse across sparse-head and dense-tail axes");
  i64 p1 = a1 - P, p2 = a2 - P;
  if (p2 != p1 + 1) {
    std::vector<i64> pb(N2);
    std::iota(pb.begin(), pb.end(), 1);
    i64 x = pb[p2 - 1];
    pb.erase(pb.begin() + (p2 - 1));
    pb.insert(pb.begin() + p1, x);
    std::vector<i64> perm(N);
    std::iota(perm.begin(), perm.end(), 1);
    for (int i = 0; i < N2; ++i) perm[P + i] = P + pb[i];
    bs_permute_inplace(ctx, A, perm);
    a2 = a1 + 1;
  }
  std::vector<i64> nd;
  for (int i = 1; i <= N - 1; ++i)
    nd.push_back(i < a2 ? (i == a1 ? A.dims[a1 - 1] * A.dims[a2 - 1] : A.dims[i - 1]) : A.dims[i]);
  TP B = bs_make(ctx, nd, N2 - 1);
  B->n = A.n;
  B->dirty = A.dirty;
  if (A.n) {
    B->keys.alloc(ctx, (size_t)A.n * (P ? P : 1));
    B->ids.alloc(ctx, A.n);
    B->data.alloc(ctx, (size_t)A.n * A.blksize);
    if (P)
      HIP_CHECK(hipMemcpyAsync(B->keys.get(), A.keys.get(), A.n * P * sizeof(i32), hipMemcpyDeviceToDevice,
                               ctx->stream));
    HIP_CHECK(hipMemcpyAsync(B->ids.get(), A.ids.get(), A.n * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(B->data.get(), A.data.get(), A.n * A.blksize * sizeof(double), hipMemcpyDeviceToDevice,
                             ctx->stream));
  }
  return B;
}

// fuse_axes_coo (coo/conversions.jl:174-223)
__global__ void coo_fuse_kernel(const i32* in, i64 n, int N, int ax1, int ax2, i64 dkeep, int ax1_out, i32* out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const i32* c = in + i * N;
  i32 fused = (i32)((c[ax2 - 1] - 1) * dkeep + c[ax1 - 1]);
  for (int j = 1; j <= N - 1; ++j) {
    int oa = j < ax2 ? j : j + 1;
    out[i * (N - 1) + j - 1] = j == ax1_out ? fused : c[oa - 1];
  }
}
TP coo_fuse_axes(Ctx* ctx, const Tensor& A, i64 ax1, i64 ax2) {
  int N = A.N();
  if (ax1 == ax2) fail(SB_E_ARGUMENT, "ax1 and ax2 must be distinct");
  if (!(1 <= ax1 && ax1 <= N) || !(1 <= ax2 && ax2 <= N)) fail(SB_E_BOUNDS, "BoundsError");
  SB_ASSERT(std::llabs(ax1 - ax2) == 1, "Currently only adjacent axes can be fused");
  i64 dkeep = A.dims[ax1 - 1], ddrop = A.dims[ax2 - 1];
  i64 ax1_out = ax1 - (ax2 < ax1 ? 1 : 0);
  std::vector<i64> nd(N - 1);
  for (i64 j = 1; j <= N - 1; ++j) nd[j - 1] = (j == ax1_out) ? dkeep * ddrop : A.dims[(j < ax2 ? j : j + 1) - 1];
  TP C = coo_new(ctx, nd);
  C->n = A.n;
  C->dirty = true;
  if (A.n) {
    C->keys.alloc(ctx, (size_t)A.n * (N - 1));
    C->data.alloc(ctx, A.n);
    KLAUNCH(coo_fuse_kernel, A.n, A.keys.get(), A.n, N, (int)ax1, (int)ax2, dkeep, (int)ax1_out, C->keys.get());
    HIP_CHECK(hipMemcpyAsync(C->data.get(), A.data.get(), A.n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  }
  coo_sort(ctx, *C);
  return C;
}

// ---------------------------------------------------------------- element access
// _find_key (storage.jl:49-65) / searchsortedfirst(ColMajorOrder) (coo/storage.jl:49-60)
__global__ void search_kernel(const i32* keys, i64 n, int w, IArr k, IArr pd, int colmajor_full, i64* out) {
  // out[0] = position (0-based lower bound), out[1] = found
  u64 target = 0, st = 1;
  for (int d = 0; d < w; ++d) {
    target += (u64)(k.v[d] - 1) * st;
    st *= (u64)pd.v[d];
  }
  i64 lo = 0, hi = n;
  while (lo < hi) {
    i64 mid = (lo + hi) >> 1;
    u64 v = 0, s = 1;
    for (int d = 0; d < w; ++d) {
      v += (u64)(keys[mid * w + d] - 1) * s;
      s *= (u64)pd.v[d];
    }
    if (v < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  bool found = lo < n;
  for (int d = 0; found && d < w; ++d) found = keys[lo * w + d] == (i32)k.v[d];
  out[0] = lo;
  out[1] = found ? 1 : 0;
  (void)colmajor_full;
}

static std::pair<i64, bool> dev_search(Ctx* ctx, const i32* keys, i64 n, int w, const i64* k, const i64* dims) {
  IArr kk{}, pd{};
  for (int d = 0; d < w; ++d) {
    kk.v[d] = k[d];
    pd.v[d] = dims[d];
  }
  DBuf<i64> r(ctx, 2);
  hipLaunchKernelGGL(search_kernel, dim3(1), dim3(1), 0, ctx->stream, keys, n, w, kk, pd, 1, r.get());
  HIP_CHECK(hipGetLastError());
  i64 h[2];
  HIP_CHECK(hipMemcpyAsync(h, r.get(), sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return {h[0], h[1] != 0};
}

static void check_index(const Tensor& A, const i64* I) {
  for (int d = 0; d < A.N(); ++d)
    if (I[d] < 1 || I[d] > A.dims[d]) fail(SB_E_BOUNDS, "BoundsError");
}

static i64 tail_index(const Tensor& A, const i64* I) {
  int P = A.P();
  i64 lin = 0, st = 1;
  for (int j = 0; j < A.N2; ++j) {
    lin += (I[P + j] - 1) * st;
    st *= A.dims[P + j];
  }
  return lin;
}

static double read1(Ctx* ctx, const double* p) {
  double v;
  HIP_CHECK(hipMemcpyAsync(&v, p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return v;
}
static i32 read1i(Ctx* ctx, const i32* p) {
  i32 v;
  HIP_CHECK(hipMemcpyAsync(&v, p, sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return v;
}

double bs_getindex(Ctx* ctx, const Tensor& A, const i64* I) {
  check_index(A, I);
  int P = A.P();
  if (A.n == 0) return 0.0;
  i64 pos = 0;
  if (P > 0) {
    auto [p, found] = dev_search(ctx, A.keys.get(), A.n, P, I, A.dims.data());
    if (!found) return 0.0;
    pos = p;
  }
  i32 id = read1i(ctx, A.ids.get() + pos);
  return read1(ctx, A.data.get() + (i64)(id - 1) * A.blksize + tail_index(A, I));
}

// getindex(::AliasedBlockSparse) (aliased/storage.jl:138-147)
double al_getindex(Ctx* ctx, const Tensor& A, const i64* I) {
  check_index(A, I);
  int P = A.P();
  if (A.n == 0) return 0.0;
  i64 pos = 0;
  if (P > 0) {
    auto [p, found] = dev_search(ctx, A.keys.get(), A.n, P, I, A.dims.data());
    if (!found) return 0.0;
    pos = p;
  }
  i32 tid = read1i(ctx, A.ids.get() + pos);
  double a = read1(ctx, A.scalars.get() + pos);
  return a * read1(ctx, A.data.get() + (i64)(tid - 1) * A.blksize + tail_index(A, I));
}

// to_blocksparse (aliased/conversions.jl:18-32): ids 1:n, payload scalar * template
TP al_to_blocksparse(Ctx* ctx, const Tensor& A) {
  TP B = bs_make(ctx, A.dims, A.N2);
  i64 n = A.n;
  int P = A.P();
  B->n = n;
  B->keys.alloc(ctx, (size_t)n * (P ? P : 1));
  B->ids.alloc(ctx, n ? n : 1);
  B->data.alloc(ctx, (size_t)n * A.blksize);
  if (n) {
    if (P)
      HIP_CHECK(hipMemcpyAsync(B->keys.get(), A.keys.get(), (size_t)n * P * sizeof(i32), hipMemcpyDeviceToDevice,
                               ctx->stream));
    k_iota(ctx, B->ids.get(), n, 1);
    k_al_expand(ctx, n, A.blksize, A.ids.get(), A.scalars.get(), A.data.get(), B->data.get());
  }
  return B;
}

__global__ void set1_kernel(double* p, double v) { *p = v; }
__global__ void set1i_kernel(i32* p, i32 v) { *p = v; }

// setindex! (storage.jl:114-124) via _ensure_block! (:73-83): a new block is
// appended to the payload and its key/id inserted at the sorted position
void bs_setindex(Ctx* ctx, Tensor& A, double v, const i64* I) {
  check_index(A, I);
  int P = A.P();
  i64 pos = 0;
  bool found = A.n > 0;
  if (P > 0 && A.n > 0) {
    auto r = dev_search(ctx, A.keys.get(), A.n, P, I, A.dims.data());
    pos = r.first;
    found = r.second;
  } else if (P > 0) {
    pos = 0;
    found = false;
  }
  i32 id;
  if (found) {
    id = read1i(ctx, A.ids.get() + pos);
  } else {
    i64 n = A.n, bs = A.blksize;
    id = (i32)(n + 1);
    DBuf<i32> k2(ctx, (size_t)(n + 1) * (P ? P : 1)), id2(ctx, n + 1);
    DBuf<double> d2(ctx, (size_t)(n + 1) * bs);
    if (n) {
      if (P) {
        HIP_CHECK(hipMemcpyAsync(k2.get(), A.keys.get(), pos * P * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(k2.get() + (pos + 1) * P, A.keys.get() + pos * P, (n - pos) * P * sizeof(i32),
                                 hipMemcpyDeviceToDevice, ctx->stream));
      }
      HIP_CHECK(hipMemcpyAsync(id2.get(), A.ids.get(), pos * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
      HIP_CHECK(hipMemcpyAsync(id2.get() + pos + 1, A.ids.get() + pos, (n - pos) * sizeof(i32),
                               hipMemcpyDeviceToDevice, ctx->stream));
      HIP_CHECK(hipMemcpyAsync(d2.get(), A.data.get(), n * bs * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    std::vector<i32> hk(P ? P : 1);
    for (int d = 0; d < P; ++d) hk[d] = (i32)I[d];
    if (P)
      HIP_CHECK(hipMemcpyAsync(k2.get() + pos * P, hk.data(), P * sizeof(i32), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(set1i_kernel, dim3(1), dim3(1), 0, ctx->stream, id2.get() + pos, id);
    k_fill_f64(ctx, d2.get() + n * bs, bs, 0.0);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    A.keys = std::move(k2);
    A.ids = std::move(id2);
    A.data = std::move(d2);
    A.n = n + 1;
  }
  hipLaunchKernelGGL(set1_kernel, dim3(1), dim3(1), 0, ctx->stream, A.data.get() + (i64)(id - 1) * A.blksize + tail_index(A, I), v);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

double coo_getindex(Ctx* ctx, const Tensor& A, const i64* I) {
  check_index(A, I);
  if (A.n == 0) return 0.0;
  auto [pos, found] = dev_search(ctx, A.keys.get(), A.n, A.N(), I, A.dims.data());
  return found ? read1(ctx, A.data.get() + pos) : 0.0;
}

__global__ void coo_match_kernel(const i32* keys, i64 n, int N, IArr k, unsigned char* keepnot, i32* iota) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool eq = true;
  for (int d = 0; d < N; ++d) eq = eq && keys[i * N + d] == (i32)k.v[d];
  keepnot[i] = eq ? 0 : 1;
  iota[i] = (i32)i;
}

// setindex! (coo/storage.jl:84-134)
void coo_setindex(Ctx* ctx, Tensor& A, double vv, const i64* I, bool dirty_kw) {
  int N = A.N();
  check_index(A, I);
  auto append = [&]() {
    i64 n = A.n;
    DBuf<i32> k2(ctx, (size_t)(n + 1) * N);
    DBuf<double> v2(ctx, n + 1);
    if (n) {
      HIP_CHECK(hipMemcpyAsync(k2.get(), A.keys.get(), n * N * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
      HIP_CHECK(hipMemcpyAsync(v2.get(), A.data.get(), n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    std::vector<i32> hk(N);
    for (int d = 0; d < N; ++d) hk[d] = (i32)I[d];
    HIP_CHECK(hipMemcpyAsync(k2.get() + n * N, hk.data(), N * sizeof(i32), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(v2.get() + n, &vv, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    A.keys = std::move(k2);
    A.data = std::move(v2);
    A.n = n + 1;
  };
  auto erase_matching = [&]() {
    i64 n = A.n;
    if (!n) return;
    IArr kk{};
    for (int d = 0; d < N; ++d) kk.v[d] = I[d];
    DBuf<unsigned char> keep(ctx, n);
    DBuf<i32> rows(ctx, n), none;
    KLAUNCH(coo_match_kernel, n, A.keys.get(), n, N, kk, keep.get(), rows.get());
    i64 m = compact_pairs(ctx, keep.get(), rows, none, n);
    DBuf<i32> k2(ctx, (size_t)(m ? m : 1) * N);
    DBuf<double> v2(ctx, m ? m : 1);
    if (m) {
      k_gather_rows(ctx, A.keys.get(), N, rows.get(), m, k2.get());
      KLAUNCH(gather_f64_kernel, m, A.data.get(), rows.get(), m, v2.get());
    }
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    A.keys = std::move(k2);
    A.data = std::move(v2);
    A.n = m;
  };
  if (A.dirty) {
    if (vv == 0.0) {
      erase_matching();
      return;
    }
    append();
    A.dirty = true;
    return;
  }
  i64 pos = 0;
  bool found = false;
  if (A.n) {
    auto r = dev_search(ctx, A.keys.get(), A.n, N, I, A.dims.data());
    pos = r.first;
    found = r.second;
  }
  if (found) {
    if (vv == 0.0) {
      erase_matching();  // keys unique in canonical mode: removes exactly this entry
      A.dirty = true;
    } else {
      hipLaunchKernelGGL(set1_kernel, dim3(1), dim3(1), 0, ctx->stream, A.data.get() + pos, vv);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return;
  }
  if (vv == 0.0) return;
  if (dirty_kw) {
    append();
    A.dirty = true;
    return;
  }
  // insert at the sorted position
  i64 n = A.n;
  DBuf<i32> k2(ctx, (size_t)(n + 1) * N);
  DBuf<double> v2(ctx, n + 1);
  if (n) {
    HIP_CHECK(hipMemcpyAsync(k2.get(), A.keys.get(), pos * N * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(k2.get() + (pos + 1) * N, A.keys.get() + pos * N, (n - pos) * N * sizeof(i32),
                             hipMemcpyDeviceToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(v2.get(), A.data.get(), pos * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(v2.get() + pos + 1, A.data.get() + pos, (n - pos) * sizeof(double),
                             hipMemcpyDeviceToDevice, ctx->stream));
  }
  std::vector<i32> hk(N);
  for (int d = 0; d < N; ++d) hk[d] = (i32)I[d];
  HIP_CHECK(hipMemcpyAsync(k2.get() + pos * N, hk.data(), N * sizeof(i32), hipMemcpyHostToDevice, ctx->stream));
  HIP_CHECK(hipMemcpyAsync(v2.get() + pos, &vv, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  A.keys = std::move(k2);
  A.data = std::move(v2);
  A.n = n + 1;
}

}  // namespace sb
