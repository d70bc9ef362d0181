// Integer kernels of the symbolic phase: key remapping, column-major
// linearisation, stable radix sorts, merge-join ranges, pair emission and
// output-block allocation with the reference's _ensure_block! semantics
// (SparseBackends/src/blocksparse/storage.jl:67-83): keys sorted by the
// column-major lin of the prefix, ids = order of first creation.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "sb_kernels.h"

namespace sb {

static inline unsigned nblk(i64 n, int t = 256) { return (unsigned)((n + t - 1) / t); }

#define LAUNCH(kern, n, ...)                                                          \
  do {                                                                                \
    if ((n) > 0) {                                                                    \
      hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(256), 0, ctx->stream, __VA_ARGS__); \
      HIP_CHECK(hipGetLastError());                                                   \
    }                                                                                 \
  } while (0)

struct SelArr {
  int v[MAXR];
};

__global__ void remap_keys_kernel(const i32* __restrict__ in, i64 n, int Pin, SelArr sel, int Pout, IArr pd,
                                  i32* __restrict__ out, u64* __restrict__ lin, i32* __restrict__ iota) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u64 l = 0, st = 1;
  for (int j = 0; j < Pout; ++j) {
    i32 k = in[i * Pin + sel.v[j] - 1];
    if (out) out[i * Pout + j] = k;
    l += (u64)(k - 1) * st;
    st *= (u64)pd.v[j];
  }
  if (lin) lin[i] = l;
  if (iota) iota[i] = (i32)i;
}

void k_remap_keys(Ctx* ctx, const i32* in_keys, i64 n, int Pin, const int* sel, int Pout, const i64* pdims_out,
                  i32* out_keys, u64* lin, i32* iota) {
  SelArr s{};
  IArr pd{};
  for (int j = 0; j < Pout; ++j) {
    s.v[j] = sel[j];
    pd.v[j] = pdims_out[j];
  }
  LAUNCH(remap_keys_kernel, n, in_keys, n, Pin, s, Pout, pd, out_keys, lin, iota);
}

__global__ void pair_keys_kernel(const i32* __restrict__ a, const i32* __restrict__ b, i64 np,
                                 const i32* __restrict__ kA, int PA, const i32* __restrict__ kB, int PB, SelArr src,
                                 int PC, IArr pd, i32* __restrict__ ck, u64* __restrict__ lin) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  i64 ia = a ? a[p] : 0, ib = b ? b[p] : 0;
  u64 l = 0, st = 1;
  for (int j = 0; j < PC; ++j) {
    int s = src.v[j];
    i32 k = s > 0 ? kA[ia * PA + s - 1] : kB[ib * PB + (-s) - 1];
    ck[p * PC + j] = k;
    l += (u64)(k - 1) * st;
    st *= (u64)pd.v[j];
  }
  lin[p] = l;
}

void k_pair_keys(Ctx* ctx, const i32* a_idx, const i32* b_idx, i64 np, const i32* keysA, int PA, const i32* keysB,
                 int PB, const int* src, int PC, const i64* pdimsC, i32* ck, u64* lin) {
  SelArr s{};
  IArr pd{};
  for (int j = 0; j < PC; ++j) {
    s.v[j] = src[j];
    pd.v[j] = pdimsC[j];
  }
  LAUNCH(pair_keys_kernel, np, a_idx, b_idx, np, keysA, PA, keysB, PB, s, PC, pd, ck, lin);
}

__global__ void cols_lin_kernel(const i32* __restrict__ keys, i64 n, int P, SelArr cols, int nc, IArr cd,
                                u64* __restrict__ lin) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u64 l = 0, st = 1;
  for (int j = 0; j < nc; ++j) {
    l += (u64)(keys[i * P + cols.v[j] - 1] - 1) * st;
    st *= (u64)cd.v[j];
  }
  lin[i] = l;
}

void k_cols_lin(Ctx* ctx, const i32* keys, i64 n, int P, const int* cols, int ncols, const i64* cdims, u64* lin) {
  SelArr s{};
  IArr cd{};
  for (int j = 0; j < ncols; ++j) {
    s.v[j] = cols[j];
    cd.v[j] = cdims[j];
  }
  LAUNCH(cols_lin_kernel, n, keys, n, P, s, ncols, cd, lin);
}

__global__ void gather_rows_kernel(const i32* __restrict__ in, int w, const i32* __restrict__ idx, i64 n,
                                   i32* __restrict__ out) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * w) return;
  i64 i = t / w;
  int j = (int)(t - i * w);
  out[t] = in[(i64)idx[i] * w + j];
}
void k_gather_rows(Ctx* ctx, const i32* in, int w, const i32* idx, i64 n, i32* out) {
  if (w == 0) return;
  LAUNCH(gather_rows_kernel, n * w, in, w, idx, n, out);
}

__global__ void iota_kernel(i32* p, i64 n, i32 base) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = base + (i32)i;
}
void k_iota(Ctx* ctx, i32* p, i64 n, i32 base) { LAUNCH(iota_kernel, n, p, n, base); }

__global__ void fill_kernel(double* p, i64 n, double v) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void k_fill_f64(Ctx* ctx, double* p, i64 n, double v) { LAUNCH(fill_kernel, n, p, n, v); }

__global__ void zero_kernel(unsigned* p, i64 nw) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) p[i] = 0u;
}
__global__ void zero16_kernel(uint4* p, i64 n) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) p[i] = z;
}
void k_zero(Ctx* ctx, void* p, size_t bytes) {
  // a fill kernel rather than hipMemsetAsync, so it can be captured in a graph
  if (!bytes) return;
  SB_ASSERT(bytes % 4 == 0, "k_zero: size must be a multiple of 4 bytes");
  if (bytes % 16 == 0 && reinterpret_cast<uintptr_t>(p) % 16 == 0 && bytes >= (1u << 20)) {
    // large fills: 16-byte stores, a grid of a few waves per CU striding the buffer
    const i64 n = (i64)(bytes / 16);
    const i64 grid = std::min<i64>((n + 255) / 256, (i64)ctx->num_cu * 8);
    hipLaunchKernelGGL(zero16_kernel, dim3((unsigned)grid), dim3(256), 0, ctx->stream, (uint4*)p, n);
    HIP_CHECK(hipGetLastError());
    return;
  }
  i64 nw = (i64)(bytes / 4);
  LAUNCH(zero_kernel, nw, (unsigned*)p, nw);
}

// ---------------------------------------------------------------- rocPRIM
int bits_for(u64 maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b == 0 ? 1 : b;
}

// keys and values back from the sort's alternate buffers: one launch instead
// of two blit copies
template <class K>
__global__ void copy_pairs_kernel(K* __restrict__ kd, const K* __restrict__ ks, i32* __restrict__ vd,
                                  const i32* __restrict__ vs, i64 n) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  kd[i] = ks[i];
  vd[i] = vs[i];
}

void sort_pairs(Ctx* ctx, u64* keys, i32* vals, i64 n, int bits) {
  if (n <= 1) return;
  DBuf<u64> k2(ctx, n);
  DBuf<i32> v2(ctx, n);
  rocprim::double_buffer<u64> kb(keys, k2.get());
  rocprim::double_buffer<i32> vb(vals, v2.get());
  size_t tb = 0;
  HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
  DBuf<unsigned char> tmp(ctx, tb ? tb : 1);
  HIP_CHECK(rocprim::radix_sort_pairs(tmp.get(), tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
  if (kb.current() != keys) {
    hipLaunchKernelGGL(copy_pairs_kernel<u64>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, keys,
                       kb.current(), vals, vb.current(), n);
    HIP_CHECK(hipGetLastError());
  }
}

void sort_pairs32(Ctx* ctx, unsigned* keys, i32* vals, i64 n, int bits) {
  if (n <= 1) return;
  DBuf<unsigned> k2(ctx, n);
  DBuf<i32> v2(ctx, n);
  rocprim::double_buffer<unsigned> kb(keys, k2.get());
  rocprim::double_buffer<i32> vb(vals, v2.get());
  size_t tb = 0;
  HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
  DBuf<unsigned char> tmp(ctx, tb ? tb : 1);
  HIP_CHECK(rocprim::radix_sort_pairs(tmp.get(), tb, kb, vb, (size_t)n, 0, bits, ctx->stream));
  if (kb.current() != keys) {
    hipLaunchKernelGGL(copy_pairs_kernel<unsigned>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       keys, kb.current(), vals, vb.current(), n);
    HIP_CHECK(hipGetLastError());
  }
}

i64 exclusive_scan_total(Ctx* ctx, const i64* in, i64* out, i64 n) {
  if (n == 0) return 0;
  size_t tb = 0;
  HIP_CHECK(rocprim::exclusive_scan(nullptr, tb, in, out, (i64)0, (size_t)n, rocprim::plus<i64>(), ctx->stream));
  DBuf<unsigned char> tmp(ctx, tb ? tb : 1);
  HIP_CHECK(rocprim::exclusive_scan(tmp.get(), tb, in, out, (i64)0, (size_t)n, rocprim::plus<i64>(), ctx->stream));
  i64 last[2];
  HIP_CHECK(hipMemcpyAsync(&last[0], in + n - 1, sizeof(i64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipMemcpyAsync(&last[1], out + n - 1, sizeof(i64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return last[0] + last[1];
}

// ---------------------------------------------------------------- joins
__global__ void equal_ranges_kernel(const u64* __restrict__ ja, i64 nA, const u64* __restrict__ jb, i64 nB,
                                    i32* __restrict__ lo, i32* __restrict__ hi) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nA) return;
  u64 v = ja[i];
  i64 a = 0, b = nB;
  while (a < b) {
    i64 m = (a + b) >> 1;
    if (jb[m] < v)
      a = m + 1;
    else
      b = m;
  }
  i64 l = a;
  b = nB;
  while (a < b) {
    i64 m = (a + b) >> 1;
    if (jb[m] <= v)
      a = m + 1;
    else
      b = m;
  }
  lo[i] = (i32)l;
  hi[i] = (i32)a;
}
void k_equal_ranges(Ctx* ctx, const u64* ja, i64 nA, const u64* jb, i64 nB, i32* lo, i32* hi) {
  LAUNCH(equal_ranges_kernel, nA, ja, nA, jb, nB, lo, hi);
}

// Sequential merge walks.  One thread walks both key lists exactly as the
// reference's while-loops do and writes, for every A row it visits inside a
// matched run, the B range it is paired with.  Rows never paired get an
// empty range.  Used where the reference's walk is not an ordered equi-join
// (unsorted COO input kept by the fast exit of ensure_rlab_last_and_sorted,
// contract_coo.jl:54-58; the multi-label compare order, contract_bs_bs.jl:430-442).
__device__ int walk_cmp(const i32* ka, const i32* kb, SelArr ca, SelArr cb, int nc) {
  for (int t = 0; t < nc; ++t) {
    i32 av = ka[ca.v[t] - 1], bv = kb[cb.v[t] - 1];
    if (av < bv) return -1;
    if (av > bv) return 1;
  }
  return 0;
}
__device__ bool walk_eq(const i32* k0, const i32* k1, SelArr c, int nc) {
  for (int t = 0; t < nc; ++t)
    if (k0[c.v[t] - 1] != k1[c.v[t] - 1]) return false;
  return true;
}
__global__ void walk_ranges_kernel(int mode, const i32* __restrict__ kA, int wA, i64 nA, const i32* __restrict__ kB,
                                   int wB, i64 nB, SelArr ca, SelArr cb, int nc, i32* __restrict__ lo,
                                   i32* __restrict__ hi) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (i64 i = 0; i < nA; ++i) lo[i] = hi[i] = 0;
  i64 i = 0, j = 0;
  if (mode == 1) {  // contract_coo.jl:139-167 (single column; element steps)
    while (i < nA && j < nB) {
      int c = walk_cmp(kA + i * wA, kB + j * wB, ca, cb, nc);
      if (c < 0)
        ++i;
      else if (c > 0)
        ++j;
      else {
        i64 jl = j;
        while (j < nB && walk_cmp(kA + i * wA, kB + j * wB, ca, cb, nc) == 0) ++j;
        i64 i0 = i;
        while (i < nA && walk_eq(kA + i0 * wA, kA + i * wA, ca, nc)) {
          lo[i] = (i32)jl;
          hi[i] = (i32)j;
          ++i;
        }
      }
    }
    return;
  }
  // mode 0 / 2: run-wise walk (contract_bs_bs.jl:218-261 and :687-728)
  while (i < nA && j < nB) {
    int c = walk_cmp(kA + i * wA, kB + j * wB, ca, cb, nc);
    if (c < 0) {
      i64 i0 = i;
      while (i < nA && walk_eq(kA + i0 * wA, kA + i * wA, ca, nc)) ++i;
      continue;
    } else if (c > 0) {
      i64 j0 = j;
      while (j < nB && walk_eq(kB + j0 * wB, kB + j * wB, cb, nc)) ++j;
      continue;
    }
    i64 i0 = i, j0 = j;
    while (j < nB && walk_eq(kB + j0 * wB, kB + j * wB, cb, nc)) ++j;
    while (i < nA && walk_eq(kA + i0 * wA, kA + i * wA, ca, nc)) {
      lo[i] = (i32)j0;
      hi[i] = (i32)j;
      ++i;
    }
  }
}
void k_walk_ranges(Ctx* ctx, int mode, const i32* keysA, int wA, i64 nA, const i32* keysB, int wB, i64 nB,
                   const int* colsA, const int* colsB, int ncols, i32* lo, i32* hi) {
  if (nA == 0) return;
  SelArr ca{}, cb{};
  for (int t = 0; t < ncols; ++t) {
    ca.v[t] = colsA[t];
    cb.v[t] = colsB[t];
  }
  hipLaunchKernelGGL(walk_ranges_kernel, dim3(1), dim3(64), 0, ctx->stream, mode, keysA, wA, nA, keysB, wB, nB, ca,
                     cb, ncols, lo, hi);
  HIP_CHECK(hipGetLastError());
}

__global__ void range_counts_kernel(const i32* lo, const i32* hi, i64 n, i64* c) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = (i64)(hi[i] - lo[i]);
}
void k_range_counts(Ctx* ctx, const i32* lo, const i32* hi, i64 n, i64* counts) {
  LAUNCH(range_counts_kernel, n, lo, hi, n, counts);
}

// thread per A row (rows own at most a B run each)
__global__ void emit_pairs_kernel(const i32* __restrict__ lo, const i64* __restrict__ off,
                                  const i64* __restrict__ cnt, i64 nA, i32* __restrict__ pa, i32* __restrict__ pb) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nA) return;
  i64 o = off[i], c = cnt[i];
  i32 l = lo[i];
  for (i64 t = 0; t < c; ++t) {
    pa[o + t] = (i32)i;
    pb[o + t] = l + (i32)t;
  }
}
void k_emit_pairs(Ctx* ctx, const i32* lo, const i64* off, const i64* counts, i64 nA, i32* pa, i32* pb) {
  LAUNCH(emit_pairs_kernel, nA, lo, off, counts, nA, pa, pb);
}

// ---------------------------------------------------------------- allocation
__global__ void heads_kernel(const u64* __restrict__ lin, i64 n, i32* __restrict__ head) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || lin[i] != lin[i - 1]) ? 1 : 0;
}
__global__ void seg_start_kernel(const i32* __restrict__ head, const i32* __restrict__ incl, i64 n,
                                 i32* __restrict__ seg_start, i64 nC) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && head[i]) seg_start[incl[i] - 1] = (i32)i;
  if (i == 0) seg_start[nC] = (i32)n;
}
__global__ void first_pair_kernel(const i32* __restrict__ seg_start, const i32* __restrict__ order, i64 nC,
                                  unsigned* __restrict__ first, i32* __restrict__ sidx) {
  i64 s = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nC) return;
  first[s] = (unsigned)order[seg_start[s]];
  sidx[s] = (i32)s;
}
__global__ void rank_ids_kernel(const i32* __restrict__ sidx_sorted, i64 nC, i32* __restrict__ ids,
                                i32* __restrict__ slot) {
  i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nC) return;
  i32 s = sidx_sorted[r];
  ids[s] = (i32)r + 1;
  slot[s] = (i32)r;
}
__global__ void seg_keys_kernel(const i32* __restrict__ seg_start, const i32* __restrict__ order,
                                const i32* __restrict__ ck, int PC, i64 nC, i32* __restrict__ keys) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nC * PC) return;
  i64 s = t / PC;
  int j = (int)(t - s * PC);
  keys[t] = ck[(i64)order[seg_start[s]] * PC + j];
}

void allocate_blocks(Ctx* ctx, i64 np, int PC, DBuf<i32>& ck, DBuf<u64>& lin, u64 maxlin, Alloc& out) {
  out.nC = 0;
  if (np == 0) {
    out.seg_start.alloc(ctx, 1);
    k_zero(ctx, out.seg_start.get(), sizeof(i32));
    return;
  }
  out.order.alloc(ctx, np);
  k_iota(ctx, out.order.get(), np, 0);
  sort_pairs(ctx, lin.get(), out.order.get(), np, bits_for(maxlin));
  DBuf<i32> head(ctx, np), incl(ctx, np);
  LAUNCH(heads_kernel, np, lin.get(), np, head.get());
  size_t tb = 0;
  HIP_CHECK(rocprim::inclusive_scan(nullptr, tb, head.get(), incl.get(), (size_t)np, rocprim::plus<i32>(),
                                    ctx->stream));
  DBuf<unsigned char> tmp(ctx, tb ? tb : 1);
  HIP_CHECK(rocprim::inclusive_scan(tmp.get(), tb, head.get(), incl.get(), (size_t)np, rocprim::plus<i32>(),
                                    ctx->stream));
  i32 nC = 0;
  HIP_CHECK(hipMemcpyAsync(&nC, incl.get() + np - 1, sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  out.nC = nC;
  out.seg_start.alloc(ctx, nC + 1);
  LAUNCH(seg_start_kernel, np, head.get(), incl.get(), np, out.seg_start.get(), (i64)nC);
  DBuf<unsigned> first(ctx, nC);
  DBuf<i32> sidx(ctx, nC);
  LAUNCH(first_pair_kernel, nC, out.seg_start.get(), out.order.get(), (i64)nC, first.get(), sidx.get());
  sort_pairs32(ctx, first.get(), sidx.get(), nC, bits_for((u64)np));
  out.ids.alloc(ctx, nC);
  out.seg_slot.alloc(ctx, nC);
  LAUNCH(rank_ids_kernel, nC, sidx.get(), (i64)nC, out.ids.get(), out.seg_slot.get());
  out.keys.alloc(ctx, (size_t)nC * (PC > 0 ? PC : 1));
  if (PC > 0) LAUNCH(seg_keys_kernel, (i64)nC * PC, out.seg_start.get(), out.order.get(), ck.get(), PC, (i64)nC,
                     out.keys.get());
}

// ---------------------------------------------------------------- offsets
__global__ void block_offsets_kernel(const i32* ids, i64 n, i64 bs, i64* off) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) off[i] = (ids ? (i64)(ids[i] - 1) : i) * bs;
}
void k_block_offsets(Ctx* ctx, const i32* ids, i64 n, i64 bs, i64* off) {
  LAUNCH(block_offsets_kernel, n, ids, n, bs, off);
}
__global__ void src_offsets_kernel(const i32* ids, const i32* idx, i64 n, i64 bs, i64* off) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) off[i] = (i64)(ids[idx[i]] - 1) * bs;
}
void k_src_offsets(Ctx* ctx, const i32* ids, const i32* idx, i64 n, i64 bs, i64* off) {
  LAUNCH(src_offsets_kernel, n, ids, idx, n, bs, off);
}
__global__ void take_kernel(const i64* base, const i32* idx, i64 n, i64* out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = base[idx[i]];
}
void k_take_i64(Ctx* ctx, const i64* base, const i32* idx, i64 n, i64* out) {
  LAUNCH(take_kernel, n, base, idx, n, out);
}

}  // namespace sb

namespace sb {

__global__ void lin_scale_kernel(const u64* lin, i64 n, i64 s, i64* off) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) off[i] = (i64)lin[i] * s;
}
void k_lin_scale(Ctx* ctx, const u64* lin, i64 n, i64 scale, i64* off) { LAUNCH(lin_scale_kernel, n, lin, n, scale, off); }

__global__ void cross_pairs_kernel(i64 nA, i64 nB, i32* pa, i32* pb) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nA * nB) return;
  pa[p] = (i32)(p / nB);
  pb[p] = (i32)(p % nB);
}
void k_cross_pairs(Ctx* ctx, i64 nA, i64 nB, i32* pa, i32* pb) { LAUNCH(cross_pairs_kernel, nA * nB, nA, nB, pa, pb); }

__global__ void rank1_exist_kernel(const unsigned char* fA, const unsigned char* fB, const i32* pa, const i32* pb,
                                   i64 np, i64 R, unsigned char* keep, unsigned char* kmask) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const unsigned char* a = fA + (i64)pa[p] * R;
  const unsigned char* b = fB + (pb ? (i64)pb[p] * R : 0);
  unsigned char any = 0;
  for (i64 r = 0; r < R; ++r) {
    unsigned char f = a[r] & b[r];
    any |= f;
    if (kmask) kmask[p * R + r] = f;
  }
  keep[p] = any;
}
void k_rank1_exist(Ctx* ctx, const unsigned char* fA, const unsigned char* fB, const i32* pa, const i32* pb, i64 np,
                   i64 R, unsigned char* keep, unsigned char* kmask) {
  LAUNCH(rank1_exist_kernel, np, fA, fB, pa, pb, np, R, keep, kmask);
}

i64 compact_pairs(Ctx* ctx, const unsigned char* keep, DBuf<i32>& pa, DBuf<i32>& pb, i64 np) {
  if (np == 0) return 0;
  DBuf<i32> oa(ctx, np), ob(ctx, np);
  DBuf<unsigned> cnt(ctx, 1);
  size_t tb = 0;
  HIP_CHECK(rocprim::select(nullptr, tb, pa.get(), keep, oa.get(), cnt.get(), (size_t)np, ctx->stream));
  DBuf<unsigned char> tmp(ctx, tb ? tb : 1);
  HIP_CHECK(rocprim::select(tmp.get(), tb, pa.get(), keep, oa.get(), cnt.get(), (size_t)np, ctx->stream));
  if (pb.get()) {
    HIP_CHECK(rocprim::select(tmp.get(), tb, pb.get(), keep, ob.get(), cnt.get(), (size_t)np, ctx->stream));
  }
  unsigned c = 0;
  HIP_CHECK(hipMemcpyAsync(&c, cnt.get(), sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  pa = std::move(oa);
  if (pb.get()) pb = std::move(ob);
  return (i64)c;
}

__global__ void take_f64_kernel(const double* v, const i32* idx, i64 n, double* out) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v[idx[i]];
}
void k_take_f64(Ctx* ctx, const double* v, const i32* idx, i64 n, double* out) { LAUNCH(take_f64_kernel, n, v, idx, n, out); }

#pragma clang fp contract(off)
__global__ void coo_products_kernel(const double* va, const double* vb, const i32* pa, const i32* pb, i64 np,
                                    double* prod, unsigned char* keep) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double v = va[pa[p]] * vb[pb[p]];
  prod[p] = v;
  keep[p] = v != 0.0;
}
__global__ void scaled_slice_nonzero_kernel(const double* alpha, const i32* keysA, int NA, const double* X,
                                            const i64* blkoff, const i32* pa, const i32* pb, i64 np, i64 chunk,
                                            unsigned char* keep) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  i32 a = pa[p];
  double al = alpha[a];
  i64 rv = keysA[(i64)a * NA + NA - 1];
  const double* x = X + blkoff[pb[p]] + (rv - 1) * chunk;
  unsigned char f = 0;
  for (i64 t = 0; t < chunk; ++t)
    if (al * x[t] != 0.0) {
      f = 1;
      break;
    }
  keep[p] = f;
}
__global__ void coo_dense_nonzero_kernel(const double* alpha, const i32* keys, int w, int col, i64 n, const double* X,
                                         i64 chunk, unsigned char* keep) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  double al = alpha[i];
  i64 rv = keys[i * w + col - 1];
  const double* x = X + (rv - 1) * chunk;
  unsigned char f = 0;
  for (i64 t = 0; t < chunk; ++t)
    if (al * x[t] != 0.0) {
      f = 1;
      break;
    }
  keep[i] = f;
}
#pragma clang fp contract(on)

void k_coo_products(Ctx* ctx, const double* va, const double* vb, const i32* pa, const i32* pb, i64 np, double* prod,
                    unsigned char* keep) {
  LAUNCH(coo_products_kernel, np, va, vb, pa, pb, np, prod, keep);
}
void k_scaled_slice_nonzero(Ctx* ctx, const double* alpha, const i32* keysA, int NA, const double* X,
                            const i64* blkoff, const i32* pa, const i32* pb, i64 np, i64 chunk, unsigned char* keep) {
  LAUNCH(scaled_slice_nonzero_kernel, np, alpha, keysA, NA, X, blkoff, pa, pb, np, chunk, keep);
}
void k_coo_dense_nonzero(Ctx* ctx, const double* alpha, const i32* keys, int w, int col, i64 n, const double* X,
                         i64 chunk, unsigned char* keep) {
  LAUNCH(coo_dense_nonzero_kernel, n, alpha, keys, w, col, n, X, chunk, keep);
}

__global__ void rv_offsets_kernel(const i32* keysA, int w, int col, const i32* pa, const i64* blkoff, const i32* pb,
                                  i64 np, i64 chunk, i64* off) {
  i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  i64 a = pa ? pa[p] : p;
  i64 rv = keysA[a * w + col - 1];
  off[p] = (blkoff ? blkoff[pb[p]] : 0) + (rv - 1) * chunk;
}
void k_rv_offsets(Ctx* ctx, const i32* keysA, int w, int col, const i32* pa, const i64* blkoff, const i32* pb, i64 np,
                  i64 chunk, i64* off) {
  LAUNCH(rv_offsets_kernel, np, keysA, w, col, pa, blkoff, pb, np, chunk, off);
}

__global__ void rv_inrange_kernel(const i32* keys, int w, int col, i64 n, i64 R, unsigned char* keep) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  i64 rv = keys[i * w + col - 1];
  keep[i] = (rv >= 1 && rv <= R) ? 1 : 0;
}
void k_rv_inrange(Ctx* ctx, const i32* keys, int w, int col, i64 n, i64 R, unsigned char* keep) {
  LAUNCH(rv_inrange_kernel, n, keys, w, col, n, R, keep);
}

__global__ void col_sorted_kernel(const i32* keys, int w, int col, i64 n, unsigned* bad) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  if (keys[(i + 1) * w + col - 1] < keys[i * w + col - 1]) atomicOr(bad, 1u);
}
bool column_sorted(Ctx* ctx, const i32* keys, int w, int col, i64 n) {
  if (n <= 1) return true;
  DBuf<unsigned> bad(ctx, 1);
  k_zero(ctx, bad.get(), sizeof(unsigned));
  LAUNCH(col_sorted_kernel, n, keys, w, col, n, bad.get());
  unsigned b = 0;
  HIP_CHECK(hipMemcpyAsync(&b, bad.get(), sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return b == 0;
}

}  // namespace sb
