// Vector kernels of the Krylov eigensolver and of the Krylov vector
// operations on block-sparse payloads.  All reductions use a fixed blocking
// and a fixed tree, so a solve repeats bit for bit on the same inputs.
#include "sb_kernels.h"

namespace sb {

constexpr int NB = 1024;  // partial sums per reduction
constexpr int RT = 256;   // threads per block
constexpr int MCH = 8;    // vectors handled per multi-dot / combine pass

int linalg_nb() { return NB; }

__device__ inline double block_sum(double a, double* s) {
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = RT / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  double r = s[0];
  __syncthreads();
  return r;
}

// part[i*NB + b] = sum over block b's range of V_i * w  (i < k), and, with
// norm, part[k*NB + b] = sum of w*w.  Block b covers [b*per, (b+1)*per).
template <int K>
__global__ __launch_bounds__(RT) void mdot_kernel(const double* __restrict__ V, i64 ldv, int k,
                                                  const double* __restrict__ w, i64 n, int norm,
                                                  double* __restrict__ part) {
  __shared__ double s[RT];
  const i64 per = (n + NB - 1) / NB;
  const i64 lo = (i64)blockIdx.x * per, hi = min(n, lo + per);
  double acc[K + 1];
#pragma unroll
  for (int i = 0; i <= K; ++i) acc[i] = 0.0;
  if (k == K) {  // a full chunk: the K loads of an element issued together, no guard per vector
    for (i64 e = lo + threadIdx.x; e < hi; e += RT) {
      const double we = w[e];
      double v[K];
#pragma unroll
      for (int i = 0; i < K; ++i) v[i] = V[(i64)i * ldv + e];
#pragma unroll
      for (int i = 0; i < K; ++i) acc[i] = fma(v[i], we, acc[i]);
      acc[K] = fma(we, we, acc[K]);
    }
  } else {
    for (i64 e = lo + threadIdx.x; e < hi; e += RT) {
      const double we = w[e];
#pragma unroll
      for (int i = 0; i < K; ++i)
        if (i < k) acc[i] = fma(V[(i64)i * ldv + e], we, acc[i]);
      acc[K] = fma(we, we, acc[K]);
    }
  }
#pragma unroll
  for (int i = 0; i < K; ++i)
    if (i < k) {
      double t = block_sum(acc[i], s);
      if (threadIdx.x == 0) part[(i64)i * NB + blockIdx.x] = t;
    }
  if (norm) {
    double t = block_sum(acc[K], s);
    if (threadIdx.x == 0) part[(i64)k * NB + blockIdx.x] = t;
  }
}

// out[i] = sum of the NB partials of row i (one block per row)
__global__ __launch_bounds__(RT) void reduce_rows_kernel(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s[RT];
  const double* p = part + (i64)blockIdx.x * NB;
  double a = 0.0;
  for (int i = threadIdx.x; i < NB; i += RT) a += p[i];
  double t = block_sum(a, s);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

void k_mdot(Ctx* ctx, const double* V, i64 ldv, int k, const double* w, i64 n, bool norm, double* part, double* out) {
  // rows: the k dot products in chunks of MCH vectors, then the norm row
  for (int c = 0; c < k || (c == 0 && norm); c += MCH) {
    const int kc = std::min(MCH, k - c);
    const bool last = c + MCH >= k;
    // a kernel of the chunk's own width, so that its loads take the unguarded path
    const void* fn = (const void*)mdot_kernel<MCH>;
    switch (kc) {
      case 1: fn = (const void*)mdot_kernel<1>; break;
      case 2: fn = (const void*)mdot_kernel<2>; break;
      case 3: fn = (const void*)mdot_kernel<3>; break;
      case 4: fn = (const void*)mdot_kernel<4>; break;
      case 5: fn = (const void*)mdot_kernel<5>; break;
      case 6: fn = (const void*)mdot_kernel<6>; break;
      case 7: fn = (const void*)mdot_kernel<7>; break;
      default: break;
    }
    const double* vc = V + (i64)c * ldv;
    int kk = std::max(kc, 0), nrm = (norm && last) ? 1 : 0;
    double* pc = part + (i64)c * NB;
    void* args[] = {(void*)&vc, (void*)&ldv, (void*)&kk, (void*)&w, (void*)&n, (void*)&nrm, (void*)&pc};
    HIP_CHECK(hipLaunchKernel(fn, dim3(NB), dim3(RT), args, 0, ctx->stream));
    HIP_CHECK(hipGetLastError());
    if (last) break;
  }
  const int rows = k + (norm ? 1 : 0);
  if (rows) {
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(rows), dim3(RT), 0, ctx->stream, part, out);
    HIP_CHECK(hipGetLastError());
  }
}

// w -= sum_i V_i h[i] (i ascending, one fma each); with norm, the block
// partials of |w|^2 after the update go to part[b]
__global__ __launch_bounds__(RT) void orth_update_kernel(const double* __restrict__ V, i64 ldv, int k,
                                                         const double* __restrict__ h, double* __restrict__ w, i64 n,
                                                         int norm, double* __restrict__ part) {
  __shared__ double s[RT];
  const i64 per = (n + NB - 1) / NB;
  const i64 lo = (i64)blockIdx.x * per, hi = min(n, lo + per);
  double a = 0.0;
  for (i64 e = lo + threadIdx.x; e < hi; e += RT) {
    double we = w[e];
    int i = 0;
    for (; i + 8 <= k; i += 8) {  // eight loads in flight, the updates in ascending i as below
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = V[(i64)(i + u) * ldv + e];
#pragma unroll
      for (int u = 0; u < 8; ++u) we = fma(-v[u], h[i + u], we);
    }
    for (; i < k; ++i) we = fma(-V[(i64)i * ldv + e], h[i], we);
    w[e] = we;
    a = fma(we, we, a);
  }
  if (norm) {
    double t = block_sum(a, s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}
// w -= sum_i V_i h[i] (i ascending, as orth_update_kernel) and, in the same
// pass over the basis, part[i*NB + b] = block b's partial of V_i . w (the new
// w; the element order and blocking of mdot_kernel), so that the second
// Gram-Schmidt pass's dot products need no read of their own: the results
// are those of the two kernels one after the other, bit for bit.  k <= 8 KC;
// the basis values of an element stay in registers (index k-1 stands in for
// the absent ones, whose terms are skipped).
template <int KC>
__global__ __launch_bounds__(RT) void orth_update_dot_kernel(const double* __restrict__ V, i64 ldv, int k,
                                                             const double* __restrict__ h, double* __restrict__ w,
                                                             i64 n, double* __restrict__ part) {
  constexpr int K = 8 * KC;
  __shared__ double s[RT];
  const i64 per = (n + NB - 1) / NB;
  const i64 lo = (i64)blockIdx.x * per, hi = min(n, lo + per);
  double hv[K], acc[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    hv[i] = h[min(i, k - 1)];
    acc[i] = 0.0;
  }
  for (i64 e = lo + threadIdx.x; e < hi; e += RT) {
    double v[K];
#pragma unroll
    for (int i = 0; i < K; ++i) v[i] = V[(i64)min(i, k - 1) * ldv + e];
    double we = w[e];
#pragma unroll
    for (int i = 0; i < K; ++i) we = i < k ? fma(-v[i], hv[i], we) : we;
    w[e] = we;
#pragma unroll
    for (int i = 0; i < K; ++i) acc[i] = i < k ? fma(v[i], we, acc[i]) : acc[i];
  }
#pragma unroll
  for (int i = 0; i < K; ++i)
    if (i < k) {
      double t = block_sum(acc[i], s);
      if (threadIdx.x == 0) part[(i64)i * NB + blockIdx.x] = t;
    }
}
// the fused pass for k <= 32 (false: the caller runs the two kernels)
bool k_orth_update_dot(Ctx* ctx, const double* V, i64 ldv, int k, const double* h, double* w, i64 n, double* part,
                       double* out) {
  if (k < 1 || k > 32) return false;
  const void* fn = k <= 8 ? (const void*)orth_update_dot_kernel<1>
                   : k <= 16 ? (const void*)orth_update_dot_kernel<2>
                   : k <= 24 ? (const void*)orth_update_dot_kernel<3>
                             : (const void*)orth_update_dot_kernel<4>;
  void* args[] = {(void*)&V, (void*)&ldv, (void*)&k, (void*)&h, (void*)&w, (void*)&n, (void*)&part};
  HIP_CHECK(hipLaunchKernel(fn, dim3(NB), dim3(RT), args, 0, ctx->stream));
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(k), dim3(RT), 0, ctx->stream, part, out);
  HIP_CHECK(hipGetLastError());
  return true;
}
void k_orth_update(Ctx* ctx, const double* V, i64 ldv, int k, const double* h, double* w, i64 n, double* part,
                   double* norm_out) {
  hipLaunchKernelGGL(orth_update_kernel, dim3(NB), dim3(RT), 0, ctx->stream, V, ldv, k, h, w, n,
                     norm_out ? 1 : 0, part);
  HIP_CHECK(hipGetLastError());
  if (norm_out) {
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(RT), 0, ctx->stream, part, norm_out);
    HIP_CHECK(hipGetLastError());
  }
}

// out_j = sum_l V_l U[l + j*kin]  (j < kout, l ascending); out must not overlap V
__global__ __launch_bounds__(RT) void combine_kernel(const double* __restrict__ V, i64 ldv, int kin,
                                                     const double* __restrict__ U, int j0, int kc,
                                                     double* __restrict__ out, i64 ldo, i64 n) {
  const i64 e = (i64)blockIdx.x * RT + threadIdx.x;
  if (e >= n) return;
  double acc[MCH];
#pragma unroll
  for (int j = 0; j < MCH; ++j) acc[j] = 0.0;
  for (int l = 0; l < kin; ++l) {
    const double v = V[(i64)l * ldv + e];
#pragma unroll
    for (int j = 0; j < MCH; ++j)
      if (j < kc) acc[j] = fma(v, U[l + (i64)(j0 + j) * kin], acc[j]);
  }
#pragma unroll
  for (int j = 0; j < MCH; ++j)
    if (j < kc) out[(i64)(j0 + j) * ldo + e] = acc[j];
}
void k_combine(Ctx* ctx, const double* V, i64 ldv, int kin, const double* U, int kout, double* out, i64 ldo, i64 n) {
  if (!n) return;
  for (int j0 = 0; j0 < kout; j0 += MCH) {
    hipLaunchKernelGGL(combine_kernel, dim3((unsigned)((n + RT - 1) / RT)), dim3(RT), 0, ctx->stream, V, ldv, kin, U,
                       j0, std::min(MCH, kout - j0), out, ldo, n);
    HIP_CHECK(hipGetLastError());
  }
}

// y = a*x + b*y (b == 0: y = a*x, y not read)
__global__ void axpby_kernel(i64 n, double a, const double* x, double b, double* y) {  // x may be y
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = b == 0.0 ? a * x[i] : fma(a, x[i], b * y[i]);
}
void k_axpby(Ctx* ctx, i64 n, double a, const double* x, double b, double* y) {
  if (!n) return;
  hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + RT - 1) / RT)), dim3(RT), 0, ctx->stream, n, a, x, b, y);
  HIP_CHECK(hipGetLastError());
}

// out block (xids[i]-1) = y's block at offset off[i] (off < 0: zeros); bs values per block
__global__ void gather_blocks_by_key_kernel(i64 n, i64 bs, const i32* __restrict__ xids, const i64* __restrict__ off,
                                            const double* __restrict__ y, double* __restrict__ out) {
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * bs) return;
  const i64 i = t / bs, e = t - i * bs;
  const i64 o = off[i];
  out[(i64)(xids[i] - 1) * bs + e] = o >= 0 ? y[o + e] : 0.0;
}
void k_gather_blocks_by_key(Ctx* ctx, i64 n, i64 bs, const i32* xids, const i64* off, const double* y, double* out) {
  if (!n || !bs) return;
  hipLaunchKernelGGL(gather_blocks_by_key_kernel, dim3((unsigned)((n * bs + RT - 1) / RT)), dim3(RT), 0, ctx->stream,
                     n, bs, xids, off, y, out);
  HIP_CHECK(hipGetLastError());
}

// dst[doff[b] + e] += src[soff[b] + e] for every block b (destinations distinct)
__global__ void add_blocks_kernel(i64 n, i64 bs, const i64* __restrict__ soff, const double* __restrict__ src,
                                  const i64* __restrict__ doff, double* __restrict__ dst) {
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * bs) return;
  const i64 b = t / bs, e = t - b * bs;
  dst[doff[b] + e] += src[soff[b] + e];
}
void k_add_blocks(Ctx* ctx, i64 n, i64 bs, const i64* soff, const double* src, const i64* doff, double* dst) {
  if (!n || !bs) return;
  hipLaunchKernelGGL(add_blocks_kernel, dim3((unsigned)((n * bs + RT - 1) / RT)), dim3(RT), 0, ctx->stream, n, bs,
                     soff, src, doff, dst);
  HIP_CHECK(hipGetLastError());
}

// off[i] = (yids[j]-1)*bs for the row j of y whose key equals row i of x
// (both key lists sorted by the column-major lin), -1 where y has none
__global__ void match_keys_kernel(const i32* __restrict__ xk, i64 nx, const i32* __restrict__ yk,
                                  const i32* __restrict__ yids, i64 ny, int P, IArr pd, i64 bs, i64* __restrict__ off) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  auto lin = [&](const i32* k) {
    u64 l = 0, s = 1;
    for (int d = 0; d < P; ++d) {
      l += (u64)(k[d] - 1) * s;
      s *= (u64)pd.v[d];
    }
    return l;
  };
  const u64 v = lin(xk + i * P);
  i64 a = 0, b = ny;
  while (a < b) {
    const i64 m = (a + b) >> 1;
    if (lin(yk + m * P) < v)
      a = m + 1;
    else
      b = m;
  }
  off[i] = (a < ny && lin(yk + a * P) == v) ? (i64)(yids[a] - 1) * bs : -1;
}
void k_match_keys(Ctx* ctx, const i32* xk, i64 nx, const i32* yk, const i32* yids, i64 ny, int P, const i64* pdims,
                  i64 bs, i64* off) {
  if (!nx) return;
  IArr pd{};
  for (int d = 0; d < P; ++d) pd.v[d] = pdims[d];
  hipLaunchKernelGGL(match_keys_kernel, dim3((unsigned)((nx + RT - 1) / RT)), dim3(RT), 0, ctx->stream, xk, nx, yk,
                     yids, ny, P, pd, bs, off);
  HIP_CHECK(hipGetLastError());
}


// row of y holding the key of each row of x (-1 if none); keys sorted by the
// column-major lin in both
__global__ void match_rows_kernel(const i32* __restrict__ xk, i64 nx, const i32* __restrict__ yk, i64 ny, int P,
                                  IArr pd, i32* __restrict__ row) {
  const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  auto lin = [&](const i32* k) {
    u64 l = 0, s = 1;
    for (int d = 0; d < P; ++d) {
      l += (u64)(k[d] - 1) * s;
      s *= (u64)pd.v[d];
    }
    return l;
  };
  const u64 v = lin(xk + i * P);
  i64 a = 0, b = ny;
  while (a < b) {
    const i64 m = (a + b) >> 1;
    if (lin(yk + m * P) < v)
      a = m + 1;
    else
      b = m;
  }
  row[i] = (a < ny && lin(yk + a * P) == v) ? (i32)a : -1;
}
void k_match_rows(Ctx* ctx, const i32* xk, i64 nx, const i32* yk, i64 ny, int P, const i64* pdims, i32* row) {
  if (!nx) return;
  IArr pd{};
  for (int d = 0; d < P; ++d) pd.v[d] = pdims[d];
  hipLaunchKernelGGL(match_rows_kernel, dim3((unsigned)((nx + RT - 1) / RT)), dim3(RT), 0, ctx->stream, xk, nx, yk,
                     ny, P, pd, row);
  HIP_CHECK(hipGetLastError());
}

// S[i, a] = max(xbox[i, a], ybox[row[i], a]) (row[i] < 0: xbox only)
__global__ void support_kernel(i64 n, int N2, const i32* __restrict__ xbox, const i32* __restrict__ ybox,
                               const i32* __restrict__ row, i32* __restrict__ S) {
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * N2) return;
  const i64 i = t / N2;
  const int a = (int)(t - i * N2);
  int v = xbox[t];
  if (row[i] >= 0) v = max(v, ybox[(i64)row[i] * N2 + a]);
  S[t] = v;
}
void k_support(Ctx* ctx, i64 n, int N2, const i32* xbox, const i32* ybox, const i32* row, i32* S) {
  if (!n || !N2) return;
  hipLaunchKernelGGL(support_kernel, dim3((unsigned)((n * N2 + RT - 1) / RT)), dim3(RT), 0, ctx->stream, n, N2, xbox,
                     ybox, row, S);
  HIP_CHECK(hipGetLastError());
}

// compact layout: row i's box S[i, :] (column-major, axis 0 fastest) packed at
// coff[i]; element t of the compact vector <-> entry of the row's block at
// boff[i] (boff < 0: absent, reads as zero).  mode 0: gather full -> compact;
// mode 1: scatter compact -> full (only inside the boxes).  One wave per
// chunk of kCompactChunk compact entries: it finds the row of the chunk's
// first entry once (binary search), then walks the rows the chunk spans, each
// row's extents, strides and offsets the wave's alike; an index inside a box
// of fewer than 2^24 entries is split in 32 bits (div24).
constexpr int kCompactChunk = 256;
int compact_chunk() { return kCompactChunk; }
__global__ __launch_bounds__(64) void compact_kernel(int mode, i64 nrows, int N2, IArr td, const i32* __restrict__ S,
                                                     const i64* __restrict__ coff, const i32* __restrict__ crow,
                                                     const i64* __restrict__ boff, const double* __restrict__ src,
                                                     double* __restrict__ dst) {
  const int lane = threadIdx.x;
  const i64 total = coff[nrows];
  for (i64 c = blockIdx.x; c * kCompactChunk < total; c += gridDim.x) {
    const i64 t0 = c * kCompactChunk, t1 = min(total, t0 + kCompactChunk);
    i64 a = 0, b = nrows - 1;  // last row with coff <= t0: from the layout's table, else searched
    if (crow) {
      a = crow[c];
    } else {
      while (a < b) {
        const i64 m = (a + b + 1) >> 1;
        if (coff[m] <= t0)
          a = m;
        else
          b = m - 1;
      }
    }
    for (i64 i = a, t = t0; t < t1; ++i) {
      const i64 r0 = coff[i], r1 = min(coff[i + 1], t1);
      if (r1 <= t) continue;  // an empty row
      const i64 bo = boff[i];
      if (mode == 0 || bo >= 0) {
        int e[MAXR];
        float inv[MAXR];
#pragma unroll
        for (int q = 0; q < MAXR; ++q)
          if (q < N2) {
            e[q] = S[i * N2 + q];  // >= 1: the row holds entries
            inv[q] = 1.0f / (float)e[q];
          }
        const bool small = coff[i + 1] - r0 < ((i64)1 << 24);
#pragma unroll 4
        for (i64 tt = t + lane; tt < r1; tt += 64) {
          i64 off = 0, st = 1;
          if (small) {
            int x = (int)(tt - r0);
#pragma unroll
            for (int q = 0; q < MAXR; ++q)
              if (q < N2) {
                const int qq = div24(x, e[q], inv[q]);
                off += (i64)(x - qq * e[q]) * st;
                st *= td.v[q];
                x = qq;
              }
          } else {
            i64 x = tt - r0;
#pragma unroll
            for (int q = 0; q < MAXR; ++q)
              if (q < N2) {
                const i64 qq = x / e[q];
                off += (x - qq * e[q]) * st;
                st *= td.v[q];
                x = qq;
              }
          }
          if (mode == 0)
            dst[tt] = bo >= 0 ? src[bo + off] : 0.0;
          else
            dst[bo + off] = src[tt];
        }
      }
      t = r1;
    }
  }
}
void k_compact(Ctx* ctx, int mode, i64 nrows, int N2, const i64* tdims, const i32* S, const i64* coff, i64 total,
               const i64* boff, const double* src, double* dst, const i32* crow) {
  if (!total || !nrows) return;
  IArr td{};
  for (int q = 0; q < N2; ++q) td.v[q] = tdims[q];
  const i64 chunks = (total + kCompactChunk - 1) / kCompactChunk;
  hipLaunchKernelGGL(compact_kernel, dim3((unsigned)std::min<i64>(chunks, 65536)), dim3(64), 0, ctx->stream, mode,
                     nrows, N2, td, S, coff, crow, boff, src, dst);
  HIP_CHECK(hipGetLastError());
}


// flag = 1 if some y box reaches beyond S (matched rows); written with a
// plain store so it is cleared first by a one-thread kernel
__global__ void box_exceeds_kernel(i64 n, int N2, const i32* __restrict__ S, const i32* __restrict__ ybox,
                                   const i32* __restrict__ row, i32* __restrict__ flag) {
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * N2) return;
  const i64 i = t / N2;
  const int a = (int)(t - i * N2);
  if (row[i] >= 0 && ybox[(i64)row[i] * N2 + a] > S[t]) *flag = 1;
}
__global__ void clear_flag_kernel(i32* flag) { *flag = 0; }
void k_box_exceeds(Ctx* ctx, i64 n, int N2, const i32* S, const i32* ybox, const i32* row, i32* flag) {
  hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, ctx->stream, flag);
  HIP_CHECK(hipGetLastError());
  if (!n || !N2) return;
  hipLaunchKernelGGL(box_exceeds_kernel, dim3((unsigned)((n * N2 + RT - 1) / RT)), dim3(RT), 0, ctx->stream, n, N2,
                     S, ybox, row, flag);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sb
