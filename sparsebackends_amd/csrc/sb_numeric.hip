// Numeric phase: the per-block-pair arithmetic of every contraction family.
//
// grouped_gemm replaces the reference's one-BLAS-call-per-pair loops
// (mul! at contract_bs_bs.jl:718,722 and contract_bs_dense.jl:575,578,607,610;
// BLAS.ger! at contract_bs_bs.jl:118-124 called from :391,:414 and
// contract_bs_dense.jl:357,383).  One workgroup owns one C tile and walks the
// pairs that feed its C block in the reference's traversal order; partial
// sums stay in MFMA accumulators, so there are no atomics and results are
// deterministic.  FP64 uses v_mfma_f64_16x16x4_f64.
//
// kahan_outer / axpy_groups / outer_groups reproduce loops whose exact
// floating-point operation order the reference fixes (Kahan loop
// contract_bs_dense.jl:167-196, _axpy_chunk! contract_coo_dense.jl:21-33,
// _scale_add_block! contract_coo_bs.jl:4-10, _outer_add! contract_bs_bs.jl:7-19):
// one thread per output element runs the same sequence of IEEE operations,
// with floating-point contraction disabled, so results are bit-identical.
#include "sb_kernels.h"

namespace sb {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

struct GArgs {
  i64 M, N, K;
  i64 sam, sak, sbk, sbn, scm, scn;
  i64 tiles_m, tiles_n;
};

// ---------------------------------------------------------------- bf16 split
__device__ inline unsigned short bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  unsigned r = ((u >> 16) & 1u) + 0x7fffu;
  return (unsigned short)((u + r) >> 16);
}
__device__ inline float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }
// x ~= h + m + l with each term a bf16 (x first rounded to FP32)
__device__ inline void split3(double xd, unsigned short& h, unsigned short& m, unsigned short& l) {
  float x = (float)xd;
  h = bf16_rne(x);
  float r1 = x - bf16_to_f(h);
  m = bf16_rne(r1);
  float r2 = r1 - bf16_to_f(m);
  l = bf16_rne(r2);
}

// ---------------------------------------------------------------- MFMA GEMM
// PREC 0: FP64 MFMA; 1: FP32 MFMA (operands rounded to FP32); 2: bf16x3.
template <int BM, int BN, int BK, int WM, int WN, int PREC>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(GArgs g, const double* __restrict__ A,
                                                        const i64* __restrict__ aoff, const double* __restrict__ B,
                                                        const i64* __restrict__ boff, double* __restrict__ C,
                                                        const i64* __restrict__ coff, const i32* __restrict__ seg,
                                                        const i32* __restrict__ order,
                                                        const unsigned char* __restrict__ kmask) {
  constexpr int FM = BM / (16 * WM), FN = BN / (16 * WN);
  static_assert(FM >= 1 && FN >= 1, "tile");
  static_assert(PREC == 0 || PREC == 1 || (PREC == 2 && BK % 16 == 0), "bf16x3 needs BK % 16 == 0");
  using T = typename std::conditional<PREC == 0, double, float>::type;
  __shared__ T As[BK][BM + 2];
  __shared__ T Bs[BK][BN + 2];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave % WM, wn = wave / WM;
  const i64 ntile = g.tiles_m * g.tiles_n;
  const i64 grp = blockIdx.x / ntile;
  const i64 t = blockIdx.x - grp * ntile;
  const i64 tm = t % g.tiles_m, tn = t / g.tiles_m;
  const i64 m0 = tm * BM, n0 = tn * BN;

  d4 accd[FM][FN];
  f4 accf[FM][FN];
  for (int i = 0; i < FM; ++i)
    for (int j = 0; j < FN; ++j) {
      accd[i][j] = d4{0, 0, 0, 0};
      accf[i][j] = f4{0, 0, 0, 0};
    }

  const i32 p0 = seg[grp], p1 = seg[grp + 1];
  for (i32 pp = p0; pp < p1; ++pp) {
    const i32 pi = order[pp];
    const double* Ab = A + aoff[pi];
    const double* Bb = B + boff[pi];
    const unsigned char* km = kmask ? kmask + (i64)pi * g.K : nullptr;
    for (i64 k0 = 0; k0 < g.K; k0 += BK) {
      __syncthreads();
      for (int idx = tid; idx < BM * BK; idx += 256) {
        int m, k;
        if (g.sam == 1) {
          m = idx % BM;
          k = idx / BM;
        } else {
          k = idx % BK;
          m = idx / BK;
        }
        i64 gm = m0 + m, gk = k0 + k;
        double v = 0.0;
        if (gm < g.M && gk < g.K && (!km || km[gk])) v = Ab[gm * g.sam + gk * g.sak];
        As[k][m] = (T)v;
      }
      for (int idx = tid; idx < BK * BN; idx += 256) {
        int n, k;
        if (g.sbk == 1) {
          k = idx % BK;
          n = idx / BK;
        } else {
          n = idx % BN;
          k = idx / BN;
        }
        i64 gn = n0 + n, gk = k0 + k;
        double v = 0.0;
        if (gn < g.N && gk < g.K) v = Bb[gk * g.sbk + gn * g.sbn];
        Bs[k][n] = (T)v;
      }
      __syncthreads();
      if constexpr (PREC == 0) {
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
          double a[FM], b[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i) a[i] = As[kk * 4 + (lane >> 4)][(wm * FM + i) * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < FN; ++j) b[j] = Bs[kk * 4 + (lane >> 4)][(wn * FN + j) * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) accd[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], accd[i][j], 0, 0, 0);
        }
      } else if constexpr (PREC == 1) {
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
          float a[FM], b[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i) a[i] = As[kk * 4 + (lane >> 4)][(wm * FM + i) * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < FN; ++j) b[j] = Bs[kk * 4 + (lane >> 4)][(wn * FN + j) * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) accf[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], accf[i][j], 0, 0, 0);
        }
      } else {
        // bf16x3: both operands split in three bf16 terms; the six products of
        // order <= 2 (hh, hm, mh, hl, lh, mm) accumulate in FP32
#pragma unroll
        for (int kk = 0; kk < BK / 16; ++kk) {
          s4 ah[FM], am[FM], al[FM], bh[FN], bm[FN], bl[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              unsigned short h, m, l;
              split3(As[kk * 16 + 4 * (lane >> 4) + v][(wm * FM + i) * 16 + (lane & 15)], h, m, l);
              ah[i][v] = (short)h;
              am[i][v] = (short)m;
              al[i][v] = (short)l;
            }
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              unsigned short h, m, l;
              split3(Bs[kk * 16 + 4 * (lane >> 4) + v][(wn * FN + j) * 16 + (lane & 15)], h, m, l);
              bh[j][v] = (short)h;
              bm[j][v] = (short)m;
              bl[j][v] = (short)l;
            }
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
              f4 c = accf[i][j];
              c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am[i], bm[j], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bl[j], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[i], bh[j], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bm[j], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am[i], bh[j], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bh[j], c, 0, 0, 0);
              accf[i][j] = c;
            }
        }
      }
    }
  }
  double* Cb = C + coff[grp];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // output row of register v: f64 16x16x4 holds rows v*4 + lane/16 (four
        // groups of one row); the f32 / bf16 16x16 forms hold rows 4*(lane/16) + v
        i64 m = m0 + (wm * FM + i) * 16 + (PREC == 0 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v);
        i64 n = n0 + (wn * FN + j) * 16 + (lane & 15);
        if (m < g.M && n < g.N) Cb[m * g.scm + n * g.scn] = PREC == 0 ? accd[i][j][v] : (double)accf[i][j][v];
      }
}

// one thread per C element (small blocks): sum over pairs and k in order
template <int PREC>
__global__ void gemm_elem_kernel(GArgs g, const double* __restrict__ A, const i64* __restrict__ aoff,
                                 const double* __restrict__ B, const i64* __restrict__ boff, double* __restrict__ C,
                                 const i64* __restrict__ coff, const i32* __restrict__ seg,
                                 const i32* __restrict__ order, const unsigned char* __restrict__ kmask, i64 ngroups) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  i64 mn = g.M * g.N;
  if (t >= ngroups * mn) return;
  i64 grp = t / mn, r = t - grp * mn;
  i64 m = r % g.M, n = r / g.M;
  double acc = 0.0;
  float accf = 0.0f;
  for (i32 pp = seg[grp]; pp < seg[grp + 1]; ++pp) {
    i32 pi = order[pp];
    const double* Ab = A + aoff[pi] + m * g.sam;
    const double* Bb = B + boff[pi] + n * g.sbn;
    const unsigned char* km = kmask ? kmask + (i64)pi * g.K : nullptr;
    for (i64 k = 0; k < g.K; ++k) {
      if (km && !km[k]) continue;
      double a = Ab[k * g.sak], b = Bb[k * g.sbk];
      if (PREC == 0)
        acc = fma(a, b, acc);
      else if (PREC == 1)
        accf = fmaf((float)a, (float)b, accf);
      else {
        unsigned short ah, am, al, bh, bm, bl;
        split3(a, ah, am, al);
        split3(b, bh, bm, bl);
        float s = bf16_to_f(am) * bf16_to_f(bm) + bf16_to_f(ah) * bf16_to_f(bl) + bf16_to_f(al) * bf16_to_f(bh) +
                  bf16_to_f(ah) * bf16_to_f(bm) + bf16_to_f(am) * bf16_to_f(bh);
        accf = fmaf(bf16_to_f(ah), bf16_to_f(bh), accf + s);
      }
    }
  }
  C[coff[grp] + m * g.scm + n * g.scn] = PREC == 0 ? acc : (double)accf;
}

template <int BM, int BN, int BK, int WM, int WN, int PREC>
static void launch_mfma(Ctx* ctx, GArgs g, const double* A, const i64* aoff, const double* B, const i64* boff,
                        double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups,
                        const unsigned char* kmask) {
  g.tiles_m = (g.M + BM - 1) / BM;
  g.tiles_n = (g.N + BN - 1) / BN;
  i64 nwg = ngroups * g.tiles_m * g.tiles_n;
  SB_ASSERT(nwg < (1ll << 31), "grouped_gemm: grid too large");
  hipLaunchKernelGGL((gemm_mfma_kernel<BM, BN, BK, WM, WN, PREC>), dim3((unsigned)nwg), dim3(256), 0, ctx->stream,
                     g, A, aoff, B, boff, C, coff, seg, order, kmask);
  HIP_CHECK(hipGetLastError());
}

template <int PREC>
static void gemm_dispatch(Ctx* ctx, GArgs g, const double* A, const i64* aoff, const double* B, const i64* boff,
                          double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups,
                          const unsigned char* kmask) {
  const i64 M = g.M, N = g.N, K = g.K;
  if (M * N <= 64 || (PREC != 2 && M * N * K <= 256)) {
    i64 tot = ngroups * M * N;
    hipLaunchKernelGGL((gemm_elem_kernel<PREC>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, g,
                       A, aoff, B, boff, C, coff, seg, order, kmask, ngroups);
    HIP_CHECK(hipGetLastError());
    return;
  }
  constexpr int BK0 = PREC == 2 ? 16 : 4;
  if (K <= 4 || PREC == 2) {
    if (M <= 16)
      launch_mfma<16, 128, (PREC == 2 ? 16 : 4), 1, 4, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
    else if (N <= 16)
      launch_mfma<128, 16, (PREC == 2 ? 16 : 4), 4, 1, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
    else if (M <= 32 && N <= 32)
      launch_mfma<32, 32, (PREC == 2 ? 16 : 4), 2, 2, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
    else
      launch_mfma<64, 64, (PREC == 2 ? 16 : 4), 2, 2, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
    (void)BK0;
    return;
  }
  if (M <= 16)
    launch_mfma<16, 128, 16, 1, 4, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
  else if (N <= 16)
    launch_mfma<128, 16, 16, 4, 1, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
  else if (M <= 32 && N <= 32)
    launch_mfma<32, 32, 16, 2, 2, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
  else
    launch_mfma<64, 64, 16, 2, 2, PREC>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
}

void grouped_gemm(Ctx* ctx, unsigned prec, const GemmDesc& d, const double* A, const i64* aoff, const double* B,
                  const i64* boff, double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups,
                  const unsigned char* kmask) {
  if (ngroups == 0 || d.M == 0 || d.N == 0) return;
  GArgs g{d.M, d.N, d.K, d.sam, d.sak, d.sbk, d.sbn, d.scm, d.scn, 0, 0};
  switch (prec & SB_PREC_MASK) {
    case SB_PREC_FP64:
      gemm_dispatch<0>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
      break;
    case SB_PREC_FP32:
      gemm_dispatch<1>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
      break;
    case SB_PREC_BF16X3:
      gemm_dispatch<2>(ctx, g, A, aoff, B, boff, C, coff, seg, order, ngroups, kmask);
      break;
    default:
      fail(SB_E_ARGUMENT, "unknown precision form");
  }
}

// ---------------------------------------------------------------- exact-order kernels
#pragma clang fp contract(off)

__global__ void kahan_outer_kernel(bool AB, i64 cA, i64 cB, const double* __restrict__ A,
                                   const i64* __restrict__ aoff, const double* __restrict__ B,
                                   const i64* __restrict__ boff, double* __restrict__ C,
                                   const i64* __restrict__ coff, const i32* __restrict__ seg,
                                   const i32* __restrict__ order, i64 ngroups) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  i64 bs = cA * cB;
  if (t >= ngroups * bs) return;
  i64 grp = t / bs, e = t - grp * bs;
  // element e of C: AB layout C[ii + jj*cA]; BA layout C[ii + jj*cB]
  i64 ia, jb;
  if (AB) {
    ia = e % cA;
    jb = e / cA;
  } else {
    jb = e % cB;
    ia = e / cB;
  }
  double c = 0.0, comp = 0.0;
  for (i32 pp = seg[grp]; pp < seg[grp + 1]; ++pp) {
    i32 pi = order[pp];
    double p = AB ? A[aoff[pi] + ia] * B[boff[pi] + jb] : B[boff[pi] + jb] * A[aoff[pi] + ia];
    double y = p - comp;
    double tt = c + y;
    comp = (tt - c) - y;
    c = tt;
  }
  C[coff[grp] + e] = c;
}

__global__ void axpy_groups_kernel(i64 L, const double* __restrict__ alpha, const double* __restrict__ X,
                                   const i64* __restrict__ xoff, double* __restrict__ C, const i64* __restrict__ coff,
                                   const i32* __restrict__ seg, const i32* __restrict__ order, i64 ngroups) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ngroups * L) return;
  i64 grp = t / L, e = t - grp * L;
  double c = 0.0;
  const i32 end = seg[grp + 1];
  i32 pp = seg[grp];
  // eight terms' loads issued together, then added in order: the same IEEE
  // operation sequence as one term at a time, without a load latency per term
  for (; pp + 8 <= end; pp += 8) {
    i32 pi[8];
    double a[8], x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) pi[u] = order[pp + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = alpha[pi[u]];
      x[u] = X[xoff[pi[u]] + e];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      double prod = a[u] * x[u];
      c = c + prod;
    }
  }
  for (; pp < end; ++pp) {
    i32 pi = order[pp];
    double prod = alpha[pi] * X[xoff[pi] + e];
    c = c + prod;
  }
  C[coff[grp] + e] = c;
}

__global__ void outer_groups_kernel(i64 m, i64 n, const double* __restrict__ A, const i64* __restrict__ aoff,
                                    const double* __restrict__ B, const i64* __restrict__ boff,
                                    double* __restrict__ C, const i64* __restrict__ coff,
                                    const i32* __restrict__ seg, const i32* __restrict__ order, i64 ngroups) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  i64 bs = m * n;
  if (t >= ngroups * bs) return;
  i64 grp = t / bs, e = t - grp * bs;
  i64 i = e % m, j = e / m;
  double c = 0.0;
  const i32 end = seg[grp + 1];
  i32 pp = seg[grp];
  for (; pp + 8 <= end; pp += 8) {  // loads of eight terms together, adds in order
    i32 pi[8];
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) pi[u] = order[pp + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = A[aoff[pi[u]] + i];
      b[u] = B[boff[pi[u]] + j];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      double prod = a[u] * b[u];
      c = c + prod;
    }
  }
  for (; pp < end; ++pp) {
    i32 pi = order[pp];
    double prod = A[aoff[pi] + i] * B[boff[pi] + j];
    c = c + prod;
  }
  C[coff[grp] + e] = c;
}

#pragma clang fp contract(on)

static unsigned g1(i64 tot) { return (unsigned)((tot + 255) / 256); }

void kahan_outer(Ctx* ctx, bool AB, i64 chunkA, i64 chunkB, const double* A, const i64* aoff, const double* B,
                 const i64* boff, double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups) {
  i64 tot = ngroups * chunkA * chunkB;
  if (!tot) return;
  hipLaunchKernelGGL(kahan_outer_kernel, dim3(g1(tot)), dim3(256), 0, ctx->stream, AB, chunkA, chunkB, A, aoff, B,
                     boff, C, coff, seg, order, ngroups);
  HIP_CHECK(hipGetLastError());
}
void axpy_groups(Ctx* ctx, i64 L, const double* alpha, const double* X, const i64* xoff, double* C, const i64* coff,
                 const i32* seg, const i32* order, i64 ngroups) {
  i64 tot = ngroups * L;
  if (!tot) return;
  hipLaunchKernelGGL(axpy_groups_kernel, dim3(g1(tot)), dim3(256), 0, ctx->stream, L, alpha, X, xoff, C, coff, seg,
                     order, ngroups);
  HIP_CHECK(hipGetLastError());
}
void outer_groups(Ctx* ctx, i64 m, i64 n, const double* A, const i64* aoff, const double* B, const i64* boff,
                  double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups) {
  i64 tot = ngroups * m * n;
  if (!tot) return;
  hipLaunchKernelGGL(outer_groups_kernel, dim3(g1(tot)), dim3(256), 0, ctx->stream, m, n, A, aoff, B, boff, C, coff,
                     seg, order, ngroups);
  HIP_CHECK(hipGetLastError());
}

__global__ void slice_nonzero_kernel(const double* __restrict__ X, const i64* __restrict__ off, i64 nrows, i64 R,
                                     i64 chunk, unsigned char* __restrict__ flag) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrows * R) return;
  i64 i = t / R, r = t - i * R;
  const double* x = X + (off ? off[i] : 0) + r * chunk;
  unsigned char f = 0;
  for (i64 q = 0; q < chunk; ++q)
    if (x[q] != 0.0) {
      f = 1;
      break;
    }
  flag[t] = f;
}
void slice_nonzero(Ctx* ctx, const double* X, const i64* off, i64 nrows, i64 R, i64 chunk, unsigned char* flag) {
  i64 tot = nrows * R;
  if (!tot) return;
  hipLaunchKernelGGL(slice_nonzero_kernel, dim3(g1(tot)), dim3(256), 0, ctx->stream, X, off, nrows, R, chunk, flag);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sb
