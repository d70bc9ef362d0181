// Per-group SVDs of the two-site block-diagonal factorisation: one wave per
// group runs one-sided Jacobi (Hestenes) on the d1 rows of its
// d1 x ncols matrix, which keeps every singular value to high relative
// accuracy, and writes the singular values (descending), the accumulated
// left rotation U and the order of the orthogonalised rows.
#include "sb_kernels.h"

namespace sb {

constexpr int SVD_MAXD = 32;  // largest d1 (rows per group)

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// W rows of group g: row r at woff[g] + r * ncols_g, ncols_g = (gstart[g+1]-gstart[g]) * d2,
// column j*d2 + t = A[bsrc[j] + r + d1*t] (block j of the group, column-major d1 x d2 block)
__global__ __launch_bounds__(64) void svd_jacobi_kernel(const double* __restrict__ A, int d1, int d2,
                                                       const i64* __restrict__ gstart, const i64* __restrict__ bsrc,
                                                       const i64* __restrict__ woff, double* __restrict__ W,
                                                       double* __restrict__ sig, double* __restrict__ Uout,
                                                       int* __restrict__ perm) {
  __shared__ double U[SVD_MAXD * SVD_MAXD];
  __shared__ double s_sig[SVD_MAXD];
  const int g = blockIdx.x, lane = threadIdx.x;
  const i64 b0 = gstart[g], nbk = gstart[g + 1] - b0;
  const i64 nc = nbk * d2;
  double* Wg = W + woff[g];
  for (i64 e = lane; e < (i64)d1 * nc; e += 64) {
    const int r = (int)(e / nc);
    const i64 col = e - (i64)r * nc;
    const i64 j = col / d2;
    const int t = (int)(col - j * d2);
    Wg[e] = A[bsrc[b0 + j] + r + (i64)d1 * t];
  }
  for (int e = lane; e < d1 * d1; e += 64) U[e] = (e % d1 == e / d1) ? 1.0 : 0.0;
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < d1 - 1; ++p)
      for (int q = p + 1; q < d1; ++q) {
        double* rp = Wg + (i64)p * nc;
        double* rq = Wg + (i64)q * nc;
        double a = 0, b = 0, c = 0;
        for (i64 j = lane; j < nc; j += 64) {
          const double x = rp[j], y = rq[j];
          a = fma(x, x, a);
          b = fma(y, y, b);
          c = fma(x, y, c);
        }
        a = wave_sum(a);
        b = wave_sum(b);
        c = wave_sum(c);
        if (c == 0.0 || fabs(c) <= 1e-15 * sqrt(a) * sqrt(b)) continue;
        const double zeta = (b - a) / (2.0 * c);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        for (i64 j = lane; j < nc; j += 64) {
          const double x = rp[j], y = rq[j];
          rp[j] = cs * x - sn * y;
          rq[j] = sn * x + cs * y;
        }
        if (lane < d1) {
          const double up = U[lane + p * d1], uq = U[lane + q * d1];
          U[lane + p * d1] = cs * up - sn * uq;
          U[lane + q * d1] = sn * up + cs * uq;
        }
        rotated = true;
        __syncthreads();  // row and U updates visible before the next pair
      }
    if (!rotated) break;
  }
  // row norms = singular values
  for (int r = 0; r < d1; ++r) {
    const double* rr = Wg + (i64)r * nc;
    double a = 0;
    for (i64 j = lane; j < nc; j += 64) a = fma(rr[j], rr[j], a);
    a = wave_sum(a);
    if (lane == 0) s_sig[r] = sqrt(a);
  }
  __syncthreads();
  if (lane == 0) {
    int ord[SVD_MAXD];
    for (int r = 0; r < d1; ++r) ord[r] = r;
    for (int i = 1; i < d1; ++i) {  // stable insertion sort, descending
      const int v = ord[i];
      int k = i - 1;
      while (k >= 0 && s_sig[ord[k]] < s_sig[v]) {
        ord[k + 1] = ord[k];
        --k;
      }
      ord[k + 1] = v;
    }
    for (int c = 0; c < d1; ++c) {
      sig[(i64)g * d1 + c] = s_sig[ord[c]];
      perm[(i64)g * d1 + c] = ord[c];
    }
  }
  __syncthreads();
  for (int e = lane; e < d1 * d1; e += 64) {
    const int r = e % d1, c = e / d1;
    Uout[(i64)g * d1 * d1 + e] = U[r + perm[(i64)g * d1 + c] * d1];
  }
}

void k_svd_jacobi(Ctx* ctx, const double* A, int d1, int d2, i64 ng, const i64* gstart, const i64* bsrc,
                  const i64* woff, double* W, double* sig, double* Uout, int* perm) {
  if (!ng) return;
  SB_ASSERT(d1 >= 1 && d1 <= SVD_MAXD, "blocksparse_svd: first dense dim above 32");
  hipLaunchKernelGGL(svd_jacobi_kernel, dim3((unsigned)ng), dim3(64), 0, ctx->stream, A, d1, d2, gstart, bsrc, woff,
                     W, sig, Uout, perm);
  HIP_CHECK(hipGetLastError());
}

// one wave per kept singular triplet nb: U block nb (d1 values, column c of
// the group's U, times sigma for ortho right), SV blocks of nb (the group's
// row perm[c] of W, which is sigma * v: divided by sigma for ortho right)
__global__ __launch_bounds__(64) void svd_emit_kernel(int d1, const i32* __restrict__ kg, const i32* __restrict__ kc,
                                                     const i64* __restrict__ soff, const i64* __restrict__ gstart,
                                                     int d2, const i64* __restrict__ woff, const double* __restrict__ W,
                                                     const double* __restrict__ sig, const double* __restrict__ Uout,
                                                     const int* __restrict__ perm, int right,
                                                     double* __restrict__ Ud, double* __restrict__ SVd) {
  const i64 nb = blockIdx.x;
  const int g = kg[nb], c = kc[nb], lane = threadIdx.x;
  const double s = sig[(i64)g * d1 + c];
  for (int r = lane; r < d1; r += 64) Ud[nb * d1 + r] = Uout[(i64)g * d1 * d1 + r + (i64)c * d1] * (right ? s : 1.0);
  const i64 nc = (gstart[g + 1] - gstart[g]) * d2;
  const double* row = W + woff[g] + (i64)perm[(i64)g * d1 + c] * nc;
  const double f = right ? (s > 0 ? 1.0 / s : 0.0) : 1.0;
  for (i64 e = lane; e < nc; e += 64) SVd[soff[nb] + e] = right ? row[e] * f : row[e];
}

void k_svd_emit(Ctx* ctx, i64 k, int d1, const i32* kg, const i32* kc, const i64* soff, const i64* gstart, int d2,
                const i64* woff, const double* W, const double* sig, const double* Uout, const int* perm, bool right,
                double* Ud, double* SVd) {
  if (!k) return;
  hipLaunchKernelGGL(svd_emit_kernel, dim3((unsigned)k), dim3(64), 0, ctx->stream, d1, kg, kc, soff, gstart, d2, woff,
                     W, sig, Uout, perm, right ? 1 : 0, Ud, SVd);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sb
