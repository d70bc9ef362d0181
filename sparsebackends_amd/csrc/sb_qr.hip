// Thin QR factorisations of dense column-major matrices on the device: one
// workgroup per matrix, Householder reflections with LAPACK's conventions
// (dgeqrf / dlarfg: beta = -sign(alpha) |x|, R's diagonal holds beta; dorgqr
// forms Q from the reflectors), so Q and R carry the signs LAPACK (and the
// reference's qr through libblastrampoline) gives them.
//
// Used by blocksparse_left_qr (SparseBackends/src/ops_factorize.jl:8-43) on
// its dense (n_l d_s) x n_r matrix, and by the sector QR of a sector-format
// MPS site tensor (one matrix per right charge sector).
#include "sb_kernels.h"

namespace sb {

namespace {

constexpr int QT = 256;  // threads of a factorisation

__device__ double qr_block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = QT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// per matrix: m rows, n columns, A at aoff (leading dimension m, overwritten by
// the reflectors), Q (m x k) at qoff, R (k x n) at roff, tau scratch at toff (k)
__global__ __launch_bounds__(QT) void qr_kernel(const QrJob* __restrict__ jobs, double* __restrict__ A,
                                                double* __restrict__ Q, double* __restrict__ R,
                                                double* __restrict__ tau) {
  __shared__ double red[QT];
  __shared__ double sbeta, stau, sscale;
  const QrJob J = jobs[blockIdx.x];
  const int m = J.m, n = J.n, k = min(m, n);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double* a = A + J.aoff;
  double* t = tau + J.toff;
  for (int j = 0; j < k; ++j) {
    double* col = a + (i64)j * m;
    // dlarfg on col[j..m)
    double ss = 0.0;
    for (int i = j + 1 + tid; i < m; i += QT) ss += col[i] * col[i];
    const double xnorm2 = qr_block_sum(ss, red);
    if (tid == 0) {
      const double alpha = col[j];
      if (xnorm2 == 0.0) {
        stau = 0.0;
        sbeta = alpha;
        sscale = 0.0;
      } else {
        const double beta = -copysign(hypot(alpha, sqrt(xnorm2)), alpha);
        stau = (beta - alpha) / beta;
        sscale = 1.0 / (alpha - beta);
        sbeta = beta;
      }
    }
    __syncthreads();
    const double tj = stau, sc = sscale;
    if (tj != 0.0)
      for (int i = j + 1 + tid; i < m; i += QT) col[i] *= sc;
    __syncthreads();
    if (tid == 0) {
      col[j] = sbeta;
      t[j] = tj;
    }
    // apply H_j = I - tau v v^T (v = [1; col[j+1..m)]) to columns j+1..n, one wave per column
    if (tj != 0.0)
      for (int c = j + 1 + wave; c < n; c += QT / 64) {
        double* y = a + (i64)c * m;
        double d = lane == 0 ? y[j] : 0.0;
        for (int i = j + 1 + lane; i < m; i += 64) d += col[i] * y[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
        const double f = tj * d;
        if (lane == 0) y[j] -= f;
        for (int i = j + 1 + lane; i < m; i += 64) y[i] -= f * col[i];
      }
    __syncthreads();
  }
  // R: the upper triangle of the first k rows
  double* r = R + J.roff;
  for (int e = tid; e < k * n; e += QT) {
    const int c = e / k, i = e - c * k;
    r[e] = i <= c ? a[(i64)c * m + i] : 0.0;
  }
  // Q = H_0 ... H_{k-1} I[:, :k] (dorgqr: reflectors applied backwards)
  double* q = Q + J.qoff;
  for (int e = tid; e < m * k; e += QT) {
    const int c = e / m, i = e - c * m;
    q[e] = i == c ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const double tj = t[j];
    const double* col = a + (i64)j * m;
    if (tj != 0.0)
      for (int c = j + wave; c < k; c += QT / 64) {
        double* y = q + (i64)c * m;
        double d = lane == 0 ? y[j] : 0.0;
        for (int i = j + 1 + lane; i < m; i += 64) d += col[i] * y[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
        const double f = tj * d;
        if (lane == 0) y[j] -= f;
        for (int i = j + 1 + lane; i < m; i += 64) y[i] -= f * col[i];
      }
    __syncthreads();
  }
}

// dst[dmap[t]] = src[smap[t]]
__global__ void copy_map_kernel(i64 n, const i64* __restrict__ smap, const double* __restrict__ src,
                                const i64* __restrict__ dmap, double* __restrict__ dst) {
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[dmap[t]] = src[smap[t]];
}

}  // namespace

void k_copy_map(Ctx* ctx, i64 n, const i64* smap, const double* src, const i64* dmap, double* dst) {
  if (n <= 0) return;
  hipLaunchKernelGGL(copy_map_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n, smap, src,
                     dmap, dst);
  HIP_CHECK(hipGetLastError());
}

void k_qr_batch(Ctx* ctx, const QrJob* jobs, int njobs, double* A, double* Q, double* R, double* tau) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(qr_kernel, dim3((unsigned)njobs), dim3(QT), 0, ctx->stream, jobs, A, Q, R, tau);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sb
