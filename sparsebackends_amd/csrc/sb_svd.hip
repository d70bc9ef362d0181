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

// v + the v of the lane that the DPP pattern CTRL names (both 32-bit halves
// moved by the data-parallel primitives, within a row of 16 lanes)
template <int CTRL>
__device__ inline double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return v + __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// lane l's v, wave-uniform
__device__ inline double read_lane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)lo);
}

// The wave's sum, wave-uniform, for the Jacobi pair step whose latency is its
// reductions: within each row of 16 lanes by DPP (lane ^ 1, lane ^ 2, then the
// half-row and row mirrors, each pairing the sums of two disjoint lane groups;
// partner lanes add the same two values, so a row's lanes hold the same bits),
// then the four row sums read from lanes 0, 16, 32, 48 and added pairwise.
__device__ inline double wave_sum_dpp(double v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141>(v);  // row_half_mirror
  v = dpp_add<0x140>(v);  // row_mirror
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// W rows of group g: row r at woff[g] + r * ncols_g, ncols_g = (gstart[g+1]-gstart[g]) * d2,
// column j*d2 + t = A[bsrc[j] + r + d1*t] (block j of the group, column-major d1 x d2 block)
// Convergence of one-sided Jacobi: a pair of rows counts as orthogonal when
// |x.y| <= tol |x| |y|.  The computed dot product of rows of length len
// carries a rounding error up to about len * eps |x| |y|, so a tolerance
// below that never holds for long rows and every sweep would rotate again;
// tol = max(1e-15, 2 len eps) (eps = 2^-53).
__device__ inline double jacobi_tol(i64 len) { return fmax(1e-15, 2.0 * (double)len * 1.1102230246251565e-16); }
// A row whose norm is below 1e-15 |X|_F (|X|_F is invariant under the
// rotations) is a zero row to working precision: a rank-deficient matrix
// rotates its null rows into rounding noise, whose direction is random, so
// the relative test above would fail for them at every sweep.  Pairs with
// such a row are left alone; the singular values they stand for lie far
// below any truncation cutoff.
__device__ inline double jacobi_null2(double fro2) { return 1e-30 * fro2; }

// One one-sided Jacobi rotation of the rows xp, xq (length len), applied to
// the columns up, uq (length n) of the accumulated rotation; one wave.
// Returns whether it rotated.
__device__ inline bool jacobi_pair(double* xp, double* xq, int len, double* up, double* uq, int n, int lane,
                                   double ctol, double null2) {
  double a = 0, b = 0, c = 0;
  for (int j = lane; j < len; j += 64) {
    const double x = xp[j], y = xq[j];
    a = fma(x, x, a);
    b = fma(y, y, b);
    c = fma(x, y, c);
  }
  a = wave_sum_dpp(a);
  b = wave_sum_dpp(b);
  c = wave_sum_dpp(c);
  if (c == 0.0 || a <= null2 || b <= null2 || fabs(c) <= ctol * sqrt(a) * sqrt(b)) return false;
  const double zeta = (b - a) / (2.0 * c);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
  for (int j = lane; j < len; j += 64) {
    const double x = xp[j], y = xq[j];
    xp[j] = cs * x - sn * y;
    xq[j] = sn * x + cs * y;
  }
  for (int r = lane; r < n; r += 64) {
    const double u = up[r], v = uq[r];
    up[r] = cs * u - sn * v;
    uq[r] = sn * u + cs * v;
  }
  return true;
}

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
  const double ctol = jacobi_tol(nc);
  double fro2 = 0;
  for (i64 e = lane; e < (i64)d1 * nc; e += 64) fro2 = fma(Wg[e], Wg[e], fro2);
  const double null2 = jacobi_null2(wave_sum(fro2));
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
        if (c == 0.0 || a <= null2 || b <= null2 || fabs(c) <= ctol * sqrt(a) * sqrt(b)) continue;
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

// ================================================================ sector SVD
// Block-diagonal SVD of a two-site tensor psi [Ql, Qr | s1, s2, l, r] by fused
// charge sector m: rows (Ql, s1, l) with (Ql, s1) in m, columns (Qr, s2, r).
// Each sector's matrix X (M, or M^T when M has more rows than columns, so X
// has the fewer rows) is factorised by one workgroup with one-sided Jacobi
// in LDS, pairs of rows rotated in parallel in round-robin order.

constexpr int SS_THREADS = 512;

// copy psi sub-blocks into the sector matrices: task t = (psi block offset,
// s1, s2, dl, dr, sector g, row offset, column offset)
__global__ __launch_bounds__(64) void sect_gather_kernel(const double* __restrict__ psi, int d1, int d2, i64 Dl,
                                                        const i64* __restrict__ task, const i64* __restrict__ xoff,
                                                        const int* __restrict__ R, const int* __restrict__ Cc,
                                                        const int* __restrict__ trans, double* __restrict__ X) {
  const i64* t = task + (i64)blockIdx.x * 8;
  const i64 base = t[0];
  const int s1 = (int)t[1], s2 = (int)t[2];
  const i64 dl = t[3], dr = t[4];
  const int g = (int)t[5];
  const i64 ro = t[6], co = t[7];
  const i64 Rg = R[g], Cg = Cc[g];
  const bool tr = trans[g] != 0;
  double* Xg = X + xoff[g];
  for (i64 e = threadIdx.x; e < dl * dr; e += 64) {
    const i64 l = e % dl, r = e / dl;
    const double v = psi[base + s1 + (i64)d1 * s2 + (i64)d1 * d2 * (l + Dl * r)];
    const i64 row = ro + l, col = co + r;
    Xg[tr ? col * Rg + row : row * Cg + col] = v;
  }
}

void k_sect_gather(Ctx* ctx, const double* psi, int d1, int d2, i64 Dl, i64 ntask, const i64* task, const i64* xoff,
                   const int* R, const int* Cc, const int* trans, double* X) {
  if (!ntask) return;
  hipLaunchKernelGGL(sect_gather_kernel, dim3((unsigned)ntask), dim3(64), 0, ctx->stream, psi, d1, d2, Dl, task, xoff,
                     R, Cc, trans, X);
  HIP_CHECK(hipGetLastError());
}

// The triplets of a converged sector (rows of X mutually orthogonal, Q the
// accumulated rotations, nrm / ord scratch of n entries in LDS): sigma = row
// norms in descending order (stable), U (Rg x n, column-major), sigma V^T
// (n x Cg, row-major).  One workgroup of SS_THREADS.
__device__ void sect_emit(const double* Xs, const double* Q, int n, int len, bool tr, int Rg, int Cg, double* nrm,
                          int* ord, double* sig, double* Ug, double* Sg) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = SS_THREADS / 64;
  for (int i = wave; i < n; i += nwave) {
    const double* xi = Xs + (i64)i * len;
    double a = 0;
    for (int j = lane; j < len; j += 64) a = fma(xi[j], xi[j], a);
    a = wave_sum(a);
    if (lane == 0) nrm[i] = sqrt(a);
  }
  __syncthreads();
  if (tid == 0) {
    for (int i = 0; i < n; ++i) ord[i] = i;
    for (int i = 1; i < n; ++i) {  // stable insertion sort, descending
      const int v = ord[i];
      int k = i - 1;
      while (k >= 0 && nrm[ord[k]] < nrm[v]) {
        ord[k + 1] = ord[k];
        --k;
      }
      ord[k + 1] = v;
    }
  }
  __syncthreads();
  for (int c = tid; c < n; c += SS_THREADS) sig[c] = nrm[ord[c]];
  // one wave per output column of U / row of sigma V^T (no index division)
  if (!tr) {  // X = M: U = Q, sigma V^T = rows of X
    for (int c = wave; c < n; c += nwave) {
      const double* qc = Q + (i64)n * ord[c];
      const double* xc = Xs + (i64)ord[c] * len;
      for (int r = lane; r < Rg; r += 64) Ug[r + (i64)Rg * c] = qc[r];
      for (int j = lane; j < Cg; j += 64) Sg[(i64)c * Cg + j] = xc[j];
    }
  } else {  // X = M^T: U columns = rows of X / sigma, sigma V^T row c = sigma * column ord[c] of Q
    for (int c = wave; c < n; c += nwave) {
      const double s = nrm[ord[c]];
      const double* qc = Q + (i64)n * ord[c];
      const double* xc = Xs + (i64)ord[c] * len;
      for (int r = lane; r < Rg; r += 64) Ug[r + (i64)Rg * c] = s > 0 ? xc[r] / s : 0.0;
      for (int j = lane; j < Cg; j += 64) Sg[(i64)c * Cg + j] = s * qc[j];
    }
  }
}

// |X|_F^2 of the n x len rows at Xg (one workgroup; every thread gets it)
__device__ double sect_fro2(const double* Xg, i64 cnt) {
  __shared__ double part[SS_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double f = 0;
  for (i64 e = tid; e < cnt; e += SS_THREADS) f = fma(Xg[e], Xg[e], f);
  f = wave_sum(f);
  if (lane == 0) part[wave] = f;
  __syncthreads();
  double fro2 = 0;
  for (int w = 0; w < SS_THREADS / 64; ++w) fro2 += part[w];
  __syncthreads();
  return fro2;
}

// Sectors whose rows and rotations fit the LDS of a workgroup: one workgroup
// per sector (glist[blockIdx.x]) runs the whole one-sided Jacobi in LDS.
__global__ __launch_bounds__(SS_THREADS) void sect_svd_kernel(const double* __restrict__ X, const i64* __restrict__ xoff,
                                                             const int* __restrict__ R, const int* __restrict__ Cc,
                                                             const int* __restrict__ trans,
                                                             const i64* __restrict__ uoff, const i64* __restrict__ soff,
                                                             const i64* __restrict__ sgoff, double* __restrict__ sig,
                                                             double* __restrict__ Uo, double* __restrict__ SVo,
                                                             const int* __restrict__ glist) {
  extern __shared__ double lds[];
  const int g = glist[blockIdx.x];
  const bool tr = trans[g] != 0;
  const int Rg = R[g], Cg = Cc[g];
  const int n = tr ? Cg : Rg, len = tr ? Rg : Cg;  // rows of X, row length
  const int n2 = n + (n & 1), m = n2 - 1, np = n2 / 2;  // round-robin pairs per round
  double* Xs = lds;
  double* Q = Xs + (i64)n * len;  // Q[r + n*c]
  double* nrm = Q + (i64)n * n;
  int* ord = reinterpret_cast<int*>(nrm + n);
  int* flag = ord + n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = SS_THREADS / 64;
  const double* Xg = X + xoff[g];
  for (i64 e = tid; e < (i64)n * len; e += SS_THREADS) Xs[e] = Xg[e];
  for (int c = wave; c < n; c += nwave)
    for (int r = lane; r < n; r += 64) Q[r + (i64)n * c] = r == c ? 1.0 : 0.0;
  const double null2 = jacobi_null2(sect_fro2(Xg, (i64)n * len));
  const double ctol = jacobi_tol(len);
  for (int sweep = 0; sweep < 40 && n > 1; ++sweep) {
    if (tid == 0) *flag = 0;
    __syncthreads();
    for (int rd = 0; rd < m; ++rd) {
      for (int pi = wave; pi < np; pi += nwave) {
        const int p = pi == 0 ? 0 : 1 + (rd + pi) % m;
        const int q = pi == 0 ? 1 + rd : 1 + (rd - pi + m) % m;
        if (p >= n || q >= n) continue;
        if (jacobi_pair(Xs + (i64)p * len, Xs + (i64)q * len, len, Q + (i64)n * p, Q + (i64)n * q, n, lane, ctol,
                        null2) &&
            lane == 0)
          *flag = 1;
      }
      __syncthreads();
    }
    const int f = *flag;
    __syncthreads();
    if (!f) break;
  }
  sect_emit(Xs, Q, n, len, tr, Rg, Cg, nrm, ord, sig + sgoff[g], Uo + uoff[g], SVo + soff[g]);
}

size_t sect_svd_lds(int n, int len) {  // LDS bytes of a sector in sect_svd_kernel
  return ((size_t)n * len + (size_t)n * n + n) * sizeof(double) + (size_t)(n + 1) * sizeof(int);
}

void k_sect_svd(Ctx* ctx, i64 ng, size_t lds, const double* X, const i64* xoff, const int* R, const int* Cc,
                const int* trans, const i64* uoff, const i64* soff, const i64* sgoff, double* sig, double* Uo,
                double* SVo, const int* glist) {
  if (!ng) return;
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_svd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(sect_svd_kernel, dim3((unsigned)ng), dim3(SS_THREADS), lds, ctx->stream, X, xoff, R, Cc, trans,
                     uoff, soff, sgoff, sig, Uo, SVo, glist);
  HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- block Jacobi
// Sectors with 32 rows or more to orthogonalise, or too large for one
// workgroup's LDS: block one-sided Jacobi over many workgroups.  The n rows of a sector's X (in place in the gathered sector
// matrix) are cut into B blocks of bs (8) rows (B even, the last ones short or
// empty); a sweep is B - 1 rounds of the round-robin pairing of blocks, one
// launch per round over every large sector, one workgroup per block pair: it
// brings the pair's rows and the matching columns of Q into LDS, rotates the
// row pairs of the sweep that fall to it, and writes both back.  Every row pair
// is visited once per sweep: in the first round (where every block meets one
// partner) a workgroup sweeps all pairs of its rows, in the later rounds only
// the pairs across its two blocks (a row of one block with a row of the
// other), in max(rows of the two blocks) steps of one pair per wave.  The host
// reads the sectors' rotation flags after each sweep.

struct BjSector {  // one large sector
  int g;           // sector number
  int n, len, bs, B;
  i64 qoff;        // offset of its Q in the workspace
};

__global__ __launch_bounds__(SS_THREADS) void sect_bj_init_kernel(const BjSector* __restrict__ sec,
                                                                 const double* __restrict__ X,
                                                                 const i64* __restrict__ xoff, double* __restrict__ Qw,
                                                                 double* __restrict__ null2, int* __restrict__ flag) {
  const BjSector s = sec[blockIdx.x];
  double* Q = Qw + s.qoff;
  for (int c = threadIdx.x >> 6; c < s.n; c += SS_THREADS / 64)
    for (int r = threadIdx.x & 63; r < s.n; r += 64) Q[r + (i64)s.n * c] = r == c ? 1.0 : 0.0;
  const double f = sect_fro2(X + xoff[s.g], (i64)s.n * s.len);
  if (threadIdx.x == 0) {
    null2[blockIdx.x] = jacobi_null2(f);
    flag[blockIdx.x] = 0;
  }
}

// eight waves: one row pair per wave at every step of a pair of 8-row blocks
// (measured: 16-row blocks on sixteen waves take 25 % longer per sweep of the
// chi = 256 Heisenberg chain's sectors, though with half the launches)
constexpr int BJ_THREADS = 512;

__global__ __launch_bounds__(BJ_THREADS) void sect_bj_round_kernel(const BjSector* __restrict__ sec, int maxpairs,
                                                                  int rd, double* __restrict__ X,
                                                                  const i64* __restrict__ xoff, double* __restrict__ Qw,
                                                                  const double* __restrict__ null2,
                                                                  int* __restrict__ flag) {
  extern __shared__ double lds[];
  const int b = blockIdx.x / maxpairs, pi = blockIdx.x % maxpairs;
  const BjSector s = sec[b];
  const int mB = s.B - 1;
  if (pi >= s.B / 2 || rd >= mB) return;
  const int pb = pi == 0 ? 0 : 1 + (rd + pi) % mB;
  const int qb = pi == 0 ? 1 + rd : 1 + (rd - pi + mB) % mB;
  // the pair's rows: block pb then block qb (either may be short or empty)
  const int p0 = pb * s.bs, p1 = min(s.n, p0 + s.bs), q0 = qb * s.bs, q1 = min(s.n, q0 + s.bs);
  const int np_ = max(0, p1 - p0), nq_ = max(0, q1 - q0), nl = np_ + nq_;
  if (nl < 2) return;
  const int n = s.n, len = s.len;
  double* Xg = X + xoff[s.g];
  double* Q = Qw + s.qoff;
  double* Xl = lds;                    // nl rows of len
  double* Ql = Xl + (i64)nl * len;     // nl columns of n
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = BJ_THREADS / 64;
  auto grow = [&](int l) { return l < np_ ? p0 + l : q0 + (l - np_); };
  // one wave per row of the pair (no index division)
  for (int l = wave; l < nl; l += nwave) {
    const double* xg = Xg + (i64)grow(l) * len;
    const double* qg = Q + (i64)n * grow(l);
    for (int j = lane; j < len; j += 64) Xl[(i64)l * len + j] = xg[j];
    for (int r = lane; r < n; r += 64) Ql[(i64)l * n + r] = qg[r];
  }
  __syncthreads();
  const double ctol = jacobi_tol(len), nu = null2[b];
  bool any = false;
  if (rd == 0) {  // every pair of the 2 bs rows (round-robin)
    const int n2 = nl + (nl & 1), m = n2 - 1, npair = n2 / 2;
    for (int r = 0; r < m; ++r) {
      for (int qi = wave; qi < npair; qi += nwave) {
        const int p = qi == 0 ? 0 : 1 + (r + qi) % m;
        const int q = qi == 0 ? 1 + r : 1 + (r - qi + m) % m;
        if (p >= nl || q >= nl) continue;
        if (jacobi_pair(Xl + (i64)p * len, Xl + (i64)q * len, len, Ql + (i64)n * p, Ql + (i64)n * q, n, lane, ctol,
                        nu))
          any = true;
      }
      __syncthreads();
    }
  } else {  // the pairs across the blocks: step k pairs row i of pb with row (i + k) mod L of qb
    const int L = max(np_, nq_);
    for (int k = 0; k < L; ++k) {
      for (int i = wave; i < L; i += nwave) {
        const int j = (i + k) % L;
        if (i >= np_ || j >= nq_) continue;
        const int p = i, q = np_ + j;
        if (jacobi_pair(Xl + (i64)p * len, Xl + (i64)q * len, len, Ql + (i64)n * p, Ql + (i64)n * q, n, lane, ctol,
                        nu))
          any = true;
      }
      __syncthreads();
    }
  }
  for (int l = wave; l < nl; l += nwave) {
    double* xg = Xg + (i64)grow(l) * len;
    double* qg = Q + (i64)n * grow(l);
    for (int j = lane; j < len; j += 64) xg[j] = Xl[(i64)l * len + j];
    for (int r = lane; r < n; r += 64) qg[r] = Ql[(i64)l * n + r];
  }
  if (any && lane == 0) flag[b] = 1;
}

__global__ __launch_bounds__(SS_THREADS) void sect_bj_finish_kernel(const BjSector* __restrict__ sec,
                                                                   const double* __restrict__ X,
                                                                   const i64* __restrict__ xoff,
                                                                   const double* __restrict__ Qw,
                                                                   const int* __restrict__ R, const int* __restrict__ Cc,
                                                                   const int* __restrict__ trans,
                                                                   const i64* __restrict__ uoff,
                                                                   const i64* __restrict__ soff,
                                                                   const i64* __restrict__ sgoff,
                                                                   double* __restrict__ sig, double* __restrict__ Uo,
                                                                   double* __restrict__ SVo) {
  extern __shared__ double lds[];
  const BjSector s = sec[blockIdx.x];
  const int g = s.g;
  double* nrm = lds;
  int* ord = reinterpret_cast<int*>(nrm + s.n);
  sect_emit(X + xoff[g], Qw + s.qoff, s.n, s.len, trans[g] != 0, R[g], Cc[g], nrm, ord, sig + sgoff[g],
            Uo + uoff[g], SVo + soff[g]);
}

// rows per block of a large sector: the pair's rows and Q columns fit the LDS
int sect_bj_rows(int n, int len) {
  const size_t lim = 159 * 1024;
  // eight rows per block, one row pair per wave of the workgroup at every
  // step; fewer where the pair's rows and Q columns would not fit the LDS
  int bs = BJ_THREADS / 64;
  while (bs > 1 && (size_t)2 * bs * (len + n) * sizeof(double) > lim) --bs;
  return bs;
}

// The block Jacobi of the large sectors gl (sector numbers), their (n, len):
// factorises them in place in X (workspace for Q allocated here) and writes
// their triplets like k_sect_svd.
void k_sect_svd_blocks(Ctx* ctx, const std::vector<int>& gl, const std::vector<int>& nn,
                       const std::vector<int>& ll, double* X, const i64* xoff, const int* R, const int* Cc,
                       const int* trans, const i64* uoff, const i64* soff, const i64* sgoff, double* sig, double* Uo,
                       double* SVo) {
  const int nb = (int)gl.size();
  if (!nb) return;
  std::vector<BjSector> hs(nb);
  i64 qtot = 0;
  int maxpairs = 1, maxrounds = 0, maxn = 1;
  size_t lds = 0;
  for (int b = 0; b < nb; ++b) {
    BjSector& s = hs[b];
    s.g = gl[b];
    s.n = nn[b];
    s.len = ll[b];
    s.bs = sect_bj_rows(s.n, s.len);
    s.B = (s.n + s.bs - 1) / s.bs;
    s.B += s.B & 1;
    if (s.B < 2) s.B = 2;
    s.qoff = qtot;
    qtot += (i64)s.n * s.n;
    maxpairs = std::max(maxpairs, s.B / 2);
    maxrounds = std::max(maxrounds, s.B - 1);
    maxn = std::max(maxn, s.n);
    lds = std::max(lds, (size_t)2 * s.bs * (s.len + s.n) * sizeof(double));
  }
  DBuf<BjSector> d_sec(ctx, nb);
  HIP_CHECK(hipMemcpyAsync(d_sec.get(), hs.data(), nb * sizeof(BjSector), hipMemcpyHostToDevice, ctx->stream));
  DBuf<double> Qw(ctx, (size_t)qtot), nul(ctx, nb);
  DBuf<int> flag(ctx, nb);
  hipLaunchKernelGGL(sect_bj_init_kernel, dim3(nb), dim3(SS_THREADS), 0, ctx->stream, d_sec.get(), X, xoff, Qw.get(),
                     nul.get(), flag.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_bj_round_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  std::vector<int> hf(nb);
  for (int sweep = 0; sweep < 40; ++sweep) {
    for (int rd = 0; rd < maxrounds; ++rd) {
      hipLaunchKernelGGL(sect_bj_round_kernel, dim3((unsigned)(nb * maxpairs)), dim3(BJ_THREADS), lds, ctx->stream,
                         d_sec.get(), maxpairs, rd, X, xoff, Qw.get(), nul.get(), flag.get());
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipMemcpyAsync(hf.data(), flag.get(), nb * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    bool any = false;
    for (int f : hf) any = any || f;
    if (!any) break;
    k_zero(ctx, flag.get(), nb * sizeof(int));
  }
  const size_t flds = (size_t)maxn * (sizeof(double) + sizeof(int));
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_bj_finish_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)flds));
  hipLaunchKernelGGL(sect_bj_finish_kernel, dim3(nb), dim3(SS_THREADS), flds, ctx->stream, d_sec.get(), X, xoff,
                     Qw.get(), R, Cc, trans, uoff, soff, sgoff, sig, Uo, SVo);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the workspace and the sector table go out of scope
}

// factors of the kept triplets.  A block b: (il, sector g) -> [s1, l, mm]
// (d1 x Dl x Dm); B block b: (g, ir) -> [s2, mm, r] (d2 x Dm x Dr).
__global__ void sect_emit_A_kernel(i64 nblk, int d1, i64 Dl, i64 Dm, i64 nl, const i64* __restrict__ blk,
                                   const i64* __restrict__ rowsec, const i64* __restrict__ rowoff,
                                   const i64* __restrict__ dl, const i64* __restrict__ kept, const int* __restrict__ R,
                                   const i64* __restrict__ uoff, const i64* __restrict__ sgoff,
                                   const double* __restrict__ sig, const double* __restrict__ Uo, int right,
                                   double* __restrict__ out) {
  const i64 bsz = (i64)d1 * Dl * Dm;
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * bsz) return;
  const i64 b = t / bsz, e = t - b * bsz;
  const i64 il = blk[2 * b], g = blk[2 * b + 1];
  const int s1 = (int)(e % d1);
  const i64 l = (e / d1) % Dl, mm = e / (d1 * Dl);
  double v = 0.0;
  if (rowsec[il + nl * s1] == g && l < dl[il] && mm < kept[g]) {
    v = Uo[uoff[g] + rowoff[il + nl * s1] + l + (i64)R[g] * mm];
    if (right) v *= sig[sgoff[g] + mm];
  }
  out[t] = v;
}

__global__ void sect_emit_B_kernel(i64 nblk, int d2, i64 Dm, i64 Dr, i64 nr, const i64* __restrict__ blk,
                                   const i64* __restrict__ colsec, const i64* __restrict__ coloff,
                                   const i64* __restrict__ dr, const i64* __restrict__ kept, const int* __restrict__ Cc,
                                   const i64* __restrict__ soff, const i64* __restrict__ sgoff,
                                   const double* __restrict__ sig, const double* __restrict__ SVo, int right,
                                   double* __restrict__ out) {
  const i64 bsz = (i64)d2 * Dm * Dr;
  const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblk * bsz) return;
  const i64 b = t / bsz, e = t - b * bsz;
  const i64 g = blk[2 * b], ir = blk[2 * b + 1];
  const int s2 = (int)(e % d2);
  const i64 mm = (e / d2) % Dm, r = e / (d2 * Dm);
  double v = 0.0;
  if (colsec[ir + nr * s2] == g && r < dr[ir] && mm < kept[g]) {
    v = SVo[soff[g] + mm * Cc[g] + coloff[ir + nr * s2] + r];
    if (right) {
      const double s = sig[sgoff[g] + mm];
      v = s > 0 ? v / s : 0.0;
    }
  }
  out[t] = v;
}

void k_sect_emit(Ctx* ctx, bool isA, i64 nblk, int d, i64 D1, i64 D2, i64 nside, const i64* blk, const i64* sec,
                 const i64* off, const i64* dims, const i64* kept, const int* RC, const i64* uoff_or_soff,
                 const i64* sgoff, const double* sig, const double* F, bool right, double* out) {
  const i64 tot = nblk * d * D1 * D2;
  if (!tot) return;
  const unsigned grid = (unsigned)((tot + 255) / 256);
  if (isA)
    hipLaunchKernelGGL(sect_emit_A_kernel, dim3(grid), dim3(256), 0, ctx->stream, nblk, d, D1, D2, nside, blk, sec,
                       off, dims, kept, RC, uoff_or_soff, sgoff, sig, F, right ? 1 : 0, out);
  else
    hipLaunchKernelGGL(sect_emit_B_kernel, dim3(grid), dim3(256), 0, ctx->stream, nblk, d, D1, D2, nside, blk, sec,
                       off, dims, kept, RC, uoff_or_soff, sgoff, sig, F, right ? 1 : 0, out);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sb
