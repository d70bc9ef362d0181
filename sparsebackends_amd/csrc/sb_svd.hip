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
                                                       int* __restrict__ perm, int max_sweeps,
                                                       int* __restrict__ noconv) {
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
  bool conv = false;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
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
    if (!rotated) {
      conv = true;
      break;
    }
  }
  if (!conv && d1 > 1 && lane == 0) atomicOr(noconv, 1);
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
                  const i64* woff, double* W, double* sig, double* Uout, int* perm, int* noconv) {
  if (!ng) return;
  SB_ASSERT(d1 >= 1 && d1 <= SVD_MAXD, "blocksparse_svd: first dense dim above 32");
  hipLaunchKernelGGL(svd_jacobi_kernel, dim3((unsigned)ng), dim3(64), 0, ctx->stream, A, d1, d2, gstart, bsrc, woff,
                     W, sig, Uout, perm, ctx->svd_sweeps, noconv);
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
  // stable sort, descending: row i goes to its rank, the number of rows with a
  // larger norm plus those of equal norm before it (a NaN ranks as -inf, so
  // the ranks stay a permutation)
  for (int i = tid; i < n; i += SS_THREADS) {
    const double v = isnan(nrm[i]) ? -INFINITY : nrm[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double u = isnan(nrm[j]) ? -INFINITY : nrm[j];
      rank += (u > v) || (u == v && j < i);
    }
    ord[rank] = i;
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
                                                             const int* __restrict__ glist, int max_sweeps, int* __restrict__ noconv) {
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
  bool conv = n <= 1;
  for (int sweep = 0; sweep < max_sweeps && n > 1; ++sweep) {
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
    if (!f) {
      conv = true;
      break;
    }
  }
  if (!conv && tid == 0) atomicOr(noconv, 1);
  sect_emit(Xs, Q, n, len, tr, Rg, Cg, nrm, ord, sig + sgoff[g], Uo + uoff[g], SVo + soff[g]);
}

size_t sect_svd_lds(int n, int len) {  // LDS bytes of a sector in sect_svd_kernel
  return ((size_t)n * len + (size_t)n * n + n) * sizeof(double) + (size_t)(n + 1) * sizeof(int);
}

void k_sect_svd(Ctx* ctx, i64 ng, size_t lds, const double* X, const i64* xoff, const int* R, const int* Cc,
                const int* trans, const i64* uoff, const i64* soff, const i64* sgoff, double* sig, double* Uo,
                double* SVo, const int* glist, int* noconv) {
  if (!ng) return;
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_svd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(sect_svd_kernel, dim3((unsigned)ng), dim3(SS_THREADS), lds, ctx->stream, X, xoff, R, Cc, trans,
                     uoff, soff, sgoff, sig, Uo, SVo, glist, ctx->svd_sweeps, noconv);
  HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- block Jacobi
// Sectors with 32 rows or more to orthogonalise, or too large for one
// workgroup's LDS: block one-sided Jacobi over many workgroups.  The n rows of
// a sector's X (in place in the gathered sector matrix) are cut into B blocks
// of bs rows (B even, the last ones short or empty); a sweep is B - 1 rounds of
// the round-robin pairing of blocks, one launch per round over every large
// sector, one workgroup per block pair: it brings the pair's rows into LDS,
// rotates the row pairs of the sweep that fall to it, and writes them back.
// Every row pair is visited once per sweep: in the first round (where every
// block meets one partner) a workgroup sweeps all pairs of its rows, in the
// later rounds only the pairs across its two blocks (a row of one block with a
// row of the other), in max(rows of the two blocks) steps of one pair per wave.
//
// The accumulated rotation Q (n x n, column c rotated like row c of X) stays
// out of the rounds: a workgroup records the product of its rotations as the
// (2 bs) x (2 bs) matrix T (row j of the new rows = sum_i T[j][i] old row i)
// and, once per sweep, one launch applies every round's T to Q in round order,
// each workgroup on a slice of Q's entries (the rotations act on the column
// index only, so slices are independent).  The rounds then move only X's rows
// and fit larger blocks.  The host reads the sectors' rotation flags after each
// sweep.

struct BjSector {  // one large sector
  int g;           // sector number
  int n, len, bs, B;
  i64 qoff;        // offset of its Q in the workspace
  i64 toff;        // offset of its round matrices T (one sweep)
  int foff;        // offset of its (round, pair) rotation flags
  int S;           // Q entries per workgroup of the apply launch
  int slice0;      // its first workgroup in the apply launch
};

// the Jacobi of a set of sectors: its sector table and accumulated rotations
struct BjRun {
  std::vector<BjSector> hs;
  DBuf<BjSector> d_sec;
  DBuf<double> Qw;
  bool unconverged = false;  // the sweep cap was reached with rows still rotating
};

constexpr int BJ_MAXBS = 8;  // rows per block: eight waves, one row pair each per step

// the two blocks of pair pi in round rd of the round-robin pairing of B blocks
__device__ inline void bj_blocks(int B, int rd, int pi, int& pb, int& qb) {
  const int mB = B - 1;
  pb = pi == 0 ? 0 : 1 + (rd + pi) % mB;
  qb = pi == 0 ? 1 + rd : 1 + (rd - pi + mB) % mB;
}

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

constexpr int BJ_THREADS = 64 * BJ_MAXBS;

// jacobi_pair with the rotation recorded in the rows tp, tq (length nt) of T
__device__ inline bool jacobi_pair_t(double* xp, double* xq, int len, double* tp, double* tq, int nt, int lane,
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
  if (lane < nt) {
    const double u = tp[lane], v = tq[lane];
    tp[lane] = cs * u - sn * v;
    tq[lane] = sn * u + cs * v;
  }
  return true;
}

__global__ __launch_bounds__(BJ_THREADS) void sect_bj_round_kernel(const BjSector* __restrict__ sec, int maxpairs,
                                                                  int rd, double* __restrict__ X,
                                                                  const i64* __restrict__ xoff,
                                                                  const double* __restrict__ null2,
                                                                  int* __restrict__ flag, double* __restrict__ Tb,
                                                                  int* __restrict__ tflag) {
  extern __shared__ double lds[];
  __shared__ int wg_any;
  const int b = blockIdx.x / maxpairs, pi = blockIdx.x % maxpairs;
  const BjSector s = sec[b];
  const int mB = s.B - 1;
  if (pi >= s.B / 2 || rd >= mB) return;
  int pb, qb;
  bj_blocks(s.B, rd, pi, pb, qb);
  // the pair's rows: block pb then block qb (either may be short or empty)
  const int p0 = pb * s.bs, p1 = min(s.n, p0 + s.bs), q0 = qb * s.bs, q1 = min(s.n, q0 + s.bs);
  const int np_ = max(0, p1 - p0), nq_ = max(0, q1 - q0), nl = np_ + nq_;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = BJ_THREADS / 64;
  int* tf = tflag + s.foff + rd * (s.B / 2) + pi;
  if (nl < 2) {  // nothing to rotate: the apply launch skips this pair
    if (tid == 0) *tf = 0;
    return;
  }
  const int len = s.len, td = 2 * s.bs;
  double* Xg = X + xoff[s.g];
  double* Xl = lds;                 // nl rows of len
  double* Tl = Xl + (i64)nl * len;  // T[j * td + i]
  auto grow = [&](int l) { return l < np_ ? p0 + l : q0 + (l - np_); };
  // one wave per row of the pair (no index division)
  for (int l = wave; l < nl; l += nwave) {
    const double* xg = Xg + (i64)grow(l) * len;
    for (int j = lane; j < len; j += 64) Xl[(i64)l * len + j] = xg[j];
  }
  for (int e = tid; e < td * td; e += BJ_THREADS) Tl[e] = (e / td == e % td) ? 1.0 : 0.0;
  if (tid == 0) wg_any = 0;
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
        if (jacobi_pair_t(Xl + (i64)p * len, Xl + (i64)q * len, len, Tl + p * td, Tl + q * td, nl, lane, ctol, nu))
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
        if (jacobi_pair_t(Xl + (i64)p * len, Xl + (i64)q * len, len, Tl + p * td, Tl + q * td, nl, lane, ctol, nu))
          any = true;
      }
      __syncthreads();
    }
  }
  if (any && lane == 0) wg_any = 1;
  for (int l = wave; l < nl; l += nwave) {
    double* xg = Xg + (i64)grow(l) * len;
    for (int j = lane; j < len; j += 64) xg[j] = Xl[(i64)l * len + j];
  }
  double* tg = Tb + s.toff + (i64)(rd * (s.B / 2) + pi) * td * td;
  for (int e = tid; e < td * td; e += BJ_THREADS) tg[e] = Tl[e];
  __syncthreads();
  if (tid == 0) {
    *tf = wg_any;
    if (wg_any) flag[b] = 1;
  }
}

// The later rounds (rd >= 1) of rows up to 64 CH long: wave i keeps row i of
// block pb in registers for the whole round and meets row (i + k) mod L of
// block qb (in LDS) at step k, so a rotation reads and writes only the LDS row
// (the LDS traffic per rotation, which bounds a step, falls from six row
// passes to two).  Same pairs, same order and same arithmetic as the rd >= 1
// branch of sect_bj_round_kernel.
template <int CH>
__global__ __launch_bounds__(BJ_THREADS) void sect_bj_cross_kernel(const BjSector* __restrict__ sec, int maxpairs,
                                                                  int rd, double* __restrict__ X,
                                                                  const i64* __restrict__ xoff,
                                                                  const double* __restrict__ null2,
                                                                  int* __restrict__ flag, double* __restrict__ Tb,
                                                                  int* __restrict__ tflag) {
  extern __shared__ double lds[];
  __shared__ int wg_any;
  const int b = blockIdx.x / maxpairs, pi = blockIdx.x % maxpairs;
  const BjSector s = sec[b];
  const int mB = s.B - 1;
  if (pi >= s.B / 2 || rd >= mB) return;
  int pb, qb;
  bj_blocks(s.B, rd, pi, pb, qb);
  const int p0 = pb * s.bs, p1 = min(s.n, p0 + s.bs), q0 = qb * s.bs, q1 = min(s.n, q0 + s.bs);
  const int np_ = max(0, p1 - p0), nq_ = max(0, q1 - q0), nl = np_ + nq_;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = BJ_THREADS / 64;
  int* tf = tflag + s.foff + rd * (s.B / 2) + pi;
  if (np_ == 0 || nq_ == 0) {  // no pair across the blocks: the apply launch skips this pair
    if (tid == 0) *tf = 0;
    return;
  }
  const int len = s.len, td = 2 * s.bs;
  double* Xg = X + xoff[s.g];
  double* Yl = lds;                  // the nq_ rows of block qb
  double* Tl = Yl + (i64)nq_ * len;  // T[j * td + i], rows 0..np_-1 of pb, np_.. of qb
  for (int l = wave; l < nq_; l += nwave) {
    const double* yg = Xg + (i64)(q0 + l) * len;
    for (int j = lane; j < len; j += 64) Yl[(i64)l * len + j] = yg[j];
  }
  double xr[CH];
  const bool hasp = wave < np_;  // np_ <= bs <= nwave
  const double* xg = Xg + (i64)(p0 + wave) * len;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int j = lane + 64 * c;
    xr[c] = hasp && j < len ? xg[j] : 0.0;
  }
  for (int e = tid; e < td * td; e += BJ_THREADS) Tl[e] = (e / td == e % td) ? 1.0 : 0.0;
  if (tid == 0) wg_any = 0;
  __syncthreads();
  const double ctol = jacobi_tol(len), nu = null2[b];
  bool any = false;
  const int L = max(np_, nq_);
  for (int k = 0; k < L; ++k) {
    const int j = (wave + k) % L;
    if (hasp && j < nq_) {
      double* y = Yl + (i64)j * len;
      double yr[CH];
      double a = 0, bb = 0, cc = 0;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int e = lane + 64 * c;
        yr[c] = e < len ? y[e] : 0.0;
        a = fma(xr[c], xr[c], a);
        bb = fma(yr[c], yr[c], bb);
        cc = fma(xr[c], yr[c], cc);
      }
      a = wave_sum_dpp(a);
      bb = wave_sum_dpp(bb);
      cc = wave_sum_dpp(cc);
      if (!(cc == 0.0 || a <= nu || bb <= nu || fabs(cc) <= ctol * sqrt(a) * sqrt(bb))) {
        const double zeta = (bb - a) / (2.0 * cc);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int e = lane + 64 * c;
          const double x = xr[c], yv = yr[c];
          xr[c] = cs * x - sn * yv;
          if (e < len) y[e] = sn * x + cs * yv;
        }
        double* tp = Tl + wave * td;
        double* tq = Tl + (np_ + j) * td;
        if (lane < nl) {
          const double u = tp[lane], v = tq[lane];
          tp[lane] = cs * u - sn * v;
          tq[lane] = sn * u + cs * v;
        }
        any = true;
      }
    }
    __syncthreads();
  }
  if (any && lane == 0) wg_any = 1;
  if (hasp) {
    double* xo = Xg + (i64)(p0 + wave) * len;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int e = lane + 64 * c;
      if (e < len) xo[e] = xr[c];
    }
  }
  for (int l = wave; l < nq_; l += nwave) {
    double* yg = Xg + (i64)(q0 + l) * len;
    for (int jj = lane; jj < len; jj += 64) yg[jj] = Yl[(i64)l * len + jj];
  }
  double* tg = Tb + s.toff + (i64)(rd * (s.B / 2) + pi) * td * td;
  for (int e = tid; e < td * td; e += BJ_THREADS) tg[e] = Tl[e];
  __syncthreads();
  if (tid == 0) {
    *tf = wg_any;
    if (wg_any) flag[b] = 1;
  }
}

// Q <- the sweep's rotations: for every round in order, the rows of each pair
// that rotated get T times their old values, on this workgroup's slice of S
// entries of every row of Q (Q row c = column c of the accumulated rotation).
// The round's T matrices are staged in LDS; one wave per pair forms the new
// rows (up to 16) x (up to 32 entries) as two 16 x 16 tiles on the matrix
// cores (T's rows against the old rows, K = 16 in four steps).
constexpr int BJ_APPLY_THREADS = 1024;
typedef double bj_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(BJ_APPLY_THREADS) void sect_bj_apply_kernel(const BjSector* __restrict__ sec, int nsec,
                                                                        double* __restrict__ Qw,
                                                                        const double* __restrict__ Tb,
                                                                        const int* __restrict__ tflag) {
  extern __shared__ double lds[];
  int b = 0;
  while (b + 1 < nsec && sec[b + 1].slice0 <= (int)blockIdx.x) ++b;
  const BjSector s = sec[b];
  const int n = s.n, S = s.S, td = 2 * s.bs, half = s.B / 2;
  const int r0 = ((int)blockIdx.x - s.slice0) * S, rs = min(S, n - r0);
  if (rs <= 0) return;
  double* Q = Qw + s.qoff;
  double* Qs = lds;                  // Qs[c * S + rl]: entries r0 + rl of Q row c
  double* Ts = Qs + (i64)n * S;      // the round's T matrices, half x td x td
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = BJ_APPLY_THREADS / 64;
  for (int e = tid; e < n * S; e += BJ_APPLY_THREADS) {
    const int c = e / S, rl = e - c * S;
    if (rl < rs) Qs[e] = Q[(i64)n * c + r0 + rl];
  }
  const int j = lane & 15, kk = lane >> 4;
  for (int rd = 0; rd < s.B - 1; ++rd) {
    const double* Tg = Tb + s.toff + (i64)rd * half * td * td;
    for (int e = tid; e < half * td * td; e += BJ_APPLY_THREADS) Ts[e] = Tg[e];
    __syncthreads();
    for (int pi = wave; pi < half; pi += nwave) {
      if (!tflag[s.foff + rd * half + pi]) continue;
      int pb, qb;
      bj_blocks(s.B, rd, pi, pb, qb);
      const int p0 = pb * s.bs, p1 = min(n, p0 + s.bs), q0 = qb * s.bs, q1 = min(n, q0 + s.bs);
      const int np_ = max(0, p1 - p0), nq_ = max(0, q1 - q0), nl = np_ + nq_;
      const double* T = Ts + pi * td * td;
      bj_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int l = 4 * k + kk;  // A[j][l] = T[j][l], B[l][col] = old row l, entry col of the tile
        const double a = (j < nl && l < nl) ? T[j * td + l] : 0.0;
        const int row = l < nl ? (l < np_ ? p0 + l : q0 + (l - np_)) : -1;
        const double b0 = (row >= 0 && j < rs) ? Qs[row * S + j] : 0.0;
        const double b1 = (row >= 0 && 16 + j < rs) ? Qs[row * S + 16 + j] : 0.0;
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc1, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {  // register v holds new row 4 v + lane / 16, entry lane & 15 of the tile
        const int jr = 4 * v + kk;
        if (jr >= nl) continue;
        const int row = jr < np_ ? p0 + jr : q0 + (jr - np_);
        if (j < rs) Qs[row * S + j] = acc0[v];
        if (16 + j < rs) Qs[row * S + 16 + j] = acc1[v];
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < n * S; e += BJ_APPLY_THREADS) {
    const int c = e / S, rl = e - c * S;
    if (rl < rs) Q[(i64)n * c + r0 + rl] = Qs[e];
  }
}

constexpr size_t kBjLds = 159 * 1024;

// rows per block of a large sector: the pair's rows and its T fit the LDS
int sect_bj_rows(int n, int len) {
  (void)n;
  int bs = BJ_MAXBS;
  while (bs > 1 && (size_t)(2 * bs * (size_t)len + 4 * bs * bs) * sizeof(double) > kBjLds) --bs;
  return bs;
}

// Q entries per workgroup of the apply launch (at most 32, two matrix-core
// tiles): n rows of S entries and one round's T matrices in LDS
static size_t sect_bj_apply_lds(int n, int S, int B, int bs) {
  return ((size_t)n * S + (size_t)(B / 2) * 4 * bs * bs) * sizeof(double);
}
static int sect_bj_slice(int n, int B, int bs) {
  int S = 32;
  while (S > 1 && sect_bj_apply_lds(n, S, B, bs) > kBjLds) S >>= 1;
  return S;
}

// The block Jacobi of the sectors gl (indices into xoff), their (n rows, len):
// orthogonalises their rows in place in X and leaves the accumulated rotations
// (Q[r + n c], c = the row it follows) in run.Qw at run.hs[b].qoff.
void bj_jacobi(Ctx* ctx, const std::vector<int>& gl, const std::vector<int>& nn, const std::vector<int>& ll,
               double* X, const i64* xoff, BjRun& run) {
  const int nb = (int)gl.size();
  if (!nb) return;
  std::vector<BjSector>& hs = run.hs;
  hs.assign(nb, BjSector{});
  i64 qtot = 0, ttot = 0;
  int ftot = 0, slices = 0;
  int maxpairs = 1, maxrounds = 0;
  size_t lds = 0, alds = 0;
  for (int b = 0; b < nb; ++b) {
    BjSector& s = hs[b];
    s.g = gl[b];
    s.n = nn[b];
    s.len = ll[b];
    if (s.n > (1 << 20) || s.len > (1 << 24))
      fail(SB_E_LIMIT, "sector SVD: a sector matrix beyond 2^20 rows or 2^24 columns");
    s.bs = sect_bj_rows(s.n, s.len);
    if (2 * (size_t)s.bs * s.len * sizeof(double) + 4 * (size_t)s.bs * s.bs * sizeof(double) > kBjLds)
      fail(SB_E_LIMIT, "sector SVD: a sector's rows are too long for the LDS of a workgroup");
    s.B = (s.n + s.bs - 1) / s.bs;
    s.B += s.B & 1;
    if (s.B < 2) s.B = 2;
    s.qoff = qtot;
    qtot += (i64)s.n * s.n;
    s.toff = ttot;
    ttot += (i64)(s.B - 1) * (s.B / 2) * 4 * s.bs * s.bs;
    s.foff = ftot;
    ftot += (s.B - 1) * (s.B / 2);
    s.S = sect_bj_slice(s.n, s.B, s.bs);
    if (sect_bj_apply_lds(s.n, s.S, s.B, s.bs) > kBjLds)
      fail(SB_E_LIMIT, "sector SVD: a sector's rotation rows and round matrices are too large for the LDS of a "
                       "workgroup (more than about 1200 rows to orthogonalise)");
    s.slice0 = slices;
    slices += (s.n + s.S - 1) / s.S;
    maxpairs = std::max(maxpairs, s.B / 2);
    maxrounds = std::max(maxrounds, s.B - 1);
    lds = std::max(lds, (size_t)(2 * s.bs * (size_t)s.len + 4 * s.bs * s.bs) * sizeof(double));
    alds = std::max(alds, sect_bj_apply_lds(s.n, s.S, s.B, s.bs));
  }
  run.d_sec.alloc(ctx, nb);
  HIP_CHECK(hipMemcpyAsync(run.d_sec.get(), hs.data(), nb * sizeof(BjSector), hipMemcpyHostToDevice, ctx->stream));
  run.Qw.alloc(ctx, (size_t)std::max<i64>(qtot, 1));
  DBuf<double> nul(ctx, nb), Tb(ctx, (size_t)std::max<i64>(ttot, 1));
  DBuf<int> flag(ctx, nb), tflag(ctx, (size_t)std::max(ftot, 1));
  hipLaunchKernelGGL(sect_bj_init_kernel, dim3(nb), dim3(SS_THREADS), 0, ctx->stream, run.d_sec.get(), X, xoff,
                     run.Qw.get(), nul.get(), flag.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_bj_round_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_bj_apply_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)alds));
  // the later rounds with one block's rows in registers when every row fits 64 CH doubles per wave
  int maxlen = 0;
  for (auto& s : hs) maxlen = std::max(maxlen, s.len);
  const int ch = maxlen <= 256 ? 4 : maxlen <= 512 ? 8 : maxlen <= 1024 ? 16 : 0;
  size_t clds = 0;
  for (auto& s : hs) clds = std::max(clds, (size_t)(s.bs * (size_t)s.len + 4 * s.bs * s.bs) * sizeof(double));
  if (ch == 4) HIP_CHECK(hipFuncSetAttribute((const void*)sect_bj_cross_kernel<4>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)clds));
  if (ch == 8) HIP_CHECK(hipFuncSetAttribute((const void*)sect_bj_cross_kernel<8>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)clds));
  if (ch == 16) HIP_CHECK(hipFuncSetAttribute((const void*)sect_bj_cross_kernel<16>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)clds));
  std::vector<int> hf(nb);
  run.unconverged = true;
  for (int sweep = 0; sweep < ctx->svd_sweeps; ++sweep) {
    // every (round, pair) flag of the sweep is written by its round launch
    for (int rd = 0; rd < maxrounds; ++rd) {
      const dim3 grid((unsigned)(nb * maxpairs)), block(BJ_THREADS);
      if (rd > 0 && ch == 4)
        hipLaunchKernelGGL(sect_bj_cross_kernel<4>, grid, block, clds, ctx->stream, run.d_sec.get(), maxpairs, rd, X,
                           xoff, nul.get(), flag.get(), Tb.get(), tflag.get());
      else if (rd > 0 && ch == 8)
        hipLaunchKernelGGL(sect_bj_cross_kernel<8>, grid, block, clds, ctx->stream, run.d_sec.get(), maxpairs, rd, X,
                           xoff, nul.get(), flag.get(), Tb.get(), tflag.get());
      else if (rd > 0 && ch == 16)
        hipLaunchKernelGGL(sect_bj_cross_kernel<16>, grid, block, clds, ctx->stream, run.d_sec.get(), maxpairs, rd,
                           X, xoff, nul.get(), flag.get(), Tb.get(), tflag.get());
      else
        hipLaunchKernelGGL(sect_bj_round_kernel, grid, block, lds, ctx->stream, run.d_sec.get(), maxpairs, rd, X,
                           xoff, nul.get(), flag.get(), Tb.get(), tflag.get());
      HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(sect_bj_apply_kernel, dim3((unsigned)slices), dim3(BJ_APPLY_THREADS), alds, ctx->stream,
                       run.d_sec.get(), nb, run.Qw.get(), Tb.get(), tflag.get());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(hf.data(), flag.get(), nb * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    bool any = false;
    for (int f : hf) any = any || f;
    if (!any) {
      run.unconverged = false;
      break;
    }
    k_zero(ctx, flag.get(), nb * sizeof(int));
  }
  HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the round workspaces go out of scope
}

// ------------------------------------------------ QR-preconditioned sector SVD
// Sectors solved by block Jacobi are first factorised M P = Q1 R (Householder
// QR with column pivoting, LAPACK's dgeqp3 conventions: the column of largest
// remaining norm first, ties to the lowest index, norms downdated and
// recomputed when cancellation leaves less than sqrt(eps) of them).  The
// factorisation stops at the numerical rank r, where every remaining column
// has norm <= 1e-15 |M|_F (the threshold below which the Jacobi rotations
// leave a row alone).  The rows of R (r of them, length C) are then
// orthogonalised by the block Jacobi: R's rows are graded, which halves the
// sweeps, and only r <= min(R, C) rows take part.  With R = Qacc Rf:
// M = Q1 diag(Qacc, I) [Rf; 0] P^T, so U = Q1 diag(Qacc, I) (its first
// min(R, C) columns, orthonormal also beyond the rank) and
// sigma V^T = [Rf; 0] P^T; the triplets beyond the rank have sigma = 0.
//
// One launch per reflector over all sectors, each workgroup on a chunk of
// QP_CHUNK columns (one wave per column): every workgroup of a sector picks
// the same pivot and forms the same reflector from the pivot column (the same
// reductions in the same order), applies it to its own columns and downdates
// their norms; the sector's first workgroup stores the reflector.  Columns are
// not moved: a column keeps its place, its position in the pivot order is
// recorded, and the pivot column is left as it is from its step on (its rows
// above the step hold its column of R, the diagonal entry is kept apart).
struct QpSector {
  int g;           // sector number (xoff, uoff, soff, sgoff)
  int R, C, k;     // rows, columns, min(R, C)
  int wg0;         // first workgroup of its column chunks in the step launches
  i64 toff;        // offset of its tau / beta (k)
  i64 poff;        // offset of its per-column arrays (C): norms, position, perm
  i64 voff;        // offset of its reflectors V (R x k, column-major, unit diagonal)
  i64 yoff;        // offset of its Jacobi rows Y (r x C, row-major)
  i64 ufoff;       // offset of its U (R x k, column-major)
  i64 qoff;        // offset of its accumulated rotation (r x r) in the Jacobi workspace, -1: identity
  double thr;      // 1e-15 |M|_F
  int r;           // numerical rank (k until the factorisation stops)
};

constexpr int QP_THREADS = 1024;
constexpr int QP_CHUNK = QP_THREADS / 64;  // columns per workgroup of a step launch

__device__ inline int qp_find(const QpSector* sec, int nsec) {
  int b = 0;
  while (b + 1 < nsec && sec[b + 1].wg0 <= (int)blockIdx.x) ++b;
  return b;
}

// column norms, positions -1 (not yet a pivot)
__global__ __launch_bounds__(QP_THREADS) void sect_qp_init_kernel(const QpSector* __restrict__ sec, int nsec,
                                                                 const double* __restrict__ X,
                                                                 const i64* __restrict__ xoff,
                                                                 double* __restrict__ vn, int* __restrict__ pos) {
  const QpSector s = sec[qp_find(sec, nsec)];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = ((int)blockIdx.x - s.wg0) * QP_CHUNK + wave;
  if (j >= s.C) return;
  const double* col = X + xoff[s.g] + (i64)j * s.R;
  double a = 0;
  for (int i = lane; i < s.R; i += 64) a = fma(col[i], col[i], a);
  a = sqrt(wave_sum(a));
  if (lane == 0) {
    vn[4 * s.poff + j] = a;
    vn[4 * s.poff + s.C + j] = a;
    pos[s.poff + j] = -1;
  }
}

// |M|_F (from the column norms) and the rank threshold; one workgroup per sector
__global__ __launch_bounds__(SS_THREADS) void sect_qp_thr_kernel(QpSector* __restrict__ sec,
                                                                const double* __restrict__ vn) {
  __shared__ double part[SS_THREADS / 64];
  QpSector& s = sec[blockIdx.x];
  double f = 0;
  for (int j = threadIdx.x; j < s.C; j += SS_THREADS) f = fma(vn[4 * s.poff + j], vn[4 * s.poff + j], f);
  f = wave_sum(f);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = f;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < SS_THREADS / 64; ++w) t += part[w];
    s.thr = 1e-15 * sqrt(t);
    s.r = s.k;
  }
}

// step i of every sector's factorisation
__global__ __launch_bounds__(QP_THREADS) void sect_qp_step_kernel(QpSector* __restrict__ sec, int nsec, int i,
                                                                 double* __restrict__ X, const i64* __restrict__ xoff,
                                                                 double* __restrict__ vn, int* __restrict__ pos,
                                                                 double* __restrict__ V, double* __restrict__ tb) {
  __shared__ double redv[QP_THREADS / 64];
  __shared__ int redi[QP_THREADS / 64];
  const int b = qp_find(sec, nsec);
  const QpSector s = sec[b];
  if (i >= s.k || i > s.r) return;  // finished (or stopped at its rank in an earlier step)
  const int R = s.R, C = s.C;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = QP_THREADS / 64;
  double* M = X + xoff[s.g];
  // the norms of step i are read from buffer i % 2 and the downdated ones
  // written to the other, so that no workgroup reads a norm another one has
  // already changed in this launch; a column made a pivot in this launch
  // (position i) still counts as a candidate here, pivots of earlier steps not
  const double* vn1 = vn + 4 * s.poff + 2 * (i & 1) * C;
  const double* vn2 = vn1 + C;
  double* wn1 = vn + 4 * s.poff + 2 * ((i + 1) & 1) * C;
  double* wn2 = wn1 + C;
  int* ps = pos + s.poff;
  // the pivot: largest norm among the columns not yet pivots, the first of equal ones
  double bv = -1.0;
  int bi = C;
  for (int j = tid; j < C; j += QP_THREADS)
    if ((ps[j] < 0 || ps[j] >= i) && vn1[j] > bv) {
      bv = vn1[j];
      bi = j;
    }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ov > bv || (ov == bv && oi < bi)) {
      bv = ov;
      bi = oi;
    }
  }
  if (lane == 0) {
    redv[wave] = bv;
    redi[wave] = bi;
  }
  __syncthreads();
  bv = redv[0];
  bi = redi[0];
  for (int w = 1; w < nwave; ++w)
    if (redv[w] > bv || (redv[w] == bv && redi[w] < bi)) {
      bv = redv[w];
      bi = redi[w];
    }
  __syncthreads();
  const bool first = (int)blockIdx.x == s.wg0;
  if (!(bv > s.thr)) {  // every remaining column is zero to working precision: rank i
    if (first && tid == 0) sec[b].r = i;
    return;
  }
  const int c = bi;
  const double* pc = M + (i64)c * R;
  // the reflector of the pivot column's rows i.. (dlarfg), formed alike in every workgroup
  double x2 = 0;
  for (int e = i + 1 + tid; e < R; e += QP_THREADS) x2 = fma(pc[e], pc[e], x2);
  x2 = wave_sum(x2);
  if (lane == 0) redv[wave] = x2;
  __syncthreads();
  x2 = 0;
  for (int w = 0; w < nwave; ++w) x2 += redv[w];
  const double alpha = pc[i];
  double ti = 0.0, sc = 0.0, beta = alpha;
  if (x2 != 0.0) {
    beta = -copysign(hypot(alpha, sqrt(x2)), alpha);
    ti = (beta - alpha) / beta;
    sc = 1.0 / (alpha - beta);
  }
  if (first) {
    double* v = V + s.voff + (i64)i * R;
    for (int e = tid; e < R; e += QP_THREADS) v[e] = e < i ? 0.0 : e == i ? 1.0 : (ti != 0.0 ? pc[e] * sc : 0.0);
    if (tid == 0) {
      tb[s.toff + i] = ti;
      tb[s.toff + s.k + i] = beta;
      ps[c] = i;
    }
  }
  // H_i to this workgroup's columns (not pivots, not the pivot of this step), then their norms
  const int j = ((int)blockIdx.x - s.wg0) * QP_CHUNK + wave;
  if (j >= C || j == c || (ps[j] >= 0 && ps[j] < i)) return;
  double* y = M + (i64)j * R;
  if (ti != 0.0) {
    double d = lane == 0 ? y[i] : 0.0;
    for (int e = i + 1 + lane; e < R; e += 64) d = fma(pc[e] * sc, y[e], d);
    d = wave_sum(d);
    const double f = ti * d;
    if (lane == 0) y[i] -= f;
    for (int e = i + 1 + lane; e < R; e += 64) y[e] -= f * (pc[e] * sc);
  }
  const double n1 = vn1[j], n2 = vn2[j];
  double o1 = 0.0, o2 = 0.0;
  if (n1 != 0.0) {  // every lane reads the same values: the choice is wave-uniform
    const double yi = __shfl(lane == 0 ? y[i] : 0.0, 0, 64);
    double t = fabs(yi) / n1;
    t = fmax(0.0, (1.0 + t) * (1.0 - t));
    const double q = n1 / n2;
    if (t * q * q <= sqrt(1.1102230246251565e-16)) {
      double a = 0;
      for (int e = i + 1 + lane; e < R; e += 64) a = fma(y[e], y[e], a);
      o1 = o2 = sqrt(wave_sum(a));
    } else {
      o1 = n1 * sqrt(t);
      o2 = n2;
    }
  }
  if (lane == 0) {
    wn1[j] = o1;
    wn2[j] = o2;
  }
}

// positions of the columns that were never pivots (r, r+1, .. in column order),
// perm[position] = column; one workgroup per sector
__global__ void sect_qp_perm_kernel(const QpSector* __restrict__ sec, int* __restrict__ pos,
                                    int* __restrict__ perm) {
  const QpSector s = sec[blockIdx.x];
  int* ps = pos + s.poff;
  int* pm = perm + s.poff;
  if (threadIdx.x == 0) {
    int nxt = s.r;
    for (int j = 0; j < s.C; ++j)
      if (ps[j] < 0) ps[j] = nxt++;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < s.C; j += blockDim.x) pm[ps[j]] = j;
}

// Y (r x C, row-major, columns in pivot order): row a of R
__global__ void sect_qp_rows_kernel(const QpSector* __restrict__ sec, const double* __restrict__ X,
                                    const i64* __restrict__ xoff, const int* __restrict__ perm,
                                    const double* __restrict__ tb, double* __restrict__ Y) {
  const QpSector s = sec[blockIdx.y];
  const i64 n = (i64)s.r * s.C;
  const double* M = X + xoff[s.g];
  const int* pm = perm + s.poff;
  double* y = Y + s.yoff;
  for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (i64)gridDim.x * blockDim.x) {
    const int a = (int)(e / s.C), p = (int)(e - (i64)a * s.C);
    y[e] = p < a ? 0.0 : p == a ? tb[s.toff + s.k + a] : M[(i64)pm[p] * s.R + a];
  }
}

// U (R x k, column-major) = Q1 diag(Qacc, I): one wave per column, QP_BT_COLS
// columns per workgroup, the column in LDS; the reflectors H_{r-1} .. H_0
// applied in turn (each column on its own: no barrier between reflectors)
constexpr int QP_BT_COLS = 8;

__global__ __launch_bounds__(64 * QP_BT_COLS) void sect_qp_back_kernel(const QpSector* __restrict__ sec, int nsec,
                                                                      const int* __restrict__ wg0,
                                                                      const double* __restrict__ Vr,
                                                                      const double* __restrict__ tau,
                                                                      const double* __restrict__ Qw,
                                                                      double* __restrict__ Uf) {
  extern __shared__ double lds[];
  int b = 0;
  while (b + 1 < nsec && wg0[b + 1] <= (int)blockIdx.x) ++b;
  const QpSector s = sec[b];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = ((int)blockIdx.x - wg0[b]) * QP_BT_COLS + wave;
  if (c >= s.k) return;
  const int R = s.R, r = s.r;
  double* u = lds + (i64)wave * R;
  for (int e = lane; e < R; e += 64) {
    double v0;
    if (c < r) v0 = e < r ? (s.qoff >= 0 ? Qw[s.qoff + e + (i64)r * c] : (e == c ? 1.0 : 0.0)) : 0.0;
    else v0 = e == c ? 1.0 : 0.0;
    u[e] = v0;
  }
  const double* tu = tau + s.toff;
  for (int j = r - 1; j >= 0; --j) {
    const double tj = tu[j];
    if (tj == 0.0) continue;
    const double* v = Vr + s.voff + (i64)j * R;  // v = [0..; 1; v[j+1..]]
    double d = lane == 0 ? u[j] : 0.0;
    for (int e = j + 1 + lane; e < R; e += 64) d = fma(v[e], u[e], d);
    d = wave_sum(d);
    const double f = tj * d;
    if (lane == 0) u[j] -= f;
    for (int e = j + 1 + lane; e < R; e += 64) u[e] -= f * v[e];
  }
  double* out = Uf + s.ufoff + (i64)R * c;
  for (int e = lane; e < R; e += 64) out[e] = u[e];
}

// sigma (descending, stable), U columns and sigma V^T rows of a sector
__global__ __launch_bounds__(SS_THREADS) void sect_qp_emit_kernel(const QpSector* __restrict__ sec,
                                                                 const double* __restrict__ Y,
                                                                 const double* __restrict__ Uf,
                                                                 const int* __restrict__ perm,
                                                                 const i64* __restrict__ uoff,
                                                                 const i64* __restrict__ soff,
                                                                 const i64* __restrict__ sgoff,
                                                                 double* __restrict__ sig, double* __restrict__ Uo,
                                                                 double* __restrict__ SVo) {
  extern __shared__ double lds[];
  const QpSector s = sec[blockIdx.x];
  const int R = s.R, C = s.C, k = s.k, r = s.r;
  double* nrm = lds;
  int* ord = reinterpret_cast<int*>(nrm + k);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = SS_THREADS / 64;
  const double* y = Y + s.yoff;
  for (int i = wave; i < k; i += nwave) {
    double a = 0;
    if (i < r)
      for (int j = lane; j < C; j += 64) a = fma(y[(i64)i * C + j], y[(i64)i * C + j], a);
    a = wave_sum(a);
    if (lane == 0) nrm[i] = sqrt(a);
  }
  __syncthreads();
  for (int i = tid; i < k; i += SS_THREADS) {  // stable descending rank (a NaN ranks as -inf)
    const double v = isnan(nrm[i]) ? -INFINITY : nrm[i];
    int rank = 0;
    for (int j = 0; j < k; ++j) {
      const double w = isnan(nrm[j]) ? -INFINITY : nrm[j];
      rank += (w > v) || (w == v && j < i);
    }
    ord[rank] = i;
  }
  __syncthreads();
  double* sg = sig + sgoff[s.g];
  for (int c = tid; c < k; c += SS_THREADS) sg[c] = nrm[ord[c]];
  const int* pm = perm + s.poff;
  double* Ug = Uo + uoff[s.g];
  double* Sg = SVo + soff[s.g];
  for (int c = wave; c < k; c += nwave) {
    const int o = ord[c];
    const double* uc = Uf + s.ufoff + (i64)R * o;
    for (int e = lane; e < R; e += 64) Ug[e + (i64)R * c] = uc[e];
    double* sv = Sg + (i64)c * C;
    if (o < r)
      for (int j = lane; j < C; j += 64) sv[pm[j]] = y[(i64)o * C + j];
    else
      for (int j = lane; j < C; j += 64) sv[j] = 0.0;
  }
}

// The sectors gl (sector numbers, R x C column-major at xoff): their triplets
// like k_sect_svd, through the QR preconditioning and the block Jacobi.
bool k_sect_svd_qrcp(Ctx* ctx, const std::vector<int>& gl, const std::vector<int>& Rs, const std::vector<int>& Cs,
                     double* X, const i64* xoff, const i64* uoff, const i64* soff, const i64* sgoff, double* sig,
                     double* Uo, double* SVo) {
  const int nb = (int)gl.size();
  if (!nb) return false;
  std::vector<QpSector> hq(nb);
  i64 ttot = 0, ptot = 0, vtot = 0, uftot = 0;
  int maxk = 1, maxR = 1, wgs = 0;
  for (int b = 0; b < nb; ++b) {
    QpSector& q = hq[b];
    q = QpSector{};
    q.g = gl[b];
    q.R = Rs[b];
    q.C = Cs[b];
    q.k = std::min(q.R, q.C);
    q.wg0 = wgs;
    wgs += (q.C + QP_CHUNK - 1) / QP_CHUNK;
    q.toff = ttot;
    ttot += 2 * (i64)q.k;
    q.poff = ptot;
    ptot += q.C;
    q.voff = vtot;
    vtot += (i64)q.R * q.k;
    q.ufoff = uftot;
    uftot += (i64)q.R * q.k;
    q.qoff = -1;
    q.r = q.k;
    maxk = std::max(maxk, q.k);
    maxR = std::max(maxR, q.R);
  }
  if ((size_t)QP_BT_COLS * maxR * sizeof(double) > kBjLds)
    fail(SB_E_LIMIT, "sector SVD: a sector with more rows than the LDS of a workgroup holds for eight columns");
  DBuf<QpSector> d_q(ctx, nb);
  HIP_CHECK(hipMemcpyAsync(d_q.get(), hq.data(), nb * sizeof(QpSector), hipMemcpyHostToDevice, ctx->stream));
  DBuf<double> tb(ctx, (size_t)std::max<i64>(ttot, 1)), vn(ctx, (size_t)std::max<i64>(4 * ptot, 1)),
      V(ctx, (size_t)std::max<i64>(vtot, 1));
  DBuf<int> pos(ctx, (size_t)std::max<i64>(ptot, 1)), perm(ctx, (size_t)std::max<i64>(ptot, 1));
  hipLaunchKernelGGL(sect_qp_init_kernel, dim3((unsigned)wgs), dim3(QP_THREADS), 0, ctx->stream, d_q.get(), nb, X,
                     xoff, vn.get(), pos.get());
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(sect_qp_thr_kernel, dim3(nb), dim3(SS_THREADS), 0, ctx->stream, d_q.get(), vn.get());
  HIP_CHECK(hipGetLastError());
  // the steps, in batches between which the host reads the ranks: a sector
  // that reached its rank stops, and so do the launches once all have
  std::vector<QpSector> hr(nb);
  for (int i0 = 0; i0 < maxk; i0 += 32) {
    for (int i = i0; i < std::min(maxk, i0 + 32); ++i) {
      hipLaunchKernelGGL(sect_qp_step_kernel, dim3((unsigned)wgs), dim3(QP_THREADS), 0, ctx->stream, d_q.get(), nb,
                         i, X, xoff, vn.get(), pos.get(), V.get(), tb.get());
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipMemcpyAsync(hr.data(), d_q.get(), nb * sizeof(QpSector), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    bool more = false;
    for (int b = 0; b < nb; ++b) more = more || (hr[b].r > i0 + 31 && hr[b].k > i0 + 32);
    if (!more) break;
  }
  // the Jacobi rows of every sector of rank 2 or more
  i64 ytot = 0;
  std::vector<i64> hyoff;
  std::vector<int> jg, jn, jl, jb;
  for (int b = 0; b < nb; ++b) {
    QpSector& q = hq[b];
    q.r = hr[b].r;
    q.thr = hr[b].thr;
    if (q.r < 0 || q.r > q.k) fail(SB_E_DEVICE, "sector SVD: the QR factorisation returned an invalid rank");
    q.yoff = ytot;
    ytot += (i64)q.r * q.C;
    if (q.r >= 2) {
      jg.push_back((int)jb.size());
      jn.push_back(q.r);
      jl.push_back(q.C);
      jb.push_back(b);
      hyoff.push_back(q.yoff);
    }
  }
  HIP_CHECK(hipMemcpyAsync(d_q.get(), hq.data(), nb * sizeof(QpSector), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(sect_qp_perm_kernel, dim3(nb), dim3(256), 0, ctx->stream, d_q.get(), pos.get(), perm.get());
  HIP_CHECK(hipGetLastError());
  DBuf<double> Y(ctx, (size_t)std::max<i64>(ytot, 1));
  hipLaunchKernelGGL(sect_qp_rows_kernel, dim3(64, nb), dim3(256), 0, ctx->stream, d_q.get(), X, xoff, perm.get(),
                     tb.get(), Y.get());
  HIP_CHECK(hipGetLastError());
  BjRun run;
  DBuf<i64> d_yoff;
  if (!jb.empty()) {
    d_yoff.alloc(ctx, hyoff.size());
    HIP_CHECK(hipMemcpyAsync(d_yoff.get(), hyoff.data(), hyoff.size() * sizeof(i64), hipMemcpyHostToDevice,
                             ctx->stream));
    bj_jacobi(ctx, jg, jn, jl, Y.get(), d_yoff.get(), run);
    for (size_t t = 0; t < jb.size(); ++t) hq[jb[t]].qoff = run.hs[t].qoff;
    HIP_CHECK(hipMemcpyAsync(d_q.get(), hq.data(), nb * sizeof(QpSector), hipMemcpyHostToDevice, ctx->stream));
  }
  std::vector<int> wg0(nb + 1, 0);
  for (int b = 0; b < nb; ++b) wg0[b + 1] = wg0[b] + (hq[b].k + QP_BT_COLS - 1) / QP_BT_COLS;
  DBuf<int> d_wg0(ctx, nb + 1);
  HIP_CHECK(hipMemcpyAsync(d_wg0.get(), wg0.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  DBuf<double> Uf(ctx, (size_t)std::max<i64>(uftot, 1));
  const size_t blds = (size_t)QP_BT_COLS * maxR * sizeof(double);
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_qp_back_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)blds));
  hipLaunchKernelGGL(sect_qp_back_kernel, dim3((unsigned)wg0[nb]), dim3(64 * QP_BT_COLS), blds, ctx->stream,
                     d_q.get(), nb, d_wg0.get(), V.get(), tb.get(), jb.empty() ? nullptr : run.Qw.get(), Uf.get());
  HIP_CHECK(hipGetLastError());
  const size_t elds = (size_t)maxk * (sizeof(double) + sizeof(int));
  HIP_CHECK(hipFuncSetAttribute((const void*)sect_qp_emit_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)elds));
  hipLaunchKernelGGL(sect_qp_emit_kernel, dim3(nb), dim3(SS_THREADS), elds, ctx->stream, d_q.get(), Y.get(),
                     Uf.get(), perm.get(), uoff, soff, sgoff, sig, Uo, SVo);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the workspaces go out of scope
  return !jb.empty() && run.unconverged;
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
