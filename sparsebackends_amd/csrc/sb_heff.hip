// Two-site H_eff*psi of DMRG as one plan with one fused kernel.
//
// ITensorMPS applies the two-site effective Hamiltonian as four contract!
// calls, psi*L*W1*W2*R (product(ProjMPO, psi); test_product_dmrg_state.jl
// contracts it this way), each materialising an intermediate that carries
// the physical and MPO-link indices densely: at chi = 2048 with 520 (n, 2Sz)
// sectors the intermediates are ~14x the size of psi.  Here every output
// block (Ql', Qr') is computed in one workgroup, straight from psi, L, R and
// the two-site MPO:
//
//   Hpsi[p', l', r'] = sum over link pairs (wl, wr), sector pairs (Ql, Qr):
//       sum_p W12[wl,wr][p', p] sum_l L[wl, l', l] sum_r psi[p, l, r] R[wr, r', r]
//
// with p = (s1, s2) and W12[wl,wr][p', p] = sum_wm W1[wl, s1', s1, wm] W2[wm, s2', s2, wr]
// (nonzero entries only).  Per sector pair the psi block is staged in LDS,
// A = L_wl * psi per left link, then every output element accumulates
// sum_r R_wr[r', r] (W12 row p' . A[:, l', r]) in registers.  No
// intermediate touches HBM: the kernel reads psi, L, R (all small enough to
// stay in the caches) and writes Hpsi once.
//
// Tensors (the layouts of the four-call chain's operands):
//   psi [Ql, Qr | s1, s2, l, r]   L [Ql', Ql | wl, l', l]   R [Qr', Qr | wr, r', r]
//   W1 [ | wl, s1', s1, wm]       W2 [ | wm, s2', s2, wr]
//   Hpsi [Ql', Qr' | s1', s2', l', r']
// The block list of Hpsi is the one the chain's last contract! (T3*R,
// contract_shared!, contract_bs_bs.jl:684-728) creates: every (Ql', Qr')
// reached by a pair, ids in the order its merge walk first meets them.
#include <algorithm>
#include <array>
#include <map>
#include <vector>

#include "sb_contract.h"

namespace sb {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kTilesPerWave = 4;                    // 16 x 16 output tiles a wave accumulates
constexpr int kMaxTiles = kWaves * kTilesPerWave;   // of one output block
constexpr size_t kMaxLds = 96 * 1024;               // dynamic LDS of one workgroup

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// per environment block: bit w set when the slice [w, :, :] holds a nonzero,
// and the extents (one past the last nonzero index) of the two bond axes
__global__ __launch_bounds__(256) void env_mask_kernel(const double* __restrict__ data, const i32* __restrict__ ids,
                                                       i64 n, int w, i64 da, i64 db, i32* __restrict__ out) {
  __shared__ int sm[3];
  const i64 bs = (i64)w * da * db;
  for (i64 i = blockIdx.x; i < n; i += gridDim.x) {
    if (threadIdx.x < 3) sm[threadIdx.x] = 0;
    __syncthreads();
    const double* b = data + (i64)(ids[i] - 1) * bs;
    int mask = 0, ea = 0, eb = 0;
    for (i64 e = threadIdx.x; e < bs; e += 256) {
      if (b[e] != 0.0) {  // NaN counts as nonzero
        const i64 q = e / w;
        const int ww = (int)(e - q * w);
        const int a = (int)(q % da), c = (int)(q / da);
        mask |= 1 << ww;
        ea = max(ea, a + 1);
        eb = max(eb, c + 1);
      }
    }
    if (mask) {
      atomicOr(&sm[0], mask);
      atomicMax(&sm[1], ea);
      atomicMax(&sm[2], eb);
    }
    __syncthreads();
    if (threadIdx.x < 3) out[i * 3 + threadIdx.x] = sm[threadIdx.x];
    __syncthreads();
  }
}

struct HeffPair {
  i64 psi, L, R;  // payload offsets of the psi, L and R blocks
  int dl, dr;     // extents of l and r
  int wlm, wrm;   // left links to apply, right links to stage from R
};

struct HeffK {
  const double *psi, *L, *R;
  double* out;
  const int2* items;  // (output block, first tile) of this launch's workgroups
  int tpw;            // tiles per wave of one workgroup (at most kTilesPerWave)
  const i32 *o_dlp, *o_drp, *o_pb;
  const i64* o_off;
  const HeffPair* pairs;
  const i32* lidx;      // link pair (wl, wr) -> its W12 tile, or -1
  const double* w16;    // per link pair a 16 x 16 tile: W12 on the diagonal blocks of P x P
  int lp, w, Dl, Dr;  // lp = log2(d1 d2)
  int sPs, sAs, sLs, sRs;  // LDS elements of the psi, A, L and R tiles (this launch)
};

__host__ __device__ __forceinline__ int up4(int x) { return (x + 3) & ~3; }
__host__ __device__ __forceinline__ int up16(int x) { return (x + 15) & ~15; }
__device__ __forceinline__ size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }

template <class T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 acc;
  // 16x16x4: lane holds A[lane & 15][lane >> 4], B[lane >> 4][lane & 15];
  // accumulator register v holds row 4 v + lane / 16, column lane & 15
  static __device__ __forceinline__ acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int v, int lane) { return 4 * v + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef f4 acc;
  // the FP32 16x16x4 form: register v holds row 4 (lane / 16) + v
  static __device__ __forceinline__ acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int v, int lane) { return 4 * (lane >> 4) + v; }
};

// One workgroup per output block (Ql', Qr'), its elements as rows (p', l')
// by columns r' in 16 x 16 tiles on the matrix cores, up to 4 tiles per wave
// held in registers over all of the block's work.  Per sector pair: the psi
// block as a (l x (p, r)) matrix, and the L and R slices of the pair's links,
// staged in LDS (zero padded) with one wait for all their loads; per left
// link A = L_wl (l' x l) * psi (l x (p, r)) on the matrix cores into LDS; per
// right link the output tiles accumulate V (r) * R_wr^T with
// V[(p', l'), r] = sum_p W12[p', p] A[l', (p, r)] formed as each lane loads
// its operand.  No atomics: each output element is written once.
template <class T>
__global__ __launch_bounds__(kThreads) void heff_kernel(HeffK k) {
  using M = Mfma<T>;
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lp = k.lp, P = 1 << lp, w = k.w;
  T* Ps = reinterpret_cast<T*>(lds);
  T* As = Ps + k.sPs;
  T* Ls = As + k.sAs;
  T* Rs = Ls + k.sLs;
  const int2 item = k.items[blockIdx.x];
  const int o = item.x, t0 = item.y, tpw = k.tpw;
  const int dlp = k.o_dlp[o], drp = k.o_drp[o];
  const int rows = dlp << lp;
  const int RT = up16(rows) >> 4;
  const int lLM = dlp > 16 ? 5 : 4, LM = 1 << lLM, lML = lLM - 4;  // l' padded to 16 or 32
  const int lNC = drp > 16 ? 5 : 4, NC = 1 << lNC;                    // r' padded to 16 or 32
  const int T16 = min(RT << (lNC - 4), t0 + kWaves * tpw);           // this workgroup's tiles: t0 + wave + kWaves i
  // this wave's output tiles: (row tile, column tile) and the lane's row (p', l')
  int tci[kTilesPerWave], tlq[kTilesPerWave], tpq[kTilesPerWave];
  bool tok[kTilesPerWave];
  typename M::acc acc[kTilesPerWave];
#pragma unroll
  for (int i = 0; i < kTilesPerWave; ++i) {
    const int t = t0 + wave + kWaves * i;
    tok[i] = i < tpw && t < T16;
    const int ri = t % RT;
    tci[i] = t / RT;
    const int row = ri * 16 + (lane & 15);
    tlq[i] = row >> lp;
    tpq[i] = row & (P - 1);
    acc[i] = typename M::acc{0, 0, 0, 0};
  }
  const int Dl = k.Dl, Dr = k.Dr;
  // staging of one sector pair: Ps[l][c] (c = p + P r), Ls[s][l][l'] for the
  // s-th left link and Rs[s][r][r'] for the s-th right link.  The first batch
  // of entries (BP, BL, BR per thread) goes through registers: its loads are
  // issued one pair ahead, while the previous pair computes, and stored at the
  // start of the pair; entries beyond it (large pairs) load at the start.
  constexpr int BP = 4, BL = 2, BR = 2;
  struct Geo {
    int dl, dr, K1, N1, K2, PDR, SL, SR, nP, nL, nR;
  };
  auto geo = [&](const HeffPair& pr) {
    Geo g;
    g.dl = pr.dl;
    g.dr = pr.dr;
    g.K1 = up4(g.dl);
    g.N1 = up16(g.dr << lp);
    g.K2 = up4(g.dr);
    g.PDR = g.dr << lp;
    g.SL = LM * g.K1;
    g.SR = g.K2 * NC;
    g.nP = g.K1 * g.N1;
    g.nL = __popc(pr.wlm) * g.SL;
    g.nR = __popc(pr.wrm) * g.SR;
    return g;
  };
  auto ps_at = [&](const HeffPair& pr, const Geo& g, int t) -> T {  // t = c + N1 l
    const int l = t / g.N1, c = t - l * g.N1;
    return (l < g.dl && c < g.PDR) ? (T)k.psi[pr.psi + (c & (P - 1)) + (((i64)l + (i64)Dl * (c >> lp)) << lp)]
                                   : (T)0;
  };
  auto ls_at = [&](const HeffPair& pr, const Geo& g, int t) -> T {
    int s = 0, u = t;
    while (u >= g.SL) {
      u -= g.SL;
      ++s;
    }
    int bits = pr.wlm;
    for (int j = 0; j < s; ++j) bits &= bits - 1;
    const int m = u & (LM - 1), l = u >> lLM;
    return (m < dlp && l < g.dl) ? (T)k.L[pr.L + (__ffs(bits) - 1) + (i64)w * (m + (i64)Dl * l)] : (T)0;
  };
  auto rs_at = [&](const HeffPair& pr, const Geo& g, int t) -> T {
    int s = 0, u = t;
    while (u >= g.SR) {
      u -= g.SR;
      ++s;
    }
    int bits = pr.wrm;
    for (int j = 0; j < s; ++j) bits &= bits - 1;
    const int n = u & (NC - 1), r = u >> lNC;
    return (n < drp && r < g.dr) ? (T)k.R[pr.R + (__ffs(bits) - 1) + (i64)w * (n + (i64)Dr * r)] : (T)0;
  };
  T vp[BP], vl[BL], vr[BR];
  auto stage_load = [&](const HeffPair& pr, const Geo& g) {
#pragma unroll
    for (int u = 0; u < BP; ++u) {
      const int t = tid + kThreads * u;
      vp[u] = t < g.nP ? ps_at(pr, g, t) : (T)0;
    }
#pragma unroll
    for (int u = 0; u < BL; ++u) {
      const int t = tid + kThreads * u;
      vl[u] = t < g.nL ? ls_at(pr, g, t) : (T)0;
    }
#pragma unroll
    for (int u = 0; u < BR; ++u) {
      const int t = tid + kThreads * u;
      vr[u] = t < g.nR ? rs_at(pr, g, t) : (T)0;
    }
  };
  auto stage_store = [&](const HeffPair& pr, const Geo& g) {
#pragma unroll
    for (int u = 0; u < BP; ++u) {
      const int t = tid + kThreads * u;
      if (t < g.nP) Ps[t] = vp[u];
    }
#pragma unroll
    for (int u = 0; u < BL; ++u) {
      const int t = tid + kThreads * u;
      if (t < g.nL) Ls[t] = vl[u];
    }
#pragma unroll
    for (int u = 0; u < BR; ++u) {
      const int t = tid + kThreads * u;
      if (t < g.nR) Rs[t] = vr[u];
    }
    for (int t = tid + kThreads * BP; t < g.nP; t += kThreads) Ps[t] = ps_at(pr, g, t);
    for (int t = tid + kThreads * BL; t < g.nL; t += kThreads) Ls[t] = ls_at(pr, g, t);
    for (int t = tid + kThreads * BR; t < g.nR; t += kThreads) Rs[t] = rs_at(pr, g, t);
  };
  const int qe = k.o_pb[o + 1];
  int q = k.o_pb[o];
  HeffPair pr{};
  Geo gq{};
  if (q < qe) {
    pr = k.pairs[q];
    gq = geo(pr);
    stage_load(pr, gq);
  }
  for (; q < qe; ++q) {
    __syncthreads();  // the previous pair is done with the tiles
    stage_store(pr, gq);
    __syncthreads();
    const HeffPair cur = pr;
    const Geo g = gq;
    if (q + 1 < qe) {  // the next pair's loads fly while this one computes
      pr = k.pairs[q + 1];
      gq = geo(pr);
      stage_load(pr, gq);
    }
    const int dr = g.dr, K1 = g.K1, N1 = g.N1, K2 = g.K2;
    int sl = 0;
    for (int wlm = cur.wlm; wlm; wlm &= wlm - 1, ++sl) {
      const int wl = __ffs(wlm) - 1;
      // A[l'][c] = sum_l L_wl[l'][l] Ps[l][c]
      const T* Lt = Ls + sl * LM * K1;
      const int nt = N1 >> (4 - lML);  // (LM / 16) * (N1 / 16) tiles
      for (int tt = wave; tt < nt; tt += kWaves) {
        const int mi = tt & ((1 << lML) - 1), ni = tt >> lML;
        typename M::acc c = typename M::acc{0, 0, 0, 0};
        for (int k0 = 0; k0 < K1; k0 += 4) {
          const int kk = k0 + (lane >> 4);
          c = M::run(Lt[kk * LM + mi * 16 + (lane & 15)], Ps[kk * N1 + ni * 16 + (lane & 15)], c);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) As[(mi * 16 + M::row(v, lane)) * N1 + ni * 16 + (lane & 15)] = c[v];
      }
      __syncthreads();
      int sr = 0;
      for (int wrm = cur.wrm; wrm; wrm &= wrm - 1, ++sr) {
        const int wr = __ffs(wrm) - 1;
        const int lk = k.lidx[wl * w + wr];
        if (lk < 0) continue;
        const T* Rt = Rs + sr * K2 * NC;
        // the 16 x 16 W12 tile as the A operand of four k-steps whose B
        // operands are the accumulator registers of Z = A R^T (the k index
        // of step s in lane group lane / 16 is the accumulator row of
        // register s, so Z feeds the matrix cores without leaving registers)
        const double* Wt = k.w16 + lk * 256 + (lane & 15) * 16;
        T wa[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) wa[v] = (T)Wt[M::row(v, lane)];
#pragma unroll
        for (int i = 0; i < kTilesPerWave; ++i) {
          if (tok[i]) {
            const T* Ar = As + tlq[i] * N1 + tpq[i];
            const T* Rc = Rt + tci[i] * 16 + (lane & 15);
            typename M::acc z = typename M::acc{0, 0, 0, 0};
            for (int k0 = 0; k0 < K2; k0 += 4) {
              const int kk = k0 + (lane >> 4);
              const T a = kk < dr ? Ar[kk << lp] : (T)0;
              z = M::run(a, Rc[kk * NC], z);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i] = M::run(wa[v], z[v], acc[i]);
          }
        }
      }
      __syncthreads();  // As is rewritten for the next left link
    }
  }
  double* ob = k.out + k.o_off[o];
#pragma unroll
  for (int i = 0; i < kTilesPerWave; ++i) {
    if (tok[i]) {
      const int col = tci[i] * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = (tlq[i] << lp) + tpq[i] - (lane & 15) + M::row(v, lane);  // the tile's row of register v
        if (row < rows && col < drp) ob[(row & (P - 1)) + ((i64)((row >> lp) + (i64)Dl * col) << lp)] = (double)acc[i][v];
      }
    }
  }
}

template <class T>
size_t heff_lds_bytes(int sPs, int sAs, int sLs, int sRs) {
  return sizeof(T) * ((size_t)sPs + (size_t)sAs + (size_t)sLs + (size_t)sRs);
}

// everything the fused plan keeps on the device, and the host-side checks of its contract
struct HeffState {
  const Tensor *psi, *L, *R, *W1, *W2;
  Tensor* out;
  i64 vL, vR, vW1, vW2;  // versions of the operator tensors at planning
  DBuf<int2> items;  // (output block, first tile), launch class by launch class
  DBuf<i32> o_dlp, o_drp, o_pb, lidx;
  DBuf<i64> o_off;
  DBuf<HeffPair> pairs;
  DBuf<double> w16;
  struct Launch {
    HeffK k{};
    size_t lds = 0;
    i64 n = 0;
    hipStream_t stream = nullptr;  // the side stream of a launch other than the first (nullptr: the context's)
    hipEvent_t done = nullptr;
  };
  std::vector<Launch> launches;  // one per size class with work
  hipEvent_t fork = nullptr;
  ~HeffState() {
    for (auto& la : launches) {
      if (la.stream) (void)hipStreamSynchronize(la.stream);
      if (la.done) (void)hipEventDestroy(la.done);
      if (la.stream) (void)hipStreamDestroy(la.stream);
    }
    if (fork) (void)hipEventDestroy(fork);
  }
  bool fp32 = false;
  i64 nout = 0;
};

template <class T>
std::vector<T> down(Ctx* ctx, const T* d, size_t n) {
  std::vector<T> h(n);
  if (n) HIP_CHECK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  return h;
}
template <class T>
void up(Ctx* ctx, DBuf<T>& d, const std::vector<T>& h) {
  d.alloc(ctx, h.empty() ? 1 : h.size());
  if (!h.empty()) HIP_CHECK(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}

}  // namespace

std::unique_ptr<Plan> heff_plan(Ctx* ctx, Tensor& out, const Tensor& psi, const Tensor& L, const Tensor& W1,
                                const Tensor& W2, const Tensor& R, unsigned flags) {
  for (const Tensor* t : {(const Tensor*)&out, &psi, &L, &W1, &W2, &R})
    if (t->kind != BS) fail(SB_E_METHOD, "MethodError: H_eff operands must be block-sparse");
  if (psi.N() != 6 || psi.N2 != 4) fail(SB_E_ARGUMENT, "ArgumentError: psi must be [Ql, Qr | s1, s2, l, r]");
  if (L.N() != 5 || L.N2 != 3 || R.N() != 5 || R.N2 != 3)
    fail(SB_E_ARGUMENT, "ArgumentError: L and R must be [Q', Q | w, x', x]");
  if (W1.N() != 4 || W1.N2 != 4 || W2.N() != 4 || W2.N2 != 4)
    fail(SB_E_ARGUMENT, "ArgumentError: W1 and W2 must be dense blocks [ | w, s', s, w']");
  const i64 nl = psi.dims[0], nr = psi.dims[1], d1 = psi.dims[2], d2 = psi.dims[3], Dl = psi.dims[4],
            Dr = psi.dims[5];
  const i64 w = L.dims[2], wm = W1.dims[3];
  auto want = [](bool ok, const char* m) {
    if (!ok) fail(SB_E_ASSERT, m);
  };
  want(L.dims[0] == nl && L.dims[1] == nl && L.dims[3] == Dl && L.dims[4] == Dl, "L dims must match psi's left bond");
  want(R.dims[0] == nr && R.dims[1] == nr && R.dims[3] == Dr && R.dims[4] == Dr, "R dims must match psi's right bond");
  want(R.dims[2] == w && W1.dims[0] == w && W2.dims[3] == w && W2.dims[0] == wm, "MPO link dims must agree");
  want(W1.dims[1] == d1 && W1.dims[2] == d1 && W2.dims[1] == d2 && W2.dims[2] == d2, "site dims must agree");
  want(out.dims == psi.dims && out.N2 == 4, "Hpsi must have psi's dims and N2 = 4");
  const i64 P = d1 * d2;
  if (P > 16 || w > 31) fail(SB_E_LIMIT, "H_eff plan: d1*d2 above 16 or MPO link dimension above 31");
  const bool fp32 = (flags & SB_PREC_MASK) == SB_PREC_FP32;
  if ((flags & SB_PREC_MASK) != SB_PREC_FP64 && !fp32)
    fail(SB_E_ARGUMENT, "ArgumentError: the fused H_eff plan computes in FP64 or FP32");

  // ---- structure from the key lists, link masks and extents of the environments
  auto kpsi = down(ctx, psi.keys.get(), (size_t)psi.n * 2), kL = down(ctx, L.keys.get(), (size_t)L.n * 2),
       kR = down(ctx, R.keys.get(), (size_t)R.n * 2);
  auto ipsi = down(ctx, psi.ids.get(), (size_t)psi.n), iL = down(ctx, L.ids.get(), (size_t)L.n),
       iR = down(ctx, R.ids.get(), (size_t)R.n);
  DBuf<i32> mL(ctx, (size_t)std::max<i64>(L.n, 1) * 3), mR(ctx, (size_t)std::max<i64>(R.n, 1) * 3);
  if (L.n)
    hipLaunchKernelGGL(env_mask_kernel, dim3((unsigned)std::min<i64>(L.n, 65536)), dim3(256), 0, ctx->stream,
                       L.dptr(), L.ids.get(), L.n, (int)w, Dl, Dl, mL.get());
  if (R.n)
    hipLaunchKernelGGL(env_mask_kernel, dim3((unsigned)std::min<i64>(R.n, 65536)), dim3(256), 0, ctx->stream,
                       R.dptr(), R.ids.get(), R.n, (int)w, Dr, Dr, mR.get());
  HIP_CHECK(hipGetLastError());
  auto hmL = down(ctx, mL.get(), (size_t)L.n * 3), hmR = down(ctx, mR.get(), (size_t)R.n * 3);
  auto hW1 = down(ctx, W1.dptr(), (size_t)(W1.n ? W1.blksize : 0)),
       hW2 = down(ctx, W2.dptr(), (size_t)(W2.n ? W2.blksize : 0));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));

  // W12[wl, wr][p', p] = sum_wm W1[wl, s1', s1, wm] W2[wm, s2', s2, wr], rows p' in CSR per link pair
  std::vector<i32> lidx((size_t)(w * w), -1), wptr(1, 0), wcol;
  std::vector<double> wval;
  int nlink = 0;
  for (i64 wl = 0; wl < w; ++wl)
    for (i64 wr = 0; wr < w; ++wr) {
      std::vector<double> blk((size_t)(P * P), 0.0);
      bool any = false;
      if (!hW1.empty() && !hW2.empty())
        for (i64 s1p = 0; s1p < d1; ++s1p)
          for (i64 s1 = 0; s1 < d1; ++s1)
            for (i64 m = 0; m < wm; ++m) {
              const double a = hW1[(size_t)(wl + w * (s1p + d1 * (s1 + d1 * m)))];
              if (a == 0.0) continue;
              for (i64 s2p = 0; s2p < d2; ++s2p)
                for (i64 s2 = 0; s2 < d2; ++s2) {
                  const double b = hW2[(size_t)(m + wm * (s2p + d2 * (s2 + d2 * wr)))];
                  if (b == 0.0) continue;
                  blk[(size_t)((s1p + d1 * s2p) + P * (s1 + d1 * s2))] += a * b;
                  any = true;
                }
            }
      if (!any) continue;
      lidx[(size_t)(wl * w + wr)] = nlink++;
      for (i64 pp = 0; pp < P; ++pp) {
        for (i64 p = 0; p < P; ++p) {
          const double c = blk[(size_t)(pp + P * p)];
          if (c != 0.0) {
            wcol.push_back((i32)p);
            wval.push_back(c);
          }
        }
        wptr.push_back((i32)wcol.size());
      }
    }
  // the W12 tiles of the matrix cores: rows p' + P a, columns p + P a (a < 16 / P)
  if (16 % P) fail(SB_E_LIMIT, "H_eff plan: d1*d2 must divide 16 for the fused kernel");
  std::vector<double> w16((size_t)std::max(nlink, 1) * 256, 0.0);
  for (int lk = 0; lk < nlink; ++lk)
    for (i64 pp = 0; pp < P; ++pp)
      for (int x = wptr[(size_t)(lk * P + pp)]; x < wptr[(size_t)(lk * P + pp + 1)]; ++x)
        for (i64 a = 0; a < 16 / P; ++a)
          w16[(size_t)(lk * 256 + (pp + P * a) * 16 + (wcol[(size_t)x] + P * a))] = wval[(size_t)x];
  std::vector<int> wl_useful((size_t)w, 0);  // right links each left link pairs with
  for (i64 wl = 0; wl < w; ++wl)
    for (i64 wr = 0; wr < w; ++wr)
      if (lidx[(size_t)(wl * w + wr)] >= 0) wl_useful[(size_t)wl] |= 1 << wr;

  // extents of the bond sectors: the environments' nonzero reach
  std::vector<int> elp((size_t)nl, 0), el((size_t)nl, 0), erp((size_t)nr, 0), er((size_t)nr, 0);
  std::vector<std::vector<std::pair<i32, i64>>> Lrow((size_t)nl), Rrow((size_t)nr);  // by Q': (Q, block row)
  for (i64 i = 0; i < L.n; ++i) {
    const int a = kL[(size_t)(2 * i)] - 1, b = kL[(size_t)(2 * i + 1)] - 1;
    elp[(size_t)a] = std::max(elp[(size_t)a], hmL[(size_t)(3 * i + 1)]);
    el[(size_t)b] = std::max(el[(size_t)b], hmL[(size_t)(3 * i + 2)]);
    Lrow[(size_t)a].push_back({b, i});
  }
  for (i64 i = 0; i < R.n; ++i) {
    const int a = kR[(size_t)(2 * i)] - 1, b = kR[(size_t)(2 * i + 1)] - 1;
    erp[(size_t)a] = std::max(erp[(size_t)a], hmR[(size_t)(3 * i + 1)]);
    er[(size_t)b] = std::max(er[(size_t)b], hmR[(size_t)(3 * i + 2)]);
    Rrow[(size_t)a].push_back({b, i});
  }
  std::map<std::pair<int, int>, i64> psi_at;  // (Ql, Qr) -> block row
  for (i64 i = 0; i < psi.n; ++i) psi_at[{kpsi[(size_t)(2 * i)] - 1, kpsi[(size_t)(2 * i + 1)] - 1}] = i;

  // ---- Hpsi's block list: the T3*R merge walk of the chain.  T3 holds the
  // (Ql', Qr) met by a psi*L pair; T3 and R keys in column-major order (Qr most
  // significant); per Qr every T3 block (ascending Ql') with every R block
  // (ascending Qr'), the first meeting of (Ql', Qr') creating its block.
  std::vector<std::vector<int>> t3_by_qr((size_t)nr), r_by_qr((size_t)nr);
  {
    std::vector<std::vector<char>> t3((size_t)nr, std::vector<char>((size_t)nl, 0));
    std::vector<std::vector<i64>> Lcol((size_t)nl);  // L blocks by Ql
    for (i64 i = 0; i < L.n; ++i) Lcol[(size_t)(kL[(size_t)(2 * i + 1)] - 1)].push_back(kL[(size_t)(2 * i)] - 1);
    for (auto& kv : psi_at)
      for (i64 a : Lcol[(size_t)kv.first.first]) t3[(size_t)kv.first.second][(size_t)a] = 1;
    for (i64 qr = 0; qr < nr; ++qr)
      for (i64 a = 0; a < nl; ++a)
        if (t3[(size_t)qr][(size_t)a]) t3_by_qr[(size_t)qr].push_back((int)a);
    for (i64 i = 0; i < R.n; ++i) r_by_qr[(size_t)(kR[(size_t)(2 * i + 1)] - 1)].push_back(kR[(size_t)(2 * i)] - 1);
    for (auto& v : r_by_qr) std::sort(v.begin(), v.end());
  }
  std::map<std::pair<int, int>, int> created;  // (Ql', Qr') -> id (1-based)
  std::vector<std::pair<int, int>> by_id;
  // T2 = T1*W1 and T3 = T2*W2 pair every block with the single MPO block: without one, T3 has no blocks
  if (W1.n && W2.n)
    for (i64 qr = 0; qr < nr; ++qr)
      for (int a : t3_by_qr[(size_t)qr])
        for (int c : r_by_qr[(size_t)qr]) {
          auto key = std::make_pair(a, c);
          if (created.emplace(key, (int)by_id.size() + 1).second) by_id.push_back(key);
        }
  const i64 nout = (i64)by_id.size();
  // keys in column-major order (Qr' most significant), ids from the walk
  std::vector<std::pair<int, int>> sorted_keys(by_id);
  std::sort(sorted_keys.begin(), sorted_keys.end(),
            [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
              return x.second != y.second ? x.second < y.second : x.first < y.first;
            });
  std::vector<i32> okeys((size_t)(2 * nout)), oids((size_t)nout);
  for (i64 i = 0; i < nout; ++i) {
    okeys[(size_t)(2 * i)] = sorted_keys[(size_t)i].first + 1;
    okeys[(size_t)(2 * i + 1)] = sorted_keys[(size_t)i].second + 1;
    oids[(size_t)i] = created[sorted_keys[(size_t)i]];
  }

  // ---- work list: per output block (by id) its sector pairs
  const i64 obs = out.blksize;
  std::vector<i32> o_dlp((size_t)nout), o_drp((size_t)nout), o_pb((size_t)nout + 1, 0);
  std::vector<i64> o_off((size_t)nout);
  std::vector<HeffPair> pairs;
  // per output block: its LDS tile needs (elements) and tile count
  std::vector<std::array<int, 4>> need((size_t)nout, std::array<int, 4>{1, 1, 1, 1});
  int max_tiles = 0;
  double flops = 0, bytes_out = 0;
  std::vector<char> psi_used((size_t)std::max<i64>(psi.n, 1), 0);
  for (i64 o = 0; o < nout; ++o) {
    const int a = by_id[(size_t)o].first, c = by_id[(size_t)o].second;
    const int dlp = elp[(size_t)a], drp = erp[(size_t)c];
    o_dlp[(size_t)o] = dlp;
    o_drp[(size_t)o] = drp;
    o_off[(size_t)o] = o * obs;  // ids 1..n: block id o+1 at o*blksize
    bytes_out += 8.0 * P * dlp * drp;
    if (dlp > 32 || drp > 32)
      fail(SB_E_LIMIT, "H_eff plan: a bond sector above 32 states (use the contract! chain)");
    const int LM = up16(dlp), NC = up16(drp);
    max_tiles = std::max<int>(max_tiles, (up16((int)(P * dlp)) / 16) * (NC / 16));
    for (auto& lb : Lrow[(size_t)a]) {
      const int il = lb.first;
      const int lmask = hmL[(size_t)(3 * lb.second)];
      for (auto& rb : Rrow[(size_t)c]) {
        const int ir = rb.first;
        auto it = psi_at.find({il, ir});
        if (it == psi_at.end()) continue;
        const int rmask = hmR[(size_t)(3 * rb.second)];
        int wlm = 0;
        for (i64 wl = 0; wl < w; ++wl)
          if ((lmask >> wl & 1) && (wl_useful[(size_t)wl] & rmask)) wlm |= 1 << wl;
        const int dl = el[(size_t)il], dr = er[(size_t)ir];
        if (!wlm || !dl || !dr || !dlp || !drp) continue;
        HeffPair hp;
        hp.psi = (i64)(ipsi[(size_t)it->second] - 1) * psi.blksize;
        hp.L = (i64)(iL[(size_t)lb.second] - 1) * L.blksize;
        hp.R = (i64)(iR[(size_t)rb.second] - 1) * R.blksize;
        hp.dl = dl;
        hp.dr = dr;
        hp.wlm = wlm;
        hp.wrm = rmask;
        pairs.push_back(hp);
        psi_used[(size_t)it->second] = 1;
        {
          const int K1 = up4(dl), N1 = up16((int)(P * dr)), K2 = up4(dr);
          auto& nd = need[(size_t)o];
          nd[0] = std::max(nd[0], K1 * N1);
          nd[1] = std::max(nd[1], LM * N1);
          nd[2] = std::max(nd[2], __builtin_popcount(wlm) * LM * K1);
          nd[3] = std::max(nd[3], __builtin_popcount(rmask) * K2 * NC);
        }
        for (i64 wl = 0; wl < w; ++wl) {
          if (!(wlm >> wl & 1)) continue;
          flops += 2.0 * P * dlp * dl * dr;  // A = L_wl psi
          for (i64 wr = 0; wr < w; ++wr) {
            const int lk = lidx[(size_t)(wl * w + wr)];
            if (lk < 0 || !(rmask >> wr & 1)) continue;
            double rowsum = 0;
            for (i64 pp = 0; pp < P; ++pp) {
              const int n = wptr[(size_t)(lk * P + pp + 1)] - wptr[(size_t)(lk * P + pp)];
              if (n) rowsum += 2.0 * n + 2.0;  // W12 row . A, then times R
            }
            flops += rowsum * dlp * drp * dr;
          }
        }
      }
    }
    o_pb[(size_t)o + 1] = (i32)pairs.size();
  }
  if (max_tiles > kMaxTiles)
    fail(SB_E_LIMIT, "H_eff plan: an output block needs more than 16 tiles of 16 x 16 (use the contract! chain)");
  if ((i64)pairs.size() >= (1ll << 31)) fail(SB_E_LIMIT, "H_eff plan: too many sector pairs");
  auto lds_of = [&](const std::array<int, 4>& nd) {
    return fp32 ? heff_lds_bytes<float>(nd[0], nd[1], nd[2], nd[3]) : heff_lds_bytes<double>(nd[0], nd[1], nd[2], nd[3]);
  };
  // launch classes by LDS need (a launch reserves its largest block's LDS for
  // every workgroup, so small blocks go together to keep occupancy up); in a
  // class, blocks with the most work first
  const size_t class_cap[3] = {24 * 1024, 48 * 1024, kMaxLds};
  std::vector<std::vector<i32>> cls(3);
  std::vector<i64> work((size_t)nout, 0);
  for (i64 o = 0; o < nout; ++o) {
    if (!o_dlp[(size_t)o] || !o_drp[(size_t)o]) continue;  // no entry to write
    const size_t b = lds_of(need[(size_t)o]);
    if (b > kMaxLds) fail(SB_E_LIMIT, "H_eff plan: sector blocks too large for the fused kernel's LDS tiles");
    int c = 0;
    while (b > class_cap[c]) ++c;
    cls[(size_t)c].push_back((i32)o);
    work[(size_t)o] = (i64)(o_pb[(size_t)o + 1] - o_pb[(size_t)o] + 1) * o_dlp[(size_t)o] * o_drp[(size_t)o];
  }
  // workgroups: small blocks whole (4 tiles a wave); in the largest class a
  // block's tiles spread over several workgroups (1 tile a wave), each
  // recomputing the pair products the tiles need, for more parallel work
  std::vector<int2> items;
  struct Span {
    size_t count;
    std::array<int, 4> mx;
    int tpw;
  };
  std::vector<Span> cls_span;  // per class with work
  for (size_t c = 0; c < cls.size(); ++c) {
    auto& v = cls[c];
    if (v.empty()) continue;
    std::stable_sort(v.begin(), v.end(), [&](i32 x, i32 y) { return work[(size_t)x] > work[(size_t)y]; });
    std::array<int, 4> mx{1, 1, 1, 1};
    for (i32 o : v)
      for (int j = 0; j < 4; ++j) mx[(size_t)j] = std::max(mx[(size_t)j], need[(size_t)o][(size_t)j]);
    const int tpw = c + 1 == cls.size() ? 1 : kTilesPerWave;
    size_t n = 0;
    for (i32 o : v) {
      const int tiles = (up16((int)(P * o_dlp[(size_t)o])) / 16) * (up16(o_drp[(size_t)o]) / 16);
      for (int t0 = 0; t0 < tiles; t0 += kWaves * tpw, ++n) items.push_back(make_int2(o, t0));
    }
    cls_span.push_back({n, mx, tpw});
  }
  if (items.size() >= (size_t)1 << 31) fail(SB_E_LIMIT, "H_eff plan: too many workgroups");
  // matrix-core work one execution issues (16 x 16 x 4 products, 2048 FLOPs each)
  double mfma = 0;
  {
    size_t it = 0;
    for (auto& cs : cls_span)
      for (size_t j = 0; j < cs.count; ++j, ++it) {
        const int o = items[it].x, t0 = items[it].y;
        const int dlp = o_dlp[(size_t)o], drp = o_drp[(size_t)o];
        const int RT = up16((int)(P * dlp)) / 16, ML = up16(dlp) / 16, CT = up16(drp) / 16;
        const int ntile = std::min(RT * CT - t0, kWaves * cs.tpw);
        for (int q = o_pb[(size_t)o]; q < o_pb[(size_t)o + 1]; ++q) {
          const HeffPair& hp = pairs[(size_t)q];
          const int K1 = up4(hp.dl), N1 = up16((int)(P * hp.dr)), K2 = up4(hp.dr);
          for (int wl = 0; wl < w; ++wl) {
            if (!(hp.wlm >> wl & 1)) continue;
            mfma += (double)ML * (N1 / 16) * (K1 / 4);
            for (int wr = 0; wr < w; ++wr)
              if ((hp.wrm >> wr & 1) && lidx[(size_t)(wl * w + wr)] >= 0) mfma += (double)ntile * (K2 / 4 + 4);
          }
        }
      }
  }

  // ---- install Hpsi's block list (overwrite semantics), payload zero
  out.n = nout;
  out.dirty = false;
  up(ctx, out.keys, okeys);
  up(ctx, out.ids, oids);
  out.data.alloc(ctx, (size_t)std::max<i64>(nout * obs, 1));
  if (nout) k_zero(ctx, out.data.get(), (size_t)(nout * obs) * sizeof(double));
  // boxes of Hpsi: the reach of every block (entries beyond stay zero)
  {
    std::vector<i32> hb((size_t)std::max<i64>(nout * 4, 1)), hf((size_t)std::max<i64>(nout, 1), 0);
    for (i64 i = 0; i < nout; ++i) {
      const int id = oids[(size_t)i] - 1;
      hb[(size_t)(4 * i)] = (i32)d1;
      hb[(size_t)(4 * i + 1)] = (i32)d2;
      hb[(size_t)(4 * i + 2)] = o_dlp[(size_t)id];
      hb[(size_t)(4 * i + 3)] = o_drp[(size_t)id];
    }
    up(ctx, out.box, hb);
    up(ctx, out.bflag, hf);
  }
  out.external = false;
  out.touch();

  auto st = std::make_shared<HeffState>();
  st->psi = &psi;
  st->L = &L;
  st->R = &R;
  st->W1 = &W1;
  st->W2 = &W2;
  st->out = &out;
  st->vL = L.version;
  st->vR = R.version;
  st->vW1 = W1.version;
  st->vW2 = W2.version;
  st->fp32 = fp32;
  st->nout = nout;
  up(ctx, st->items, items);
  up(ctx, st->o_dlp, o_dlp);
  up(ctx, st->o_drp, o_drp);
  up(ctx, st->o_pb, o_pb);
  up(ctx, st->o_off, o_off);
  up(ctx, st->pairs, pairs);
  up(ctx, st->lidx, lidx);
  up(ctx, st->w16, w16);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host vectors go out of scope
  HeffK k{};
  k.o_dlp = st->o_dlp.get();
  k.o_drp = st->o_drp.get();
  k.o_pb = st->o_pb.get();
  k.o_off = st->o_off.get();
  k.pairs = st->pairs.get();
  k.lidx = st->lidx.get();
  k.w16 = st->w16.get();
  k.lp = __builtin_ctz((unsigned)P);
  k.w = (int)w;
  k.Dl = (int)Dl;
  k.Dr = (int)Dr;
  size_t lds_max = 0;
  // the class of the largest blocks first, on the context's stream; the others on
  // side streams of their own, forked from and joined to it at every execution
  std::vector<size_t> corder(cls_span.size());
  for (size_t i = 0; i < corder.size(); ++i) corder[i] = corder.size() - 1 - i;  // largest blocks start first
  std::vector<size_t> cfirst(cls_span.size(), 0);
  for (size_t i = 1; i < cls_span.size(); ++i) cfirst[i] = cfirst[i - 1] + cls_span[i - 1].count;
  for (size_t ci : corder) {
    const Span& cs = cls_span[ci];
    HeffState::Launch la;
    la.k = k;
    la.k.items = st->items.get() + cfirst[ci];
    la.k.tpw = cs.tpw;
    la.k.sPs = cs.mx[0];
    la.k.sAs = cs.mx[1];
    la.k.sLs = cs.mx[2];
    la.k.sRs = cs.mx[3];
    la.lds = lds_of(cs.mx);
    la.n = (i64)cs.count;
    lds_max = std::max(lds_max, la.lds);
    st->launches.push_back(la);
  }
  for (size_t i = 1; i < st->launches.size(); ++i) {
    auto& la = st->launches[i];
    HIP_CHECK(hipStreamCreateWithFlags(&la.stream, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&la.done, hipEventDisableTiming));
  }
  if (st->launches.size() > 1) HIP_CHECK(hipEventCreateWithFlags(&st->fork, hipEventDisableTiming));
  if (lds_max) {
    if (fp32)
      HIP_CHECK(hipFuncSetAttribute((const void*)heff_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds_max));
    else
      HIP_CHECK(hipFuncSetAttribute((const void*)heff_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds_max));
  }

  auto pl = std::make_unique<Plan>();
  pl->ctx = ctx;
  pl->family = "heff_fused";
  pl->hold.push_back(st);
  HeffState* s = st.get();
  pl->steps.push_back({"heff_fused", [ctx, s]() {
                         if (s->L->version != s->vL || s->R->version != s->vR || s->W1->version != s->vW1 ||
                             s->W2->version != s->vW2)
                           fail(SB_E_ERROR, "H_eff plan: L, R, W1 or W2 changed since the plan was made");
                         if (s->fork) HIP_CHECK(hipEventRecord(s->fork, ctx->stream));
                         for (auto& la : s->launches) {
                           HeffK k = la.k;
                           k.psi = s->psi->dptr();
                           k.L = s->L->dptr();
                           k.R = s->R->dptr();
                           k.out = s->out->dptr();
                           hipStream_t st = la.stream ? la.stream : ctx->stream;
                           if (la.stream) HIP_CHECK(hipStreamWaitEvent(la.stream, s->fork, 0));
                           if (s->fp32)
                             hipLaunchKernelGGL(heff_kernel<float>, dim3((unsigned)la.n), dim3(kThreads), la.lds,
                                                st, k);
                           else
                             hipLaunchKernelGGL(heff_kernel<double>, dim3((unsigned)la.n), dim3(kThreads), la.lds,
                                                st, k);
                           HIP_CHECK(hipGetLastError());
                           if (la.done) HIP_CHECK(hipEventRecord(la.done, la.stream));
                         }
                         for (auto& la : s->launches)
                           if (la.done) HIP_CHECK(hipStreamWaitEvent(ctx->stream, la.done, 0));
                       }});
  pl->flops = flops;
  // algorithmic bytes: the psi blocks the pairs read, L and R inside their reach, Hpsi written
  double bpsi = 0, benv = 0;
  for (i64 i = 0; i < psi.n; ++i)
    if (psi_used[(size_t)i])
      bpsi += 8.0 * P * el[(size_t)(kpsi[(size_t)(2 * i)] - 1)] * er[(size_t)(kpsi[(size_t)(2 * i + 1)] - 1)];
  for (i64 i = 0; i < L.n; ++i)
    benv += 8.0 * w * hmL[(size_t)(3 * i + 1)] * hmL[(size_t)(3 * i + 2)];
  for (i64 i = 0; i < R.n; ++i)
    benv += 8.0 * w * hmR[(size_t)(3 * i + 1)] * hmR[(size_t)(3 * i + 2)];
  pl->gemm_bytes = bpsi + benv + bytes_out;
  pl->a_bytes = bpsi;
  pl->b_bytes = benv;
  pl->c_bytes = bytes_out;
  pl->npairs = (i64)pairs.size();
  pl->mfma_flops = 2048.0 * mfma;
  pl->nblocks_out = nout;
  pl->out = &out;
  pl->out_box = true;
  pl->in_a = &psi;
  pl->in_b = &L;
  return pl;
}

}  // namespace sb
