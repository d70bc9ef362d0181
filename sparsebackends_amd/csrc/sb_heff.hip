// Two-site H_eff*psi of DMRG as one plan: a pack kernel and one fused kernel.
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
// with p = (s1, s2) and W12[wl,wr][p', p] = sum_wm W1[wl, s1', s1, wm] W2[wm, s2', s2, wr].
//
// Physical rows.  A psi block (Ql, Qr) of a charge-conserving state holds
// nonzeros only in the rows p whose charge closes q(Ql) + q(p) = q(Qr): 1, 2
// or 4 of the 16 of a Hubbard two-site block.  Every execution first runs a
// pack kernel that finds, per psi block, the rows p holding a nonzero and
// copies them, 4 rows to a chunk, into matrix-core panels; the fused kernel
// then multiplies only those rows, and skips a (pair, link pair) item whose
// MPO block W12[wl,wr][:, rows] is zero (an output block of forbidden charge
// is then only written with zeros).  Nothing is assumed about psi's values:
// the rows are found from them at every execution.
//
// One launch of 4-wave workgroups covers every output block: a block of more
// than 4 rows l' takes a workgroup, smaller blocks go four to a workgroup, one
// per wave.  Per block, its active (sector pair, chunk of 4 rows) entries are
// listed in LDS; per entry A_wl = L_wl^T psi (l' x (p_c, r)) runs on the
// matrix cores with its operands read straight from the panels (row strides
// chosen so that 16 lanes read 16 consecutive values), only its live rows
// kept in LDS; per link pair the wave's output tiles take Z = A R_wr^T
// ((p_c, l') x r') and then W12[:, rows] Z, with Z fed to the matrix cores
// straight from its accumulator registers.  No atomics: each output element
// is written once.  A block with nothing to multiply is written with zeros,
// unless it is zero from the plan's previous execution already.  The
// environment panels are built once, at planning (L, R and the MPO must keep
// their values, which the plan checks).
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
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "sb_contract.h"

namespace sb {

namespace {

constexpr size_t kMaxLds = 160 * 1024;  // dynamic LDS of one workgroup (gfx950)
constexpr size_t kLdsBudget = 80 * 1024;  // above this, A panels are computed one at a time

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

__host__ __device__ __forceinline__ int up4(int x) { return (x + 3) & ~3; }
__host__ __device__ __forceinline__ int up16(int x) { return (x + 15) & ~15; }

// row stride (elements) of an LDS panel of n columns (n a multiple of 16):
// 16 modulo 32 doubles or 64 floats, so that the 4 k-rows x 16 columns a
// matrix-core operand read touches fall on distinct banks
template <class T>
__host__ __device__ __forceinline__ int lds_stride(int n) {
  constexpr int mod = sizeof(T) == 8 ? 32 : 64;
  return n + ((16 - n % mod) % mod + mod) % mod;
}

// one sector pair (Ql, Qr) of an output block
struct HPair {
  i64 pp;          // packed psi: the first chunk panel of the psi block
  i64 lp, rp;      // packed L / R: the first link panel of the environment block
  i32 blk;         // the psi block's row in the pack tables
  i32 item0;       // first link item: (iL | iR << 5 | lk << 10, W12 columns of lk), sorted by iL
  int16_t nitem;
  uint8_t nl, nr;  // link panels of the L and R blocks
  uint8_t dl, dr;  // reach of l and r
  uint16_t pc[4];  // per L panel (nl <= 4): the W12 columns its items reach
};

// one psi block to pack: payload offset, panel base, reach
struct PsiJob {
  i64 src, dst;
  i32 blk, dl, dr, pstr;
  i32 u0, nu;  // the pairs that read this block: uses[u0, u0 + nu) = (pair, output block)
};

// one environment link panel to build: payload offset, panel base, link,
// rows (inner bond index) and columns (outer bond index) inside the reach
struct EnvJob {
  i64 src, dst;
  i32 link, rows, cols, stride;
};

template <class T>
struct HK {
  const T *pp, *lp, *rp, *w12;  // psi chunks, L and R panels, W12 [lk][p'][p]
  const i32* pmask;             // per psi block: row mask | row count << 16
  const int4* pinfo;            // per pair: its chunks with work, their rows and L panels (the pack kernel's)
  double* out;
  const int4* jobs;   // per workgroup: one block (y == -2) or up to 4 blocks, one per wave (-1: none)
  const i32 *o_dlp, *o_drp, *o_pb;
  const i64* o_off;
  const HPair* pairs;
  const int2* items;
  int Dl;
  int apb;  // L panels whose A products LDS holds at once
  // LDS region of a block (bytes; the entry list at 0): its link items (int2,
  // copied at the start) and the A products; a wave on its own has its own
  // region, wave_lds bytes apart
  int off_it, off_as, wave_lds;
  i32* dirty;          // per output block: 0 when its payload is known to be zero
  const i32* active;   // per output block: == epoch when a pair of it has work this execution
  int epoch;
  const i32* o_ib;     // per output block: its first link item (the items of a block are contiguous)
};

// one active (sector pair, row chunk) of an output block, listed in LDS at
// the start of its team's work with what the walk needs of its pair (no
// global load of the pair per entry): the chunk's psi panel, the pair's L
// and R panels, its link items, its rows p and the L panels an active item uses
struct HEnt {
  i64 pp;           // the chunk's psi panel
  i64 lp, rp;       // the pair's first L / R link panel
  int item0, nitem; // the pair's link items
  int geo;          // dl | dr << 8 | nl << 16
  int cm, amask;
  int pad;
};
static_assert(sizeof(HEnt) == 48, "entry records are 48 bytes");

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
  // rows m of a Z tile as (row chunk position, l' offset), ordered so that
  // accumulator register v holds l' offset v with the k index = position
  static __device__ __forceinline__ int zpos(int m) { return m & 3; }
  static __device__ __forceinline__ int zoff(int m) { return m >> 2; }
};
template <>
struct Mfma<float> {
  typedef f4 acc;
  // the FP32 16x16x4 form: register v holds row 4 (lane / 16) + v
  static __device__ __forceinline__ acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int v, int lane) { return 4 * (lane >> 4) + v; }
  static __device__ __forceinline__ int zpos(int m) { return m >> 2; }
  static __device__ __forceinline__ int zoff(int m) { return m & 3; }
};

// a value every lane of the wave holds alike, in a scalar register
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ i64 uni(i64 v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)v), hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (i64)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int o = 32; o; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// the position of the j-th set bit of a row list packed as 4-bit fields
__device__ __forceinline__ int nib(uint64_t lst, int j) { return (int)((lst >> (4 * j)) & 15); }

__device__ __forceinline__ uint64_t row_list(int msk, int P) {
  uint64_t lst = 0;
  int c = 0;
  for (int b = 0; b < P; ++b)
    if (msk >> b & 1) lst |= (uint64_t)b << (4 * c++);
  return lst;
}

// ---- pack kernels

template <class T>
struct PackK {
  const double* psi;
  const PsiJob* jobs;
  int Dl;
  T* pp;
  i32* pmask;
  const int2* uses;  // (pair, output block) of the pairs reading each psi block
  const HPair* pairs;
  const int2* items;
  i32* active;  // per output block: the execution's epoch when a pair of it has work
  int epoch;
  int4* pinfo;  // per pair: chunks with work (bits), rows of chunks 0 and 1 (16 bits each), L panels (8 bits per chunk), psi's row mask
};

// one workgroup per psi block: find the rows p holding a nonzero inside the
// reach, copy them 4 to a chunk into panels [l][p_c + 4 r] (row stride pstr;
// entries beyond the reach are zero from planning and never written), and
// mark the output blocks of the pairs that have work with these rows (a
// chunk of rows meets the W12 columns of one of the pair's link items)
template <class T, int P>
__global__ __launch_bounds__(64) void heff_pack_psi(PackK<T> k) {
  // one wave per psi block: the row mask is a wave reduction (no LDS, no
  // barrier), and the pairs reading the block are fetched while psi is read
  const PsiJob jb = k.jobs[blockIdx.x];
  const int lane = threadIdx.x;
  const int n = jb.dl * jb.dr;
  const i64 Dl = k.Dl;
  int2 u0 = make_int2(0, 0);
  HPair pr0{};
  if (lane < jb.nu) {
    u0 = k.uses[jb.u0 + lane];
    pr0 = k.pairs[u0.x];
  }
  int bits = 0;
  for (int t = lane; t < n; t += 64) {
    const int r = t / jb.dl, l = t - r * jb.dl;
    // P consecutive values, 16-byte aligned (block payloads and rows are multiples of P >= 4 values)
    const double2* s = reinterpret_cast<const double2*>(k.psi + jb.src + (i64)P * (l + Dl * r));
    double2 v[P / 2];
#pragma unroll
    for (int h = 0; h < P / 2; ++h) v[h] = s[h];
#pragma unroll
    for (int h = 0; h < P / 2; ++h) {
      if (v[h].x != 0.0) bits |= 1 << (2 * h);  // NaN counts as nonzero
      if (v[h].y != 0.0) bits |= 1 << (2 * h + 1);
    }
  }
  const int msk = wave_or(bits), np = __popc(msk);
  if (lane == 0) k.pmask[jb.blk] = msk | (np << 16);
  const uint64_t lst = row_list(msk, P);
  const int K1 = up4(jb.dl);
  const i64 csz = (i64)K1 * jb.pstr;
  for (int t = lane; t < n; t += 64) {  // the same values again, from this lane's cache lines
    const int r = t / jb.dl, l = t - r * jb.dl;
    const double* s = k.psi + jb.src + (i64)P * (l + Dl * r);
    T* d = k.pp + jb.dst + (i64)l * jb.pstr + 4 * r;
    for (int ch = 0; 4 * ch < np; ++ch) {
      T v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = 4 * ch + j < np ? (T)s[nib(lst, 4 * ch + j)] : (T)0;
      T* dc = d + ch * csz;
#pragma unroll
      for (int j = 0; j < 4; ++j) dc[j] = v[j];
    }
  }
  for (int t = lane; t < jb.nu; t += 64) {
    int2 u = u0;
    HPair pr = pr0;
    if (t != lane) {
      u = k.uses[jb.u0 + t];
      pr = k.pairs[u.x];
    }
    // per chunk of rows: the L panels of the items meeting it (those with work)
    int cnt = 0, cms = 0, ams = 0;
    for (int ch = 0; 4 * ch < np; ++ch) {
      int cm = 0;
      for (int j = 4 * ch; j < min(np, 4 * ch + 4); ++j) cm |= 1 << nib(lst, j);
      int am = 0;
      if (pr.nl <= 4) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (i < pr.nl && (pr.pc[i] & cm)) am |= 1 << i;
      } else {
        for (int it = 0; it < pr.nitem; ++it) {
          const int2 c2 = k.items[pr.item0 + it];
          if (c2.y & cm) am |= 1 << (c2.x & 31);
        }
      }
      if (am) {
        if (ch < 2) cms |= cm << (16 * ch);
        ams |= (int)((unsigned)(am & 255) << (8 * ch));
        cnt |= 1 << ch;
      }
    }
    k.pinfo[u.x] = make_int4(cnt, cms, ams, msk | (np << 16));
    if (cnt) k.active[u.y] = k.epoch;
  }
}

// one workgroup per environment link panel: panel[k][m] = X[link, m, k]
// (k the inner bond index, m the outer) inside the reach
template <class T>
__global__ __launch_bounds__(256) void heff_pack_env(const double* __restrict__ src, const EnvJob* __restrict__ jobs,
                                                     int w, i64 D, T* __restrict__ dst) {
  const EnvJob jb = jobs[blockIdx.x];
  const int n = jb.rows * jb.cols;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int k = t / jb.cols, m = t - k * jb.cols;
    dst[jb.dst + (i64)k * jb.stride + m] = (T)src[jb.src + jb.link + (i64)w * (m + D * k)];
  }
}

// copy n16 16-byte units global -> LDS, all loads of a thread's batch first
__device__ __forceinline__ void copy16(void* __restrict__ dst, const void* __restrict__ src, int n16, int tid,
                                       int nt) {
  uint4* d = reinterpret_cast<uint4*>(dst);
  const uint4* s = reinterpret_cast<const uint4*>(src);
  int t = tid;
  for (; t + 3 * nt < n16; t += 4 * nt) {
    const uint4 a = s[t], b = s[t + nt], c = s[t + 2 * nt], e = s[t + 3 * nt];
    d[t] = a;
    d[t + nt] = b;
    d[t + 2 * nt] = c;
    d[t + 3 * nt] = e;
  }
  for (; t < n16; t += nt) d[t] = s[t];
}

// ---- the fused kernel
//
// A workgroup of 4 waves takes either one output block of more than 4 rows l'
// (a team of 4 waves: A products split over the waves, output row groups of
// 4 l' owned by waves w, w + 4, ...; workgroup barriers) or up to 4 blocks of
// at most 4 rows l', one per wave, each wave on its own in its own LDS region
// (no workgroup barriers).  Output rows (p', l') come in groups of 4 l'; per
// group P / 4 accumulator tiles of 16 rows (p' fastest), and column tiles of
// 16 r' (at most CT), held in registers over the block's whole work.
//
// A block's team first lists its active (pair, chunk) entries in LDS (a chunk
// is active when an item's W12 columns meet its rows) and copies its link
// items there.  Per entry, A_i = L_i^T psi_chunk on the matrix cores with
// the operands read straight from the panels (L1/L2-resident, 16 consecutive
// values per 16 lanes) goes to LDS, its live rows only; per link item the
// group's tiles take Z = A R^T and then W12[:, rows] Z, Z fed to the matrix
// cores straight from its accumulator registers.
template <class T, int P, int GPW, int CT, int KS>
__global__ __launch_bounds__(256, KS == 4 ? 4 : 2) void heff_kernel(HK<T> k) {
  using M = Mfma<T>;
  constexpr int TPG = P / 4;   // accumulator tiles per group of 4 l'
  constexpr int LPT = 16 / P;  // l' per accumulator tile
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int sm_ne[4];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m = lane & 15, kq = lane >> 4;
  const int4 job = k.jobs[blockIdx.x];
  const bool team4 = job.y == -2;  // the same for the whole workgroup
  const int NW = team4 ? 4 : 1;     // waves on this block
  const int tw = team4 ? wave : 0;  // this wave's place in its team
  const int o = team4 ? job.x : (wave == 0 ? job.x : wave == 1 ? job.y : wave == 2 ? job.z : job.w);
  if (o < 0) return;  // a wave without a block (never in a team of 4)
  unsigned char* reg = lds_raw + (team4 ? 0 : wave * k.wave_lds);
  HEnt* ents = reinterpret_cast<HEnt*>(reg);
  int2* its = reinterpret_cast<int2*>(reg + k.off_it);
  T* As = reinterpret_cast<T*>(reg + k.off_as);
  // ordering of the team's LDS traffic: a workgroup barrier for a team of 4,
  // for a wave on its own the wave's LDS operations in program order
  auto bar = [&]() {
    if (team4) {
      __syncthreads();
    } else {
      __threadfence_block();
      __builtin_amdgcn_wave_barrier();
    }
  };
  const int dlp = k.o_dlp[o], drp = k.o_drp[o];
  const int ng = (dlp + 3) >> 2, nct = (drp + 15) >> 4;
  const int LM = up16(dlp), LSTR = lds_stride<T>(LM), RSTR = lds_stride<T>(nct * 16);
  const int nth = 64 * NW, ttid = 64 * tw + lane;
  double* ob = k.out + k.o_off[o];
  const i64 Dl = k.Dl;
  // nothing to multiply: zeros over the reach, (P dlp) contiguous values per
  // r', unless the block is zero from the previous execution already
  auto zero_block = [&]() {
    if (!k.dirty[o]) return;
    const int run = P * dlp / 2;  // 16-byte units per r'
    for (int t = ttid; t < run * drp; t += nth) {
      const int rq = t / run, u = t - rq * run;
      reinterpret_cast<uint4*>(ob + (i64)P * Dl * rq)[u] = make_uint4(0, 0, 0, 0);
    }
    if (ttid == 0) k.dirty[o] = 0;
  };
  if (k.active[o] != k.epoch) {  // no pair of the block has work with psi's rows
    zero_block();
    return;
  }
  const int q0 = k.o_pb[o], qe = k.o_pb[o + 1];
  const int ib0 = k.o_ib[o], nib_ = k.o_ib[o + 1] - ib0;
  for (int t = ttid; t < nib_; t += nth) its[t] = k.items[ib0 + t];

  // ---- the entry list (the team's first wave; a deterministic order: pairs, then chunks)
  if (tw == 0) {
    int base = 0;
    for (int qb = q0; qb < qe; qb += 64) {
      const int q = qb + lane;
      int cnt = 0, cms = 0, ams = 0;  // per chunk: rows (16 bits, chunks 0 and 1), L panels (8 bits)
      HPair pr{};
      int pm = 0;
      if (q < qe) {
        pr = k.pairs[q];
        const int4 pi = k.pinfo[q];
        cnt = pi.x;
        cms = pi.y;
        ams = pi.z;
        pm = pi.w;
      }
      const int ncnt = __popc(cnt);
      int inc = ncnt;  // inclusive scan of the counts over the wave
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(inc, d);
        if (lane >= d) inc += y;
      }
      int pos = base + inc - ncnt;
      if (cnt) {
        for (int ch = 0; ch < 4; ++ch) {
          if (!(cnt >> ch & 1)) continue;
          int cm;
          if (ch < 2) {
            cm = (cms >> (16 * ch)) & 0xffff;
          } else {  // chunks 2 and 3 (16 rows): the rows again
            const int np = pm >> 16;
            const uint64_t lst = row_list(pm & 0xffff, P);
            cm = 0;
            for (int j = 4 * ch; j < min(np, 4 * ch + 4); ++j) cm |= 1 << nib(lst, j);
          }
          int am = (int)(((unsigned)ams >> (8 * ch)) & 255u);
          if (pr.nl > 8) {  // more L panels than the packed mask holds
            am = 0;
            for (int it = 0; it < pr.nitem; ++it) {
              const int2 c2 = k.items[pr.item0 + it];
              if (c2.y & cm) am |= 1 << (c2.x & 31);
            }
          }
          HEnt* en = ents + pos++;
          en->pp = pr.pp + (i64)ch * up4(pr.dl) * lds_stride<T>(4 * up4(pr.dr));
          en->lp = pr.lp;
          en->rp = pr.rp;
          en->item0 = pr.item0;
          en->nitem = pr.nitem;
          en->geo = pr.dl | (pr.dr << 8) | (pr.nl << 16);
          en->cm = cm;
          en->amask = am;
        }
      }
      base += __shfl(inc, 63);
    }
    if (lane == 0) sm_ne[wave] = base;
  }
  bar();
  const int nE = sm_ne[team4 ? 0 : wave];

  typename M::acc acc[GPW][CT][TPG];
#pragma unroll
  for (int a = 0; a < GPW; ++a)
#pragma unroll
    for (int b = 0; b < CT; ++b)
#pragma unroll
      for (int c = 0; c < TPG; ++c) acc[a][b][c] = typename M::acc{0, 0, 0, 0};

  if (nE == 0) {
    zero_block();
    return;
  }
  if (ttid == 0) k.dirty[o] = 1;

  const int AR = 4 * ng;  // live rows of the A products
  const int zp = M::zpos(m), zl = M::zoff(m);
  const T* const gpp = k.pp;
  const T* const glp = k.lp;
  const T* const grp = k.rp;
  const T* const gw12 = k.w12;
  // this lane's row of an entry's chunk (the k index of the W12 operand): the
  // kq-th set bit of the chunk's row mask, -1 past the chunk's rows
  auto row_of = [=](int cm) {
    int x = cm;
    x = kq >= 1 ? x & (x - 1) : x;
    x = kq >= 2 ? x & (x - 1) : x;
    x = kq >= 3 ? x & (x - 1) : x;
    return x ? __builtin_ctz(x) : -1;
  };
  const int lm_sh = LM > 16 ? 1 : 0;  // row tiles of A: LM / 16 = 1 or 2
  for (int e = 0; e < nE; ++e) {
    const HEnt& en = ents[e];
    const int cm = uni(en.cm), amask = uni(en.amask), geo = uni(en.geo);
    const int item0 = uni(en.item0), nitem = uni(en.nitem);
    const i64 lp0 = uni(en.lp), rp0 = uni(en.rp);
    const int dl = geo & 255, dr = (geo >> 8) & 255, nl = geo >> 16;
    const int K1 = up4(dl), K2 = up4(dr), N1 = 4 * K2, PSTR = lds_stride<T>(N1);
    const int prow = row_of(cm);
    const T* Pg = gpp + uni(en.pp) + m;
    for (int b0 = 0; b0 < nl; b0 += k.apb) {
      const int b1 = min(nl, b0 + k.apb);
      if (!(((unsigned)amask >> b0) & ((1u << (b1 - b0)) - 1u))) continue;
      // the link items of the batch meeting the chunk's rows: one ballot per 64
      // items, each lane decoding one item, the walk reading them with
      // readlane; the first 64 here, so that (kernels for sectors of up to 16)
      // the first item's W12 and R operands are fetched before the A products
      int code0 = 0;
      bool on0 = false;
      if (lane < nitem) {
        const int2 c2 = its[item0 - ib0 + lane];
        const int iL_l = c2.x & 31;
        code0 = c2.x;
        on0 = iL_l >= b0 && iL_l < b1 && (c2.y & cm);
      }
      const uint64_t im0 = __ballot(on0);
      constexpr bool PRE = KS == 4 && CT == 1;
      T wv0 = (T)0, rb0[KS];
      if constexpr (PRE) {
        if (tw < ng && im0) {
          const int code = __builtin_amdgcn_readlane(code0, __builtin_ctzll(im0));
          const int iR = (code >> 5) & 31, lk = code >> 10;
          wv0 = prow >= 0 ? gw12[lk * P * P + (m % P) * P + prow] : (T)0;
          const T* Rg = grp + rp0 + (i64)iR * K2 * RSTR + m;
#pragma unroll
          for (int st = 0; st < KS; ++st) rb0[st] = 4 * st < K2 ? Rg[(4 * st + kq) * RSTR] : (T)0;
        }
      }
      bar();  // the previous user of As is done
      // A_i[l'][c] = sum_l L_i[l][l'] psi[l][c], c = p_c + 4 r: operands straight from the panels
      // tiles of the active panels of the batch, dealt to the team's waves in turn
      const int ntl = N1 >> 4, per = ntl << lm_sh;
      unsigned pbits = ((unsigned)amask >> b0) & ((1u << (b1 - b0)) - 1u);
      for (int tk = 0; pbits; pbits &= pbits - 1) {
        const int ii = __builtin_ctz(pbits), i = b0 + ii;
        for (int rem = 0; rem < per; ++rem, ++tk) {
        if (NW > 1 && (tk & 3) != tw) continue;
        const int mi = rem & lm_sh, ni = rem >> lm_sh;
        const T* Lg = glp + lp0 + (i64)i * K1 * LSTR + mi * 16 + m;
        const T* Pt = Pg + ni * 16;
        T la[KS], pa[KS];
#pragma unroll
        for (int st = 0; st < KS; ++st) {
          const int kk = 4 * st + kq;
          la[st] = 4 * st < K1 ? Lg[kk * LSTR] : (T)0;
          pa[st] = 4 * st < K1 ? Pt[kk * PSTR] : (T)0;
        }
        typename M::acc c = typename M::acc{0, 0, 0, 0};
#pragma unroll
        for (int st = 0; st < KS; ++st)
          if (4 * st < K1) c = M::run(la[st], pa[st], c);
        T* At = As + (i64)ii * AR * PSTR + ni * 16 + m;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int lq = mi * 16 + M::row(v, lane);
          if (lq < AR) At[lq * PSTR] = c[v];
        }
        }
      }
      bar();
      if (tw >= ng) continue;  // a wave of a team past the block's row groups has no output tiles
      bool first = PRE && im0 != 0;  // the first item's operands are in wv0 / rb0
      for (int c0 = 0; c0 < nitem; c0 += 64) {
        int code_l = code0;
        uint64_t imask = im0;
        if (c0) {
          bool on = false;
          if (c0 + lane < nitem) {
            const int2 c2 = its[item0 - ib0 + c0 + lane];
            const int iL_l = c2.x & 31;
            code_l = c2.x;
            on = iL_l >= b0 && iL_l < b1 && (c2.y & cm);
          }
          imask = __ballot(on);
        }
        for (; imask; imask &= imask - 1) {
        const int code = __builtin_amdgcn_readlane(code_l, __builtin_ctzll(imask));
        const int iL = code & 31, iR = (code >> 5) & 31, lk = code >> 10;
        const T wv = first ? wv0 : prow >= 0 ? gw12[lk * P * P + (m % P) * P + prow] : (T)0;
        const T* Ai = As + (i64)(iL - b0) * AR * PSTR + zp;
        const T* Rg = grp + rp0 + (i64)iR * K2 * RSTR + m;
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          if (ct >= nct) break;
          T rb[KS];
          if (first) {
#pragma unroll
            for (int st = 0; st < KS; ++st) rb[st] = rb0[st];
          } else {
#pragma unroll
            for (int st = 0; st < KS; ++st) rb[st] = 4 * st < K2 ? Rg[(4 * st + kq) * RSTR + ct * 16] : (T)0;
          }
#pragma unroll
          for (int gi = 0; gi < GPW; ++gi) {
            const int gq = tw + NW * gi;
            if (gq >= ng) break;
            const T* Ag = Ai + (4 * gq + zl) * PSTR;
            typename M::acc z = typename M::acc{0, 0, 0, 0};
#pragma unroll
            for (int st = 0; st < KS; ++st)
              if (4 * st < K2) z = M::run(Ag[4 * (4 * st + kq)], rb[st], z);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const T wa = (m / P) == (v % LPT) ? wv : (T)0;
              acc[gi][ct][v / LPT] = M::run(wa, z[v], acc[gi][ct][v / LPT]);
            }
          }
        }
        first = false;
        }
      }
    }
  }
#pragma unroll
  for (int gi = 0; gi < GPW; ++gi) {
    const int gq = tw + NW * gi;
    if (gq >= ng) break;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      if (ct >= nct) break;
      const int rq = ct * 16 + m;
#pragma unroll
      for (int t = 0; t < TPG; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = M::row(v, lane);
          const int lq = 4 * gq + t * LPT + row / P, pq = row % P;
          if (lq < dlp && rq < drp) ob[pq + (i64)P * (lq + Dl * rq)] = (double)acc[gi][ct][t][v];
        }
    }
  }
}

// kernels by (precision, P, groups per wave, column tiles)
// (big: sectors of up to 32 states; else up to 16, with fewer registers)
template <class T>
const void* heff_fn(int P, int big) {
  if (P == 4) return big ? (const void*)heff_kernel<T, 4, 2, 2, 8> : (const void*)heff_kernel<T, 4, 1, 1, 4>;
  if (P == 8) return big ? (const void*)heff_kernel<T, 8, 2, 2, 8> : (const void*)heff_kernel<T, 8, 1, 1, 4>;
  return big ? (const void*)heff_kernel<T, 16, 2, 2, 8> : (const void*)heff_kernel<T, 16, 1, 1, 4>;
}

// everything the fused plan keeps on the device, and the host-side checks of its contract
struct HeffState {
  const Tensor *psi, *L, *R, *W1, *W2;
  Tensor* out;
  i64 vL, vR, vW1, vW2;  // versions of the operator tensors at planning
  bool fp32 = false;
  int P = 16;
  i64 nout = 0;
  DBuf<i32> o_dlp, o_drp, o_pb, o_ib, pmask, dirty, active;
  DBuf<int2> uses;
  DBuf<int4> pinfo;  // per pair: its chunks with work, written by every pack
  int epoch = 0;  // executions so far: the pack kernel marks the blocks with work with it
  DBuf<int4> jobs;
  i64 vout = -1;  // Hpsi's version after this plan's last execution (another value: Hpsi written outside)
  DBuf<int2> items;
  DBuf<i64> o_off;
  DBuf<HPair> pairs;
  DBuf<PsiJob> pjobs;
  DBuf<unsigned char> pp, lp, rp, w12;  // panels and W12 in the plan's precision
  i64 njobs = 0;
  struct Launch {
    const void* fn = nullptr;
    i64 first = 0, n = 0;  // workgroups: jobs [first, first + n)
    int apb = 1;           // L panels whose A products LDS holds at once
    int off_it = 0, off_as = 0, wave_lds = 0;  // LDS region of a block (bytes)
    size_t lds = 0;
  };
  std::vector<Launch> launches;  // one per kernel shape with work, in order on the context's stream
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

template <class T>
void launch_pack_psi_t(HeffState* s, hipStream_t st) {
  PackK<T> k{};
  k.psi = s->psi->dptr();
  k.jobs = s->pjobs.get();
  k.Dl = (int)s->psi->dims[4];
  k.pp = (T*)s->pp.get();
  k.pmask = s->pmask.get();
  k.uses = s->uses.get();
  k.pairs = s->pairs.get();
  k.items = s->items.get();
  k.active = s->active.get();
  k.epoch = s->epoch;
  k.pinfo = s->pinfo.get();
  const dim3 grid((unsigned)s->njobs), block(64);
  if (s->P == 4)
    hipLaunchKernelGGL((heff_pack_psi<T, 4>), grid, block, 0, st, k);
  else if (s->P == 8)
    hipLaunchKernelGGL((heff_pack_psi<T, 8>), grid, block, 0, st, k);
  else
    hipLaunchKernelGGL((heff_pack_psi<T, 16>), grid, block, 0, st, k);
  HIP_CHECK(hipGetLastError());
}

void launch_pack_psi(HeffState* s, hipStream_t st) {
  if (!s->njobs) return;
  if (s->fp32)
    launch_pack_psi_t<float>(s, st);
  else
    launch_pack_psi_t<double>(s, st);
}

template <class T>
void launch_main(HeffState* s, const HeffState::Launch& la, hipStream_t st) {
  HK<T> k{};
  k.pp = (const T*)s->pp.get();
  k.lp = (const T*)s->lp.get();
  k.rp = (const T*)s->rp.get();
  k.w12 = (const T*)s->w12.get();
  k.pmask = s->pmask.get();
  k.pinfo = s->pinfo.get();
  k.out = s->out->dptr();
  k.jobs = s->jobs.get() + la.first;
  k.o_dlp = s->o_dlp.get();
  k.o_drp = s->o_drp.get();
  k.o_pb = s->o_pb.get();
  k.o_off = s->o_off.get();
  k.pairs = s->pairs.get();
  k.items = s->items.get();
  k.Dl = (int)s->psi->dims[4];
  k.apb = la.apb;
  k.off_it = la.off_it;
  k.off_as = la.off_as;
  k.wave_lds = la.wave_lds;
  k.o_ib = s->o_ib.get();
  k.dirty = s->dirty.get();
  k.active = s->active.get();
  k.epoch = s->epoch;
  void* args[] = {&k};
  HIP_CHECK(hipLaunchKernel(la.fn, dim3((unsigned)la.n), dim3(256), args, la.lds, st));
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
  if (!(P == 4 || P == 8 || P == 16) || w > 31)
    fail(SB_E_LIMIT, "H_eff plan: d1*d2 must be 4, 8 or 16 and the MPO link dimension at most 31");
  const bool fp32 = (flags & SB_PREC_MASK) == SB_PREC_FP32;
  if ((flags & SB_PREC_MASK) != SB_PREC_FP64 && !fp32)
    fail(SB_E_ARGUMENT, "ArgumentError: the fused H_eff plan computes in FP64 or FP32");
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  auto stride = [&](int n) { return fp32 ? lds_stride<float>(n) : lds_stride<double>(n); };

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

  // W12[wl, wr][p', p] = sum_wm W1[wl, s1', s1, wm] W2[wm, s2', s2, wr], P x P per nonzero link pair
  std::vector<i32> lidx((size_t)(w * w), -1), csupp;
  std::vector<double> w12;
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
                  blk[(size_t)((s1p + d1 * s2p) * P + (s1 + d1 * s2))] += a * b;
                  any = true;
                }
            }
      if (!any) continue;
      lidx[(size_t)(wl * w + wr)] = nlink++;
      int cs = 0;
      for (i64 pp = 0; pp < P; ++pp)
        for (i64 p = 0; p < P; ++p)
          if (blk[(size_t)(pp * P + p)] != 0.0) cs |= 1 << p;
      csupp.push_back(cs);
      w12.insert(w12.end(), blk.begin(), blk.end());
    }

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

  // ---- panels: per used L / R block its link panels, per used psi block its chunk panels
  std::vector<i64> lbase((size_t)std::max<i64>(L.n, 1), -1), rbase((size_t)std::max<i64>(R.n, 1), -1),
      pbase((size_t)std::max<i64>(psi.n, 1), -1);
  std::vector<i32> pblk((size_t)std::max<i64>(psi.n, 1), -1);
  std::vector<EnvJob> ljobs, rjobs;
  std::vector<PsiJob> pjobs;
  i64 lsize = 0, rsize = 0, psize = 0;
  auto env_panels = [&](i64 i, const std::vector<i32>& key, const std::vector<i32>& ids, const std::vector<i32>& hm,
                        const std::vector<int>& eo, const std::vector<int>& ei, const Tensor& X, i64& size,
                        std::vector<i64>& base, std::vector<EnvJob>& jobs) {
    if (base[(size_t)i] >= 0) return;
    const int a = key[(size_t)(2 * i)] - 1, b = key[(size_t)(2 * i + 1)] - 1;
    const int rows = up4(ei[(size_t)b]), str = stride(up16(eo[(size_t)a]));
    base[(size_t)i] = size;
    int j = 0;
    for (int lk = 0; lk < (int)w; ++lk) {
      if (!(hm[(size_t)(3 * i)] >> lk & 1)) continue;
      EnvJob jb;
      jb.src = (i64)(ids[(size_t)i] - 1) * X.blksize;
      jb.dst = size + (i64)j * rows * str;
      jb.link = lk;
      jb.rows = ei[(size_t)b];
      jb.cols = eo[(size_t)a];
      jb.stride = str;
      jobs.push_back(jb);
      ++j;
    }
    size += (i64)j * rows * str;
  };

  // ---- work list: per output block (by id) its sector pairs and their link items
  const i64 obs = out.blksize;
  std::vector<i32> o_dlp((size_t)nout), o_drp((size_t)nout), o_pb((size_t)nout + 1, 0), o_ib((size_t)nout + 1, 0);
  std::vector<int2> items;
  std::vector<i64> o_off((size_t)nout);
  std::vector<HPair> pairs;
  std::vector<std::vector<int2>> uses_by_job;  // per psi job: (pair, output block) of the pairs reading it
  std::vector<std::array<int, 4>> pgeo;  // per pair: K1, N1 (psi chunk), nl, nr, for the LDS need
  for (i64 o = 0; o < nout; ++o) {
    const int a = by_id[(size_t)o].first, c = by_id[(size_t)o].second;
    const int dlp = elp[(size_t)a], drp = erp[(size_t)c];
    o_dlp[(size_t)o] = dlp;
    o_drp[(size_t)o] = drp;
    o_off[(size_t)o] = o * obs;  // ids 1..n: block id o+1 at o*blksize
    if (dlp > 32 || drp > 32)
      fail(SB_E_LIMIT, "H_eff plan: a bond sector above 32 states (use the contract! chain)");
    for (auto& lb : Lrow[(size_t)a]) {
      const int il = lb.first;
      const int lmask = hmL[(size_t)(3 * lb.second)];
      for (auto& rb : Rrow[(size_t)c]) {
        const int ir = rb.first;
        auto itp = psi_at.find({il, ir});
        if (itp == psi_at.end()) continue;
        const int rmask = hmR[(size_t)(3 * rb.second)];
        const int dl = el[(size_t)il], dr = er[(size_t)ir];
        if (!dl || !dr || !dlp || !drp) continue;
        std::vector<int2> its;
        for (int wl = 0; wl < (int)w; ++wl) {
          if (!(lmask >> wl & 1)) continue;
          for (int wr = 0; wr < (int)w; ++wr) {
            if (!(rmask >> wr & 1)) continue;
            const int lk = lidx[(size_t)(wl * w + wr)];
            if (lk < 0) continue;
            const int xl = __builtin_popcount(lmask & ((1 << wl) - 1)), xr = __builtin_popcount(rmask & ((1 << wr) - 1));
            its.push_back(make_int2(xl | (xr << 5) | (lk << 10), csupp[(size_t)lk]));
          }
        }
        if (its.empty()) continue;
        const i64 ib = itp->second;
        env_panels(lb.second, kL, iL, hmL, elp, el, L, lsize, lbase, ljobs);
        env_panels(rb.second, kR, iR, hmR, erp, er, R, rsize, rbase, rjobs);
        const int K1 = up4(dl), N1 = 4 * up4(dr), pstr = stride(N1);
        if (pbase[(size_t)ib] < 0) {
          pbase[(size_t)ib] = psize;
          pblk[(size_t)ib] = (i32)pjobs.size();
          PsiJob jb;
          jb.src = (i64)(ipsi[(size_t)ib] - 1) * psi.blksize;
          jb.dst = psize;
          jb.blk = (i32)pjobs.size();
          jb.dl = dl;
          jb.dr = dr;
          jb.pstr = pstr;
          pjobs.push_back(jb);
          psize += (P / 4) * (i64)K1 * pstr;
        }
        HPair hp;
        hp.pp = pbase[(size_t)ib];
        hp.lp = lbase[(size_t)lb.second];
        hp.rp = rbase[(size_t)rb.second];
        hp.blk = pblk[(size_t)ib];
        hp.item0 = (i32)items.size();
        hp.nitem = (int16_t)its.size();
        hp.nl = (uint8_t)__builtin_popcount(lmask);
        hp.nr = (uint8_t)__builtin_popcount(rmask);
        hp.dl = (uint8_t)dl;
        hp.dr = (uint8_t)dr;
        for (int i = 0; i < 4; ++i) hp.pc[i] = 0;
        for (const int2& c2 : its)
          if ((c2.x & 31) < 4) hp.pc[c2.x & 31] |= (uint16_t)c2.y;
        items.insert(items.end(), its.begin(), its.end());
        pairs.push_back(hp);
        if (uses_by_job.size() < pjobs.size()) uses_by_job.resize(pjobs.size());
        uses_by_job[(size_t)hp.blk].push_back(make_int2((int)pairs.size() - 1, (int)o));
        pgeo.push_back({K1, N1, hp.nl, hp.nr});
      }
    }
    o_pb[(size_t)o + 1] = (i32)pairs.size();
    o_ib[(size_t)o + 1] = (i32)items.size();
  }
  if ((i64)pairs.size() >= (1ll << 31) || (i64)items.size() >= (1ll << 31))
    fail(SB_E_LIMIT, "H_eff plan: too many sector pairs");

  auto st = std::make_shared<HeffState>();
  st->fp32 = fp32;
  st->P = (int)P;

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
  st->nout = nout;
  up(ctx, st->dirty, std::vector<i32>((size_t)std::max<i64>(nout, 1), 0));  // Hpsi starts zero
  st->vout = out.version;
  st->njobs = (i64)pjobs.size();
  up(ctx, st->o_dlp, o_dlp);
  up(ctx, st->o_drp, o_drp);
  up(ctx, st->o_pb, o_pb);
  up(ctx, st->o_ib, o_ib);
  up(ctx, st->o_off, o_off);
  up(ctx, st->pairs, pairs);
  up(ctx, st->items, items);
  {
    std::vector<int2> uses;
    uses_by_job.resize(pjobs.size());
    for (size_t j = 0; j < pjobs.size(); ++j) {
      pjobs[j].u0 = (i32)uses.size();
      pjobs[j].nu = (i32)uses_by_job[j].size();
      uses.insert(uses.end(), uses_by_job[j].begin(), uses_by_job[j].end());
    }
    up(ctx, st->uses, uses);
  }
  up(ctx, st->pjobs, pjobs);
  up(ctx, st->active, std::vector<i32>((size_t)std::max<i64>(nout, 1), -1));
  st->pmask.alloc(ctx, (size_t)std::max<i64>(st->njobs, 1));
  st->pinfo.alloc(ctx, std::max<size_t>(pairs.size(), 1));
  // panels: zero (the padding stays zero), then filled from the operands
  auto zero_buf = [&](DBuf<unsigned char>& b, i64 elems) {
    const size_t bytes = (size_t)std::max<i64>(elems, 4) * esz;
    b.alloc(ctx, (bytes + 15) & ~(size_t)15);
    k_zero(ctx, b.get(), (bytes + 15) & ~(size_t)15);
  };
  zero_buf(st->pp, psize);
  zero_buf(st->lp, lsize);
  zero_buf(st->rp, rsize);
  {
    std::vector<unsigned char> hw(std::max<size_t>(w12.size(), 1) * esz, 0);
    for (size_t i = 0; i < w12.size(); ++i) {
      if (fp32) {
        const float f = (float)w12[i];
        std::memcpy(hw.data() + i * esz, &f, esz);
      } else {
        std::memcpy(hw.data() + i * esz, &w12[i], esz);
      }
    }
    up(ctx, st->w12, hw);
  }
  {
    DBuf<EnvJob> lj, rj;
    up(ctx, lj, ljobs);
    up(ctx, rj, rjobs);
    if (!ljobs.empty()) {
      if (fp32)
        hipLaunchKernelGGL(heff_pack_env<float>, dim3((unsigned)ljobs.size()), dim3(256), 0, ctx->stream, L.dptr(),
                           lj.get(), (int)w, Dl, (float*)st->lp.get());
      else
        hipLaunchKernelGGL(heff_pack_env<double>, dim3((unsigned)ljobs.size()), dim3(256), 0, ctx->stream, L.dptr(),
                           lj.get(), (int)w, Dl, (double*)st->lp.get());
    }
    if (!rjobs.empty()) {
      if (fp32)
        hipLaunchKernelGGL(heff_pack_env<float>, dim3((unsigned)rjobs.size()), dim3(256), 0, ctx->stream, R.dptr(),
                           rj.get(), (int)w, Dr, (float*)st->rp.get());
      else
        hipLaunchKernelGGL(heff_pack_env<double>, dim3((unsigned)rjobs.size()), dim3(256), 0, ctx->stream, R.dptr(),
                           rj.get(), (int)w, Dr, (double*)st->rp.get());
    }
    HIP_CHECK(hipGetLastError());
    // psi's rows at planning: the accounting below (every execution finds them anew)
    launch_pack_psi(st.get(), ctx->stream);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host vectors and job lists go out of scope
  }
  std::vector<i32> hpm = down(ctx, st->pmask.get(), (size_t)st->njobs);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));

  // ---- work: a block of more than 4 rows l' takes a workgroup (a team of 4
  // waves), blocks of at most 4 rows go 4 to a workgroup, one per wave;
  // blocks with sectors above 16 states or more than 16 l' per wave run a
  // kernel with more registers.  Per kernel shape the largest blocks first.
  auto a16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
  struct Shape {
    std::vector<i32> team, solo;
    size_t list = 16, its = 16, a1_team = 0, aall_team = 0, a1_solo = 0, aall_solo = 0;  // bytes
  };
  Shape shapes[2];
  std::vector<i64> work((size_t)nout, 0);
  for (i64 o = 0; o < nout; ++o) {
    const int dlp = o_dlp[(size_t)o], drp = o_drp[(size_t)o];
    if (!dlp || !drp) continue;  // no entry to write
    const int ng = (dlp + 3) / 4, nct = (drp + 15) / 16;
    int kmax = 0;  // the largest inner sector of the block's pairs
    for (int q = o_pb[(size_t)o]; q < o_pb[(size_t)o + 1]; ++q)
      kmax = std::max<int>(kmax, std::max(pairs[(size_t)q].dl, pairs[(size_t)q].dr));
    const int big = (ng > 4 || nct > 1 || kmax > 16) ? 1 : 0;
    Shape& sh = shapes[big];
    const bool team = ng > 1;
    (team ? sh.team : sh.solo).push_back((i32)o);
    sh.list = std::max(sh.list, a16(sizeof(HEnt) * (size_t)(P / 4) * (size_t)(o_pb[(size_t)o + 1] - o_pb[(size_t)o])));
    sh.its = std::max(sh.its, a16(sizeof(int2) * (size_t)(o_ib[(size_t)o + 1] - o_ib[(size_t)o])));
    const int AR = 4 * ng;  // live rows of the A products
    size_t a1 = 0, aall = 0;
    for (int q = o_pb[(size_t)o]; q < o_pb[(size_t)o + 1]; ++q) {
      const auto& gq = pgeo[(size_t)q];
      const int K1 = gq[0], N1 = gq[1], pstr = stride(N1);
      a1 = std::max(a1, esz * AR * pstr);
      aall = std::max(aall, esz * gq[2] * AR * pstr);
      // work with psi's rows at planning (the order of the jobs only)
      const HPair& hp = pairs[(size_t)q];
      const int pm = hpm[(size_t)hp.blk], np = pm >> 16;
      int nch = 0;
      for (int ch = 0; 4 * ch < np; ++ch) {
        int cm = 0;
        for (int b = 0, c = 0; b < P; ++b)
          if (pm >> b & 1) {
            if (c >= 4 * ch && c < 4 * ch + 4) cm |= 1 << b;
            ++c;
          }
        for (int it = 0; it < hp.nitem; ++it)
          if (items[(size_t)(hp.item0 + it)].y & cm) {
            ++nch;
            break;
          }
      }
      work[(size_t)o] += (i64)nch * (gq[2] + hp.nitem) * K1 * N1;
    }
    if (team) {
      sh.a1_team = std::max(sh.a1_team, a1);
      sh.aall_team = std::max(sh.aall_team, aall);
    } else {
      sh.a1_solo = std::max(sh.a1_solo, a1);
      sh.aall_solo = std::max(sh.aall_solo, aall);
    }
  }
  std::vector<int4> jobs;
  for (int big = 0; big < 2; ++big) {
    Shape& sh = shapes[big];
    if (sh.team.empty() && sh.solo.empty()) continue;
    auto by_work = [&](i32 x, i32 y) { return work[(size_t)x] > work[(size_t)y]; };
    std::stable_sort(sh.team.begin(), sh.team.end(), by_work);
    std::stable_sort(sh.solo.begin(), sh.solo.end(), by_work);
    HeffState::Launch la;
    la.first = (i64)jobs.size();
    // teams and quads of solo blocks in one list, the most work first
    std::vector<std::pair<i64, int4>> js;
    for (i32 o : sh.team) js.push_back({work[(size_t)o], make_int4(o, -2, -2, -2)});
    for (size_t i = 0; i < sh.solo.size(); i += 4) {
      auto at = [&](size_t j) { return j < sh.solo.size() ? sh.solo[j] : -1; };
      js.push_back({work[(size_t)sh.solo[i]], make_int4(at(i), at(i + 1), at(i + 2), at(i + 3))});
    }
    std::stable_sort(js.begin(), js.end(),
                     [](const std::pair<i64, int4>& x, const std::pair<i64, int4>& y) { return x.first > y.first; });
    for (auto& j : js) jobs.push_back(j.second);
    la.n = (i64)jobs.size() - la.first;
    la.off_it = (int)sh.list;
    la.off_as = (int)(sh.list + sh.its);
    auto lds_with = [&](size_t at, size_t as_) {
      const size_t wl = a16(la.off_as + as_);
      return std::make_pair(wl, std::max(sh.solo.empty() ? 0 : 4 * wl, (size_t)la.off_as + at));
    };
    la.apb = 31;
    auto l = lds_with(sh.aall_team, sh.aall_solo);
    if (l.second > kLdsBudget) {
      la.apb = 1;
      l = lds_with(sh.a1_team, sh.a1_solo);
    }
    la.wave_lds = (int)l.first;
    la.lds = std::max<size_t>(l.second, 16);
    if (la.lds + 64 > kMaxLds)
      fail(SB_E_LIMIT, "H_eff plan: sector blocks too large for the fused kernel's LDS tiles");
    la.fn = fp32 ? heff_fn<float>((int)P, big) : heff_fn<double>((int)P, big);
    HIP_CHECK(hipFuncSetAttribute(la.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)la.lds));
    st->launches.push_back(la);
  }

  up(ctx, st->jobs, jobs);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));


  // ---- accounting, for psi's rows at planning.  Useful FLOPs: per active
  // chunk the A products of its L panels (2 dlp dl (rows dr)) and per active
  // item Z (2 rows dlp dr drp) and the W12 step (2 nnz(W12[:, rows]) dlp drp);
  // matrix-core FLOPs: the 16 x 16 x 4 products those issue, padding included
  double flops = 0, mfma = 0;
  for (i64 o = 0; o < nout; ++o) {
    const int dlp = o_dlp[(size_t)o], drp = o_drp[(size_t)o];
    if (!dlp || !drp) continue;
    const int ng = (dlp + 3) / 4, nct = (drp + 15) / 16, LM = up16(dlp);
    for (int q = o_pb[(size_t)o]; q < o_pb[(size_t)o + 1]; ++q) {
      const HPair& hp = pairs[(size_t)q];
      const int pm = hpm[(size_t)hp.blk], np = pm >> 16;
      std::vector<int> rowsl;
      for (int b = 0; b < P; ++b)
        if (pm >> b & 1) rowsl.push_back(b);
      const int K1 = up4(hp.dl), K2 = up4(hp.dr);
      for (int ch = 0; 4 * ch < np; ++ch) {
        int cm = 0, nrow = 0;
        for (int j = 4 * ch; j < std::min(np, 4 * ch + 4); ++j) cm |= 1 << rowsl[(size_t)j], ++nrow;
        int amask = 0;
        for (int it = 0; it < hp.nitem; ++it) {
          const int code = items[(size_t)(hp.item0 + it)].x, lk = code >> 10;
          if (!(csupp[(size_t)lk] & cm)) continue;
          amask |= 1 << (code & 31);
          int nnz = 0;
          for (int pp = 0; pp < P; ++pp)
            for (int b = 0; b < P; ++b)
              if ((cm >> b & 1) && w12[(size_t)(lk * P * P + pp * P + b)] != 0.0) ++nnz;
          flops += 2.0 * nrow * dlp * hp.dr * drp + 2.0 * nnz * dlp * drp;
          mfma += (double)ng * nct * (K2 / 4 + 4);
        }
        const int na = __builtin_popcount(amask);
        flops += 2.0 * na * dlp * hp.dl * nrow * hp.dr;
        mfma += (double)na * (LM / 16) * (4 * K2 / 16) * (K1 / 4);
      }
    }
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
                         ++s->epoch;
                         launch_pack_psi(s, ctx->stream);
                         // Hpsi written outside this plan: its payload is cleared, padding
                         // included (the kernel writes the blocks' reach only), every block zero
                         if (s->out->version != s->vout && s->nout) {
                           k_zero(ctx, s->out->dptr(), (size_t)(s->nout * s->out->blksize) * sizeof(double));
                           k_zero(ctx, s->dirty.get(), (size_t)s->nout * sizeof(i32));
                         }
                         s->vout = s->out->version + 1;  // the plan's finish() marks Hpsi written
                         for (auto& la : s->launches) {
                           if (s->fp32)
                             launch_main<float>(s, la, ctx->stream);
                           else
                             launch_main<double>(s, la, ctx->stream);
                         }
                       }});
  pl->flops = flops;
  // algorithmic bytes: the psi blocks the pairs read, L and R inside their reach, Hpsi written
  double bpsi = 0, benv = 0, bytes_out = 0;
  for (const PsiJob& jb : pjobs) bpsi += 8.0 * P * jb.dl * jb.dr;
  for (i64 i = 0; i < L.n; ++i)
    benv += 8.0 * w * hmL[(size_t)(3 * i + 1)] * hmL[(size_t)(3 * i + 2)];
  for (i64 i = 0; i < R.n; ++i)
    benv += 8.0 * w * hmR[(size_t)(3 * i + 1)] * hmR[(size_t)(3 * i + 2)];
  for (i64 o = 0; o < nout; ++o) bytes_out += 8.0 * P * o_dlp[(size_t)o] * o_drp[(size_t)o];
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
  if (trace_on()) {
    std::printf("[SB_TRACE] heff_plan  psi BS(blocks=%lld) L BS(blocks=%lld) R BS(blocks=%lld)  -> fused H_eff*psi  "
                "pairs=%lld  items=%lld  Hpsi.blocks=%lld  launches=%zu  flops=%.6g (psi's rows at planning)\n",
                (long long)psi.n, (long long)L.n, (long long)R.n, (long long)pairs.size(), (long long)items.size(),
                (long long)nout, st->launches.size(), flops);
    for (size_t i = 0; i < st->launches.size(); ++i)
      std::printf("[SB_TRACE]   heff launch %zu: %lld workgroups, %zu bytes of LDS each, %d A panels per batch\n", i,
                  (long long)st->launches[i].n, st->launches[i].lds, st->launches[i].apb);
    std::fflush(stdout);
  }
  return pl;
}

}  // namespace sb
