// Box-aware grouped block products for BS x BS contract_shared!
// (the reference: permute A and B, then mul! per pair, contract_bs_bs.jl:572-730).
//
// Two ideas replace the reference's permute-then-GEMM:
//  * no operand permutes: the block product reads A, B and writes C through
//    per-axis strides of their own tail layouts (index decode tables in LDS),
//    so the payloads are never copied;
//  * boxes: for every block and tail axis, one past the last index that holds
//    a nonzero.  Entries beyond a box are zero, so their products are zero and
//    are skipped: each C block is computed only inside the box spanned by its
//    operands, and the rest of the block is kept at zero.  This is the common
//    case of symmetry sectors stored in a block padded to the largest sector.
//    A block holding a non-finite value widens the partner's extents to the
//    full block, so NaN/Inf propagate exactly as through the full products.
//
// Per execution: box_prepare (group boxes, per-pair K extents, tile counts,
// clearing of C entries that fall outside a shrunken box) then box_gemm, a
// persistent kernel that walks the (group, tile) work list.
#include <atomic>
#include <cstdlib>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "sb_kernels.h"

namespace sb {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s4 __attribute__((ext_vector_type(4)));

static inline unsigned nb256(i64 n) { return (unsigned)((n + 255) / 256); }

// ---------------------------------------------------------------- operand boxes
// One workgroup per chunk of kScanChunk values of a block (a block of 100k
// values is scanned by 13 workgroups, not one): per tail axis the largest
// index of a nonzero, reduced in LDS and then into the block's box with one
// atomic maximum per axis (box and flags cleared before the launch).  Blocks of
// fewer than 2^24 values decode the column-major index in 32 bits with a float
// reciprocal and one correction step, exact there (the quotient's rounding
// error stays below 1); larger blocks divide in 64 bits.
constexpr i64 kScanChunk = 8192;
constexpr i64 kScanSmall = (i64)1 << 24;


struct ScanInv {
  float v[MAXR];
};

__global__ __launch_bounds__(256) void box_scan_kernel(const double* __restrict__ data, const i32* __restrict__ ids,
                                                       i64 n, i64 bs, i64 cpb, int N2, IArr td, ScanInv inv,
                                                       i32* __restrict__ box, i32* __restrict__ flag) {
  __shared__ int sm[MAXR];
  __shared__ int sf;
  const int tid = threadIdx.x;
  const bool small = bs < kScanSmall;
  for (i64 w = blockIdx.x; w < n * cpb; w += gridDim.x) {
    const i64 i = w / cpb, c = w - i * cpb;
    if (tid < MAXR) sm[tid] = 0;
    if (tid == 0) sf = 0;
    __syncthreads();
    const double* b = data + (i64)(ids[i] - 1) * bs;
    const i64 hi = min(bs, (c + 1) * kScanChunk);
    int lm[MAXR];
#pragma unroll
    for (int a = 0; a < MAXR; ++a) lm[a] = 0;
    int lf = 0;
    for (i64 e = c * kScanChunk + tid; e < hi; e += 256) {
      const double v = b[e];
      if (v != 0.0) {  // NaN compares unequal, so it counts as nonzero
        if (small) {
          int x = (int)e;
#pragma unroll
          for (int a = 0; a < MAXR; ++a)
            if (a < N2) {
              const int d = (int)td.v[a];
              const int q = div24(x, d, inv.v[a]);
              lm[a] = max(lm[a], x - q * d + 1);
              x = q;
            }
        } else {
          i64 x = e;
#pragma unroll
          for (int a = 0; a < MAXR; ++a)
            if (a < N2) {
              const i64 q = x / td.v[a];
              lm[a] = max(lm[a], (int)(x - q * td.v[a]) + 1);
              x = q;
            }
        }
        if (!isfinite(v)) lf = 1;
      }
    }
#pragma unroll
    for (int a = 0; a < MAXR; ++a)
      if (a < N2 && lm[a]) atomicMax(&sm[a], lm[a]);
    if (lf) atomicOr(&sf, 1);
    __syncthreads();
    if (tid < N2 && sm[tid]) atomicMax(&box[i * N2 + tid], sm[tid]);
    if (tid == 0 && sf) atomicOr(&flag[i], 1);
    __syncthreads();
  }
}

void box_scan(Ctx* ctx, const double* data, const i32* ids, i64 n, i64 bs, int N2, const i64* tdims, i32* box,
              i32* flag) {
  if (n == 0) return;
  k_zero(ctx, box, (size_t)(n * N2) * sizeof(i32));
  k_zero(ctx, flag, (size_t)n * sizeof(i32));
  if (bs == 0) return;
  IArr td{};
  ScanInv inv{};
  for (int a = 0; a < N2; ++a) {
    td.v[a] = tdims[a];
    inv.v[a] = 1.0f / (float)std::max<i64>(tdims[a], 1);
  }
  const i64 cpb = (bs + kScanChunk - 1) / kScanChunk;
  const unsigned g = (unsigned)std::min<i64>(n * cpb, 65536);
  hipLaunchKernelGGL(box_scan_kernel, dim3(g), dim3(256), 0, ctx->stream, data, ids, n, bs, cpb, N2, td, inv, box,
                     flag);
  HIP_CHECK(hipGetLastError());
}

__global__ void fill_rows_kernel(i32* __restrict__ p, i64 n, int w, IArr row) {
  i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * w) return;
  p[t] = (i32)row.v[t % w];
}
void k_fill_rows(Ctx* ctx, i32* p, i64 n, int w, const i64* row) {
  if (n * w == 0) return;
  IArr r{};
  for (int a = 0; a < w; ++a) r.v[a] = row[a];
  hipLaunchKernelGGL(fill_rows_kernel, dim3(nb256(n * w)), dim3(256), 0, ctx->stream, p, n, w, r);
  HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- preparation
// One thread per group (taken in the locality order gorder): the group's M / N
// boxes (max over its contributing pairs), every pair's K extents (min of the
// two operands, widened where an operand holds a non-finite value), the new C
// box, whether the old C box reaches beyond it (then the GEMM kernel clears the
// difference), and the group's tile count at its position in the order.
__device__ __forceinline__ void box_group_one(const BoxDesc& d, const i64 g, const i32* __restrict__ seg,
                                              const i32* __restrict__ arow, const i32* __restrict__ brow,
                                              const i32* __restrict__ abox, const i32* __restrict__ aflag,
                                              const i32* __restrict__ bbox, const i32* __restrict__ bflag,
                                              i32* __restrict__ gbox, i32* __restrict__ kbox, i32* __restrict__ kc,
                                              i32* __restrict__ cbox, i32* __restrict__ cflag, i32* __restrict__ cnew,
                                              i32* __restrict__ clear, i32& tiles, int BM, int BN) {
  int mb[MAXAX], nb[MAXAX];
#pragma unroll
  for (int a = 0; a < MAXAX; ++a) mb[a] = nb[a] = 0;
  bool fullM = false, fullN = false;
  for (i32 pp = seg[g]; pp < seg[g + 1]; ++pp) {
    const i32 ar = arow[pp], br = brow[pp];
    const bool fa = aflag[ar] != 0, fb = bflag[br] != 0;
    int kcount = 1;
#pragma unroll
    for (int q = 0; q < MAXAX; ++q)
      if (q < d.nk) {
        int ea = abox[(i64)ar * d.NA2 + d.k_axa[q]], eb = bbox[(i64)br * d.NB2 + d.k_axb[q]];
        int e = min(ea, eb);
        if (fa) e = max(e, ea);
        if (fb) e = max(e, eb);
        kbox[(i64)pp * d.nk + q] = e;
        kcount *= e;
      }
    kc[pp] = kcount;
    if (kcount == 0) continue;  // contributes nothing
    fullM = fullM || fb;
    fullN = fullN || fa;
#pragma unroll
    for (int a = 0; a < MAXAX; ++a) {
      if (a < d.nm) mb[a] = max(mb[a], abox[(i64)ar * d.NA2 + d.m_ax[a]]);
      if (a < d.nn) nb[a] = max(nb[a], bbox[(i64)br * d.NB2 + d.n_ax[a]]);
    }
  }
  int Mc = 1, Nc = 1;
#pragma unroll
  for (int a = 0; a < MAXAX; ++a) {
    if (a < d.nm) {
      if (fullM) mb[a] = d.m_ext[a];
      Mc *= mb[a];
    }
    if (a < d.nn) {
      if (fullN) nb[a] = d.n_ext[a];
      Nc *= nb[a];
    }
  }
  if (Mc == 0 || Nc == 0) {
    Mc = Nc = 0;
#pragma unroll
    for (int a = 0; a < MAXAX; ++a) mb[a] = nb[a] = 0;
  }
  const int nmn = d.nm + d.nn;
  int cb[MAXAX];
#pragma unroll
  for (int a = 0; a < MAXAX; ++a) {
    if (a < d.nm) {
      gbox[g * nmn + a] = mb[a];
      cb[d.m_cax[a]] = mb[a];
    }
    if (a < d.nn) {
      gbox[g * nmn + d.nm + a] = nb[a];
      cb[d.n_cax[a]] = nb[a];
    }
  }
  bool old_empty = false, beyond = false;
#pragma unroll
  for (int a = 0; a < MAXAX; ++a)
    if (a < d.NC2) {
      int o = cbox[g * d.NC2 + a];
      old_empty = old_empty || o == 0;
      beyond = beyond || o > cb[a];
    }
  const bool need = !old_empty && beyond;
  clear[g] = need ? 1 : 0;
#pragma unroll
  for (int a = 0; a < MAXAX; ++a)
    if (a < d.NC2) {
      cnew[g * d.NC2 + a] = cb[a];
      if (!need) cbox[g * d.NC2 + a] = cb[a];
    }
  cflag[g] = 0;
  tiles = (Mc && Nc) ? ((Mc + BM - 1) / BM) * ((Nc + BN - 1) / BN) : 0;
}

__global__ void box_groups_kernel(BoxDesc d, i64 ng, const i32* __restrict__ gorder, const i32* __restrict__ seg,
                                  const i32* __restrict__ arow, const i32* __restrict__ brow,
                                  const i32* __restrict__ abox, const i32* __restrict__ aflag,
                                  const i32* __restrict__ bbox, const i32* __restrict__ bflag, i32* __restrict__ gbox,
                                  i32* __restrict__ kbox, i32* __restrict__ kc, i32* __restrict__ cbox,
                                  i32* __restrict__ cflag, i32* __restrict__ cnew, i32* __restrict__ clear,
                                  i32* __restrict__ tiles, int BM, int BN) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ng) return;
  const i64 g = gorder ? gorder[i] : i;
  box_group_one(d, g, seg, arow, brow, abox, aflag, bbox, bflag, gbox, kbox, kc, cbox, cflag, cnew, clear, tiles[i], BM,
                BN);
}

// work list from the device-wide scan of the per-position tile counts (toff[0] = 0,
// toff[ng] = total): trec[t] = {group, tile index within the group}, positions in order
__global__ void tile_records_kernel(const i32* __restrict__ toff, i64 ng, const i32* __restrict__ gorder,
                                    int2* __restrict__ trec) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ng) return;
  const i32 g = gorder ? gorder[i] : (i32)i;
  for (i32 t = toff[i]; t < toff[i + 1]; ++t) trec[t] = make_int2(g, t - toff[i]);
}

// the active job owning batch position i: the last a with gbase[a] <= i
__device__ __forceinline__ int job_of(const i64* __restrict__ gbase, int nact, i64 i) {
  int lo = 0, hi = nact;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (gbase[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// box_groups_kernel over the positions of every active job of a batch
__global__ void box_groups_batch_kernel(const BoxJob* __restrict__ jobs, const i32* __restrict__ act,
                                        const i64* __restrict__ gbase, int nact, i64 npos, i32* __restrict__ tiles,
                                        int BM, int BN) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npos) return;
  const int a = job_of(gbase, nact, i);
  const BoxJob& J = jobs[act[a]];
  const i64 pos = i - gbase[a];
  const i64 g = J.gorder ? J.gorder[pos] : pos;
  box_group_one(J.d, g, J.seg, J.arow, J.brow, J.abox, J.aflag, J.bbox, J.bflag, J.gbox, J.kbox, J.kc, J.cbox,
                J.cflag, J.cnew, J.clear, tiles[i], BM, BN);
}

__global__ void tile_records_batch_kernel(const i32* __restrict__ toff, i64 npos, const BoxJob* __restrict__ jobs,
                                          const i32* __restrict__ act, const i64* __restrict__ gbase, int nact,
                                          int4* __restrict__ trec) {
  i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npos) return;
  const int a = job_of(gbase, nact, i);
  const i32 j = act[a];
  const i64 pos = i - gbase[a];
  const i32 g = jobs[j].gorder ? jobs[j].gorder[pos] : (i32)pos;
  for (i32 t = toff[i]; t < toff[i + 1]; ++t) trec[t] = make_int4(j, g, t - toff[i], 0);
}

void box_prepare_batch(Ctx* ctx, const BoxBatchArgs& b) {
  if (b.npos == 0) return;
  const BoxTile t = box_tile(b.cfg);
  hipLaunchKernelGGL(box_groups_batch_kernel, dim3(nb256(b.npos)), dim3(256), 0, ctx->stream, b.jobs, b.act, b.gbase,
                     b.nact, b.npos, b.tiles, t.BM, t.BN);
  HIP_CHECK(hipGetLastError());
  size_t tb = b.scan_tmp_bytes;
  HIP_CHECK(rocprim::inclusive_scan(b.scan_tmp, tb, b.tiles, b.toff + 1, (size_t)b.npos, rocprim::plus<i32>(),
                                    ctx->stream));
  hipLaunchKernelGGL(tile_records_batch_kernel, dim3(nb256(b.npos)), dim3(256), 0, ctx->stream, b.toff, b.npos,
                     b.jobs, b.act, b.gbase, b.nact, b.trec);
  HIP_CHECK(hipGetLastError());
}

size_t box_scan_tmp_bytes(Ctx* ctx, i64 ng) {
  size_t tb = 0;
  HIP_CHECK(rocprim::inclusive_scan(nullptr, tb, (const i32*)nullptr, (i32*)nullptr, (size_t)std::max<i64>(ng, 1),
                                    rocprim::plus<i32>(), ctx->stream));
  return tb ? tb : 1;
}

void box_prepare(Ctx* ctx, const BoxGemmArgs& g, const i32* abox, const i32* aflag, const i32* bbox,
                 const i32* bflag, i32* cbox, i32* cflag, double* C) {
  (void)C;
  if (g.ng == 0) return;
  const BoxTile t = box_tile(g.cfg);
  hipLaunchKernelGGL(box_groups_kernel, dim3(nb256(g.ng)), dim3(256), 0, ctx->stream, g.d, g.ng, g.gorder, g.seg,
                     g.arow, g.brow, abox, aflag, bbox, bflag, g.gbox, g.kbox, g.kc, cbox, cflag, g.cnew, g.clear,
                     g.tiles, t.BM, t.BN);
  HIP_CHECK(hipGetLastError());
  if (g.trec) {  // the work list (absent while the tile shape is being chosen); toff[0] == 0 from planning
    size_t tb = g.scan_tmp_bytes;
    HIP_CHECK(rocprim::inclusive_scan(g.scan_tmp, tb, g.tiles, g.toff + 1, (size_t)g.ng, rocprim::plus<i32>(),
                                      ctx->stream));
    hipLaunchKernelGGL(tile_records_kernel, dim3(nb256(g.ng)), dim3(256), 0, ctx->stream, g.toff, g.ng, g.gorder,
                       g.trec);
    HIP_CHECK(hipGetLastError());
  }
}

// ---------------------------------------------------------------- block products
// index x of a compressed box -> (offset in operand 1, offset in operand 2),
// axis 0 fastest
__device__ inline void decode2(int x, int n, const int* ext, const int* s1, const int* s2, int& o1, int& o2) {
  int a1 = 0, a2 = 0;
#pragma unroll
  for (int a = 0; a < MAXAX; ++a)
    if (a < n) {
      int q = x / ext[a];
      int r = x - q * ext[a];
      a1 += r * s1[a];
      a2 += r * s2[a];
      x = q;
    }
  o1 = a1;
  o2 = a2;
}

__device__ inline unsigned short bf16_rne_b(float x) {
  unsigned u = __float_as_uint(x);
  unsigned r = ((u >> 16) & 1u) + 0x7fffu;
  return (unsigned short)((u + r) >> 16);
}
__device__ inline float bf16_to_f_b(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ inline void split3_b(float x, unsigned short& h, unsigned short& m, unsigned short& l) {
  h = bf16_rne_b(x);
  float r1 = x - bf16_to_f_b(h);
  m = bf16_rne_b(r1);
  float r2 = r1 - bf16_to_f_b(m);
  l = bf16_rne_b(r2);
}

constexpr int NXCD = 8;  // XCDs of the MI355X: consecutive workgroups go round-robin over them

// zero the entries of C block g that fell outside its shrunken box, then
// install the new box (one workgroup)
__device__ __forceinline__ void box_clear_stale(const BoxDesc& d, const i64 g, const i64* __restrict__ coff,
                                                const i32* __restrict__ cnew, i32* __restrict__ cbox,
                                                double* __restrict__ C, const int tid) {
  int ob[MAXAX], nbx[MAXAX];
  i64 vol = 1;
#pragma unroll
  for (int a = 0; a < MAXAX; ++a)
    if (a < d.NC2) {
      ob[a] = cbox[g * d.NC2 + a];
      nbx[a] = cnew[g * d.NC2 + a];
      vol *= ob[a];
    }
  double* base = C + coff[g];
  for (i64 e = tid; e < vol; e += 256) {
    i64 x = e, off = 0, st = 1;
    bool inside = true;
#pragma unroll
    for (int a = 0; a < MAXAX; ++a)
      if (a < d.NC2) {
        i64 q = x / ob[a];
        int r = (int)(x - q * ob[a]);
        inside = inside && r < nbx[a];
        off += r * st;
        st *= d.c_ext[a];
        x = q;
      }
    if (!inside) base[off] = 0.0;
  }
  __syncthreads();  // every thread has read the old box before it is replaced
  if (tid < d.NC2) cbox[g * d.NC2 + tid] = nbx[tid];
}

// one (group, tile) of the work list: C_g[tile] = sum over the group's pairs of A_p * B_p
template <int BM, int BN, int WM, int WN, int PREC, class T>
__device__ __forceinline__ void box_tile(const BoxDesc& d, const i32 g, const i32 lt, const i32* __restrict__ seg,
                                         const i64* __restrict__ aoff, const i64* __restrict__ boff,
                                         const i64* __restrict__ coff, const i32* __restrict__ gbox,
                                         const i32* __restrict__ kbox, const i32* __restrict__ kc,
                                         const double* __restrict__ A, const double* __restrict__ B,
                                         double* __restrict__ C, i32* __restrict__ cflag, const int amf, const int bnf,
                                         T (&As)[16][BM + 2], T (&Bs)[16][BN + 2], int (&rowA)[BM], int (&rowC)[BM],
                                         int (&colB)[BN], int (&colC)[BN], i64 (&tabA)[512], i64 (&tabB)[512]) {
  constexpr int BK = 16, KT = 512;
  constexpr int FM = BM / (16 * WM), FN = BN / (16 * WN);
  constexpr int AE = BM * BK / 256, BE = BN * BK / 256;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave % WM, wn = wave / WM;
  const int nmn = d.nm + d.nn;
  int mb[MAXAX], nbx[MAXAX];
  int Mc = 1, Nc = 1;
#pragma unroll
  for (int a = 0; a < MAXAX; ++a) {
    mb[a] = a < d.nm ? gbox[(i64)g * nmn + a] : 1;
    nbx[a] = a < d.nn ? gbox[(i64)g * nmn + d.nm + a] : 1;
    Mc *= mb[a];
    Nc *= nbx[a];
  }
  const int tiles_m = (Mc + BM - 1) / BM;
  const int tm = lt % tiles_m, tn = lt / tiles_m;
  const int m0 = tm * BM, n0 = tn * BN;
  __syncthreads();  // the previous tile is done with the LDS tables
  for (int r = tid; r < BM; r += 256) {
    int oa = -1, oc = -1;
    if (m0 + r < Mc) decode2(m0 + r, d.nm, mb, d.m_sa, d.m_sc, oa, oc);
    rowA[r] = oa;
    rowC[r] = oc;
  }
  for (int c = tid; c < BN; c += 256) {
    int ob = -1, oc = -1;
    if (n0 + c < Nc) decode2(n0 + c, d.nn, nbx, d.n_sb, d.n_sc, ob, oc);
    colB[c] = ob;
    colC[c] = oc;
  }
  d4 accd[FM][FN];
  f4 accf[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      accd[i][j] = d4{0, 0, 0, 0};
      accf[i][j] = f4{0, 0, 0, 0};
    }
  const i32 p1 = seg[g + 1];
  i32 pp = seg[g];
  int kin = 0;
  while (pp < p1) {
    __syncthreads();
    int fill = 0;
    while (pp < p1 && fill < KT) {
      const int kcp = kc[pp];
      const int take = min(kcp - kin, KT - fill);
      if (take > 0) {
        int kb[MAXAX];
#pragma unroll
        for (int a = 0; a < MAXAX; ++a) kb[a] = a < d.nk ? kbox[(i64)pp * d.nk + a] : 1;
        const i64 ao = aoff[pp], bo = boff[pp];
        for (int t = tid; t < take; t += 256) {
          int oa, ob;
          decode2(kin + t, d.nk, kb, d.k_sa, d.k_sb, oa, ob);
          tabA[fill + t] = ao + oa;
          tabB[fill + t] = bo + ob;
        }
      }
      fill += take;
      kin += take;
      if (kin >= kcp) {
        ++pp;
        kin = 0;
      }
    }
    __syncthreads();
    double ra[AE], rb[BE];
    auto load = [&](int k0) {
#pragma unroll
      for (int e = 0; e < AE; ++e) {
        const int idx = tid + e * 256;
        const int m = amf ? idx % BM : idx / BK, k = amf ? idx / BM : idx % BK;
        const int r = rowA[m];
        ra[e] = (r >= 0 && k0 + k < fill) ? A[tabA[k0 + k] + r] : 0.0;
      }
#pragma unroll
      for (int e = 0; e < BE; ++e) {
        const int idx = tid + e * 256;
        const int n = bnf ? idx % BN : idx / BK, k = bnf ? idx / BN : idx % BK;
        const int c = colB[n];
        rb[e] = (c >= 0 && k0 + k < fill) ? B[tabB[k0 + k] + c] : 0.0;
      }
    };
    load(0);
    for (int k0 = 0; k0 < fill; k0 += BK) {
      if (k0) __syncthreads();  // the previous chunk's products are done with As / Bs
#pragma unroll
      for (int e = 0; e < AE; ++e) {
        const int idx = tid + e * 256;
        const int m = amf ? idx % BM : idx / BK, k = amf ? idx / BM : idx % BK;
        As[k][m] = (T)ra[e];
      }
#pragma unroll
      for (int e = 0; e < BE; ++e) {
        const int idx = tid + e * 256;
        const int n = bnf ? idx % BN : idx / BK, k = bnf ? idx / BN : idx % BK;
        Bs[k][n] = (T)rb[e];
      }
      __syncthreads();
      if (k0 + BK < fill) load(k0 + BK);  // in flight while this chunk multiplies
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
            for (int j = 0; j < FN; ++j)
              accd[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], accd[i][j], 0, 0, 0);
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
            for (int j = 0; j < FN; ++j)
              accf[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], accf[i][j], 0, 0, 0);
        }
      } else {
        // bf16x3: both operands split in three bf16 terms; the six products of
        // order <= 2 accumulate in FP32
#pragma unroll
        for (int kk = 0; kk < BK / 16; ++kk) {
        s4 ah[FM], am[FM], al[FM], bh[FN], bm[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            unsigned short h, m, l;
            split3_b(As[kk * 16 + 4 * (lane >> 4) + v][(wm * FM + i) * 16 + (lane & 15)], h, m, l);
            ah[i][v] = (short)h;
            am[i][v] = (short)m;
            al[i][v] = (short)l;
          }
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            unsigned short h, m, l;
            split3_b(Bs[kk * 16 + 4 * (lane >> 4) + v][(wn * FN + j) * 16 + (lane & 15)], h, m, l);
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
  double* Cb = C + coff[g];
  int bad = 0;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // output row of register v: f64 16x16x4 holds rows 4*v + lane/16;
        // the f32 / bf16 16x16 forms hold rows 4*(lane/16) + v
        const int ml = (wm * FM + i) * 16 + (PREC == 0 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v);
        const int nl = (wn * FN + j) * 16 + (lane & 15);
        const int rc = rowC[ml], cc = colC[nl];
        if (rc >= 0 && cc >= 0) {
          const double val = PREC == 0 ? accd[i][j][v] : (double)accf[i][j][v];
          Cb[rc + cc] = val;
          bad |= !isfinite(val);
        }
      }
  if (bad) atomicOr(&cflag[g], 1);
}

// One workgroup (4 waves) per (group, tile) of the work list, persistent over
// the list.  The list is cut into NXCD contiguous ranges, one per XCD, so the
// tiles of neighbouring groups (which share operand blocks) meet in one L2.
// The K dimension of a tile is the concatenation of the K boxes of the
// group's pairs in the reference's pair order; a table in LDS maps each K
// position to its absolute A and B element offsets.  Operand tiles for the
// next K chunk are loaded into registers while the current one multiplies.
// Before the tiles, the workgroups clear the C entries that fell out of a
// shrunken box (disjoint from every tile's writes).  PREC as grouped_gemm.
// The tiles are latency-bound (short K per tile), so the register budget is held
// to five resident waves per SIMD, one more than the unconstrained allocation gives.
template <int BM, int BN, int WM, int WN, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void box_gemm_kernel(
    BoxDesc d, i64 ng, const i32* __restrict__ gorder, const i32* __restrict__ seg, const i64* __restrict__ aoff,
    const i64* __restrict__ boff, const i64* __restrict__ coff, const i32* __restrict__ gbox,
    const i32* __restrict__ kbox, const i32* __restrict__ kc, const i32* __restrict__ toff,
    const int2* __restrict__ trec, const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C,
    i32* __restrict__ cflag, const i32* __restrict__ clear, const i32* __restrict__ cnew, i32* __restrict__ cbox,
    int amf, int bnf) {
  constexpr int BK = 16, KT = 512;
  constexpr int FM = BM / (16 * WM), FN = BN / (16 * WN);
  constexpr int AE = BM * BK / 256, BE = BN * BK / 256;
  static_assert(FM >= 1 && FN >= 1 && WM * WN == 4 && AE >= 1 && BE >= 1, "tile");
  using T = typename std::conditional<PREC == 0, double, float>::type;
  __shared__ T As[BK][BM + 2];
  __shared__ T Bs[BK][BN + 2];
  __shared__ int rowA[BM], rowC[BM], colB[BN], colC[BN];
  __shared__ i64 tabA[KT], tabB[KT];
  const int tid = threadIdx.x;

  // ---- stale entries of shrunken C boxes
  for (i64 g = blockIdx.x; g < ng; g += gridDim.x)
    if (clear[g]) box_clear_stale(d, g, coff, cnew, cbox, C, tid);

  // ---- tiles
  const i32 total = toff[ng];
  const int xcd = blockIdx.x % NXCD, slot = blockIdx.x / NXCD, nper = gridDim.x / NXCD;
  const i32 chunk = (total + NXCD - 1) / NXCD;
  const i32 wlo = xcd * chunk, whi = min(total, wlo + chunk);
  for (i32 w = wlo + slot; w < whi; w += nper) {
    const int2 rec = trec[w];
    box_tile<BM, BN, WM, WN, PREC>(d, rec.x, rec.y, seg, aoff, boff, coff, gbox, kbox, kc, A, B, C, cflag, amf, bnf, As,
                                   Bs, rowA, rowC, colB, colC, tabA, tabB);
  }
}

// The same over the work list of a batch of jobs: each tile reads its job's
// descriptor (uniform per workgroup, so scalar loads), every job's stale C
// entries are cleared first.
template <int BM, int BN, int WM, int WN, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void box_gemm_batch_kernel(
    const BoxJob* __restrict__ jobs, const i32* __restrict__ act, const i64* __restrict__ gbase, int nact, i64 npos,
    const i32* __restrict__ toff, const int4* __restrict__ trec) {
  constexpr int BK = 16, KT = 512;
  using T = typename std::conditional<PREC == 0, double, float>::type;
  __shared__ T As[BK][BM + 2];
  __shared__ T Bs[BK][BN + 2];
  __shared__ int rowA[BM], rowC[BM], colB[BN], colC[BN];
  __shared__ i64 tabA[KT], tabB[KT];
  const int tid = threadIdx.x;

  for (i64 i = blockIdx.x; i < npos; i += gridDim.x) {
    const int a = job_of(gbase, nact, i);
    const BoxJob& J = jobs[act[a]];
    const i64 g = i - gbase[a];  // every group of the job once, in any order
    if (J.clear[g]) box_clear_stale(J.d, g, J.coff, J.cnew, J.cbox, J.C, tid);
  }

  const i32 total = toff[npos];
  const int xcd = blockIdx.x % NXCD, slot = blockIdx.x / NXCD, nper = gridDim.x / NXCD;
  const i32 chunk = (total + NXCD - 1) / NXCD;
  const i32 wlo = xcd * chunk, whi = min(total, wlo + chunk);
  for (i32 w = wlo + slot; w < whi; w += nper) {
    const int4 rec = trec[w];
    const BoxJob& J = jobs[rec.x];
    box_tile<BM, BN, WM, WN, PREC>(J.d, rec.y, rec.z, J.seg, J.aoff, J.boff, J.coff, J.gbox, J.kbox, J.kc, J.A, J.B,
                                   J.C, J.cflag, J.amf, J.bnf, As, Bs, rowA, rowC, colB, colC, tabA, tabB);
  }
}

// Wave tiles: every wave (64 lanes) owns one (group, tile) of FM x FN
// blocks of 16 x 16 and works alone, with no LDS and no barriers: each MFMA
// step's operands are loaded straight into the lanes that hold them (lane l
// takes A[m = l % 16][k] and B[k][n = l % 16] with k = l / 16 for the
// 16x16x4 forms, k = 4 (l / 16) + v for bf16), K boxes padded per pair to the
// MFMA depth.  Two K steps are loaded before either multiplies.  Suits the
// small blocks of symmetry sectors: many independent waves hide the latency.
template <int FM, int FN, int PREC>
__global__ __launch_bounds__(256) void box_wave_kernel(
    BoxDesc d, i64 ng, const i32* __restrict__ seg, const i64* __restrict__ aoff, const i64* __restrict__ boff,
    const i64* __restrict__ coff, const i32* __restrict__ gbox, const i32* __restrict__ kbox,
    const i32* __restrict__ kc, const i32* __restrict__ toff, const int2* __restrict__ trec,
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C, i32* __restrict__ cflag,
    const i32* __restrict__ clear, const i32* __restrict__ cnew, i32* __restrict__ cbox) {
  constexpr int TM = 16 * FM, TN = 16 * FN;
  constexpr int KS = PREC == 2 ? 16 : 4;  // K depth of one MFMA
  constexpr int KV = PREC == 2 ? 4 : 1;   // K values per lane per MFMA
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: per-wave work goes through scalar loads
  const int lr = lane & 15, lq = lane >> 4;
  const int nmn = d.nm + d.nn;

  // ---- stale entries of shrunken C boxes (whole workgroup)
  for (i64 g = blockIdx.x; g < ng; g += gridDim.x) {
    if (!clear[g]) continue;
    int ob[MAXAX], nbx[MAXAX];
    i64 vol = 1;
#pragma unroll
    for (int a = 0; a < MAXAX; ++a)
      if (a < d.NC2) {
        ob[a] = cbox[g * d.NC2 + a];
        nbx[a] = cnew[g * d.NC2 + a];
        vol *= ob[a];
      }
    double* base = C + coff[g];
    for (i64 e = tid; e < vol; e += 256) {
      i64 x = e, off = 0, st = 1;
      bool inside = true;
#pragma unroll
      for (int a = 0; a < MAXAX; ++a)
        if (a < d.NC2) {
          i64 q = x / ob[a];
          int r = (int)(x - q * ob[a]);
          inside = inside && r < nbx[a];
          off += r * st;
          st *= d.c_ext[a];
          x = q;
        }
      if (!inside) base[off] = 0.0;
    }
    __syncthreads();  // every thread has read the old box before it is replaced
    if (tid < d.NC2) cbox[g * d.NC2 + tid] = nbx[tid];
  }

  // ---- tiles, one per wave
  const i32 total = toff[ng];
  const int xcd = blockIdx.x % NXCD;
  const int slot = (blockIdx.x / NXCD) * 4 + wave, nper = (gridDim.x / NXCD) * 4;
  const i32 chunk = (total + NXCD - 1) / NXCD;
  const i32 wlo = xcd * chunk, whi = min(total, wlo + chunk);
  for (i32 w = wlo + slot; w < whi; w += nper) {
    const int2 rec = trec[w];
    const i32 g = rec.x, lt = rec.y;
    int mb[MAXAX], nbx[MAXAX];
    int Mc = 1, Nc = 1;
#pragma unroll
    for (int a = 0; a < MAXAX; ++a) {
      mb[a] = a < d.nm ? gbox[(i64)g * nmn + a] : 1;
      nbx[a] = a < d.nn ? gbox[(i64)g * nmn + d.nm + a] : 1;
      Mc *= mb[a];
      Nc *= nbx[a];
    }
    const int tiles_m = (Mc + TM - 1) / TM;
    const int m0 = (lt % tiles_m) * TM, n0 = (lt / tiles_m) * TN;
    int ra[FM], cb[FN], cc[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      int oa = -1, oc = 0;
      if (m0 + 16 * i + lr < Mc) decode2(m0 + 16 * i + lr, d.nm, mb, d.m_sa, d.m_sc, oa, oc);
      ra[i] = oa;
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      int ob = -1, oc = -1;
      if (n0 + 16 * j + lr < Nc) decode2(n0 + 16 * j + lr, d.nn, nbx, d.n_sb, d.n_sc, ob, oc);
      cb[j] = ob;
      cc[j] = oc;
    }
    d4 accd[FM][FN];
    f4 accf[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        accd[i][j] = d4{0, 0, 0, 0};
        accf[i][j] = f4{0, 0, 0, 0};
      }
    const i32 p1 = seg[g + 1];
    for (i32 pp = seg[g]; pp < p1; ++pp) {
      const int kcp = kc[pp];
      if (kcp == 0) continue;
      int kb[MAXAX];
#pragma unroll
      for (int a = 0; a < MAXAX; ++a) kb[a] = a < d.nk ? kbox[(i64)pp * d.nk + a] : 1;
      const double* Ab = A + aoff[pp];
      const double* Bb = B + boff[pp];
      const int nsteps = (kcp + KS - 1) / KS;
      for (int s0 = 0; s0 < nsteps; s0 += 2) {
        double av[2][FM][KV], bv[2][FN][KV];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int v = 0; v < KV; ++v) {
            const int k = (s0 + u) * KS + (PREC == 2 ? 4 * lq + v : lq);
            int ka = 0, kbo = 0;
            const bool ok = k < kcp;
            if (ok) decode2(k, d.nk, kb, d.k_sa, d.k_sb, ka, kbo);
#pragma unroll
            for (int i = 0; i < FM; ++i) av[u][i][v] = (ok && ra[i] >= 0) ? Ab[ka + ra[i]] : 0.0;
#pragma unroll
            for (int j = 0; j < FN; ++j) bv[u][j][v] = (ok && cb[j] >= 0) ? Bb[kbo + cb[j]] : 0.0;
          }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          if (s0 + u >= nsteps) break;
          if constexpr (PREC == 0) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
              for (int j = 0; j < FN; ++j)
                accd[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][i][0], bv[u][j][0], accd[i][j], 0, 0, 0);
          } else if constexpr (PREC == 1) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
              for (int j = 0; j < FN; ++j)
                accf[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)av[u][i][0], (float)bv[u][j][0], accf[i][j],
                                                                  0, 0, 0);
          } else {
            s4 ah[FM], am[FM], al[FM], bh[FN], bm[FN], bl[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                unsigned short h, m, l;
                split3_b((float)av[u][i][v], h, m, l);
                ah[i][v] = (short)h;
                am[i][v] = (short)m;
                al[i][v] = (short)l;
              }
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                unsigned short h, m, l;
                split3_b((float)bv[u][j][v], h, m, l);
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
    double* Cb = C + coff[g];
    int bad = 0;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // output row of register v: f64 16x16x4 holds rows 4*v + lane/16;
        // the f32 / bf16 16x16 forms hold rows 4*(lane/16) + v
        const int m = m0 + 16 * i + (PREC == 0 ? 4 * v + lq : 4 * lq + v);
        if (m >= Mc) continue;
        int oa, rc;
        decode2(m, d.nm, mb, d.m_sa, d.m_sc, oa, rc);
#pragma unroll
        for (int j = 0; j < FN; ++j)
          if (cc[j] >= 0) {
            const double val = PREC == 0 ? accd[i][j][v] : (double)accf[i][j][v];
            Cb[rc + cc[j]] = val;
            bad |= !isfinite(val);
          }
      }
    if (bad) atomicOr(&cflag[g], 1);
  }
}

static const BoxTile kTiles[] = {
    // workgroup tiles (4 waves share LDS operand tiles)
    {16, 128}, {128, 16}, {32, 32}, {64, 32}, {32, 64}, {64, 64}, {16, 64}, {64, 16},
    // wave tiles (one wave per tile)
    {16, 16}, {32, 16}, {16, 32}, {32, 32}, {64, 16}, {16, 64},
};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
BoxTile box_tile(int cfg) {
  SB_ASSERT(cfg >= 0 && cfg < kNumTiles, "box tile configuration");
  return kTiles[cfg];
}
int box_tile_configs() { return kNumTiles; }
bool box_tile_is_wave(int cfg) { return cfg >= 8; }

template <class K>
static i64 grid_for(Ctx* ctx, K kern, std::atomic<int>& per_cu_cache, int want = 0) {
  if (want > 0) return std::max<i64>(NXCD, (i64)ctx->num_cu * want / NXCD * NXCD);
  int per_cu = per_cu_cache.load(std::memory_order_relaxed);
  if (!per_cu) {  // resident workgroups per CU (the same on every device of this build)
    int nb = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, 0));
    per_cu = std::max(1, nb);
    if (const char* e = std::getenv("SB_BOX_WG_PER_CU")) per_cu = std::max(1, std::atoi(e));  // tuning override
    per_cu_cache.store(per_cu, std::memory_order_relaxed);
  }
  i64 nwg = (i64)ctx->num_cu * per_cu;
  return std::max<i64>(NXCD, nwg / NXCD * NXCD);  // a multiple of the XCD count
}

template <int BM, int BN, int WM, int WN, int PREC>
static void launch_one(Ctx* ctx, const BoxGemmArgs& g, const double* A, const double* B, double* C, i32* cflag) {
  auto kern = box_gemm_kernel<BM, BN, WM, WN, PREC>;
  static std::atomic<int> per_cu{0};
  i64 nwg = grid_for(ctx, kern, per_cu, g.wgcu);
  hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(256), 0, ctx->stream, g.d, g.ng, g.gorder, g.seg, g.aoff,
                     g.boff, g.coff, g.gbox, g.kbox, g.kc, g.toff, g.trec, A, B, C, cflag, g.clear, g.cnew, g.cbox,
                     g.amf, g.bnf);
  HIP_CHECK(hipGetLastError());
}

template <int BM, int BN, int WM, int WN, int PREC>
static void launch_one_batch(Ctx* ctx, const BoxBatchArgs& b) {
  auto kern = box_gemm_batch_kernel<BM, BN, WM, WN, PREC>;
  static std::atomic<int> per_cu{0};
  i64 nwg = grid_for(ctx, kern, per_cu);
  hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(256), 0, ctx->stream, b.jobs, b.act, b.gbase, b.nact, b.npos,
                     b.toff, b.trec);
  HIP_CHECK(hipGetLastError());
}

template <int PREC>
static void launch_box_batch(Ctx* ctx, const BoxBatchArgs& b) {
  switch (b.cfg) {
    case 0: launch_one_batch<16, 128, 1, 4, PREC>(ctx, b); break;
    case 1: launch_one_batch<128, 16, 4, 1, PREC>(ctx, b); break;
    case 2: launch_one_batch<32, 32, 2, 2, PREC>(ctx, b); break;
    case 3: launch_one_batch<64, 32, 2, 2, PREC>(ctx, b); break;
    case 4: launch_one_batch<32, 64, 2, 2, PREC>(ctx, b); break;
    case 5: launch_one_batch<64, 64, 2, 2, PREC>(ctx, b); break;
    case 6: launch_one_batch<16, 64, 1, 4, PREC>(ctx, b); break;
    case 7: launch_one_batch<64, 16, 4, 1, PREC>(ctx, b); break;
    default: fail(SB_E_ASSERT, "batched box GEMM takes workgroup tiles only");
  }
}

void box_gemm_batch(Ctx* ctx, unsigned prec, const BoxBatchArgs& b) {
  if (b.npos == 0) return;
  switch (prec & SB_PREC_MASK) {
    case SB_PREC_FP64: launch_box_batch<0>(ctx, b); break;
    case SB_PREC_FP32: launch_box_batch<1>(ctx, b); break;
    case SB_PREC_BF16X3: launch_box_batch<2>(ctx, b); break;
    default: fail(SB_E_ARGUMENT, "unknown precision form");
  }
}

template <int FM, int FN, int PREC>
static void launch_wave(Ctx* ctx, const BoxGemmArgs& g, const double* A, const double* B, double* C, i32* cflag) {
  auto kern = box_wave_kernel<FM, FN, PREC>;
  static std::atomic<int> per_cu{0};
  i64 nwg = grid_for(ctx, kern, per_cu, g.wgcu);
  hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(256), 0, ctx->stream, g.d, g.ng, g.seg, g.aoff, g.boff,
                     g.coff, g.gbox, g.kbox, g.kc, g.toff, g.trec, A, B, C, cflag, g.clear, g.cnew, g.cbox);
  HIP_CHECK(hipGetLastError());
}

template <int PREC>
static void launch_box(Ctx* ctx, const BoxGemmArgs& g, const double* A, const double* B, double* C, i32* cflag) {
  switch (g.cfg) {
    case 0: launch_one<16, 128, 1, 4, PREC>(ctx, g, A, B, C, cflag); break;
    case 1: launch_one<128, 16, 4, 1, PREC>(ctx, g, A, B, C, cflag); break;
    case 2: launch_one<32, 32, 2, 2, PREC>(ctx, g, A, B, C, cflag); break;
    case 3: launch_one<64, 32, 2, 2, PREC>(ctx, g, A, B, C, cflag); break;
    case 4: launch_one<32, 64, 2, 2, PREC>(ctx, g, A, B, C, cflag); break;
    case 5: launch_one<64, 64, 2, 2, PREC>(ctx, g, A, B, C, cflag); break;
    case 6: launch_one<16, 64, 1, 4, PREC>(ctx, g, A, B, C, cflag); break;
    case 7: launch_one<64, 16, 4, 1, PREC>(ctx, g, A, B, C, cflag); break;
    case 8: launch_wave<1, 1, PREC>(ctx, g, A, B, C, cflag); break;
    case 9: launch_wave<2, 1, PREC>(ctx, g, A, B, C, cflag); break;
    case 10: launch_wave<1, 2, PREC>(ctx, g, A, B, C, cflag); break;
    case 11: launch_wave<2, 2, PREC>(ctx, g, A, B, C, cflag); break;
    case 12: launch_wave<4, 1, PREC>(ctx, g, A, B, C, cflag); break;
    case 13: launch_wave<1, 4, PREC>(ctx, g, A, B, C, cflag); break;
    default: fail(SB_E_ASSERT, "box tile configuration");
  }
}

void box_gemm(Ctx* ctx, unsigned prec, const BoxGemmArgs& g, const double* A, const double* B, double* C,
              i32* cflag) {
  if (g.ng == 0) return;
  switch (prec & SB_PREC_MASK) {
    case SB_PREC_FP64: launch_box<0>(ctx, g, A, B, C, cflag); break;
    case SB_PREC_FP32: launch_box<1>(ctx, g, A, B, C, cflag); break;
    case SB_PREC_BF16X3: launch_box<2>(ctx, g, A, B, C, cflag); break;
    default: fail(SB_E_ARGUMENT, "unknown precision form");
  }
}

}  // namespace sb
