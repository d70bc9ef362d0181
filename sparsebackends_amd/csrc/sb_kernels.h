// Host-side launchers of the device kernels (all asynchronous on ctx->stream).
#pragma once

#include "sb_internal.h"

namespace sb {

using u64 = unsigned long long;
constexpr int MAXR = 8;  // largest rank of a key prefix / dense tail handled on device

// Small fixed-size argument blocks passed by value to kernels.
struct IArr {
  i64 v[MAXR];
};
struct SArr {  // signed source selectors for key assembly: +a -> A axis a, -b -> B axis b (1-based)
  int v[MAXR];
};

// n / d for 0 <= n < 2^24 and d >= 1 by a float reciprocal (inv = 1.0f / d)
// and one correction step: exact there, the quotient's rounding error staying
// below 1 (device code that splits column-major indices of small blocks)
#ifdef __HIPCC__
__device__ __forceinline__ int div24(int n, int d, float inv) {
  int q = (int)((float)n * inv);
  const int r = n - q * d;
  if (r >= d)
    ++q;
  else if (r < 0)
    --q;
  return q;
}
#endif

// ---- keys -------------------------------------------------------------
// out_keys[i, j] = in_keys[i, sel[j]-1]  (j < Pout); lin[i] = col-major lin
// of out_keys[i] over pdims_out (1-based); iota[i] = i
void k_remap_keys(Ctx*, const i32* in_keys, i64 n, int Pin, const int* sel, int Pout, const i64* pdims_out,
                  i32* out_keys, u64* lin, i32* iota);
// keys of pairs: ck[p, j] = src[j] > 0 ? A[a[p], src[j]-1] : B[b[p], -src[j]-1]; lin over pdimsC
void k_pair_keys(Ctx*, const i32* a_idx, const i32* b_idx, i64 np, const i32* keysA, int PA, const i32* keysB,
                 int PB, const int* src, int PC, const i64* pdimsC, i32* ck, u64* lin);
// value of the given key columns as a col-major lin (join value), for n rows
void k_cols_lin(Ctx*, const i32* keys, i64 n, int P, const int* cols, int ncols, const i64* cdims, u64* lin);
// gather rows: out[i, :] = in[idx[i], :] (row width w)
void k_gather_rows(Ctx*, const i32* in, int w, const i32* idx, i64 n, i32* out);
void k_iota(Ctx*, i32* p, i64 n, i32 base);
void k_fill_f64(Ctx*, double* p, i64 n, double v);

// ---- sort / scan (rocPRIM) --------------------------------------------
// stable radix sort of (key, val) pairs by key bits [0, bits); in place
void sort_pairs(Ctx*, u64* keys, i32* vals, i64 n, int bits);
void sort_pairs32(Ctx*, unsigned* keys, i32* vals, i64 n, int bits);
// exclusive scan of counts (int64) -> offsets; returns total (synchronises)
i64 exclusive_scan_total(Ctx*, const i64* in, i64* out, i64 n);
int bits_for(u64 maxval);

// ---- joins --------------------------------------------------------------
// For sorted (non-decreasing) ja[nA], jb[nB]: lo/hi of B rows equal to each ja
void k_equal_ranges(Ctx*, const u64* ja, i64 nA, const u64* jb, i64 nB, i32* lo, i32* hi);
// sequential merge walk of the reference (run-wise compare by cmp_cols order)
// producing lo/hi per A row; mode 0: the reference's r-run walk
// (contract_bs_bs.jl:215-261), 1: element-step walk of contract_coo.jl:139-167,
// 2: the multi-label walk with a column compare order (contract_bs_bs.jl:684-728)
void k_walk_ranges(Ctx*, int mode, const i32* keysA, int wA, i64 nA, const i32* keysB, int wB, i64 nB,
                   const int* colsA, const int* colsB, int ncols, i32* lo, i32* hi);
// counts[i] = hi[i]-lo[i] (int64)
void k_range_counts(Ctx*, const i32* lo, const i32* hi, i64 n, i64* counts);
// emit pairs: for A row i, pairs (i, lo[i] + t) at off[i] + t
void k_emit_pairs(Ctx*, const i32* lo, const i64* off, const i64* counts, i64 nA, i32* pa, i32* pb);
// run starts of equal values in a sorted u64 array: flags -> segment ids
// (returns the number of segments; synchronises)

// ---- output-block allocation (ensure_block! semantics) ------------------
struct Alloc {
  i64 nC = 0;
  DBuf<i32> keys;       // nC x PC, sorted by col-major lin
  DBuf<i32> ids;        // nC, slot = rank of first creation (1-based)
  DBuf<i32> seg_start;  // nC + 1: pairs of C block s are order[seg_start[s] .. seg_start[s+1])
  DBuf<i32> order;      // npairs: pair indices grouped by C key, pair order within a group
  DBuf<i32> seg_slot;   // nC: 0-based payload slot of C block s (= ids[s]-1)
};
// ck: np x PC keys of the pairs, lin: their col-major lin (both consumed)
void allocate_blocks(Ctx*, i64 np, int PC, DBuf<i32>& ck, DBuf<u64>& lin, u64 maxlin, Alloc& out);

// ---- permutes ------------------------------------------------------------
// Tail permute of a batch of blocks: out block o (at o*bs) takes source block
// at src_off[o] (elements) with its N2-dim tail permuted by perm (new axis j =
// old axis perm[j], 0-based), tail dims tdims (old order).  src_off == null:
// source block o at o*bs.
void permute_blocks(Ctx*, const double* in, const i64* src_off, i64 nblocks, int N2, const i64* tdims,
                    const int* perm, double* out);
// gather whole blocks without permutation: out[o] = in[src_off[o] .. +bs)
void gather_blocks(Ctx*, const double* in, const i64* src_off, i64 nblocks, i64 bs, double* out);
// source offsets (elements) of sorted blocks: off[o] = (ids[idx[o]]-1)*bs
void k_src_offsets(Ctx*, const i32* ids, const i32* idx, i64 n, i64 bs, i64* off);

// ---- numeric -------------------------------------------------------------
// Grouped GEMM: for each group g (C block), C_g = sum over its pairs p of
// A_p (M x K) * B_p (K x N).  Element addresses:
//   A_p[m,k] = A + aoff[p] + m*sam + k*sak ; B_p[k,n] = B + boff[p] + k*sbk + n*sbn
//   C_g[m,n] = C + coff[g] + m*scm + n*scn
// Pairs of group g: order[seg[g] .. seg[g+1]).  C is overwritten (starts at 0).
// kmask (optional, npairs x K bytes): k terms with kmask 0 are skipped.
struct GemmDesc {
  i64 M, N, K;
  i64 sam, sak, sbk, sbn, scm, scn;
};
void grouped_gemm(Ctx*, unsigned prec, const GemmDesc& d, const double* A, const i64* aoff, const double* B,
                  const i64* boff, double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups,
                  const unsigned char* kmask);
// Kahan-compensated sum of outer products, the reference's exact order
// (contract_bs_dense.jl:167-196): per group, pairs in order, term
// A_p[ii] * B_p[jj] added to C[ii + jj*m] (AB) or C[jj + ii*n] (BA).
void kahan_outer(Ctx*, bool AB, i64 chunkA, i64 chunkB, const double* A, const i64* aoff, const double* B,
                 const i64* boff, double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups);
// sequential per-element accumulation C[t] += alpha_p * X_p[t] (t < L), pair order
void axpy_groups(Ctx*, i64 L, const double* alpha, const double* X, const i64* xoff, double* C, const i64* coff,
                 const i32* seg, const i32* order, i64 ngroups);
// sequential outer accumulation C[i + j*m] += A_p[i] * B_p[j] (pair order)
void outer_groups(Ctx*, i64 m, i64 n, const double* A, const i64* aoff, const double* B, const i64* boff,
                  double* C, const i64* coff, const i32* seg, const i32* order, i64 ngroups);
// per (row, r) nonzero flags of slices: flag[i*R + r] = any(X[off_i + r*chunk .. +chunk] != 0)
void slice_nonzero(Ctx*, const double* X, const i64* off, i64 nrows, i64 R, i64 chunk, unsigned char* flag);

// ---- box-aware block products (sb_boxgemm.hip) ---------------------------
// A box row holds, per tail axis of a block, one past the last index with a
// nonzero entry; a flag row marks blocks holding a non-finite value.
constexpr int MAXAX = MAXR;
struct BoxDesc {
  int nm, nn, nk;        // axes of the M (A kept), N (B kept) and K (reduced) index spaces
  int NA2, NB2, NC2;     // tail ranks of A, B, C (box row widths)
  int m_ext[MAXAX], n_ext[MAXAX], k_ext[MAXAX];  // full extents
  int m_sa[MAXAX], m_sc[MAXAX];  // element strides of the M axes in an A block / a C block
  int n_sb[MAXAX], n_sc[MAXAX];  // of the N axes in a B block / a C block
  int k_sa[MAXAX], k_sb[MAXAX];  // of the K axes in an A block / a B block
  int m_ax[MAXAX], n_ax[MAXAX];  // box column (tail axis) of each M axis in A, N axis in B
  int k_axa[MAXAX], k_axb[MAXAX];
  int m_cax[MAXAX], n_cax[MAXAX];  // box column of each M / N axis in C
  int c_ext[MAXAX];                // C tail extents in C's axis order
};
struct BoxTile {
  int BM, BN;
};
BoxTile box_tile(int cfg);
int box_tile_configs();
bool box_tile_is_wave(int cfg);
struct BoxGemmArgs {
  BoxDesc d;
  i64 ng = 0, np = 0, max_tiles = 0;
  int cfg = 0, amf = 0, bnf = 0;
  int wgcu = 0;  // persistent workgroups per CU (0: as many as are resident)
  // fixed per plan: pairs in group order
  const i32* seg = nullptr;   // ng + 1
  const i64* aoff = nullptr;  // per pair: element offset of the A block
  const i64* boff = nullptr;
  const i32* arow = nullptr;  // per pair: box row of the A block
  const i32* brow = nullptr;
  const i64* coff = nullptr;  // per group: element offset of the C block
  const i32* gorder = nullptr;  // groups in locality order (null: key order)
  // per execution
  i32 *gbox = nullptr, *kbox = nullptr, *kc = nullptr, *cnew = nullptr, *clear = nullptr;
  i32 *tiles = nullptr, *toff = nullptr;  // tile counts per position; toff[0] = 0 from planning
  int2* trec = nullptr;                   // work list: {group, tile within the group}
  i32* cbox = nullptr;  // C's box rows (old until the GEMM kernel installs the new ones)
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
};
// One box GEMM of a batch, as a device-resident descriptor: the fixed plan
// arrays and per-execution buffers of a BoxGemmArgs with its operands.
struct BoxJob {
  BoxDesc d;
  i64 ng;
  int amf, bnf;
  const i32 *gorder, *seg, *arow, *brow;
  const i64 *aoff, *boff, *coff;
  const i32 *abox, *aflag, *bbox, *bflag;
  i32 *gbox, *kbox, *kc, *cnew, *clear, *cbox, *cflag;
  const double* A;
  const double* B;
  double* C;
};
// A launch over the group positions of several jobs (the active ones, act[0..nact)):
// job act[a] owns positions gbase[a] .. gbase[a+1]-1; one work list over all of them.
struct BoxBatchArgs {
  const BoxJob* jobs = nullptr;
  const i32* act = nullptr;
  const i64* gbase = nullptr;  // nact + 1 entries
  int nact = 0, cfg = 0;
  i64 npos = 0;
  i32 *tiles = nullptr, *toff = nullptr;  // per position; toff[0] = 0
  int4* trec = nullptr;                   // work list: {job, group, tile within the group, 0}
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
};
void box_prepare_batch(Ctx*, const BoxBatchArgs&);
void box_gemm_batch(Ctx*, unsigned prec, const BoxBatchArgs&);

// boxes and flags of n blocks (block i at (ids[i]-1)*bs, tail extents tdims)
void box_scan(Ctx*, const double* data, const i32* ids, i64 n, i64 bs, int N2, const i64* tdims, i32* box,
              i32* flag);
// every row of an n x w array set to row[0..w)
void k_fill_rows(Ctx*, i32* p, i64 n, int w, const i64* row);
size_t box_scan_tmp_bytes(Ctx*, i64 ng);
// group boxes, pair K extents, C boxes (cbox old -> new, entries outside a shrunken box zeroed) and,
// when trec is set, the (group, tile) work list
void box_prepare(Ctx*, const BoxGemmArgs&, const i32* abox, const i32* aflag, const i32* bbox, const i32* bflag,
                 i32* cbox, i32* cflag, double* C);
// C_g[box] = sum over pairs and K boxes of A_p * B_p (flags non-finite outputs in cflag); first
// clears the entries of C that fell outside shrunken boxes and installs those boxes
void box_gemm(Ctx*, unsigned prec, const BoxGemmArgs&, const double* A, const double* B, double* C, i32* cflag);

// element offsets of blocks: off[i] = (ids[i]-1)*bs  (ids == null: i*bs)
void k_block_offsets(Ctx*, const i32* ids, i64 n, i64 bs, i64* off);
// off[p] = base[idx[p]] (+ add)
void k_take_i64(Ctx*, const i64* base, const i32* idx, i64 n, i64* out);
void k_zero(Ctx*, void* p, size_t bytes);

}  // namespace sb

namespace sb {
// ---- extra symbolic helpers ----------------------------------------------
// off[i] = (i64)lin[i] * scale
void k_lin_scale(Ctx*, const u64* lin, i64 n, i64 scale, i64* off);
// full cross product, A-major: pair p = (p / nB, p % nB)
void k_cross_pairs(Ctx*, i64 nA, i64 nB, i32* pa, i32* pb);
// keep[p] = exists r: fA[pa[p]*R + r] && fB[(pb? pb[p] : 0)*R + r]; kmask[p*R + r] = the same
// per r (kmask may be null); fB row stride is R (a single row when pb == null)
void k_rank1_exist(Ctx*, const unsigned char* fA, const unsigned char* fB, const i32* pa, const i32* pb, i64 np,
                   i64 R, unsigned char* keep, unsigned char* kmask);
// compact pairs by keep flags (stable); returns the kept count (synchronises)
i64 compact_pairs(Ctx*, const unsigned char* keep, DBuf<i32>& pa, DBuf<i32>& pb, i64 np);
// out[i] = v[idx[i]]
void k_take_f64(Ctx*, const double* v, const i32* idx, i64 n, double* out);
// COO x COO products of pairs: prod[p] = va[pa[p]] * vb[pb[p]]; keep[p] = prod != 0
void k_coo_products(Ctx*, const double* va, const double* vb, const i32* pa, const i32* pb, i64 np, double* prod,
                    unsigned char* keep);
// COO x BS (tail r) existence: keep[p] = any(alpha[pa[p]] * X[boff[pb[p]] + (rv[pa[p]]-1)*chunk + t] != 0)
void k_scaled_slice_nonzero(Ctx*, const double* alpha, const i32* keysA, int NA, const double* X, const i64* blkoff,
                            const i32* pa, const i32* pb, i64 np, i64 chunk, unsigned char* keep);
// per-pair offsets into a tail-r matrix: off[p] = blkoff[pb[p]] + (keysA[pa[p]*NA + NA-1]-1)*chunk
// (blkoff == null: 0)
void k_rv_offsets(Ctx*, const i32* keysA, int w, int col, const i32* pa, const i64* blkoff, const i32* pb, i64 np,
                  i64 chunk, i64* off);
// keep[i] = 1 <= keys[i*w + col-1] <= R
void k_rv_inrange(Ctx*, const i32* keys, int w, int col, i64 n, i64 R, unsigned char* keep);
// keep[i] &= any(alpha[i]*X[(rv_i-1)*chunk + t] != 0)
void k_coo_dense_nonzero(Ctx*, const double* alpha, const i32* keys, int w, int col, i64 n, const double* X,
                         i64 chunk, unsigned char* keep);
// is the column non-decreasing over rows? (synchronises)
bool column_sorted(Ctx*, const i32* keys, int w, int col, i64 n);
// COO output keys: ck[p, t] = src[t] > 0 ? A[pa[p], src[t]-1] : B[pb[p], -src[t]-1]   (same as k_pair_keys w/o lin)
// ---- aliased block-sparse (sb_aliased.hip) ---------------------------------
void k_al_expand(Ctx*, i64 n, i64 bs, const i32* ids, const double* sc, const double* tm, double* out);
void k_al_xy(Ctx*, int mode, const i32* pa, const i32* pb, i64 np, const i32* idsA, const i32* keysA, int wA,
             int col, const i32* idsB, const int* cols, const i64* str, int ncols, i32* xy);
void k_al_xy_lin(Ctx*, const i32* xy, i64 np, i64 nx, u64* lin);
void k_al_seg_ids(Ctx*, i64 ns, const i32* seg, const i32* order, const i32* ids, const i32* ukeys, int w, i32* val,
                  i32* byid);
void k_al_offsets(Ctx*, const i32* u, int w, int c, i64 n, i64 scale, i64 add, i64* off);
void k_al_groups(Ctx*, i64 ng, const i32* seg, const i32* order, const i32* ctid, int demote_any, i32* is_acc,
                 i32* first_diff, i32* aid);
void k_al_final_ids(Ctx*, i64 ng, const i32* is_acc, const i64* rank, const i32* aid, i64 base, i32* out);
void k_al_acc_list(Ctx*, i64 ng, const i32* is_acc, const i64* rank, i32* list);
void k_al_alpha(Ctx*, const i32* pa, const i32* pb, i64 np, const double* sA, const double* sB, double* alpha);
void k_al_scalars(Ctx*, i64 ng, const i32* seg, const i32* order, const double* alpha, const i32* is_acc,
                  double* sc);
void k_al_accumulate(Ctx*, i64 nacc, i64 bs, const i32* list, const i32* seg, const i32* order,
                     const i32* first_diff, const double* alpha, const double* base, const i64* toff, double* out);
void k_al_ctid_off(Ctx*, const i32* ctid, i64 np, i64 bs, i64* toff);
void k_i32_to_u32(Ctx*, const i32* a, i64 n, unsigned* b);
void k_i32_to_i64(Ctx*, const i32* a, i64 n, i64* b);
void k_take_col(Ctx*, const i32* keys, int w, int col0, const i32* idx, i64 n, i32* out);
void k_rv_lin(Ctx*, const i32* v, i64 n, u64* lin);
void k_scatter_i32(Ctx*, const i32* idx, const i32* vals, i64 n, i32* out);
// thin Householder QR (LAPACK conventions) of a batch of column-major
// matrices, one workgroup each (sb_qr.hip); A is overwritten
struct QrJob {
  int m, n;
  i64 aoff, qoff, roff, toff;
};
void k_qr_batch(Ctx*, const QrJob* jobs, int njobs, double* A, double* Q, double* R, double* tau);
// dst[dmap[t]] = src[smap[t]] for t < n
void k_copy_map(Ctx*, i64 n, const i64* smap, const double* src, const i64* dmap, double* dst);
}  // namespace sb
