// contract! on the device: label analysis and dispatch exactly as the
// reference (SparseBackends/src/tensoralgebra/contract.jl:37-104), then for
// each kernel family a symbolic phase (key permutes, merge-joins, output
// block allocation; all integer kernels on the GPU, bit-exact) and a numeric
// phase (operand permutes + grouped block products).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <set>

#include "sb_contract.h"

namespace sb {

using LabelMap = std::map<Label, int>;

static LabelMap make_map(const Labels& l) {
  LabelMap m;
  for (size_t i = 0; i < l.size(); ++i) m[l[i]] = (int)i + 1;
  return m;
}
static bool has(const LabelMap& m, const Label& l) { return m.count(l) != 0; }
static int at(const LabelMap& m, const Label& l) {
  auto it = m.find(l);
  if (it == m.end()) fail(SB_E_KEY, "KeyError: label not found in tensor labels");
  return it->second;
}
static bool contains(const Labels& v, const Label& l) { return std::find(v.begin(), v.end(), l) != v.end(); }
static Labels sub(const Labels& l, size_t lo, size_t hi) { return Labels(l.begin() + lo, l.begin() + hi); }
static Labels cat(const Labels& a, const Labels& b) {
  Labels r(a);
  r.insert(r.end(), b.begin(), b.end());
  return r;
}
static bool same_set(const Labels& a, const Labels& b) {
  std::set<Label> x(a.begin(), a.end()), y(b.begin(), b.end());
  return x == y;
}
static bool is_identity(const std::vector<i64>& p) {
  for (size_t i = 0; i < p.size(); ++i)
    if (p[i] != (i64)i + 1) return false;
  return true;
}
static Labels permute_labels(const Labels& l, const std::vector<i64>& p) {
  Labels r(p.size());
  for (size_t i = 0; i < p.size(); ++i) r[i] = l[p[i] - 1];
  return r;
}

// An operand as seen by a kernel: the current tensor (maybe a permuted copy
// whose payload the plan fills), its labels and label map.
struct Oper {
  const Tensor* t;
  Labels labels;
  LabelMap map;
};

// ---------------------------------------------------------------- planned permutes
// permutedims(::NewBlockSparseSorted) split into its symbolic part (keys,
// ids, source offsets; run now) and its numeric part (payload gather; a plan step)
static const Tensor* plan_bs_permute(Plan& pl, const Tensor& A, const std::vector<i64>& perm) {
  Ctx* ctx = pl.ctx;
  int N = A.N(), P = A.P(), N2 = A.N2;
  std::vector<i64> dimsB(N);
  for (int j = 0; j < N; ++j) dimsB[j] = A.dims[perm[j] - 1];
  TP B = bs_make(ctx, dimsB, N2);
  i64 n = A.n;
  B->n = n;
  if (n > 0) {
    std::vector<int> sel(P);
    bool prefix_id = true;
    for (int j = 0; j < P; ++j) {
      sel[j] = (int)perm[j];
      prefix_id = prefix_id && perm[j] == j + 1;
    }
    B->keys.alloc(ctx, (size_t)n * (P ? P : 1));
    B->ids.alloc(ctx, n);
    B->data.alloc(ctx, (size_t)n * B->blksize);
    k_iota(ctx, B->ids.get(), n, 1);
    DBuf<i64> src(ctx, n);
    if (P > 0) {
      DBuf<i32> nk(ctx, (size_t)n * P), idx(ctx, n);
      DBuf<u64> lin(ctx, n);
      k_remap_keys(ctx, A.keys.get(), n, P, sel.data(), P, dimsB.data(), nk.get(), lin.get(), idx.get());
      if (!(prefix_id && !A.dirty)) sort_pairs(ctx, lin.get(), idx.get(), n, bits_for((u64)prod(dimsB, 0, P) - 1));
      k_gather_rows(ctx, nk.get(), P, idx.get(), n, B->keys.get());
      k_src_offsets(ctx, A.ids.get(), idx.get(), n, A.blksize, src.get());
    } else {
      k_block_offsets(ctx, A.ids.get(), n, A.blksize, src.get());
    }
    std::vector<i64> tdims(A.dims.begin() + P, A.dims.end());
    std::vector<int> pblk(N2);
    for (int i = 0; i < N2; ++i) pblk[i] = (int)(perm[P + i] - P - 1);
    const double* in = A.data.get();
    double* out = B->data.get();
    i64* so = pl.keep(std::move(src));
    i64 bs = B->blksize;
    pl.steps.push_back({"permute_blocks", [ctx, in, so, n, N2, tdims, pblk, out]() {
                          permute_blocks(ctx, in, so, n, N2, tdims.data(), pblk.data(), out);
                        }});
    pl.perm_bytes += 2.0 * n * bs * sizeof(double);
  }
  const Tensor* r = B.get();
  pl.keep_tensor(std::move(B));
  return r;
}

static const Tensor* plan_dense_permute(Plan& pl, const Tensor& Bt, const std::vector<i64>& perm) {
  Ctx* ctx = pl.ctx;
  int N = Bt.N();
  std::vector<i64> nd(N);
  std::vector<int> p0(N);
  for (int j = 0; j < N; ++j) {
    nd[j] = Bt.dims[perm[j] - 1];
    p0[j] = (int)perm[j] - 1;
  }
  if (N > MAXR) fail(SB_E_LIMIT, "dense permute of rank above 8");
  TP R = dense_new(ctx, nd);
  const double* in = Bt.dptr();
  double* out = R->data.get();
  std::vector<i64> dims = Bt.dims;
  pl.steps.push_back({"permute_dense", [ctx, in, N, dims, p0, out]() {
                        permute_blocks(ctx, in, nullptr, 1, N, dims.data(), p0.data(), out);
                      }});
  pl.perm_bytes += 2.0 * Bt.numel() * sizeof(double);
  const Tensor* r = R.get();
  pl.keep_tensor(std::move(R));
  return r;
}

// _find_permutation_for_tensor, single label (contract_bs_bs.jl:33-45)
static std::vector<i64> find_perm_for_tensor(const Tensor& T, const LabelMap& m, const Label& rlab) {
  int P = T.P(), N = T.N();
  int axisr = at(m, rlab);
  std::vector<i64> perm;
  if (axisr <= P) {
    for (int a = 1; a <= P; ++a)
      if (a != axisr) perm.push_back(a);
    perm.push_back(axisr);
    for (int a = P + 1; a <= N; ++a) perm.push_back(a);
  } else {
    for (int a = 1; a <= P; ++a) perm.push_back(a);
    for (int a = P + 1; a <= N; ++a)
      if (a != axisr) perm.push_back(a);
    perm.push_back(axisr);
  }
  return perm;
}

// _permute_r_to_last_prefix (contract_bs_bs.jl:77-95)
static void permute_r_to_last_prefix(Plan& pl, Oper& X, const Label& rlab) {
  int axisr = at(X.map, rlab);
  SB_ASSERT(axisr <= X.t->P(), "rlab must be in prefix to use prefix kernel");
  if (axisr == X.t->P()) return;
  auto perm = find_perm_for_tensor(*X.t, X.map, rlab);
  X.t = plan_bs_permute(pl, *X.t, perm);
  X.labels = permute_labels(X.labels, perm);
  X.map = make_map(X.labels);
}

// ---------------------------------------------------------------- shared pipeline pieces
struct Pairs {
  i64 n = 0;
  DBuf<i32> a, b;
};

// join-equal ranges -> pairs (A-major, B-minor within each A row: the reference traversal order)
static Pairs pairs_from_ranges(Ctx* ctx, DBuf<i32>& lo, DBuf<i32>& hi, i64 nA) {
  Pairs P;
  if (nA == 0) return P;
  DBuf<i64> cnt(ctx, nA), off(ctx, nA);
  k_range_counts(ctx, lo.get(), hi.get(), nA, cnt.get());
  P.n = exclusive_scan_total(ctx, cnt.get(), off.get(), nA);
  if (P.n >= (1ll << 31)) fail(SB_E_LIMIT, "more than 2^31 block pairs in one contraction");
  P.a.alloc(ctx, P.n ? P.n : 1);
  P.b.alloc(ctx, P.n ? P.n : 1);
  k_emit_pairs(ctx, lo.get(), off.get(), cnt.get(), nA, P.a.get(), P.b.get());
  return P;
}

// keys of C for every pair, then output-block allocation (_ensure_block! order)
static void allocate_from_pairs(Plan& pl, Tensor& C, const Pairs& pr, const i32* keysA, int PAw, const i32* keysB,
                                int PBw, const std::vector<int>& src, Alloc& al) {
  Ctx* ctx = pl.ctx;
  int PC = C.P();
  DBuf<i32> ck(ctx, (size_t)(pr.n ? pr.n : 1) * (PC ? PC : 1));
  DBuf<u64> lin(ctx, pr.n ? pr.n : 1);
  if (pr.n && PC > 0)
    k_pair_keys(ctx, pr.a.get(), pr.b.get(), pr.n, keysA, PAw, keysB, PBw, src.data(), PC, C.dims.data(), ck.get(),
                lin.get());
  else if (pr.n)
    k_zero(ctx, lin.get(), pr.n * sizeof(u64));
  u64 maxlin = PC > 0 ? (u64)prod(C.dims, 0, PC) - 1 : 0;
  allocate_blocks(ctx, pr.n, PC, ck, lin, maxlin, al);
  // install the block list in C (overwrite semantics: empty!(C.keys/ids/data))
  C.n = al.nC;
  C.dirty = false;
  C.keys = std::move(al.keys);
  C.ids.alloc(ctx, al.nC ? al.nC : 1);
  if (al.nC)
    HIP_CHECK(hipMemcpyAsync(C.ids.get(), al.ids.get(), al.nC * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
  C.data.alloc(ctx, (size_t)al.nC * C.blksize);
  pl.nblocks_out = al.nC;
  pl.npairs = pr.n;
}
static void allocate_from_pairs(Plan& pl, Tensor& C, const Pairs& pr, const Tensor* A, int PAw, const Tensor* B,
                                int PBw, const std::vector<int>& src, Alloc& al) {
  allocate_from_pairs(pl, C, pr, A ? A->keys.get() : nullptr, PAw, B ? B->keys.get() : nullptr, PBw, src, al);
}

// per-pair offsets: off[p] = blockoff(T)[idx[p]]
static i64* pair_offsets(Plan& pl, const Tensor& T, const i32* idx, i64 np) {
  Ctx* ctx = pl.ctx;
  DBuf<i64> bo = block_offsets(ctx, T);
  DBuf<i64> o(ctx, np ? np : 1);
  if (np) k_take_i64(ctx, bo.get(), idx, np, o.get());
  return pl.keep(std::move(o));
}
// per-group offsets of C blocks in group (key) order
static i64* group_offsets(Plan& pl, const Tensor& C) {
  DBuf<i64> o = block_offsets(pl.ctx, C);
  return pl.keep(std::move(o));
}

// ---------------------------------------------------------------- grouping of C dense labels
enum Mode { AthenB, BthenA, Interleaved };
// _cdense_grouping_and_orders (contract_bs_bs.jl:470-505)
static Mode cdense_grouping(const Labels& Cd, const Labels& keepA, const Labels& keepB, Labels& dA, Labels& dB) {
  std::vector<int> origin(Cd.size());
  for (size_t i = 0; i < Cd.size(); ++i) {
    if (contains(keepA, Cd[i]))
      origin[i] = 1;
    else if (contains(keepB, Cd[i]))
      origin[i] = 0;
    else
      fail(SB_E_ERROR, "Cdense label not found in keepA/keepB");
  }
  int firstB = -1, lastA = -1, firstA = -1, lastB = -1;
  for (int i = 0; i < (int)Cd.size(); ++i) {
    if (origin[i]) {
      if (firstA < 0) firstA = i;
      lastA = i;
    } else {
      if (firstB < 0) firstB = i;
      lastB = i;
    }
  }
  Mode m;
  if (firstB < 0 || firstA < 0)
    m = AthenB;
  else if (lastA < firstB)
    m = AthenB;
  else if (lastB < firstA)
    m = BthenA;
  else
    m = Interleaved;
  dA.clear();
  dB.clear();
  for (auto& l : Cd) {
    if (contains(keepA, l)) dA.push_back(l);
    if (contains(keepB, l)) dB.push_back(l);
  }
  return m;
}

static void add_gemm_step(Plan& pl, unsigned prec, const GemmDesc& d, const double* A, const i64* aoff,
                          const double* B, const i64* boff, double* C, const i64* coff, const i32* seg,
                          const i32* order, i64 ng, const unsigned char* kmask) {
  Ctx* ctx = pl.ctx;
  pl.steps.push_back({"grouped_gemm", [=]() { grouped_gemm(ctx, prec, d, A, aoff, B, boff, C, coff, seg, order, ng, kmask); }});
}

// ================================================================ BS x BS
// The keys of a block-sparse operand with its prefix reordered by pperm
// (1-based) and sorted by the column-major lin of the new prefix, exactly as
// permutedims leaves them (storage.jl:188-264); the payload stays where it
// is: view row i is row row[i] of the tensor (row == null: the identity).
struct KeyView {
  const i32* keys = nullptr;
  const i32* row = nullptr;
  Labels labels;  // prefix labels in view order, then the tail labels in tensor order
  LabelMap map;
};
static KeyView key_view(Plan& pl, const Oper& X, const std::vector<i64>& pperm) {
  Ctx* ctx = pl.ctx;
  const Tensor& T = *X.t;
  int P = T.P(), N = T.N();
  KeyView v;
  for (int j = 0; j < P; ++j) v.labels.push_back(X.labels[pperm[j] - 1]);
  for (int j = P; j < N; ++j) v.labels.push_back(X.labels[j]);
  v.map = make_map(v.labels);
  if ((is_identity(pperm) && !T.dirty) || T.n == 0 || P == 0) {
    v.keys = T.keys.get();
    return v;
  }
  i64 n = T.n;
  std::vector<int> sel(P);
  std::vector<i64> pd(P);
  for (int j = 0; j < P; ++j) {
    sel[j] = (int)pperm[j];
    pd[j] = T.dims[pperm[j] - 1];
  }
  DBuf<i32> nk(ctx, (size_t)n * P), idx(ctx, n), sorted(ctx, (size_t)n * P);
  DBuf<u64> lin(ctx, n);
  k_remap_keys(ctx, T.keys.get(), n, P, sel.data(), P, pd.data(), nk.get(), lin.get(), idx.get());
  sort_pairs(ctx, lin.get(), idx.get(), n, bits_for((u64)prod(pd, 0, P) - 1));
  k_gather_rows(ctx, nk.get(), P, idx.get(), n, sorted.get());
  v.keys = pl.keep(std::move(sorted));
  v.row = pl.keep(std::move(idx));
  return v;
}

// boxes of a block-sparse tensor, recomputed from its payload unless the last
// change of the payload left them valid
static void ensure_box(Ctx* ctx, const Tensor& T) {
  if (T.box_trusted()) return;
  if (T.bflag.n < (size_t)T.n) T.bflag.alloc(ctx, T.n ? T.n : 1);
  if (T.box.n < (size_t)(T.n * T.N2)) T.box.alloc(ctx, (size_t)(T.n * T.N2 ? T.n * T.N2 : 1));
  std::vector<i64> td(T.dims.begin() + T.P(), T.dims.end());
  box_scan(ctx, T.data.get(), T.ids.get(), T.n, T.blksize, T.N2, td.data(), T.box.get(), T.bflag.get());
  T.box_version = T.version;
}

static void shared_bs_bs_permuted(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B,
                                  const Labels& shared, unsigned flags);

// contract_shared! BS x BS (contract_bs_bs.jl:572-730) without operand
// permutes: the symbolic phase (key views, merge-join, output allocation)
// follows the reference; the numeric phase is the box-aware block product.
static void shared_bs_bs(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared,
                         unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "bs_bs_shared";
  for (auto& l : shared) SB_ASSERT(!contains(labelsC, l), "shared label must be reduced (must not appear in labelsC)");
  int PA = A.t->P(), PB = B.t->P(), NA = A.t->N(), NB = B.t->N();
  Labels sp, sd;
  for (auto& l : shared) {
    SB_ASSERT(has(A.map, l) && has(B.map, l), "shared label must exist in both A and B");
    int ap = at(A.map, l), bp = at(B.map, l);
    bool apf = ap <= PA, bpf = bp <= PB;
    if (apf && bpf)
      sp.push_back(l);
    else if (!apf && !bpf)
      sd.push_back(l);
    else
      fail(SB_E_ERROR, "shared label crosses prefix/dense; not supported");
  }
  Labels Ad0 = sub(A.labels, PA, NA), Bd0 = sub(B.labels, PB, NB);
  Labels red, keepA0, keepB0;
  for (auto& l : Ad0) (contains(sd, l) ? red : keepA0).push_back(l);
  for (auto& l : Bd0)
    if (!contains(sd, l)) keepB0.push_back(l);
  int PC = C.P(), NC = C.N();
  Labels Cd = sub(labelsC, PC, NC), dA, dB;
  Mode mode = cdense_grouping(Cd, keepA0, keepB0, dA, dB);
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims; would require permuting C (unsupported)");
  SB_ASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA covers keepA");
  SB_ASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB covers keepB");
  // the reduced extents must agree axis by axis for the strided product (the
  // reference compares their products only); blocks must index with int32
  bool strided = A.t->N2 <= MAXAX && B.t->N2 <= MAXAX && C.N2 <= MAXAX && A.t->blksize < (1ll << 31) &&
                 B.t->blksize < (1ll << 31) && C.blksize < (1ll << 31);
  for (auto& l : red) strided = strided && A.t->dims[at(A.map, l) - 1] == B.t->dims[at(B.map, l) - 1];
  if (!strided) {
    shared_bs_bs_permuted(pl, C, labelsC, A, B, shared, flags);
    return;
  }
  // prefix order of the reference's permuted operands: unshared axes, then the shared ones
  auto prefix_perm = [&](const Oper& X) {
    int P = X.t->P();
    std::vector<i64> spax, perm;
    for (auto& l : sp) spax.push_back(at(X.map, l));
    for (int a = 1; a <= P; ++a)
      if (std::find(spax.begin(), spax.end(), a) == spax.end()) perm.push_back(a);
    for (auto a : spax) perm.push_back(a);
    return perm;
  };
  KeyView VA = key_view(pl, A, prefix_perm(A)), VB = key_view(pl, B, prefix_perm(B));
  const Tensor &TA = *A.t, &TB = *B.t;
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    if (has(VA.map, l) && at(VA.map, l) <= PA)
      src[j] = at(VA.map, l);
    else if (has(VB.map, l) && at(VB.map, l) <= PB)
      src[j] = -at(VB.map, l);
    else
      fail(SB_E_ERROR, "C prefix label must come from A/B prefix");
  }
  std::vector<int> jA, jB;
  std::vector<i64> spd;
  for (auto& l : sp) {
    jA.push_back(at(VA.map, l));
    jB.push_back(at(VB.map, l));
    spd.push_back(TA.dims[at(A.map, l) - 1]);
  }
  // tail geometry: M = A-kept axes, N = B-kept axes, K = reduced axes
  auto tail_strides = [](const Tensor& T) {
    std::vector<i64> s(T.N2);
    i64 acc = 1;
    for (int a = 0; a < T.N2; ++a) {
      s[a] = acc;
      acc *= T.dims[T.P() + a];
    }
    return s;
  };
  std::vector<i64> sA = tail_strides(TA), sB = tail_strides(TB), sC = tail_strides(C);
  struct Ax {
    int ext, s1, s2, c1, c2;  // extent, strides in operand 1/2, box columns in operand 1/2
  };
  auto cpos = [&](const Label& l) { return (int)(std::find(Cd.begin(), Cd.end(), l) - Cd.begin()); };
  std::vector<Ax> mx, nx, kx;
  for (auto& l : dA) {
    int ta = at(A.map, l) - PA - 1, tc = cpos(l);
    mx.push_back({(int)TA.dims[PA + ta], (int)sA[ta], (int)sC[tc], ta, tc});
  }
  for (auto& l : dB) {
    int tb = at(B.map, l) - PB - 1, tc = cpos(l);
    nx.push_back({(int)TB.dims[PB + tb], (int)sB[tb], (int)sC[tc], tb, tc});
  }
  for (auto& l : red) {
    int ta = at(A.map, l) - PA - 1, tb = at(B.map, l) - PB - 1;
    kx.push_back({(int)TA.dims[PA + ta], (int)sA[ta], (int)sB[tb], ta, tb});
  }
  // decode order: smallest operand stride first (the fastest-varying index of a tile)
  auto by_s1 = [](const Ax& x, const Ax& y) { return x.s1 < y.s1; };
  std::stable_sort(mx.begin(), mx.end(), by_s1);
  std::stable_sort(nx.begin(), nx.end(), by_s1);
  std::stable_sort(kx.begin(), kx.end(), by_s1);
  i64 M = 1, K = 1, N = 1;
  for (auto& x : mx) M *= x.ext;
  for (auto& x : kx) K *= x.ext;
  for (auto& x : nx) N *= x.ext;
  SB_ASSERT(C.blksize == M * N, "C.blksize must equal M*N");
  i64 nA = TA.n, nB = TB.n;
  DBuf<i32> lo(ctx, nA ? nA : 1), hi(ctx, nA ? nA : 1);
  int nsp = (int)sp.size();
  if (nA && nB) {
    if (nsp == 0) {
      // join on nothing: every A block meets every B block
      DBuf<u64> z(ctx, nA), zb(ctx, nB);
      k_zero(ctx, z.get(), nA * sizeof(u64));
      k_zero(ctx, zb.get(), nB * sizeof(u64));
      k_equal_ranges(ctx, z.get(), nA, zb.get(), nB, lo.get(), hi.get());
    } else if (nsp >= 2 && !(flags & SB_FLAG_EXACT_JOIN)) {
      // the reference's merge walk compares the first shared axis first
      k_walk_ranges(ctx, 2, VA.keys, PA, nA, VB.keys, PB, nB, jA.data(), jB.data(), nsp, lo.get(), hi.get());
    } else {
      DBuf<u64> ja(ctx, nA), jb(ctx, nB);
      k_cols_lin(ctx, VA.keys, nA, PA, jA.data(), nsp, spd.data(), ja.get());
      k_cols_lin(ctx, VB.keys, nB, PB, jB.data(), nsp, spd.data(), jb.get());
      k_equal_ranges(ctx, ja.get(), nA, jb.get(), nB, lo.get(), hi.get());
    }
  }
  Pairs pr = (nA && nB) ? pairs_from_ranges(ctx, lo, hi, nA) : Pairs();
  Alloc al;
  allocate_from_pairs(pl, C, pr, VA.keys, PA, VB.keys, PB, src, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * M * K * N * np;
  pl.gemm_bytes += 8.0 * (double)np * (M * K + K * N) + 8.0 * ng * M * N;
  pl.a_bytes += 8.0 * TA.n * TA.blksize;
  pl.b_bytes += 8.0 * TB.n * TB.blksize;
  pl.c_bytes += 8.0 * ng * C.blksize;
  if (!ng) return;
  // C's fresh payload is zeroed by one contiguous fill and its boxes start
  // empty, so the first execution clears nothing (a clear of stale entries
  // walks every padded block element by element)
  C.bflag.alloc(ctx, ng);
  C.box.alloc(ctx, (size_t)(ng * C.N2 ? ng * C.N2 : 1));
  k_zero(ctx, C.data.get(), (size_t)ng * C.blksize * sizeof(double));
  k_zero(ctx, C.box.get(), (size_t)ng * C.N2 * sizeof(i32));
  C.box_version = C.version;
  // pairs in group order: tensor rows and payload offsets of their blocks
  DBuf<i32> ar(ctx, np ? np : 1), br(ctx, np ? np : 1);
  if (np) {
    DBuf<i32> t(ctx, np);
    k_gather_rows(ctx, pr.a.get(), 1, al.order.get(), np, t.get());
    if (VA.row)
      k_gather_rows(ctx, VA.row, 1, t.get(), np, ar.get());
    else
      std::swap(ar, t);
    DBuf<i32> u(ctx, np);
    k_gather_rows(ctx, pr.b.get(), 1, al.order.get(), np, u.get());
    if (VB.row)
      k_gather_rows(ctx, VB.row, 1, u.get(), np, br.get());
    else
      std::swap(br, u);
  }
  DBuf<i64> ao(ctx, np ? np : 1), bo(ctx, np ? np : 1);
  if (np) {
    k_src_offsets(ctx, TA.ids.get(), ar.get(), np, TA.blksize, ao.get());
    k_src_offsets(ctx, TB.ids.get(), br.get(), np, TB.blksize, bo.get());
  }
  BoxGemmArgs g;
  BoxDesc& d = g.d;
  std::memset(&d, 0, sizeof(d));
  d.nm = (int)mx.size();
  d.nn = (int)nx.size();
  d.nk = (int)kx.size();
  d.NA2 = TA.N2;
  d.NB2 = TB.N2;
  d.NC2 = C.N2;
  SB_ASSERT(d.nm + d.nn == C.N2, "C tail = A-kept + B-kept axes");
  for (int i = 0; i < d.nm; ++i) {
    d.m_ext[i] = mx[i].ext;
    d.m_sa[i] = mx[i].s1;
    d.m_sc[i] = mx[i].s2;
    d.m_ax[i] = mx[i].c1;
    d.m_cax[i] = mx[i].c2;
  }
  for (int i = 0; i < d.nn; ++i) {
    d.n_ext[i] = nx[i].ext;
    d.n_sb[i] = nx[i].s1;
    d.n_sc[i] = nx[i].s2;
    d.n_ax[i] = nx[i].c1;
    d.n_cax[i] = nx[i].c2;
  }
  for (int i = 0; i < d.nk; ++i) {
    d.k_ext[i] = kx[i].ext;
    d.k_sa[i] = kx[i].s1;
    d.k_sb[i] = kx[i].s2;
    d.k_axa[i] = kx[i].c1;
    d.k_axb[i] = kx[i].c2;
  }
  for (int a = 0; a < C.N2; ++a) d.c_ext[a] = (int)C.dims[PC + a];
  g.ng = ng;
  g.np = np;
  g.amf = d.nm > 0 && (d.nk == 0 || d.m_sa[0] < d.k_sa[0]);
  g.bnf = d.nn > 0 && (d.nk == 0 || d.n_sb[0] < d.k_sb[0]);
  g.seg = pl.keep(std::move(al.seg_start));
  g.arow = pl.keep(std::move(ar));
  g.brow = pl.keep(std::move(br));
  g.aoff = pl.keep(std::move(ao));
  g.boff = pl.keep(std::move(bo));
  g.coff = group_offsets(pl, C);
  g.gbox = pl.keep(DBuf<i32>(ctx, (size_t)ng * std::max(1, d.nm + d.nn)));
  g.kbox = pl.keep(DBuf<i32>(ctx, (size_t)std::max<i64>(1, np * d.nk)));
  g.kc = pl.keep(DBuf<i32>(ctx, np ? np : 1));
  g.cnew = pl.keep(DBuf<i32>(ctx, (size_t)ng * std::max(1, C.N2)));
  g.clear = pl.keep(DBuf<i32>(ctx, ng));
  g.tiles = pl.keep(DBuf<i32>(ctx, ng));
  g.toff = pl.keep(DBuf<i32>(ctx, ng + 1));
  {
    const i64 zero = 0;
    k_fill_rows(ctx, g.toff, 1, 1, &zero);  // toff[0]; the scan writes toff[1..ng] every execution
  }
  g.cbox = C.box.get();
  // locality order of the groups: by the C key with its A-sourced axes most
  // significant, so groups reading the same A blocks sit next to each other
  if (PC > 0 && ng > 1) {
    std::vector<int> sel;
    std::vector<i64> pd;
    for (int pass = 0; pass < 2; ++pass)
      for (int j = 0; j < PC; ++j)
        if ((src[j] > 0) == (pass == 1)) {
          sel.push_back(j + 1);
          pd.push_back(C.dims[j]);
        }
    DBuf<i32> tk(ctx, (size_t)ng * PC), idx(ctx, ng);
    DBuf<u64> lin(ctx, ng);
    k_remap_keys(ctx, C.keys.get(), ng, PC, sel.data(), PC, pd.data(), tk.get(), lin.get(), idx.get());
    sort_pairs(ctx, lin.get(), idx.get(), ng, bits_for((u64)prod(pd, 0, PC) - 1));
    g.gorder = pl.keep(std::move(idx));
  }
  g.scan_tmp_bytes = box_scan_tmp_bytes(ctx, ng);
  g.scan_tmp = pl.keep(DBuf<unsigned char>(ctx, g.scan_tmp_bytes));
  std::vector<i64> cfull(C.dims.begin() + PC, C.dims.end());
  const Tensor *pA = &TA, *pB = &TB;
  Tensor* pC = &C;
  // tile shape: forced by the flags, or the configuration with the least
  // padded work over the group boxes of the current operands (any shape is
  // correct; this is speed only)
  const unsigned forced = (flags & SB_TILE_MASK) >> SB_TILE_SHIFT;
  if (forced) {
    if ((int)forced > box_tile_configs()) fail(SB_E_ARGUMENT, "unknown tile shape in flags");
    g.cfg = (int)forced - 1;
  } else {
    ensure_box(ctx, TA);
    ensure_box(ctx, TB);
    box_prepare(ctx, g, TA.box.get(), TA.bflag.get(), TB.box.get(), TB.bflag.get(), C.box.get(), C.bflag.get(),
                C.data.get());
    C.box_version = C.version;  // entries outside the boxes are zero now
    std::vector<i32> hb((size_t)ng * std::max(1, d.nm + d.nn));
    HIP_CHECK(hipMemcpyAsync(hb.data(), g.gbox, hb.size() * sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double best = 1e300;
    for (int c = 0; c < box_tile_configs(); ++c) {
      if (box_tile_is_wave(c)) continue;  // wave tiles only on request
      BoxTile t = box_tile(c);
      double cost = 0;
      for (i64 q = 0; q < ng; ++q) {
        i64 mc = 1, nc = 1;
        for (int a = 0; a < d.nm; ++a) mc *= hb[q * (d.nm + d.nn) + a];
        for (int a = 0; a < d.nn; ++a) nc *= hb[q * (d.nm + d.nn) + d.nm + a];
        if (mc && nc) cost += (double)((mc + t.BM - 1) / t.BM) * ((nc + t.BN - 1) / t.BN) * (t.BM * t.BN + 1024);
      }
      if (cost < best) {
        best = cost;
        g.cfg = c;
      }
    }
    if (flags & SB_FLAG_AUTOTUNE) {
      const BoxOperands o{TA.box.get(), TA.bflag.get(), TB.box.get(), TB.bflag.get(), C.box.get(), C.bflag.get(),
                          TA.data.get(), TB.data.get(), C.data.get()};
      pl.tune = autotune_box(ctx, flags, g, hb, M, N, o);
      g.cfg = pl.tune.cfg;
      g.wgcu = pl.tune.wgcu;
    }
  }
  BoxTile t = box_tile(g.cfg);
  g.max_tiles = ng * ((M + t.BM - 1) / t.BM) * ((N + t.BN - 1) / t.BN);
  SB_ASSERT(g.max_tiles < (1ll << 31), "box_gemm: work list too large");
  g.trec = pl.keep(DBuf<int2>(ctx, (size_t)std::max<i64>(1, g.max_tiles)));
  pl.steps.push_back({"box_scan", [ctx, pA, pB]() {
                        ensure_box(ctx, *pA);
                        ensure_box(ctx, *pB);
                      }});
  pl.steps.push_back({"box_prepare", [ctx, g, pA, pB, pC, cfull]() {
                        if (!pC->box_trusted()) k_fill_rows(ctx, pC->box.get(), g.ng, pC->N2, cfull.data());
                        box_prepare(ctx, g, pA->box.get(), pA->bflag.get(), pB->box.get(), pB->bflag.get(),
                                    pC->box.get(), pC->bflag.get(), pC->data.get());
                      }});
  const double *Ad = TA.data.get(), *Bd = TB.data.get();
  double* Cdp = C.data.get();
  i32* cflag = C.bflag.get();
  pl.steps.push_back({"grouped_gemm", [ctx, flags, g, Ad, Bd, Cdp, cflag]() { box_gemm(ctx, flags, g, Ad, Bd, Cdp, cflag); }});
  pl.out_box = true;
  if (pl.steps.size() == 3) pl.box = std::make_shared<BoxPlan>(BoxPlan{g, pA, pB, pC, flags, cfull});
}

// contract_shared! BS x BS through permuted operand copies (the reference's
// own route, contract_bs_bs.jl:572-730); used where the strided product does
// not apply
static void shared_bs_bs_permuted(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B,
                                  const Labels& shared, unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "bs_bs_shared";
  for (auto& l : shared) SB_ASSERT(!contains(labelsC, l), "shared label must be reduced (must not appear in labelsC)");
  int PA = A.t->P(), PB = B.t->P(), NA = A.t->N(), NB = B.t->N();
  Labels sp, sd;
  for (auto& l : shared) {
    SB_ASSERT(has(A.map, l) && has(B.map, l), "shared label must exist in both A and B");
    int ap = at(A.map, l), bp = at(B.map, l);
    bool apf = ap <= PA, bpf = bp <= PB;
    if (apf && bpf)
      sp.push_back(l);
    else if (!apf && !bpf)
      sd.push_back(l);
    else
      fail(SB_E_ERROR, "shared label crosses prefix/dense; not supported");
  }
  Labels Ad0 = sub(A.labels, PA, NA), Bd0 = sub(B.labels, PB, NB);
  Labels red, keepA0, keepB0;
  for (auto& l : Ad0) (contains(sd, l) ? red : keepA0).push_back(l);
  for (auto& l : Bd0)
    if (!contains(sd, l)) keepB0.push_back(l);
  int PC = C.P(), NC = C.N();
  Labels Cd = sub(labelsC, PC, NC), dA, dB;
  Mode mode = cdense_grouping(Cd, keepA0, keepB0, dA, dB);
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims; would require permuting C (unsupported)");
  SB_ASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA covers keepA");
  SB_ASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB covers keepB");
  auto build_perm = [&](Oper& X, bool red_first, const Labels& keep) {
    int P = X.t->P();
    std::vector<i64> spax;
    for (auto& l : sp) spax.push_back(at(X.map, l));
    std::vector<i64> perm;
    for (int a = 1; a <= P; ++a)
      if (std::find(spax.begin(), spax.end(), a) == spax.end()) perm.push_back(a);
    for (auto a : spax) perm.push_back(a);
    std::vector<i64> ka, ra;
    for (auto& l : keep) ka.push_back(at(X.map, l) - P);
    for (auto& l : red) ra.push_back(at(X.map, l) - P);
    SB_ASSERT((int)(ka.size() + ra.size()) == X.t->N2, "dense permutation covers the tail");
    for (auto a : (red_first ? ra : ka)) perm.push_back(P + a);
    for (auto a : (red_first ? ka : ra)) perm.push_back(P + a);
    return perm;
  };
  auto pA = build_perm(A, false, dA);
  auto pB = build_perm(B, true, dB);
  if (!(is_identity(pA) && !A.t->dirty)) {
    A.t = plan_bs_permute(pl, *A.t, pA);
    A.labels = permute_labels(A.labels, pA);
    A.map = make_map(A.labels);
  }
  if (!(is_identity(pB) && !B.t->dirty)) {
    B.t = plan_bs_permute(pl, *B.t, pB);
    B.labels = permute_labels(B.labels, pB);
    B.map = make_map(B.labels);
  }
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    if (has(A.map, l) && at(A.map, l) <= PA)
      src[j] = at(A.map, l);
    else if (has(B.map, l) && at(B.map, l) <= PB)
      src[j] = -at(B.map, l);
    else
      fail(SB_E_ERROR, "C prefix label must come from A/B prefix");
  }
  std::vector<int> jA, jB;
  std::vector<i64> spd;
  for (auto& l : sp) {
    jA.push_back(at(A.map, l));
    jB.push_back(at(B.map, l));
    spd.push_back(A.t->dims[at(A.map, l) - 1]);
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 M = prod(TA.dims, PA, PA + dA.size());
  i64 K = prod(TA.dims, PA + dA.size(), NA);
  i64 N = prod(TB.dims, PB + red.size(), NB);
  if (!red.empty()) SB_ASSERT(K == prod(TB.dims, PB, PB + red.size()), "Reduced dense extents mismatch between A and B");
  SB_ASSERT(C.blksize == M * N, "C.blksize must equal M*N");
  i64 nA = TA.n, nB = TB.n;
  DBuf<i32> lo(ctx, nA ? nA : 1), hi(ctx, nA ? nA : 1);
  int nsp = (int)sp.size();
  if (nA && nB) {
    if (nsp == 0) {
      // join on nothing: every A block meets every B block
      DBuf<u64> z(ctx, nA), zb(ctx, nB);
      k_zero(ctx, z.get(), nA * sizeof(u64));
      k_zero(ctx, zb.get(), nB * sizeof(u64));
      k_equal_ranges(ctx, z.get(), nA, zb.get(), nB, lo.get(), hi.get());
    } else if (nsp >= 2 && !(flags & SB_FLAG_EXACT_JOIN)) {
      // the reference's merge walk compares the first shared axis first
      k_walk_ranges(ctx, 2, TA.keys.get(), PA, nA, TB.keys.get(), PB, nB, jA.data(), jB.data(), nsp, lo.get(),
                    hi.get());
    } else {
      DBuf<u64> ja(ctx, nA), jb(ctx, nB);
      k_cols_lin(ctx, TA.keys.get(), nA, PA, jA.data(), nsp, spd.data(), ja.get());
      k_cols_lin(ctx, TB.keys.get(), nB, PB, jB.data(), nsp, spd.data(), jb.get());
      k_equal_ranges(ctx, ja.get(), nA, jb.get(), nB, lo.get(), hi.get());
    }
  }
  Pairs pr = (nA && nB) ? pairs_from_ranges(ctx, lo, hi, nA) : Pairs();
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, PA, &TB, PB, src, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * M * K * N * np;
  pl.gemm_bytes += 8.0 * (double)np * (M * K + K * N) + 8.0 * ng * M * N;
  pl.a_bytes += 8.0 * TA.n * TA.blksize;
  pl.b_bytes += 8.0 * TB.n * TB.blksize;
  pl.c_bytes += 8.0 * ng * C.blksize;
  if (!np) return;
  i64* aoff = pair_offsets(pl, TA, pr.a.get(), np);
  i64* boff = pair_offsets(pl, TB, pr.b.get(), np);
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  pl.keep(std::move(pr.a));
  pl.keep(std::move(pr.b));
  GemmDesc d{M, N, K, 1, M, 1, K, mode == AthenB ? 1 : N, mode == AthenB ? M : 1};
  add_gemm_step(pl, flags, d, TA.data.get(), aoff, TB.data.get(), boff, C.data.get(), coff, seg, order, ng, nullptr);
}

// contract_prefix_outer_bb! (contract_bs_bs.jl:173-263)
static void prefix_outer_bb(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "bs_bs_prefix_outer";
  SB_ASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PA = A.t->P(), PB = B.t->P(), PC = C.P(), NA = A.t->N(), NB = B.t->N(), NC = C.N();
  Labels Ad = sub(A.labels, PA, NA), Bd = sub(B.labels, PB, NB), Cd = sub(labelsC, PC, NC);
  for (auto& l : Ad) SB_ASSERT(!contains(Bd, l), "outer kernel assumes disjoint dense labels");
  int dense_order;
  if (Cd == cat(Ad, Bd))
    dense_order = 1;
  else if (Cd == cat(Bd, Ad))
    dense_order = 2;
  else
    fail(SB_E_ERROR, "C dense must be [A_dense...,B_dense...] or [B_dense..., A_dense...]");
  SB_ASSERT(C.blksize == A.t->blksize * B.t->blksize, "C.blksize must equal A.blksize * B.blksize");
  int axisAr = at(A.map, rlab), axisBr = at(B.map, rlab);
  SB_ASSERT(axisAr == PA, "axisAr == PA");
  SB_ASSERT(axisBr == PB, "axisBr == PB");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    if (has(A.map, labelsC[j]))
      src[j] = at(A.map, labelsC[j]);
    else {
      SB_ASSERT(has(B.map, labelsC[j]), "Label not found in A or B");
      src[j] = -at(B.map, labelsC[j]);
    }
    if ((src[j] > 0 && src[j] > PA) || (src[j] < 0 && -src[j] > PB)) fail(SB_E_BOUNDS, "BoundsError: key tuple");
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 nA = TA.n, nB = TB.n;
  DBuf<i32> lo(ctx, nA ? nA : 1), hi(ctx, nA ? nA : 1);
  if (nA && nB) {
    int ca = PA, cb = PB;
    if (TA.dirty || TB.dirty) {
      // keys not in column-major order: the reference's run walk as written
      k_walk_ranges(ctx, 0, TA.keys.get(), PA, nA, TB.keys.get(), PB, nB, &ca, &cb, 1, lo.get(), hi.get());
    } else {
      i64 rd = TA.dims[PA - 1];
      DBuf<u64> ja(ctx, nA), jb(ctx, nB);
      k_cols_lin(ctx, TA.keys.get(), nA, PA, &ca, 1, &rd, ja.get());
      k_cols_lin(ctx, TB.keys.get(), nB, PB, &cb, 1, &rd, jb.get());
      k_equal_ranges(ctx, ja.get(), nA, jb.get(), nB, lo.get(), hi.get());
    }
  }
  Pairs pr = (nA && nB) ? pairs_from_ranges(ctx, lo, hi, nA) : Pairs();
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, PA, &TB, PB, src, al);
  i64 np = pr.n, ng = al.nC, ma = TA.blksize, nb = TB.blksize;
  pl.flops += 2.0 * ma * nb * np;
  pl.gemm_bytes += 8.0 * (double)np * (ma + nb) + 8.0 * ng * ma * nb;
  if (!np) return;
  i64* aoff = pair_offsets(pl, TA, pr.a.get(), np);
  i64* boff = pair_offsets(pl, TB, pr.b.get(), np);
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  pl.keep(std::move(pr.a));
  pl.keep(std::move(pr.b));
  const double *Ad_ = TA.data.get(), *Bd_ = TB.data.get();
  double* Cd_ = C.data.get();
  if (dense_order == 1)
    pl.steps.push_back({"outer_groups", [=]() { outer_groups(ctx, ma, nb, Ad_, aoff, Bd_, boff, Cd_, coff, seg, order, ng); }});
  else
    pl.steps.push_back({"outer_groups", [=]() { outer_groups(ctx, nb, ma, Bd_, boff, Ad_, aoff, Cd_, coff, seg, order, ng); }});
}

// contract_dense_bb! (contract_bs_bs.jl:288-423)
static void dense_bb(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab, unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "bs_bs_dense_r";
  SB_ASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PA = A.t->P(), PB = B.t->P(), NA = A.t->N(), NB = B.t->N();
  int axisAr = at(A.map, rlab);
  SB_ASSERT(axisAr > PA, "rlab must be in A dense tail");
  int axisBr = at(B.map, rlab);
  SB_ASSERT(axisBr > PB, "rlab must be in B dense tail");
  LabelMap mapA0 = A.map, mapB0 = B.map;  // prefix sourcing uses the unpermuted maps
  if (axisAr != NA) {
    auto p = find_perm_for_tensor(*A.t, A.map, rlab);
    A.t = plan_bs_permute(pl, *A.t, p);
    A.labels = permute_labels(A.labels, p);
  }
  if (axisBr != NB) {
    auto p = find_perm_for_tensor(*B.t, B.map, rlab);
    B.t = plan_bs_permute(pl, *B.t, p);
    B.labels = permute_labels(B.labels, p);
  }
  SB_ASSERT(A.labels.back() == rlab, "labelsA[end] == rlab");
  SB_ASSERT(B.labels.back() == rlab, "labelsB[end] == rlab");
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 R = TA.dims[NA - 1];
  SB_ASSERT(TB.dims[NB - 1] == R, "dimsB_dense[end] == R");
  i64 chunkA = TA.blksize / R, chunkB = TB.blksize / R;
  int PC = C.P(), NC = C.N();
  SB_ASSERT(C.N2 == (TA.N2 - 1) + (TB.N2 - 1), "N2C == (N2A-1) + (N2B-1)");
  SB_ASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(B.labels, PB, NB - 1), Cd = sub(labelsC, PC, NC);
  bool AB;
  if (Cd == cat(Awo, Bwo))
    AB = true;
  else if (Cd == cat(Bwo, Awo))
    AB = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense..., B_dense...] or [B_dense..., A_dense...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    if (has(mapA0, labelsC[j])) {
      int ax = at(mapA0, labelsC[j]);
      SB_ASSERT(ax <= PA, "ax <= PA");
      src[j] = ax;
    } else {
      int ax = at(mapB0, labelsC[j]);
      SB_ASSERT(ax <= PB, "ax <= PB");
      src[j] = -ax;
    }
  }
  i64 nA = TA.n, nB = TB.n;
  Pairs pr;
  DBuf<unsigned char> kmask;
  if (nA && nB) {
    // the payloads of the permuted operands are needed now (block existence
    // depends on values): run the permute steps recorded so far
    pl.run_pending();
    DBuf<i64> offA = block_offsets(ctx, TA), offB = block_offsets(ctx, TB);
    DBuf<unsigned char> fA(ctx, nA * R), fB(ctx, nB * R);
    slice_nonzero(ctx, TA.data.get(), offA.get(), nA, R, chunkA, fA.get());
    slice_nonzero(ctx, TB.data.get(), offB.get(), nB, R, chunkB, fB.get());
    i64 np0 = nA * nB;
    if (np0 >= (1ll << 31)) fail(SB_E_LIMIT, "more than 2^31 block pairs in one contraction");
    pr.a.alloc(ctx, np0);
    pr.b.alloc(ctx, np0);
    k_cross_pairs(ctx, nA, nB, pr.a.get(), pr.b.get());
    DBuf<unsigned char> keep(ctx, np0);
    k_rank1_exist(ctx, fA.get(), fB.get(), pr.a.get(), pr.b.get(), np0, R, keep.get(), nullptr);
    pr.n = compact_pairs(ctx, keep.get(), pr.a, pr.b, np0);
    if (pr.n) {
      kmask.alloc(ctx, pr.n * R);
      DBuf<unsigned char> keep2(ctx, pr.n);
      k_rank1_exist(ctx, fA.get(), fB.get(), pr.a.get(), pr.b.get(), pr.n, R, keep2.get(), kmask.get());
    }
  }
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, PA, &TB, PB, src, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * chunkA * chunkB * R * np;
  pl.gemm_bytes += 8.0 * (double)np * (chunkA + chunkB) * R + 8.0 * ng * chunkA * chunkB;
  if (!np) return;
  i64* aoff = pair_offsets(pl, TA, pr.a.get(), np);
  i64* boff = pair_offsets(pl, TB, pr.b.get(), np);
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  unsigned char* km = pl.keep(std::move(kmask));
  const i32* pa = pl.keep(std::move(pr.a));
  const i32* pb = pl.keep(std::move(pr.b));
  // existence was decided on the current values and replays keep that block
  // list; which r slices contribute (both slices nonzero, :374-379) follows
  // the values of every replay
  {
    const i64* oA = pl.keep(block_offsets(ctx, TA));
    const i64* oB = pl.keep(block_offsets(ctx, TB));
    unsigned char* fa = pl.keep(DBuf<unsigned char>(ctx, nA * R));
    unsigned char* fb = pl.keep(DBuf<unsigned char>(ctx, nB * R));
    unsigned char* kp = pl.keep(DBuf<unsigned char>(ctx, np));
    const double* dA = TA.data.get();
    const double* dB = TB.data.get();
    pl.steps.push_back({"rank1_mask", [=]() {
                          slice_nonzero(ctx, dA, oA, nA, R, chunkA, fa);
                          slice_nonzero(ctx, dB, oB, nB, R, chunkB, fb);
                          k_rank1_exist(ctx, fa, fb, pa, pb, np, R, kp, km);
                        }});
  }
  // A block (chunkA x R), B block (chunkB x R): C[i,j] = sum_r A[i,r] B[j,r]
  GemmDesc d{chunkA, chunkB, R, 1, chunkA, chunkB, 1, AB ? 1 : chunkB, AB ? chunkA : 1};
  add_gemm_step(pl, flags, d, TA.data.get(), aoff, TB.data.get(), boff, C.data.get(), coff, seg, order, ng, km);
}

// ================================================================ BS x Dense
// contract_prefix_outer_bd! (contract_bs_dense.jl:34-204)
static void prefix_outer_bd(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "bs_dense_prefix_outer";
  SB_ASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int axisAr = at(A.map, rlab);
  SB_ASSERT(axisAr <= A.t->P(), "rlab must be in A prefix/head");
  int NB = B.t->N();
  int axisBr = at(B.map, rlab);
  SB_ASSERT(1 <= axisBr && axisBr <= NB, "rlab must be a dim of dense B");
  permute_r_to_last_prefix(pl, A, rlab);
  int PA = A.t->P(), NA = A.t->N(), PC = C.P(), NC = C.N();
  SB_ASSERT(at(A.map, rlab) == PA, "mapA[rlab] == PA");
  const Tensor* Bt = B.t;
  i64 R = Bt->dims[axisBr - 1];
  if (axisBr != NB) {
    std::vector<i64> p;
    for (int d = 1; d <= NB; ++d)
      if (d != axisBr) p.push_back(d);
    p.push_back(axisBr);
    Bt = plan_dense_permute(pl, *Bt, p);
  }
  i64 chunkB = Bt->numel() / R;
  Labels Ad = sub(A.labels, PA, NA), Bwo, Cd = sub(labelsC, PC, NC);
  for (auto& l : B.labels)
    if (l != rlab) Bwo.push_back(l);
  bool AB;
  if (Cd == cat(Ad, Bwo))
    AB = true;
  else if (Cd == cat(Bwo, Ad))
    AB = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense..., B_wo_r...] or [B_wo_r..., A_dense...]");
  const Tensor& TA = *A.t;
  i64 chunkA = TA.blksize;
  SB_ASSERT(C.blksize == chunkA * chunkB, "C.blksize must equal A.blksize * (prod(B dims without r))");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    SB_ASSERT(has(A.map, labelsC[j]), "C prefix label not in A");
    int ax = at(A.map, labelsC[j]);
    SB_ASSERT(ax <= PA && labelsC[j] != rlab, "C prefix must be from A prefix excluding rlab");
    src[j] = ax;
  }
  i64 nA = TA.n;
  // groups in lexicographic (stable) order of the C keys (:121-126): the pair
  // list is A's blocks in that order, so first-creation order = lex order
  Pairs pr;
  pr.n = nA;
  if (nA) {
    pr.a.alloc(ctx, nA);
    k_iota(ctx, pr.a.get(), nA, 0);
    if (PC > 0) {
      std::vector<int> cols(PC);
      std::vector<i64> cd(PC);
      for (int j = 0; j < PC; ++j) {
        cols[j] = src[PC - 1 - j];
        cd[j] = C.dims[PC - 1 - j];
      }
      DBuf<u64> lex(ctx, nA);
      k_cols_lin(ctx, TA.keys.get(), nA, PA, cols.data(), PC, cd.data(), lex.get());
      sort_pairs(ctx, lex.get(), pr.a.get(), nA, bits_for((u64)prod(C.dims, 0, PC)));
    }
    // the staging buffer of the reference holds at most R blocks per group
  }
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, PA, nullptr, 0, src, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * chunkA * chunkB * np;
  pl.gemm_bytes += 8.0 * (double)np * (chunkA + chunkB) + 8.0 * ng * chunkA * chunkB;
  if (!np) return;
  {
    // nb <= R per group (the reference stages at most R blocks, :135-157)
    std::vector<i32> ss(ng + 1);
    HIP_CHECK(hipMemcpyAsync(ss.data(), al.seg_start.get(), (ng + 1) * sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (i64 s = 0; s < ng; ++s) SB_ASSERT(ss[s + 1] - ss[s] <= R, "group size <= R (staging buffer)");
  }
  i64* aoff = pair_offsets(pl, TA, pr.a.get(), np);
  DBuf<i64> bo(ctx, np);
  // B column of pair p: rv = A key's r (last prefix axis)
  k_rv_offsets(ctx, TA.keys.get(), PA, PA, pr.a.get(), nullptr, nullptr, np, chunkB, bo.get());
  i64* boff = pl.keep(std::move(bo));
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  pl.keep(std::move(pr.a));
  const double *Ad_ = TA.data.get(), *Bd_ = Bt->dptr();
  double* Cd_ = C.data.get();
  pl.steps.push_back({"kahan_outer", [=]() { kahan_outer(ctx, AB, chunkA, chunkB, Ad_, aoff, Bd_, boff, Cd_, coff, seg, order, ng); }});
}

// contract_dense_bd! (contract_bs_dense.jl:248-392)
static void dense_bd(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab, unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "bs_dense_dense_r";
  SB_ASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PA = A.t->P(), NA = A.t->N(), NB = B.t->N();
  int axisAr = at(A.map, rlab);
  SB_ASSERT(axisAr > PA, "rlab must be in A dense tail");
  int axisBr = at(B.map, rlab);
  SB_ASSERT(1 <= axisBr && axisBr <= NB, "rlab must be an axis of B");
  LabelMap mapA0 = A.map;
  if (axisAr != NA) {
    auto p = find_perm_for_tensor(*A.t, A.map, rlab);
    A.t = plan_bs_permute(pl, *A.t, p);
    A.labels = permute_labels(A.labels, p);
  }
  SB_ASSERT(A.labels.back() == rlab, "labelsA[end] == rlab");
  const Tensor* Bt = B.t;
  Labels Bl = B.labels;
  if (axisBr != NB) {
    std::vector<i64> p;
    for (int d = 1; d <= NB; ++d)
      if (d != axisBr) p.push_back(d);
    p.push_back(axisBr);
    Bt = plan_dense_permute(pl, *Bt, p);
    Bl = permute_labels(B.labels, p);
  }
  SB_ASSERT(Bl.back() == rlab, "labelsB[end] == rlab");
  const Tensor& TA = *A.t;
  i64 R = TA.dims[NA - 1];
  SB_ASSERT(Bt->dims[NB - 1] == R, "dimsB_dense[end] == R");
  i64 chunkA = TA.blksize / R;
  i64 chunkB = prod(Bt->dims, 0, NB - 1);
  int PC = C.P(), NC = C.N();
  SB_ASSERT(C.N2 == (TA.N2 - 1) + (NB - 1), "N2C == (N2A - 1) + (NB - 1)");
  SB_ASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(Bl, 0, NB - 1), Cd = sub(labelsC, PC, NC);
  bool AB;
  if (Cd == cat(Awo, Bwo))
    AB = true;
  else if (Cd == cat(Bwo, Awo))
    AB = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense..., B_dense...] or [B_dense..., A_dense...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    int ax = at(mapA0, labelsC[j]);
    SB_ASSERT(ax <= PA, "C sparse prefix label must come from A sparse head");
    src[j] = ax;
  }
  i64 nA = TA.n;
  Pairs pr;
  DBuf<unsigned char> kmask;
  if (nA) {
    pl.run_pending();  // existence depends on the permuted payloads
    DBuf<i64> offA = block_offsets(ctx, TA);
    DBuf<unsigned char> fA(ctx, nA * R), fB(ctx, R);
    slice_nonzero(ctx, TA.data.get(), offA.get(), nA, R, chunkA, fA.get());
    slice_nonzero(ctx, Bt->dptr(), nullptr, 1, R, chunkB, fB.get());
    pr.a.alloc(ctx, nA);
    k_iota(ctx, pr.a.get(), nA, 0);
    DBuf<unsigned char> keep(ctx, nA);
    k_rank1_exist(ctx, fA.get(), fB.get(), pr.a.get(), nullptr, nA, R, keep.get(), nullptr);
    pr.n = compact_pairs(ctx, keep.get(), pr.a, pr.b, nA);
    if (pr.n) {
      kmask.alloc(ctx, pr.n * R);
      DBuf<unsigned char> keep2(ctx, pr.n);
      k_rank1_exist(ctx, fA.get(), fB.get(), pr.a.get(), nullptr, pr.n, R, keep2.get(), kmask.get());
    }
  }
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, PA, nullptr, 0, src, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * chunkA * chunkB * R * np;
  pl.gemm_bytes += 8.0 * (double)np * chunkA * R + 8.0 * chunkB * R + 8.0 * ng * chunkA * chunkB;
  if (!np) return;
  i64* aoff = pair_offsets(pl, TA, pr.a.get(), np);
  DBuf<i64> bz(ctx, np);
  k_zero(ctx, bz.get(), np * sizeof(i64));
  i64* boff = pl.keep(std::move(bz));
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  unsigned char* km = pl.keep(std::move(kmask));
  const i32* pa = pl.keep(std::move(pr.a));
  // the contributing r slices follow the values of every replay (the block
  // list stays the first call's)
  {
    const i64* oA = pl.keep(block_offsets(ctx, TA));
    unsigned char* fa = pl.keep(DBuf<unsigned char>(ctx, nA * R));
    unsigned char* fb = pl.keep(DBuf<unsigned char>(ctx, R));
    unsigned char* kp = pl.keep(DBuf<unsigned char>(ctx, np));
    const double* dA = TA.data.get();
    const double* dB = Bt->dptr();
    pl.steps.push_back({"rank1_mask", [=]() {
                          slice_nonzero(ctx, dA, oA, nA, R, chunkA, fa);
                          slice_nonzero(ctx, dB, nullptr, 1, R, chunkB, fb);
                          k_rank1_exist(ctx, fa, fb, pa, nullptr, np, R, kp, km);
                        }});
  }
  GemmDesc d{chunkA, chunkB, R, 1, chunkA, chunkB, 1, AB ? 1 : chunkB, AB ? chunkA : 1};
  add_gemm_step(pl, flags, d, TA.data.get(), aoff, Bt->dptr(), boff, C.data.get(), coff, seg, order, ng, km);
}

// contract_shared! BS x Dense (contract_bs_dense.jl:411-619).  The dense
// operand is permuted to [red..., keepB..., shared_prefix...] so that every
// shared-prefix slice is one contiguous K x N matrix (the reference uses
// [shared_prefix..., red..., keepB...] and strided slices; same values).
static void shared_bs_dense(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared,
                            unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "bs_dense_shared";
  for (auto& l : shared) SB_ASSERT(!contains(labelsC, l), "shared label must be reduced (must not appear in labelsC)");
  int PA = A.t->P(), NA = A.t->N(), NB = B.t->N();
  Labels sp, sd;
  for (auto& l : shared) {
    SB_ASSERT(has(A.map, l), "shared label must exist in A");
    SB_ASSERT(has(B.map, l), "shared label must exist in B");
    (at(A.map, l) <= PA ? sp : sd).push_back(l);
  }
  Labels Ad0 = sub(A.labels, PA, NA), red, keepA0, keepB0;
  for (auto& l : Ad0) (contains(sd, l) ? red : keepA0).push_back(l);
  for (auto& l : B.labels)
    if (!contains(shared, l)) keepB0.push_back(l);
  int PC = C.P(), NC = C.N();
  Labels Cd = sub(labelsC, PC, NC), dA, dB;
  Mode mode = cdense_grouping(Cd, keepA0, keepB0, dA, dB);
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims; would require permuting C (unsupported)");
  SB_ASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA covers keepA");
  SB_ASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB covers keepB");
  std::vector<i64> pA;
  {
    std::vector<i64> spax;
    for (auto& l : sp) spax.push_back(at(A.map, l));
    for (int a = 1; a <= PA; ++a)
      if (std::find(spax.begin(), spax.end(), a) == spax.end()) pA.push_back(a);
    for (auto a : spax) pA.push_back(a);
    std::vector<i64> ka, ra;
    for (auto& l : dA) ka.push_back(at(A.map, l) - PA);
    for (auto& l : red) ra.push_back(at(A.map, l) - PA);
    SB_ASSERT((int)(ka.size() + ra.size()) == A.t->N2, "dense permutation covers the tail");
    for (auto a : ka) pA.push_back(PA + a);
    for (auto a : ra) pA.push_back(PA + a);
  }
  std::vector<i64> pBref;  // the reference's order, checked for its length assert
  for (auto& l : sp) pBref.push_back(at(B.map, l));
  for (auto& l : red) pBref.push_back(at(B.map, l));
  for (auto& l : dB) pBref.push_back(at(B.map, l));
  SB_ASSERT((int)pBref.size() == NB, "B perm length mismatch; labelsB must match B ndims");
  std::vector<i64> pB;
  for (auto& l : red) pB.push_back(at(B.map, l));
  for (auto& l : dB) pB.push_back(at(B.map, l));
  for (auto& l : sp) pB.push_back(at(B.map, l));
  if (!is_identity(pA)) {
    A.t = plan_bs_permute(pl, *A.t, pA);
    A.labels = permute_labels(A.labels, pA);
    A.map = make_map(A.labels);
  }
  const Tensor* Bt = B.t;
  if (!is_identity(pB)) Bt = plan_dense_permute(pl, *Bt, pB);
  int nsp = (int)sp.size(), nred = (int)red.size(), nkA = (int)dA.size();
  std::vector<int> csrc(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    SB_ASSERT(has(A.map, l), "C prefix label must exist in A");
    int ap = at(A.map, l);
    SB_ASSERT(ap <= PA, "C prefix label must come from A prefix");
    SB_ASSERT(!contains(sp, l), "C prefix label cannot be a reduced shared_prefix label");
    csrc[j] = ap;
  }
  const Tensor& TA = *A.t;
  i64 M = nkA == 0 ? 1 : prod(TA.dims, PA, PA + nkA);
  i64 K = nred == 0 ? 1 : prod(TA.dims, PA + nkA, NA);
  i64 N = 1;
  for (auto& l : dB) N *= B.t->dims[at(B.map, l) - 1];
  if (nred > 0) {
    i64 Kb = 1;
    for (auto& l : red) Kb *= B.t->dims[at(B.map, l) - 1];
    SB_ASSERT(K == Kb, "Reduced dense extents mismatch between A and B");
  }
  SB_ASSERT(C.blksize == M * N, "C.blksize must equal M*N");
  i64 nA = TA.n;
  // pairs = A blocks in order; B slice offset = sp_lin * K * N
  Pairs pr;
  pr.n = nA;
  if (nA) {
    pr.a.alloc(ctx, nA);
    k_iota(ctx, pr.a.get(), nA, 0);
  }
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, PA, nullptr, 0, csrc, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * M * K * N * np;
  pl.gemm_bytes += 8.0 * (double)np * M * K + 8.0 * (double)Bt->numel() + 8.0 * ng * M * N;
  pl.a_bytes += 8.0 * TA.n * TA.blksize;
  pl.b_bytes += 8.0 * (double)Bt->numel();
  pl.c_bytes += 8.0 * ng * C.blksize;
  if (!np) return;
  i64* aoff = pair_offsets(pl, TA, pr.a.get(), np);
  DBuf<i64> bo(ctx, np);
  if (nsp > 0) {
    std::vector<int> jA;
    std::vector<i64> spd;
    for (auto& l : sp) {
      jA.push_back(at(A.map, l));
      spd.push_back(TA.dims[at(A.map, l) - 1]);
    }
    DBuf<u64> sl(ctx, np);
    k_cols_lin(ctx, TA.keys.get(), np, PA, jA.data(), nsp, spd.data(), sl.get());
    k_lin_scale(ctx, sl.get(), np, K * N, bo.get());
  } else {
    k_zero(ctx, bo.get(), np * sizeof(i64));
  }
  i64* boff = pl.keep(std::move(bo));
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  pl.keep(std::move(pr.a));
  GemmDesc d{M, N, K, 1, M, 1, K, mode == AthenB ? 1 : N, mode == AthenB ? M : 1};
  add_gemm_step(pl, flags, d, TA.data.get(), aoff, Bt->dptr(), boff, C.data.get(), coff, seg, order, ng, nullptr);
}

// ================================================================ COO kernels
// permutedims + sort! of a COO operand: keys permuted and sorted now, the
// values gathered by a plan step
static const Tensor* plan_coo_permute_sort(Plan& pl, const Tensor& A, const std::vector<i64>& perm) {
  Ctx* ctx = pl.ctx;
  DBuf<i32> idx;
  TP t = coo_permute_sorted(ctx, A, perm, idx);
  i64 n = t->n;
  if (n) {
    const double* in = A.data.get();
    double* out = t->data.get();
    i32* ip = pl.keep(std::move(idx));
    pl.steps.push_back({"coo_gather", [=]() { k_take_f64(ctx, in, ip, n, out); }});
  }
  const Tensor* r = t.get();
  pl.keep_tensor(std::move(t));
  return r;
}

// ensure_rlab_last_and_sorted (contract_coo.jl:48-82)
struct CooOper {
  const Tensor* t;
  Labels labels;
  LabelMap map;
};
static void coo_rlab_last_sorted(Plan& pl, CooOper& X, const Label& rlab) {
  int N = X.t->N();
  if (X.labels.back() == rlab) {
    X.map = make_map(X.labels);
    return;  // fast exit: no permute, no sort
  }
  Labels l2;
  for (auto& l : X.labels)
    if (l != rlab) l2.push_back(l);
  l2.push_back(rlab);
  LabelMap m = make_map(X.labels);
  std::vector<i64> perm(N);
  for (int j = 0; j < N; ++j) perm[j] = at(m, l2[j]);
  X.t = plan_coo_permute_sort(pl, *X.t, perm);
  X.labels = l2;
  X.map = make_map(l2);
}

// B range of each A entry for a join on one column; mode 0: the reference's
// run-wise walk (contract_coo_bs.jl:74-122), 1: element-step walk (contract_coo.jl:139-167)
static void coo_join(Ctx* ctx, int mode, const Tensor& A, int wA, int colA, const Tensor& B, int wB, int colB,
                     DBuf<i32>& lo, DBuf<i32>& hi) {
  i64 nA = A.n, nB = B.n;
  bool sorted = column_sorted(ctx, A.keys.get(), wA, colA, nA) && column_sorted(ctx, B.keys.get(), wB, colB, nB);
  if (sorted) {
    i64 big = 1ll << 31;
    DBuf<u64> ja(ctx, nA), jb(ctx, nB);
    k_cols_lin(ctx, A.keys.get(), nA, wA, &colA, 1, &big, ja.get());
    k_cols_lin(ctx, B.keys.get(), nB, wB, &colB, 1, &big, jb.get());
    k_equal_ranges(ctx, ja.get(), nA, jb.get(), nB, lo.get(), hi.get());
  } else {
    k_walk_ranges(ctx, mode, A.keys.get(), wA, nA, B.keys.get(), wB, nB, &colA, &colB, 1, lo.get(), hi.get());
  }
}

// contract! COO x COO (contract_coo.jl:93-170): every nonzero product is
// pushed, duplicates are not coalesced, C is left dirty
static void contract_coo_coo(Plan& pl, Tensor& C, const Labels& labelsC, CooOper& A, CooOper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "coo_coo";
  coo_rlab_last_sorted(pl, A, rlab);
  coo_rlab_last_sorted(pl, B, rlab);
  int NC = C.N(), NA = A.t->N(), NB = B.t->N();
  std::vector<int> src(NC);
  for (int t = 0; t < NC; ++t) {
    if (has(A.map, labelsC[t]))
      src[t] = at(A.map, labelsC[t]);
    else
      src[t] = -at(B.map, labelsC[t]);
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 nA = TA.n, nB = TB.n;
  C.keys.release();
  C.data.release();
  C.n = 0;
  C.dirty = true;
  if (!nA || !nB) return;
  DBuf<i32> lo(ctx, nA), hi(ctx, nA);
  coo_join(ctx, 1, TA, NA, NA, TB, NB, NB, lo, hi);
  Pairs pr = pairs_from_ranges(ctx, lo, hi, nA);
  if (!pr.n) return;
  pl.run_pending();  // operand values (gathered by the sort steps) decide which products are kept
  DBuf<double> prod_(ctx, pr.n);
  DBuf<unsigned char> keep(ctx, pr.n);
  k_coo_products(ctx, TA.data.get(), TB.data.get(), pr.a.get(), pr.b.get(), pr.n, prod_.get(), keep.get());
  // keep the surviving pairs and their products (stable)
  DBuf<i32> idx(ctx, pr.n), none;
  k_iota(ctx, idx.get(), pr.n, 0);
  i64 m = compact_pairs(ctx, keep.get(), idx, none, pr.n);
  C.n = m;
  pl.npairs = pr.n;
  pl.nblocks_out = m;
  pl.flops += (double)pr.n;
  pl.gemm_bytes += 8.0 * (2.0 * pr.n) + (8.0 + 4.0 * NC) * m;
  if (!m) return;
  DBuf<i32> pa2(ctx, m), pb2(ctx, m);
  k_gather_rows(ctx, pr.a.get(), 1, idx.get(), m, pa2.get());
  k_gather_rows(ctx, pr.b.get(), 1, idx.get(), m, pb2.get());
  C.keys.alloc(ctx, (size_t)m * NC);
  C.data.alloc(ctx, m);
  DBuf<u64> lin(ctx, m);
  std::vector<i64> bigd(NC, 1);
  k_pair_keys(ctx, pa2.get(), pb2.get(), m, TA.keys.get(), NA, TB.keys.get(), NB, src.data(), NC, bigd.data(),
              C.keys.get(), lin.get());
  i32* pa = pl.keep(std::move(pa2));
  i32* pb = pl.keep(std::move(pb2));
  const double *va = TA.data.get(), *vb = TB.data.get();
  double* out = C.data.get();
  DBuf<unsigned char> kk(ctx, m);
  unsigned char* kp = pl.keep(std::move(kk));
  pl.steps.push_back({"coo_products", [=]() { k_coo_products(ctx, va, vb, pa, pb, m, out, kp); }});
}

// contract_prefix! COO x BS (contract_coo_bs.jl:51-124)
static void coo_bs_prefix(Plan& pl, Tensor& C, const Labels& labelsC, CooOper& A, Oper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "coo_bs_prefix";
  if (C.N2 != B.t->N2) fail(SB_E_METHOD, "MethodError: contract_prefix! needs C and B with the same N2");
  int NA = A.t->N(), PB = B.t->P(), PC = C.P();
  SB_ASSERT(PC == NA + PB - 2, "PC == NA + PB - 2");
  int axisAr = at(A.map, rlab), axisBr = at(B.map, rlab);
  SB_ASSERT(axisBr <= PB, "axisBr <= PB");
  const Tensor &TA = *A.t, &TB = *B.t;
  SB_ASSERT(C.blksize == TB.blksize, "length(Cvec) == length(Bvec)");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    if (has(A.map, labelsC[j]))
      src[j] = at(A.map, labelsC[j]);
    else {
      int ax = at(B.map, labelsC[j]);
      SB_ASSERT(ax <= PB, "ax <= PB");
      src[j] = -ax;
    }
  }
  i64 nA = TA.n, nB = TB.n;
  DBuf<i32> lo(ctx, nA ? nA : 1), hi(ctx, nA ? nA : 1);
  if (nA && nB) coo_join(ctx, 0, TA, NA, axisAr, TB, PB, axisBr, lo, hi);
  Pairs pr = (nA && nB) ? pairs_from_ranges(ctx, lo, hi, nA) : Pairs();
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, NA, &TB, PB, src, al);
  i64 np = pr.n, ng = al.nC, L = C.blksize;
  pl.flops += 2.0 * L * np;
  pl.gemm_bytes += 8.0 * (double)np * (L + 1) + 8.0 * ng * L;
  if (!np) return;
  DBuf<double> alpha(ctx, np);
  DBuf<i64> xo(ctx, np);
  i64* xoff = pair_offsets(pl, TB, pr.b.get(), np);
  i32* pa = pl.keep(std::move(pr.a));
  double* al_ = pl.keep(std::move(alpha));
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  pl.keep(std::move(pr.b));
  const double *va = TA.data.get(), *X = TB.data.get();
  double* Cd = C.data.get();
  pl.steps.push_back({"axpy_groups", [=]() {
                        k_take_f64(ctx, va, pa, np, al_);
                        axpy_groups(ctx, L, al_, X, xoff, Cd, coff, seg, order, ng);
                      }});
}

// contract_dense_index! COO x BS (contract_coo_bs.jl:140-241)
static void coo_bs_dense_index(Plan& pl, Tensor& C, const Labels& labelsC, CooOper& A, Oper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "coo_bs_dense_r";
  SB_ASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PB = B.t->P(), NB = B.t->N(), N2 = B.t->N2, NA = A.t->N(), PC = C.P(), NC = C.N();
  SB_ASSERT(C.N2 == N2 - 1, "N2C == N2 - 1");
  int axisBr = at(B.map, rlab);
  SB_ASSERT(axisBr > PB, "rlab must be in B dense tail");
  int redpos = axisBr - PB;
  int axisAr = at(A.map, rlab);
  SB_ASSERT(axisAr == NA, "Expected A to have rlab as last axis already (axisAr == NA)");
  Labels Bd = sub(B.labels, PB, NB), Bdp = Bd;
  LabelMap mapB0 = B.map;
  if (redpos != N2) {
    std::vector<i64> pd, pg;
    for (int d = 1; d <= N2; ++d)
      if (d != redpos) pd.push_back(d);
    pd.push_back(redpos);
    for (int a = 1; a <= PB; ++a) pg.push_back(a);
    for (auto d : pd) pg.push_back(PB + d);
    B.t = plan_bs_permute(pl, *B.t, pg);
    for (int i = 0; i < N2; ++i) Bdp[i] = Bd[pd[i] - 1];
  }
  SB_ASSERT(Bdp.back() == rlab, "labelsB_dense_p[end] == rlab");
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 dim_r = TB.dims[NB - 1];
  i64 chunk = prod(TB.dims, PB, NB - 1);
  SB_ASSERT(sub(labelsC, PC, NC) == sub(Bdp, 0, N2 - 1), "collect(labelsC[PC+1:NC]) == expectedC_dense");
  for (int i = 0; i < C.N2; ++i) SB_ASSERT(C.dims[PC + i] == TB.dims[PB + i], "C dense dims == B dense dims");
  SB_ASSERT(C.blksize == chunk, "C.blksize == chunklen");
  SB_ASSERT(TB.blksize == chunk * dim_r, "Bp.blksize == chunklen * dim_r");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    if (has(A.map, labelsC[j]))
      src[j] = at(A.map, labelsC[j]);
    else {
      int ax = at(mapB0, labelsC[j]);
      SB_ASSERT(ax <= PB, "Label in C prefix must come from B prefix (not dense)");
      src[j] = -ax;
    }
  }
  i64 nA = TA.n, nB = TB.n;
  if (nA) {
    std::vector<i32> hk((size_t)nA * NA);
    HIP_CHECK(hipMemcpyAsync(hk.data(), TA.keys.get(), hk.size() * sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (i64 i = 0; i < nA; ++i) {
      i64 rv = hk[i * NA + NA - 1];
      SB_ASSERT(1 <= rv && rv <= dim_r, "Reduction index rv out of bounds for dim_r");
    }
  }
  Pairs pr;
  DBuf<i64> blkoff = block_offsets(ctx, TB);
  if (nA && nB) {
    pl.run_pending();  // existence depends on the (permuted) payload of B
    i64 np0 = nA * nB;
    if (np0 >= (1ll << 31)) fail(SB_E_LIMIT, "more than 2^31 entry-block pairs in one contraction");
    pr.a.alloc(ctx, np0);
    pr.b.alloc(ctx, np0);
    k_cross_pairs(ctx, nA, nB, pr.a.get(), pr.b.get());
    DBuf<unsigned char> keep(ctx, np0);
    k_scaled_slice_nonzero(ctx, TA.data.get(), TA.keys.get(), NA, TB.data.get(), blkoff.get(), pr.a.get(),
                           pr.b.get(), np0, chunk, keep.get());
    pr.n = compact_pairs(ctx, keep.get(), pr.a, pr.b, np0);
  }
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, NA, &TB, PB, src, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * chunk * np;
  pl.gemm_bytes += 8.0 * (double)np * (chunk + 1) + 8.0 * ng * chunk;
  if (!np) return;
  DBuf<i64> xo(ctx, np);
  k_rv_offsets(ctx, TA.keys.get(), NA, NA, pr.a.get(), blkoff.get(), pr.b.get(), np, chunk, xo.get());
  i64* xoff = pl.keep(std::move(xo));
  DBuf<double> alpha(ctx, np);
  double* al_ = pl.keep(std::move(alpha));
  i32* pa = pl.keep(std::move(pr.a));
  pl.keep(std::move(pr.b));
  pl.keep(std::move(blkoff));
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  const double *va = TA.data.get(), *X = TB.data.get();
  double* Cd = C.data.get();
  pl.steps.push_back({"axpy_groups", [=]() {
                        k_take_f64(ctx, va, pa, np, al_);
                        axpy_groups(ctx, chunk, al_, X, xoff, Cd, coff, seg, order, ng);
                      }});
}

// contract! COO x Dense (contract_coo_dense.jl:48-125)
static void contract_coo_dense(Plan& pl, Tensor& C, const Labels& labelsC, CooOper& A, Oper& B, const Label& rlab,
                               unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "coo_dense";
  int PC = C.P(), NC = C.N(), NA = A.t->N(), NB = B.t->N();
  {  // aligned_A_to_Cprefix (contract_coo_dense.jl:4-18)
    Labels desired = sub(labelsC, 0, PC);
    desired.push_back(rlab);
    LabelMap m = make_map(A.labels);
    SB_ASSERT((int)desired.size() == NA, "length(desired) == NA");
    std::vector<i64> perm(NA);
    for (int j = 0; j < NA; ++j) perm[j] = at(m, desired[j]);
    A.t = plan_coo_permute_sort(pl, *A.t, perm);
    A.labels = desired;
    A.map = make_map(desired);
  }
  SB_ASSERT(A.labels.back() == rlab, "labelsA[end] == rlab");
  Labels dnr;
  for (int j = PC; j < NC; ++j)
    if (labelsC[j] != rlab) dnr.push_back(labelsC[j]);
  int rpos = -1;
  for (int i = 0; i < NB; ++i)
    if (B.labels[i] == rlab) {
      rpos = i + 1;
      break;
    }
  SB_ASSERT(rpos > 0, "rlab not found in labelsB");
  std::vector<i64> perm(NB, 0);
  for (size_t k = 0; k < dnr.size(); ++k) {
    int pos = -1;
    for (int i = 0; i < NB; ++i)
      if (B.labels[i] == dnr[k]) {
        pos = i + 1;
        break;
      }
    SB_ASSERT(pos > 0, "label not found in labelsB");
    if ((int)k >= NB) fail(SB_E_BOUNDS, "BoundsError: perm");
    perm[k] = pos;
  }
  perm[NB - 1] = rpos;
  const Tensor* Bt = B.t;
  Labels Bl = B.labels;
  if (!is_identity(perm)) {
    for (auto p : perm)
      if (p < 1) fail(SB_E_ARGUMENT, "invalid permutation of B");
    std::vector<i64> s(perm);
    std::sort(s.begin(), s.end());
    for (int i = 0; i < NB; ++i)
      if (s[i] != i + 1) fail(SB_E_ARGUMENT, "invalid permutation of B");
    Bt = plan_dense_permute(pl, *Bt, perm);
    Bl = permute_labels(B.labels, perm);
  }
  SB_ASSERT(Bl.back() == rlab, "labelsB[end] == rlab");
  SB_ASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  Labels ep = sub(A.labels, 0, NA - 1), ed = sub(Bl, 0, NB - 1);
  SB_ASSERT(PC == (int)ep.size(), "PC mismatch");
  SB_ASSERT(C.N2 == (int)ed.size(), "N2 mismatch");
  SB_ASSERT(sub(labelsC, 0, PC) == ep, "C prefix labels must be A labels without r");
  SB_ASSERT(sub(labelsC, PC, NC) == ed, "C dense labels must be B labels without r");
  i64 R = Bt->dims[NB - 1];
  i64 chunk = prod(Bt->dims, 0, NB - 1);
  SB_ASSERT(C.blksize == chunk, "C.blksize must equal prod(B dims without r)");
  const Tensor& TA = *A.t;
  i64 nA = TA.n;
  Pairs pr;
  if (nA) {
    pr.a.alloc(ctx, nA);
    k_iota(ctx, pr.a.get(), nA, 0);
    DBuf<unsigned char> keep(ctx, nA);
    k_rv_inrange(ctx, TA.keys.get(), NA, NA, nA, R, keep.get());
    if (flags & SB_FLAG_SKIP_ZERO_SLICES) {
      pl.run_pending();
      k_coo_dense_nonzero(ctx, TA.data.get(), TA.keys.get(), NA, NA, nA, Bt->dptr(), chunk, keep.get());
    }
    pr.n = compact_pairs(ctx, keep.get(), pr.a, pr.b, nA);
  }
  std::vector<int> src(PC);
  std::iota(src.begin(), src.end(), 1);
  Alloc al;
  allocate_from_pairs(pl, C, pr, &TA, NA, nullptr, 0, src, al);
  i64 np = pr.n, ng = al.nC;
  pl.flops += 2.0 * chunk * np;
  pl.gemm_bytes += 8.0 * (double)np * (chunk + 1) + 8.0 * ng * chunk;
  if (!np) return;
  DBuf<i64> xo(ctx, np);
  k_rv_offsets(ctx, TA.keys.get(), NA, NA, pr.a.get(), nullptr, nullptr, np, chunk, xo.get());
  i64* xoff = pl.keep(std::move(xo));
  DBuf<double> alpha(ctx, np);
  double* al_ = pl.keep(std::move(alpha));
  i32* pa = pl.keep(std::move(pr.a));
  i64* coff = group_offsets(pl, C);
  i32* seg = pl.keep(std::move(al.seg_start));
  i32* order = pl.keep(std::move(al.order));
  const double *va = TA.data.get(), *X = Bt->dptr();
  double* Cd = C.data.get();
  pl.steps.push_back({"axpy_groups", [=]() {
                        k_take_f64(ctx, va, pa, np, al_);
                        axpy_groups(ctx, chunk, al_, X, xoff, Cd, coff, seg, order, ng);
                      }});
}

// ================================================================ AliasedBlockSparse
// permutedims(::AliasedBlockSparse) (aliased/storage.jl:172-239) split into
// its symbolic part (keys remapped and stably sorted, alias ids gathered; run
// now) and its numeric part (templates transposed, scalars gathered; plan steps)
static const Tensor* plan_al_permute(Plan& pl, const Tensor& A, const std::vector<i64>& perm) {
  Ctx* ctx = pl.ctx;
  int N = A.N(), P = A.P(), N2 = A.N2;
  check_no_cross_perm(perm, N, P);
  std::vector<i64> dimsB(N);
  for (int j = 0; j < N; ++j) dimsB[j] = A.dims[perm[j] - 1];
  TP B = al_new(ctx, dimsB, N2);
  SB_ASSERT(B->blksize == A.blksize, "permutation within tail preserves block size");
  i64 n = A.n, nt = A.ntemplates;
  B->n = n;
  B->ntemplates = nt;
  if (n > 0) {
    B->keys.alloc(ctx, (size_t)n * (P ? P : 1));
    B->ids.alloc(ctx, n);
    B->scalars.alloc(ctx, n);
    DBuf<i32> idx(ctx, n);
    if (P > 0) {
      std::vector<int> sel(P);
      for (int j = 0; j < P; ++j) sel[j] = (int)perm[j];
      DBuf<i32> nk(ctx, (size_t)n * P);
      DBuf<u64> lin(ctx, n);
      k_remap_keys(ctx, A.keys.get(), n, P, sel.data(), P, dimsB.data(), nk.get(), lin.get(), idx.get());
      sort_pairs(ctx, lin.get(), idx.get(), n, bits_for((u64)prod(dimsB, 0, P) - 1));
      k_gather_rows(ctx, nk.get(), P, idx.get(), n, B->keys.get());
    } else {
      k_iota(ctx, idx.get(), n, 0);
    }
    k_gather_rows(ctx, A.ids.get(), 1, idx.get(), n, B->ids.get());
    const double* sa = A.scalars.get();
    double* sb = B->scalars.get();
    i32* ip = pl.keep(std::move(idx));
    pl.steps.push_back({"al_scalars_gather", [=]() { k_take_f64(ctx, sa, ip, n, sb); }});
  }
  if (nt > 0) {
    B->data.alloc(ctx, (size_t)nt * B->blksize);
    std::vector<i64> tdims(A.dims.begin() + P, A.dims.end());
    std::vector<int> pblk(N2);
    for (int i = 0; i < N2; ++i) pblk[i] = (int)(perm[P + i] - P - 1);
    const double* in = A.data.get();
    double* out = B->data.get();
    pl.steps.push_back({"permute_blocks", [ctx, in, nt, N2, tdims, pblk, out]() {
                          permute_blocks(ctx, in, nullptr, nt, N2, tdims.data(), pblk.data(), out);
                        }});
    pl.perm_bytes += 2.0 * nt * B->blksize * sizeof(double);
  }
  const Tensor* r = B.get();
  pl.keep_tensor(std::move(B));
  return r;
}
TP al_permute(Ctx* ctx, const Tensor& A, const std::vector<i64>& perm) {
  Plan pl;
  pl.ctx = ctx;
  const Tensor* t = plan_al_permute(pl, A, perm);
  pl.run_pending();
  for (auto& h : pl.hold)  // hand the permuted tensor over; the plan keeps only scratch
    if (h.get() == (const void*)t) {
      TP out(new Tensor(std::move(*const_cast<Tensor*>(t))));
      return out;
    }
  fail(SB_E_ASSERT, "aliased permute lost its result");
}
static void al_oper_perm(Plan& pl, Oper& X, const std::vector<i64>& perm) {
  X.t = plan_al_permute(pl, *X.t, perm);
  X.labels = permute_labels(X.labels, perm);
  X.map = make_map(X.labels);
}
// _aliased_r_to_last_prefix / _aliased_r_to_last_dense (contract_aliased.jl:33-65)
static void al_r_to_last_prefix(Plan& pl, Oper& X, const Label& rlab) {
  int axisr = at(X.map, rlab), P = X.t->P(), N = X.t->N();
  SB_ASSERT(axisr <= P, "rlab must be in sparse prefix");
  if (axisr == P) return;
  std::vector<i64> perm;
  for (int a = 1; a <= P; ++a)
    if (a != axisr) perm.push_back(a);
  perm.push_back(axisr);
  for (int a = P + 1; a <= N; ++a) perm.push_back(a);
  al_oper_perm(pl, X, perm);
}
static void al_r_to_last_dense(Plan& pl, Oper& X, const Label& rlab) {
  int axisr = at(X.map, rlab), P = X.t->P(), N = X.t->N(), N2 = X.t->N2;
  SB_ASSERT(axisr > P, "rlab must be in dense tail");
  if (axisr == N) return;
  std::vector<i64> perm;
  for (int a = 1; a <= P; ++a) perm.push_back(a);
  for (int d = 1; d <= N2; ++d)
    if (d != axisr - P) perm.push_back(P + d);
  perm.push_back(axisr);
  al_oper_perm(pl, X, perm);
}
// dense operand with rlab moved last (contract_aliased.jl:168-177)
static void dense_r_last(Plan& pl, Oper& B, const Label& rlab) {
  int NB = B.t->N(), ax = at(B.map, rlab);
  if (ax == NB) return;
  std::vector<i64> perm;
  for (int d = 1; d <= NB; ++d)
    if (d != ax) perm.push_back(d);
  perm.push_back(ax);
  B.t = plan_dense_permute(pl, *B.t, perm);
  B.labels = permute_labels(B.labels, perm);
  B.map = make_map(B.labels);
}

// The contributions of an aliased kernel in the reference's traversal order
// and what the output needs from them.
struct AlContrib {
  i64 np = 0;
  DBuf<i32> pa, pb;       // operand rows (pb empty when B is dense)
  DBuf<i32> ck;           // np x PC output keys
  DBuf<i32> xy;           // np x 2 combination keys (1-based), over (nx, ny)
  i64 nx = 1, ny = 1;
  const double* sA = nullptr;  // per A row scalar (COO: entry value)
  const double* sB = nullptr;  // per B row scalar, or null
};
// computes template t (0-based, first-seen order) from combination (ux[t], uy[t])
using TemplateFn = std::function<void(Ctx*, i64 nu, const i32* ux_uy, double* out)>;

// Output of an aliased kernel from its contributions (contract_aliased.jl:75-146):
// combined templates deduplicated in first-seen order (their values by tfn),
// output keys grouped and sorted, each key aliased (one combined template,
// scalars summed in order) or accumulated into a template of its own (scalar
// 1), accumulated templates appended in key order.  COO mode
// (contract_aliased_coo_dense.jl:96-145): templates are the B slices of the
// keys' first entries (first-seen order of rv), any second entry demotes;
// coo_B: the B slices (coo_R of them), coo_keys / coo_w: the aligned COO keys.
static void al_finish(Plan& pl, Tensor& C, AlContrib& c, const TemplateFn& tfn, const double* coo_B, i64 coo_R,
                      const i32* coo_keys, int coo_w) {
  Ctx* ctx = pl.ctx;
  int PC = C.P();
  i64 bs = C.blksize, np = c.np;
  bool coo = coo_B != nullptr;
  C.n = 0;
  C.ntemplates = 0;
  C.dirty = false;
  if (np == 0) return;
  // combined templates
  i64 ncomb = 0;
  DBuf<i32> ctid, byid;
  if (!coo) {
    DBuf<u64> lin(ctx, np);
    k_al_xy_lin(ctx, c.xy.get(), np, c.nx, lin.get());
    Alloc U;
    DBuf<i32> xy2(ctx, (size_t)np * 2);
    HIP_CHECK(hipMemcpyAsync(xy2.get(), c.xy.get(), np * 2 * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
    allocate_blocks(ctx, np, 2, xy2, lin, (u64)(c.nx * c.ny - 1), U);
    ncomb = U.nC;
    ctid.alloc(ctx, np);
    byid.alloc(ctx, (size_t)ncomb * 2);
    k_al_seg_ids(ctx, U.nC, U.seg_start.get(), U.order.get(), U.ids.get(), U.keys.get(), 2, ctid.get(), byid.get());
  }
  // output keys
  Alloc G;
  DBuf<u64> clin(ctx, np);
  if (PC > 0) {
    DBuf<i32> ck2(ctx, (size_t)np * PC);
    std::vector<int> sel(PC);
    std::iota(sel.begin(), sel.end(), 1);
    DBuf<i32> dummy(ctx, np);
    k_remap_keys(ctx, c.ck.get(), np, PC, sel.data(), PC, C.dims.data(), ck2.get(), clin.get(), dummy.get());
    allocate_blocks(ctx, np, PC, ck2, clin, (u64)prod(C.dims, 0, PC) - 1, G);
  } else {
    k_zero(ctx, clin.get(), np * sizeof(u64));
    DBuf<i32> ck2(ctx, 1);
    allocate_blocks(ctx, np, 0, ck2, clin, 0, G);
  }
  i64 ng = G.nC;
  DBuf<i32> is_acc(ctx, ng), fdiff(ctx, ng), aid(ctx, ng);
  k_al_groups(ctx, ng, G.seg_start.get(), G.order.get(), coo ? nullptr : ctid.get(), coo ? 1 : 0, is_acc.get(),
              fdiff.get(), aid.get());
  i64 nbase = ncomb;
  DBuf<i32> rvtid;  // COO: per template, its rv
  if (coo) {
    // first entry of every key, in traversal order; templates by first-seen rv
    DBuf<i32> fpos(ctx, ng), gidx(ctx, ng);
    k_gather_rows(ctx, G.order.get(), 1, G.seg_start.get(), ng, fpos.get());  // order[seg[g]]
    DBuf<unsigned> fkey(ctx, ng);
    k_i32_to_u32(ctx, fpos.get(), ng, fkey.get());
    k_iota(ctx, gidx.get(), ng, 0);
    sort_pairs32(ctx, fkey.get(), gidx.get(), ng, bits_for((u64)np));
    DBuf<i32> fsorted(ctx, ng), ent(ctx, ng), rvk(ctx, ng);
    k_gather_rows(ctx, fpos.get(), 1, gidx.get(), ng, fsorted.get());        // traversal index, increasing
    k_gather_rows(ctx, c.pa.get(), 1, fsorted.get(), ng, ent.get());         // entry rows
    k_take_col(ctx, coo_keys, coo_w, coo_w - 1, ent.get(), ng, rvk.get());  // rv of each
    DBuf<u64> rlin(ctx, ng);
    k_rv_lin(ctx, rvk.get(), ng, rlin.get());
    DBuf<i32> rk2(ctx, ng);
    HIP_CHECK(hipMemcpyAsync(rk2.get(), rvk.get(), ng * sizeof(i32), hipMemcpyDeviceToDevice, ctx->stream));
    Alloc U;
    allocate_blocks(ctx, ng, 1, rk2, rlin, (u64)std::max<i64>(1, coo_R - 1), U);
    nbase = U.nC;
    DBuf<i32> tsorted(ctx, ng);
    rvtid.alloc(ctx, U.nC);
    k_al_seg_ids(ctx, U.nC, U.seg_start.get(), U.order.get(), U.ids.get(), U.keys.get(), 1, tsorted.get(),
                 rvtid.get());
    k_scatter_i32(ctx, gidx.get(), tsorted.get(), ng, aid.get());  // aid[g] = template of g's first entry
  }
  DBuf<i64> acc64(ctx, ng), rank(ctx, ng);
  k_i32_to_i64(ctx, is_acc.get(), ng, acc64.get());
  i64 nacc = exclusive_scan_total(ctx, acc64.get(), rank.get(), ng);
  // install the structure in C
  C.n = ng;
  C.keys = std::move(G.keys);
  C.ids.alloc(ctx, ng);
  k_al_final_ids(ctx, ng, is_acc.get(), rank.get(), aid.get(), nbase, C.ids.get());
  C.scalars.alloc(ctx, ng);
  C.ntemplates = nbase + nacc;
  C.data.alloc(ctx, (size_t)C.ntemplates * bs);
  pl.nblocks_out = ng;
  pl.npairs = np;
  // numeric steps
  i32* pa = pl.keep(std::move(c.pa));
  i32* pb = c.pb.n ? pl.keep(std::move(c.pb)) : nullptr;
  double* alpha = pl.keep(DBuf<double>(ctx, np));
  const double *sA = c.sA, *sB = c.sB;
  i32* seg = pl.keep(std::move(G.seg_start));
  i32* order = pl.keep(std::move(G.order));
  i32* isacc = pl.keep(std::move(is_acc));
  i32* fd = pl.keep(std::move(fdiff));
  double* tm = C.data.get();
  double* sc = C.scalars.get();
  pl.steps.push_back({"al_alpha", [=]() { k_al_alpha(ctx, pa, pb, np, sA, sB, alpha); }});
  i64* toff = pl.keep(DBuf<i64>(ctx, np));
  const double* base;
  if (!coo) {
    i32* ub = pl.keep(std::move(byid));
    k_al_ctid_off(ctx, ctid.get(), np, bs, toff);
    if (ncomb) pl.steps.push_back({"al_templates", [=]() { tfn(ctx, ncomb, ub, tm); }});
    base = tm;
  } else {
    // toff[p] = (rv_p - 1) * bs into B; templates = the first-seen B slices
    k_rv_offsets(ctx, coo_keys, coo_w, coo_w, pa, nullptr, nullptr, np, bs, toff);
    i32* rv = pl.keep(std::move(rvtid));
    i64 nt = nbase;
    i64* so = pl.keep(DBuf<i64>(ctx, nt));
    k_al_offsets(ctx, rv, 1, 0, nt, bs, 0, so);
    pl.steps.push_back({"al_templates", [=]() { gather_blocks(ctx, coo_B, so, nt, bs, tm); }});
    base = coo_B;
  }
  pl.steps.push_back({"al_scalars", [=]() { k_al_scalars(ctx, ng, seg, order, alpha, isacc, sc); }});
  if (nacc) {
    i32* list = pl.keep(DBuf<i32>(ctx, nacc));
    k_al_acc_list(ctx, ng, isacc, rank.get(), list);
    double* out = tm + nbase * bs;
    pl.steps.push_back({"al_accumulate", [=]() {
                          k_al_accumulate(ctx, nacc, bs, list, seg, order, fd, alpha, base, toff, out);
                        }});
  }
  pl.keep(std::move(rank));
}

// template builders: outer products (exact) and block products
static TemplateFn outer_templates(const double* A, i64 blkA, const double* B, i64 blkB, bool ab) {
  return [=](Ctx* ctx, i64 nu, const i32* u, double* out) {
    DBuf<i64> ao(ctx, nu), bo(ctx, nu), co(ctx, nu);
    DBuf<i32> seq(ctx, nu + 1);
    k_al_offsets(ctx, u, 2, 0, nu, blkA, 0, ao.get());
    k_al_offsets(ctx, u, 2, 1, nu, blkB, 0, bo.get());
    k_block_offsets(ctx, nullptr, nu, blkA * blkB, co.get());
    k_iota(ctx, seq.get(), nu + 1, 0);
    if (ab)  // t[i + j*blkA] = a[i] * b[j]
      outer_groups(ctx, blkA, blkB, A, ao.get(), B, bo.get(), out, co.get(), seq.get(), seq.get(), nu);
    else  // t[i + j*blkB] = b[i] * a[j]
      outer_groups(ctx, blkB, blkA, B, bo.get(), A, ao.get(), out, co.get(), seq.get(), seq.get(), nu);
  };
}
// out_t = A_x (M x K, sam/sak) * B_y (K x N, boff = (y-1)*bscale + badd, sbk/sbn), C strides scm/scn
static TemplateFn gemm_templates(unsigned prec, const double* A, i64 blkA, const double* B, i64 bscale, i64 badd,
                                 const GemmDesc& d, i64 bsC) {
  return [=](Ctx* ctx, i64 nu, const i32* u, double* out) {
    DBuf<i64> ao(ctx, nu), bo(ctx, nu), co(ctx, nu);
    DBuf<i32> seq(ctx, nu + 1);
    k_al_offsets(ctx, u, 2, 0, nu, blkA, 0, ao.get());
    k_al_offsets(ctx, u, 2, 1, nu, bscale, badd, bo.get());
    k_block_offsets(ctx, nullptr, nu, bsC, co.get());
    k_iota(ctx, seq.get(), nu + 1, 0);
    grouped_gemm(ctx, prec, d, A, ao.get(), B, bo.get(), out, co.get(), seq.get(), seq.get(), nu, nullptr);
  };
}

// _contract_aliased_prefix_outer_ad! (contract_aliased.jl:151-233)
static void al_prefix_outer_ad(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "aliased_dense_prefix";
  int PA = A.t->P(), NA = A.t->N(), PC = C.P(), NC = C.N();
  SB_ASSERT(at(A.map, rlab) == PA, "rlab must be the last prefix axis of A");
  SB_ASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  i64 R = B.t->dims[at(B.map, rlab) - 1];
  dense_r_last(pl, B, rlab);
  int NB = B.t->N();
  i64 chunkB = prod(B.t->dims, 0, NB - 1);
  SB_ASSERT(C.blksize == A.t->blksize * chunkB, "C.blksize must equal A.blksize * chunkB");
  Labels Ad = sub(A.labels, PA, NA), Bwo = sub(B.labels, 0, NB - 1), Cd = sub(labelsC, PC, NC);
  bool ab;
  if (Cd == cat(Ad, Bwo))
    ab = true;
  else if (Cd == cat(Bwo, Ad))
    ab = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense...,B_wo_r...] or [B_wo_r...,A_dense...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    int ax = at(A.map, labelsC[j]);
    SB_ASSERT(ax <= PA && labelsC[j] != rlab, "C prefix label must come from A sparse prefix excluding rlab");
    src[j] = ax;
  }
  const Tensor& TA = *A.t;
  AlContrib c;
  i64 nA = TA.n;
  if (nA) {
    c.pa.alloc(ctx, nA);
    k_iota(ctx, c.pa.get(), nA, 0);
    DBuf<unsigned char> keep(ctx, nA);
    k_rv_inrange(ctx, TA.keys.get(), PA, PA, nA, R, keep.get());
    DBuf<i32> none;
    c.np = compact_pairs(ctx, keep.get(), c.pa, none, nA);
  }
  if (c.np) {
    c.ck.alloc(ctx, (size_t)c.np * std::max(1, PC));
    DBuf<u64> lin(ctx, c.np);
    if (PC) k_pair_keys(ctx, c.pa.get(), c.pa.get(), c.np, TA.keys.get(), PA, nullptr, 0, src.data(), PC,
                        C.dims.data(), c.ck.get(), lin.get());
    c.xy.alloc(ctx, (size_t)c.np * 2);
    k_al_xy(ctx, 0, c.pa.get(), nullptr, c.np, TA.ids.get(), TA.keys.get(), PA, PA - 1, nullptr, nullptr, nullptr,
            0, c.xy.get());
  }
  c.nx = std::max<i64>(1, TA.ntemplates);
  c.ny = std::max<i64>(1, R);
  c.sA = TA.scalars.get();
  al_finish(pl, C, c, outer_templates(TA.data.get(), TA.blksize, B.t->dptr(), chunkB, ab), nullptr, 0, nullptr, 0);
}

// _contract_aliased_dense_ad! (contract_aliased.jl:238-365): one block product
// per A template; keys, alias ids and scalars carried over
static void al_dense_ad(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab,
                        unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "aliased_dense_tail";
  int PA = A.t->P(), NA = A.t->N(), PC = C.P(), NC = C.N();
  SB_ASSERT(at(A.map, rlab) == NA, "rlab must be the last axis of A");
  SB_ASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  SB_ASSERT(PC == PA, "C sparse prefix must equal A sparse prefix for dense-tail reduction");
  dense_r_last(pl, B, rlab);
  int NB = B.t->N();
  i64 R = B.t->dims[NB - 1];
  i64 chunkB = prod(B.t->dims, 0, NB - 1);
  SB_ASSERT(A.t->dims[NA - 1] == R, "Reduction size R mismatch between A dense tail and B");
  i64 chunkA = A.t->blksize / R;
  SB_ASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(B.labels, 0, NB - 1), Cd = sub(labelsC, PC, NC);
  bool ab;
  if (Cd == cat(Awo, Bwo))
    ab = true;
  else if (Cd == cat(Bwo, Awo))
    ab = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense_wo_r...,B_wo_r...] or [B_wo_r...,A_dense_wo_r...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    int ax = at(A.map, labelsC[j]);
    SB_ASSERT(ax <= PA, "C prefix label must come from A sparse prefix");
    src[j] = ax;
  }
  const Tensor& TA = *A.t;
  i64 n = TA.n, nt = TA.ntemplates;
  C.n = n;
  C.ntemplates = nt;
  C.dirty = false;
  C.data.alloc(ctx, (size_t)nt * C.blksize);
  C.keys.alloc(ctx, (size_t)n * std::max(1, PC));
  C.ids.alloc(ctx, n);
  C.scalars.alloc(ctx, n);
  pl.nblocks_out = n;
  pl.npairs = nt;
  if (n) {
    DBuf<i32> nk(ctx, (size_t)n * std::max(1, PC)), idx(ctx, n);
    DBuf<u64> lin(ctx, n);
    if (PC) {
      k_remap_keys(ctx, TA.keys.get(), n, PA, src.data(), PC, C.dims.data(), nk.get(), lin.get(), idx.get());
      sort_pairs(ctx, lin.get(), idx.get(), n, bits_for((u64)prod(C.dims, 0, PC) - 1));
      k_gather_rows(ctx, nk.get(), PC, idx.get(), n, C.keys.get());
    } else {
      k_iota(ctx, idx.get(), n, 0);
    }
    k_gather_rows(ctx, TA.ids.get(), 1, idx.get(), n, C.ids.get());
    const double* sa = TA.scalars.get();
    double* sc = C.scalars.get();
    i32* ip = pl.keep(std::move(idx));
    pl.steps.push_back({"al_scalars_gather", [=]() { k_take_f64(ctx, sa, ip, n, sc); }});
  }
  if (nt) {
    // template t: (chunkA x R) * (chunkB x R)^T, AB -> (chunkA x chunkB), BA -> (chunkB x chunkA)
    GemmDesc d{chunkA, chunkB, R, 1, chunkA, chunkB, 1, ab ? 1 : chunkB, ab ? chunkA : 1};
    DBuf<i64> ao(ctx, nt), bo(ctx, nt), co(ctx, nt);
    DBuf<i32> seq(ctx, nt + 1);
    k_block_offsets(ctx, nullptr, nt, TA.blksize, ao.get());
    k_zero(ctx, bo.get(), nt * sizeof(i64));
    k_block_offsets(ctx, nullptr, nt, C.blksize, co.get());
    k_iota(ctx, seq.get(), nt + 1, 0);
    i64 *aop = pl.keep(std::move(ao)), *bop = pl.keep(std::move(bo)), *cop = pl.keep(std::move(co));
    i32* sq = pl.keep(std::move(seq));
    add_gemm_step(pl, flags, d, TA.data.get(), aop, B.t->dptr(), bop, C.data.get(), cop, sq, sq, nt, nullptr);
    pl.flops += 2.0 * chunkA * chunkB * R * nt;
  }
}

// _contract_aliased_prefix_outer_aa! (contract_aliased.jl:415-507)
static void al_prefix_outer_aa(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "aliased_aliased_prefix";
  int PA = A.t->P(), NA = A.t->N(), PB = B.t->P(), NB = B.t->N(), PC = C.P(), NC = C.N();
  SB_ASSERT(at(A.map, rlab) == PA && at(B.map, rlab) == PB, "rlab is the last prefix axis of A and B");
  SB_ASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  SB_ASSERT(C.blksize == A.t->blksize * B.t->blksize, "C.blksize == A.blksize * B.blksize");
  Labels Ad = sub(A.labels, PA, NA), Bd = sub(B.labels, PB, NB), Cd = sub(labelsC, PC, NC);
  bool ab;
  if (Cd == cat(Ad, Bd))
    ab = true;
  else if (Cd == cat(Bd, Ad))
    ab = false;
  else
    fail(SB_E_ERROR, "C dense must be [A_dense...,B_dense...] or [B_dense...,A_dense...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    if (has(A.map, l) && at(A.map, l) != at(A.map, rlab)) {
      int ax = at(A.map, l);
      SB_ASSERT(ax <= PA, "C prefix label from A prefix");
      src[j] = ax;
    } else {
      SB_ASSERT(has(B.map, l), "C prefix label not found in A or B");
      int ax = at(B.map, l);
      SB_ASSERT(ax <= PB && l != rlab, "C prefix label from B prefix");
      src[j] = -ax;
    }
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 nA = TA.n, nB = TB.n;
  AlContrib c;
  if (nA && nB) {
    DBuf<i32> lo(ctx, nA), hi(ctx, nA);
    int ca = PA, cb = PB;
    k_walk_ranges(ctx, 0, TA.keys.get(), PA, nA, TB.keys.get(), PB, nB, &ca, &cb, 1, lo.get(), hi.get());
    Pairs pr = pairs_from_ranges(ctx, lo, hi, nA);
    c.np = pr.n;
    c.pa = std::move(pr.a);
    c.pb = std::move(pr.b);
  }
  if (c.np) {
    c.ck.alloc(ctx, (size_t)c.np * std::max(1, PC));
    DBuf<u64> lin(ctx, c.np);
    if (PC) k_pair_keys(ctx, c.pa.get(), c.pb.get(), c.np, TA.keys.get(), PA, TB.keys.get(), PB, src.data(), PC,
                        C.dims.data(), c.ck.get(), lin.get());
    c.xy.alloc(ctx, (size_t)c.np * 2);
    k_al_xy(ctx, 1, c.pa.get(), c.pb.get(), c.np, TA.ids.get(), TA.keys.get(), PA, 0, TB.ids.get(), nullptr, nullptr,
            0, c.xy.get());
  }
  c.nx = std::max<i64>(1, TA.ntemplates);
  c.ny = std::max<i64>(1, TB.ntemplates);
  c.sA = TA.scalars.get();
  c.sB = TB.scalars.get();
  al_finish(pl, C, c, outer_templates(TA.data.get(), TA.blksize, TB.data.get(), TB.blksize, ab), nullptr, 0, nullptr,
            0);
}

// _contract_aliased_dense_aa! (contract_aliased.jl:512-643): every (A, B) block pair
static void al_dense_aa(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab,
                        unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "aliased_aliased_tail";
  int PA = A.t->P(), NA = A.t->N(), PB = B.t->P(), NB = B.t->N(), PC = C.P(), NC = C.N();
  SB_ASSERT(at(A.map, rlab) == NA && at(B.map, rlab) == NB, "rlab is the last axis of A and B");
  SB_ASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  i64 R = A.t->dims[NA - 1];
  SB_ASSERT(B.t->dims[NB - 1] == R, "Reduction size mismatch");
  i64 chunkA = A.t->blksize / R, chunkB = B.t->blksize / R;
  SB_ASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(B.labels, PB, NB - 1), Cd = sub(labelsC, PC, NC);
  bool ab;
  if (Cd == cat(Awo, Bwo))
    ab = true;
  else if (Cd == cat(Bwo, Awo))
    ab = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense_wo_r...,B_dense_wo_r...] or swapped");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    if (has(A.map, l)) {
      int ax = at(A.map, l);
      SB_ASSERT(ax <= PA, "C prefix label from A prefix");
      src[j] = ax;
    } else {
      SB_ASSERT(has(B.map, l), "C prefix label from B");
      int ax = at(B.map, l);
      SB_ASSERT(ax <= PB, "C prefix label from B prefix");
      src[j] = -ax;
    }
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 nA = TA.n, nB = TB.n;
  AlContrib c;
  if (nA && nB) {
    if (nA * nB >= (1ll << 31)) fail(SB_E_LIMIT, "more than 2^31 block pairs in one contraction");
    c.np = nA * nB;
    c.pa.alloc(ctx, c.np);
    c.pb.alloc(ctx, c.np);
    k_cross_pairs(ctx, nA, nB, c.pa.get(), c.pb.get());
    c.ck.alloc(ctx, (size_t)c.np * std::max(1, PC));
    DBuf<u64> lin(ctx, c.np);
    if (PC) k_pair_keys(ctx, c.pa.get(), c.pb.get(), c.np, TA.keys.get(), PA, TB.keys.get(), PB, src.data(), PC,
                        C.dims.data(), c.ck.get(), lin.get());
    c.xy.alloc(ctx, (size_t)c.np * 2);
    k_al_xy(ctx, 1, c.pa.get(), c.pb.get(), c.np, TA.ids.get(), TA.keys.get(), PA, 0, TB.ids.get(), nullptr, nullptr,
            0, c.xy.get());
  }
  c.nx = std::max<i64>(1, TA.ntemplates);
  c.ny = std::max<i64>(1, TB.ntemplates);
  c.sA = TA.scalars.get();
  c.sB = TB.scalars.get();
  GemmDesc d{chunkA, chunkB, R, 1, chunkA, chunkB, 1, ab ? 1 : chunkB, ab ? chunkA : 1};
  al_finish(pl, C, c, gemm_templates(flags, TA.data.get(), TA.blksize, TB.data.get(), TB.blksize, 0, d, C.blksize),
            nullptr, 0, nullptr, 0);
}

// contract_aliased! (contract_aliased.jl:383-408, 659-690) and the 9-argument
// shims (contract_aliased_shared.jl:535-562)
static void contract_aliased_ad(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab,
                                unsigned flags) {
  SB_ASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  if (at(A.map, rlab) <= A.t->P()) {
    al_r_to_last_prefix(pl, A, rlab);
    al_prefix_outer_ad(pl, C, labelsC, A, B, rlab);
  } else {
    al_r_to_last_dense(pl, A, rlab);
    al_dense_ad(pl, C, labelsC, A, B, rlab, flags);
  }
}
static void contract_aliased_aa(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab,
                                unsigned flags) {
  SB_ASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  if (at(B.map, rlab) <= B.t->P()) {
    al_r_to_last_prefix(pl, A, rlab);
    al_r_to_last_prefix(pl, B, rlab);
    al_prefix_outer_aa(pl, C, labelsC, A, B, rlab);
  } else {
    SB_ASSERT(at(A.map, rlab) > A.t->P(), "rlab must be in A dense tail when it is in B dense tail");
    al_r_to_last_dense(pl, A, rlab);
    al_r_to_last_dense(pl, B, rlab);
    al_dense_aa(pl, C, labelsC, A, B, rlab, flags);
  }
}

// contract_shared! Aliased x Dense (contract_aliased_shared.jl:103-302)
static void shared_al_dense(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared,
                            unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "aliased_dense_shared";
  for (auto& l : shared) SB_ASSERT(!contains(labelsC, l), "shared label must be reduced (not in labelsC)");
  int PA = A.t->P(), NA = A.t->N(), NB = B.t->N(), PC = C.P(), NC = C.N();
  Labels sp, sd;
  for (auto& l : shared) {
    SB_ASSERT(has(A.map, l), "shared label must exist in A");
    SB_ASSERT(has(B.map, l), "shared label must exist in B");
    (at(A.map, l) <= PA ? sp : sd).push_back(l);
  }
  Labels Ad0 = sub(A.labels, PA, NA), red, keepA0, keepB0;
  for (auto& l : Ad0) (contains(sd, l) ? red : keepA0).push_back(l);
  for (auto& l : B.labels)
    if (!contains(shared, l)) keepB0.push_back(l);
  Labels Cd = sub(labelsC, PC, NC), dA, dB;
  Mode mode = cdense_grouping(Cd, keepA0, keepB0, dA, dB);
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims; unsupported");
  SB_ASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA covers keepA");
  SB_ASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB covers keepB");
  {
    std::vector<i64> spax, perm;
    for (auto& l : sp) spax.push_back(at(A.map, l));
    for (int a = 1; a <= PA; ++a)
      if (std::find(spax.begin(), spax.end(), a) == spax.end()) perm.push_back(a);
    for (auto a : spax) perm.push_back(a);
    std::vector<i64> ka, ra;
    for (auto& l : dA) ka.push_back(at(A.map, l) - PA);
    for (auto& l : red) ra.push_back(at(A.map, l) - PA);
    SB_ASSERT((int)(ka.size() + ra.size()) == A.t->N2, "perm covers dense axes");
    for (auto a : ka) perm.push_back(PA + a);
    for (auto a : ra) perm.push_back(PA + a);
    if (!is_identity(perm)) al_oper_perm(pl, A, perm);
  }
  std::vector<i64> permB;
  for (auto& l : sp) permB.push_back(at(B.map, l));
  for (auto& l : red) permB.push_back(at(B.map, l));
  for (auto& l : dB) permB.push_back(at(B.map, l));
  SB_ASSERT((int)permB.size() == NB, "B perm length mismatch; labelsB must match B ndims");
  const Tensor* Bt = B.t;
  if (!is_identity(permB)) Bt = plan_dense_permute(pl, *Bt, permB);
  int nsp = (int)sp.size(), nred = (int)red.size(), nkA = (int)dA.size(), nkB = (int)dB.size();
  const Tensor& TA = *A.t;
  const std::vector<i64>& dBp = Bt->dims;
  i64 M = nkA == 0 ? 1 : prod(TA.dims, PA, PA + nkA);
  i64 K = nred == 0 ? 1 : prod(TA.dims, PA + nkA, NA);
  i64 Nn = nkB == 0 ? 1 : prod(dBp, nsp + nred, NB);
  if (nred > 0) SB_ASSERT(K == prod(dBp, nsp, nsp + nred), "Reduction extent mismatch between A dense tail and B");
  SB_ASSERT(C.blksize == M * Nn, "C.blksize must equal M*N");
  std::vector<int> csrc(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    SB_ASSERT(has(A.map, l), "C prefix label must exist in A");
    int ap = at(A.map, l);
    SB_ASSERT(ap <= PA, "C prefix label must come from A sparse prefix");
    SB_ASSERT(!contains(sp, l), "C prefix label cannot be a reduced shared-prefix label");
    csrc[j] = ap;
  }
  std::vector<int> jA;
  std::vector<i64> spstr;
  i64 SP = 1;
  for (int t = 0; t < nsp; ++t) {
    jA.push_back(at(A.map, sp[t]) - 1);
    spstr.push_back(SP);
    SP *= dBp[t];
  }
  AlContrib c;
  i64 nA = TA.n;
  if (nA) {  // every A block, in key order (the runs of equal shared-prefix values keep that order)
    c.np = nA;
    c.pa.alloc(ctx, nA);
    k_iota(ctx, c.pa.get(), nA, 0);
    c.ck.alloc(ctx, (size_t)nA * std::max(1, PC));
    DBuf<u64> lin(ctx, nA);
    if (PC) k_pair_keys(ctx, c.pa.get(), c.pa.get(), nA, TA.keys.get(), PA, nullptr, 0, csrc.data(), PC,
                        C.dims.data(), c.ck.get(), lin.get());
    c.xy.alloc(ctx, (size_t)nA * 2);
    k_al_xy(ctx, 2, c.pa.get(), nullptr, nA, TA.ids.get(), TA.keys.get(), PA, 0, nullptr, jA.data(), spstr.data(),
            nsp, c.xy.get());
  }
  c.nx = std::max<i64>(1, TA.ntemplates);
  c.ny = SP;
  c.sA = TA.scalars.get();
  // template (tidA, sp_lin): A_t (M x K) * Bp[sp_lin, :, :] (K x N, element (k, n) at sp_lin-1 + SP*(k + K*n))
  GemmDesc d{M, Nn, K, 1, M, SP, SP * K, mode == AthenB ? 1 : Nn, mode == AthenB ? M : 1};
  al_finish(pl, C, c, gemm_templates(flags, TA.data.get(), TA.blksize, Bt->dptr(), 1, 0, d, C.blksize), nullptr, 0,
            nullptr, 0);
}

// contract_shared! Aliased x Aliased (contract_aliased_shared.jl:325-523)
static void shared_al_al(Plan& pl, Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared,
                         unsigned flags) {
  Ctx* ctx = pl.ctx;
  pl.family = "aliased_aliased_shared";
  for (auto& l : shared) SB_ASSERT(!contains(labelsC, l), "shared label must be reduced (not in labelsC)");
  int PA = A.t->P(), NA = A.t->N(), PB = B.t->P(), NB = B.t->N(), PC = C.P(), NC = C.N();
  Labels sp, sd;
  for (auto& l : shared) {
    SB_ASSERT(has(A.map, l) && has(B.map, l), "shared label must exist in both A and B");
    bool ap = at(A.map, l) <= PA, bp = at(B.map, l) <= PB;
    if (ap && bp)
      sp.push_back(l);
    else if (!ap && !bp)
      sd.push_back(l);
    else
      fail(SB_E_ERROR, "shared label crosses prefix/dense boundary; not supported");
  }
  Labels Ad0 = sub(A.labels, PA, NA), Bd0 = sub(B.labels, PB, NB), red, keepA0, keepB0;
  for (auto& l : Ad0) (contains(sd, l) ? red : keepA0).push_back(l);
  for (auto& l : Bd0)
    if (!contains(sd, l)) keepB0.push_back(l);
  Labels Cd = sub(labelsC, PC, NC), dA, dB;
  Mode mode = cdense_grouping(Cd, keepA0, keepB0, dA, dB);
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims; unsupported");
  SB_ASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA covers keepA");
  SB_ASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB covers keepB");
  auto build_perm = [&](Oper& X, bool red_first, const Labels& keep) {
    int P = X.t->P();
    std::vector<i64> spax, perm;
    for (auto& l : sp) spax.push_back(at(X.map, l));
    for (int a = 1; a <= P; ++a)
      if (std::find(spax.begin(), spax.end(), a) == spax.end()) perm.push_back(a);
    for (auto a : spax) perm.push_back(a);
    std::vector<i64> ka, ra;
    for (auto& l : keep) ka.push_back(at(X.map, l) - P);
    for (auto& l : red) ra.push_back(at(X.map, l) - P);
    SB_ASSERT((int)(ka.size() + ra.size()) == X.t->N2, "perm covers dense axes");
    for (auto a : (red_first ? ra : ka)) perm.push_back(P + a);
    for (auto a : (red_first ? ka : ra)) perm.push_back(P + a);
    return perm;
  };
  auto pA = build_perm(A, false, dA);
  auto pB = build_perm(B, true, dB);
  if (!is_identity(pA)) al_oper_perm(pl, A, pA);
  if (!is_identity(pB)) al_oper_perm(pl, B, pB);
  const Tensor &TA = *A.t, &TB = *B.t;
  int nkA = (int)dA.size(), nkB = (int)dB.size(), nred = (int)red.size();
  i64 M = nkA == 0 ? 1 : prod(TA.dims, PA, PA + nkA);
  i64 K = nred == 0 ? 1 : prod(TA.dims, PA + nkA, NA);
  i64 Nn = nkB == 0 ? 1 : prod(TB.dims, PB + nred, NB);
  if (nred > 0) SB_ASSERT(K == prod(TB.dims, PB, PB + nred), "Reduction extent mismatch between A and B dense tails");
  SB_ASSERT(C.blksize == M * Nn, "C.blksize must equal M*N");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    if (has(A.map, l) && at(A.map, l) <= PA)
      src[j] = at(A.map, l);
    else if (has(B.map, l) && at(B.map, l) <= PB)
      src[j] = -at(B.map, l);
    else
      fail(SB_E_ERROR, "C prefix label must come from A or B sparse prefix (excluding shared_prefix)");
  }
  std::vector<int> jA, jB;
  std::vector<i64> spd;
  for (auto& l : sp) {
    jA.push_back(at(A.map, l));
    jB.push_back(at(B.map, l));
    spd.push_back(TA.dims[at(A.map, l) - 1]);
  }
  i64 nA = TA.n, nB = TB.n;
  int nsp = (int)sp.size();
  AlContrib c;
  if (nA && nB) {
    DBuf<i32> lo(ctx, nA), hi(ctx, nA);
    if (nsp == 0) {
      DBuf<u64> z(ctx, nA), zb(ctx, nB);
      k_zero(ctx, z.get(), nA * sizeof(u64));
      k_zero(ctx, zb.get(), nB * sizeof(u64));
      k_equal_ranges(ctx, z.get(), nA, zb.get(), nB, lo.get(), hi.get());
    } else if (nsp >= 2 && !(flags & SB_FLAG_EXACT_JOIN)) {
      k_walk_ranges(ctx, 2, TA.keys.get(), PA, nA, TB.keys.get(), PB, nB, jA.data(), jB.data(), nsp, lo.get(),
                    hi.get());
    } else {
      DBuf<u64> ja(ctx, nA), jb(ctx, nB);
      k_cols_lin(ctx, TA.keys.get(), nA, PA, jA.data(), nsp, spd.data(), ja.get());
      k_cols_lin(ctx, TB.keys.get(), nB, PB, jB.data(), nsp, spd.data(), jb.get());
      k_equal_ranges(ctx, ja.get(), nA, jb.get(), nB, lo.get(), hi.get());
    }
    Pairs pr = pairs_from_ranges(ctx, lo, hi, nA);
    c.np = pr.n;
    c.pa = std::move(pr.a);
    c.pb = std::move(pr.b);
  }
  if (c.np) {
    c.ck.alloc(ctx, (size_t)c.np * std::max(1, PC));
    DBuf<u64> lin(ctx, c.np);
    if (PC) k_pair_keys(ctx, c.pa.get(), c.pb.get(), c.np, TA.keys.get(), PA, TB.keys.get(), PB, src.data(), PC,
                        C.dims.data(), c.ck.get(), lin.get());
    c.xy.alloc(ctx, (size_t)c.np * 2);
    k_al_xy(ctx, 1, c.pa.get(), c.pb.get(), c.np, TA.ids.get(), TA.keys.get(), PA, 0, TB.ids.get(), nullptr, nullptr,
            0, c.xy.get());
  }
  c.nx = std::max<i64>(1, TA.ntemplates);
  c.ny = std::max<i64>(1, TB.ntemplates);
  c.sA = TA.scalars.get();
  c.sB = TB.scalars.get();
  // template (tidA, tidB): A_t (M x K) * B_t (K x N)
  GemmDesc d{M, Nn, K, 1, M, 1, K, mode == AthenB ? 1 : Nn, mode == AthenB ? M : 1};
  al_finish(pl, C, c, gemm_templates(flags, TA.data.get(), TA.blksize, TB.data.get(), TB.blksize, 0, d, C.blksize),
            nullptr, 0, nullptr, 0);
}

// contract_aliased! COO x Dense -> Aliased (contract_aliased_coo_dense.jl:34-164)
static void coo_dense_aliased(Plan& pl, Tensor& C, const Labels& labelsC, CooOper& A, Oper& B, const Label& rlab) {
  Ctx* ctx = pl.ctx;
  pl.family = "coo_dense_aliased";
  int PC = C.P(), NC = C.N(), NA = A.t->N(), NB = B.t->N();
  {  // aligned_A_to_Cprefix (contract_coo_dense.jl:4-18)
    Labels desired = sub(labelsC, 0, PC);
    desired.push_back(rlab);
    LabelMap m = make_map(A.labels);
    SB_ASSERT((int)desired.size() == NA, "length(desired) == NA");
    std::vector<i64> perm(NA);
    for (int j = 0; j < NA; ++j) perm[j] = at(m, desired[j]);
    A.t = plan_coo_permute_sort(pl, *A.t, perm);
    A.labels = desired;
    A.map = make_map(desired);
  }
  Labels dnr;
  for (int j = PC; j < NC; ++j)
    if (labelsC[j] != rlab) dnr.push_back(labelsC[j]);
  int rpos = -1;
  for (int i = 0; i < NB; ++i)
    if (B.labels[i] == rlab) {
      rpos = i + 1;
      break;
    }
  SB_ASSERT(rpos > 0, "rlab not found in labelsB");
  std::vector<i64> perm(NB, 0);
  for (size_t k = 0; k < dnr.size(); ++k) {
    int pos = -1;
    for (int i = 0; i < NB; ++i)
      if (B.labels[i] == dnr[k]) {
        pos = i + 1;
        break;
      }
    SB_ASSERT(pos > 0, "label not found in labelsB");
    if ((int)k >= NB) fail(SB_E_BOUNDS, "BoundsError: perm");
    perm[k] = pos;
  }
  perm[NB - 1] = rpos;
  const Tensor* Bt = B.t;
  if (!is_identity(perm)) {
    std::vector<i64> s(perm);
    std::sort(s.begin(), s.end());
    for (int i = 0; i < NB; ++i)
      if (s[i] != i + 1) fail(SB_E_ARGUMENT, "invalid permutation of B");
    Bt = plan_dense_permute(pl, *Bt, perm);
  }
  i64 R = Bt->dims[NB - 1];
  i64 bs = prod(Bt->dims, 0, NB - 1);
  SB_ASSERT(C.blksize == bs, "C.blksize must equal prod(B dims without r)");
  const Tensor& TA = *A.t;
  AlContrib c;
  i64 nA = TA.n;
  if (nA) {
    c.pa.alloc(ctx, nA);
    k_iota(ctx, c.pa.get(), nA, 0);
    DBuf<unsigned char> keep(ctx, nA);
    k_rv_inrange(ctx, TA.keys.get(), NA, NA, nA, R, keep.get());
    DBuf<i32> none;
    c.np = compact_pairs(ctx, keep.get(), c.pa, none, nA);
  }
  if (c.np) {
    std::vector<int> src(PC);
    std::iota(src.begin(), src.end(), 1);
    c.ck.alloc(ctx, (size_t)c.np * std::max(1, PC));
    DBuf<u64> lin(ctx, c.np);
    if (PC) k_pair_keys(ctx, c.pa.get(), c.pa.get(), c.np, TA.keys.get(), NA, nullptr, 0, src.data(), PC,
                        C.dims.data(), c.ck.get(), lin.get());
  }
  c.sA = TA.data.get();  // the aligned COO values (gathered by an earlier plan step)
  al_finish(pl, C, c, TemplateFn(), Bt->dptr(), R, TA.keys.get(), NA);
}

// ================================================================ dispatcher

// SB_TRACE=1 in the environment: one line per contract! call naming the
// storage kinds, block counts and the kernel family taken, as the reference
// prints its dispatch (contract_bs_bs.jl:138-143, contract_bs_dense.jl:225-230,
// contract_coo_dense.jl:59-63, contract_coo_bs.jl:29-34), followed by what the
// plan holds: matched block pairs, output blocks, FLOPs
bool trace_on() {
  static const bool on = [] {
    const char* v = std::getenv("SB_TRACE");
    return v && std::string(v) == "1";
  }();
  return on;
}

namespace {

std::string trace_operand(const Tensor& T) {
  char b[160];
  if (T.kind == BS)
    std::snprintf(b, sizeof b, "BS(blocks=%lld, P=%d, N2=%d)", (long long)T.n, T.P(), T.N2);
  else if (T.kind == COO)
    std::snprintf(b, sizeof b, "COO(nnz=%lld, N=%d)", (long long)T.n, (int)T.dims.size());
  else if (T.kind == ALIASED)
    std::snprintf(b, sizeof b, "AliasedBS(blocks=%lld, P=%d, N2=%d)", (long long)T.n, T.P(), T.N2);
  else
    std::snprintf(b, sizeof b, "Dense(N=%d)", (int)T.dims.size());
  return b;
}

void trace_dispatch(const char* entry, const Tensor& A, const Tensor& B, int axis, const char* family,
                    const Plan& pl, const Tensor& C) {
  std::printf("[SB_TRACE] %s  %s x %s  axis=%d  -> %s  pairs=%lld  C.blocks=%lld  flops=%.6g\n", entry,
              trace_operand(A).c_str(), trace_operand(B).c_str(), axis, family, (long long)pl.npairs,
              (long long)C.n, pl.flops);
  std::fflush(stdout);
}

}  // namespace

std::unique_ptr<Plan> contract_plan(Ctx* ctx, Tensor& C, const Labels& labelsC, const Tensor& A,
                                    const Labels& labelsA, const Tensor& B, const Labels& labelsB, unsigned flags) {
  SB_ASSERT(labelsA.size() == A.dims.size(), "labelsA length must match A rank");
  SB_ASSERT(labelsB.size() == B.dims.size(), "labelsB length must match B rank");
  SB_ASSERT(labelsC.size() == C.dims.size(), "labelsC length must match C rank");
  auto pl = std::make_unique<Plan>();
  pl->ctx = ctx;
  C.touch();  // C's block list and payload are replaced
  pl->out = &C;
  pl->in_a = &A;
  pl->in_b = &B;
  LabelMap mapA = make_map(labelsA), mapB = make_map(labelsB);
  Labels shared, reduced;
  for (auto& l : labelsA)
    if (has(mapB, l)) shared.push_back(l);
  for (auto& l : shared)
    if (!contains(labelsC, l)) reduced.push_back(l);
  Oper Ao{&A, labelsA, mapA}, Bo{&B, labelsB, mapB};
  const char* entry = "contract!";
  const char* family = "";
  int axis = 0;
  if (reduced.size() == 1) {
    SB_ASSERT(shared.size() == 1, "Expected exactly 1 shared label");
    const Label& r = reduced[0];
    i64 da = A.dims[at(mapA, r) - 1], db = B.dims[at(mapB, r) - 1];
    SB_ASSERT(da == db, "Dim mismatch on shared label");
    SB_ASSERT(!contains(labelsC, r), "Reduction label must NOT appear in output labels");
    if (C.kind == BS && A.kind == BS && B.kind == BS) {
      // contract! BS x BS (contract_bs_bs.jl:126-151)
      int axisBr = at(mapB, r);
      entry = "contract_bs_bs.contract!";
      axis = axisBr;
      if (axisBr <= B.P()) {
        family = "prefix_outer";
        permute_r_to_last_prefix(*pl, Ao, r);
        permute_r_to_last_prefix(*pl, Bo, r);
        prefix_outer_bb(*pl, C, labelsC, Ao, Bo, r);
      } else {
        family = "dense_bb";
        dense_bb(*pl, C, labelsC, Ao, Bo, r, flags);
      }
    } else if (C.kind == BS && A.kind == BS && B.kind == DENSE) {
      // contract! BS x Dense (contract_bs_dense.jl:207-245)
      SB_ASSERT(has(mapA, r), "rlab not found in mapA");
      SB_ASSERT(has(mapB, r), "rlab not found in mapB");
      entry = "contract_bs_dense.contract!";
      axis = at(mapA, r);
      if (at(mapA, r) <= A.P()) {
        family = "prefix_outer_bd";
        permute_r_to_last_prefix(*pl, Ao, r);
        prefix_outer_bd(*pl, C, labelsC, Ao, Bo, r);
      } else {
        family = "dense_bd";
        dense_bd(*pl, C, labelsC, Ao, Bo, r, flags);
      }
    } else if (C.kind == COO && A.kind == COO && B.kind == COO) {
      CooOper Ac{&A, labelsA, mapA}, Bc{&B, labelsB, mapB};
      entry = "contract_coo_coo.contract!";
      family = "coo_coo";
      contract_coo_coo(*pl, C, labelsC, Ac, Bc, r);
    } else if (C.kind == BS && A.kind == COO && B.kind == BS) {
      // contract! COO x BS (contract_coo_bs.jl:14-46)
      CooOper Ac{&A, labelsA, mapA};
      coo_rlab_last_sorted(*pl, Ac, r);
      entry = "contract_coo_bs.contract!";
      axis = at(mapB, r);
      if (at(mapB, r) <= B.P()) {
        family = "prefix";
        permute_r_to_last_prefix(*pl, Bo, r);
        coo_bs_prefix(*pl, C, labelsC, Ac, Bo, r);
      } else {
        family = "dense_index";
        coo_bs_dense_index(*pl, C, labelsC, Ac, Bo, r);
      }
    } else if (C.kind == BS && A.kind == COO && B.kind == DENSE) {
      CooOper Ac{&A, labelsA, mapA};
      entry = "contract_coo_dense.contract!";
      family = "coo_dense";
      contract_coo_dense(*pl, C, labelsC, Ac, Bo, r, flags);
    } else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == DENSE) {
      entry = "contract_aliased!";
      family = "aliased_dense";
      contract_aliased_ad(*pl, C, labelsC, Ao, Bo, r, flags);  // contract_aliased_shared.jl:535-547
    } else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == ALIASED) {
      entry = "contract_aliased!";
      family = "aliased_aliased";
      contract_aliased_aa(*pl, C, labelsC, Ao, Bo, r, flags);  // contract_aliased_shared.jl:550-562
    } else if (C.kind == ALIASED && A.kind == COO && B.kind == DENSE) {
      CooOper Ac{&A, labelsA, mapA};
      entry = "contract_aliased_coo_dense.contract_aliased!";
      family = "coo_dense_aliased";
      coo_dense_aliased(*pl, C, labelsC, Ac, Bo, r);  // tensor_wrappers_aliased.jl:263-275
    } else {
      fail(SB_E_METHOD, "MethodError: no single-label contract! method for these storage kinds");
    }
  } else {
    entry = "contract_shared!";
    axis = (int)shared.size();
    if (C.kind == BS && A.kind == BS && B.kind == BS) {
      family = "shared_bs_bs";
      shared_bs_bs(*pl, C, labelsC, Ao, Bo, shared, flags);
    } else if (C.kind == BS && A.kind == BS && B.kind == DENSE) {
      family = "shared_bs_dense";
      shared_bs_dense(*pl, C, labelsC, Ao, Bo, shared, flags);
    } else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == DENSE) {
      family = "shared_aliased_dense";
      shared_al_dense(*pl, C, labelsC, Ao, Bo, shared, flags);
    } else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == ALIASED) {
      family = "shared_aliased_aliased";
      shared_al_al(*pl, C, labelsC, Ao, Bo, shared, flags);
    } else {
      fail(SB_E_METHOD, "MethodError: no contract_shared! method for these storage kinds");
    }
  }
  if (trace_on()) trace_dispatch(entry, A, B, axis, family, *pl, C);
  return pl;
}

}  // namespace sb
