// Oracle: a CPU restatement of the reference SparseBackends algorithms
// (Ritvik1sharma/SparseBackends, Julia), written from the mathematics and
// control flow of the reference sources cited beside each function.
//
// THIS IS TEST INFRASTRUCTURE.  It exists only to check the HIP library
// (tests/, __graft_entry__.smoke()) and to time the reference algorithm on
// the host (bench.py cpu_baseline).  Nothing in the product loads it.
//
// Conventions kept from the reference so results compare bit for bit:
//   * keys and ids are 1-based; a block-sparse key is the tuple of prefix
//     coordinates, stored row-wise (nblocks x P);
//   * keys are kept sorted by the column-major linear index of the prefix
//     (axis 1 fastest), ids are payload slots in first-allocation order;
//   * all arrays are column-major; FP type is Float64.
// The per-block-pair `mul!` / `BLAS.ger!` calls of the reference go to the
// image's OpenBLAS 0.3.29 (the version the reference pins,
// Manifest.toml:512-515) when orc_set_blas() has loaded it, else to plain
// loops (tests use dyadic values, where both are exact).
//
// Parity status of each function: see the header comment of oracle/oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <dlfcn.h>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/sb_status.h"

namespace orc {

using i64 = int64_t;

struct SbError {
  int code;
  std::string msg;
};

[[noreturn]] static void fail(int code, const std::string& m) { throw SbError{code, m}; }
#define OASSERT(c, m)                         \
  do {                                        \
    if (!(c)) fail(SB_E_ASSERT, std::string(m)); \
  } while (0)

// ---------------------------------------------------------------- BLAS hook
typedef void (*dgemm_fn)(const char*, const char*, const i64*, const i64*, const i64*, const double*,
                         const double*, const i64*, const double*, const i64*, const double*, double*,
                         const i64*, size_t, size_t);
typedef void (*dger_fn)(const i64*, const i64*, const double*, const double*, const i64*, const double*,
                        const i64*, double*, const i64*);
typedef void (*setthreads_fn)(i64);
static dgemm_fn g_dgemm = nullptr;
static dger_fn g_dger = nullptr;
static setthreads_fn g_setthreads = nullptr;

// C(m x n, ldc) += op(A) * op(B), column-major; ta/tb = 'N' or 'T'
static void gemm_acc(char ta, char tb, i64 m, i64 n, i64 k, const double* A, i64 lda, const double* B, i64 ldb,
                     double* C, i64 ldc) {
  if (m == 0 || n == 0) return;
  if (g_dgemm) {
    const double one = 1.0;
    g_dgemm(&ta, &tb, &m, &n, &k, &one, A, &lda, B, &ldb, &one, C, &ldc, 1, 1);
    return;
  }
  for (i64 j = 0; j < n; ++j)
    for (i64 i = 0; i < m; ++i) {
      double s = 0.0;
      for (i64 p = 0; p < k; ++p) {
        double a = (ta == 'N') ? A[i + p * lda] : A[p + i * lda];
        double b = (tb == 'N') ? B[p + j * ldb] : B[j + p * ldb];
        s += a * b;
      }
      C[i + j * ldc] += s;
    }
}

// C(m x n) += x * y'   (BLAS.ger! with alpha = 1; contract_bs_bs.jl:118-124)
static void ger_acc(i64 m, i64 n, const double* x, const double* y, double* C) {
  if (g_dger) {
    const double one = 1.0;
    const i64 inc = 1;
    g_dger(&m, &n, &one, x, &inc, y, &inc, C, &m);
    return;
  }
  for (i64 j = 0; j < n; ++j)
    for (i64 i = 0; i < m; ++i) C[i + j * m] += x[i] * y[j];
}

// ---------------------------------------------------------------- labels
struct Label {
  uint64_t id, plev;
  bool operator==(const Label& o) const { return id == o.id && plev == o.plev; }
  bool operator!=(const Label& o) const { return !(*this == o); }
  bool operator<(const Label& o) const { return id < o.id || (id == o.id && plev < o.plev); }
};
using Labels = std::vector<Label>;
using LabelMap = std::map<Label, int>;  // label -> 1-based axis (contract.jl:51-61)

static LabelMap make_map(const Labels& l) {
  LabelMap m;
  for (size_t i = 0; i < l.size(); ++i) m[l[i]] = (int)i + 1;
  return m;
}
static bool has(const LabelMap& m, const Label& l) { return m.count(l) != 0; }
static int at(const LabelMap& m, const Label& l) {
  auto it = m.find(l);
  if (it == m.end()) fail(SB_E_KEY, "KeyError: label not found");
  return it->second;
}
static bool contains(const Labels& v, const Label& l) { return std::find(v.begin(), v.end(), l) != v.end(); }

// ---------------------------------------------------------------- tensors
enum Kind { BS = SB_KIND_BLOCKSPARSE, COO = SB_KIND_COO, DENSE = SB_KIND_DENSE, ALIASED = SB_KIND_ALIASED };

struct Tensor {
  Kind kind;
  std::vector<i64> dims;
  // block-sparse (also aliased: keys/scalars/alias ids)
  int N2 = 0;
  i64 blksize = 0;
  std::vector<i64> keys;     // BS: nblocks x P ; COO: nnz x N
  std::vector<i64> ids;      // BS: 1-based slots ; ALIASED: alias ids (1-based)
  std::vector<double> data;  // BS payload / COO vals / dense data / aliased templates
  bool dirty = false;        // COO
  std::vector<double> scalars;  // ALIASED
  i64 ntemplates = 0;           // ALIASED
  int N() const { return (int)dims.size(); }
  int P() const { return (int)dims.size() - N2; }
  i64 nblocks() const {
    int p = P();
    if (kind == COO) return dims.empty() ? (i64)data.size() : (i64)keys.size() / (i64)dims.size();
    return p == 0 ? (i64)ids.size() : (i64)keys.size() / p;
  }
};

static i64 prod(const std::vector<i64>& v, size_t lo, size_t hi) {
  i64 p = 1;
  for (size_t i = lo; i < hi; ++i) p *= v[i];
  return p;
}

// _prefix_lin (storage.jl:16-18): 1-based column-major linear index of a key
static i64 prefix_lin(const i64* k, const i64* pdims, int P) {
  i64 lin = 1, stride = 1;
  for (int d = 0; d < P; ++d) {
    lin += (k[d] - 1) * stride;
    stride *= pdims[d];
  }
  return lin;
}

// _find_key (storage.jl:49-65): binary search on the prefix lin
static std::pair<i64, bool> find_key(const Tensor& A, const i64* k) {
  int P = A.P();
  i64 n = A.nblocks();
  i64 target = prefix_lin(k, A.dims.data(), P);
  i64 lo = 0, hi = n;
  while (lo < hi) {
    i64 mid = (lo + hi) >> 1;
    i64 midv = prefix_lin((A.keys.data() + (mid * P)), A.dims.data(), P);
    if (midv < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  bool found = lo < n && std::equal(k, k + P, (A.keys.data() + (lo * P)));
  return {lo, found};
}

// _alloc_block! / _ensure_block! (storage.jl:67-83)
static i64 ensure_block(Tensor& A, const i64* k) {
  int P = A.P();
  if (P == 0) {  // single block with key ()
    if (!A.ids.empty()) return A.ids[0];
    i64 id = (i64)A.data.size() / A.blksize + 1;
    A.data.resize(A.data.size() + A.blksize, 0.0);
    A.ids.push_back(id);
    return id;
  }
  auto [pos, found] = find_key(A, k);
  if (found) return A.ids[pos];
  i64 id = (i64)A.data.size() / A.blksize + 1;
  A.data.resize(A.data.size() + A.blksize, 0.0);
  A.keys.insert(A.keys.begin() + pos * P, k, k + P);
  A.ids.insert(A.ids.begin() + pos, id);
  return id;
}

static double* block_ptr(Tensor& A, i64 id) { return A.data.data() + (id - 1) * A.blksize; }
static const double* block_ptr(const Tensor& A, i64 id) { return A.data.data() + (id - 1) * A.blksize; }

static void clear_bs(Tensor& C) {
  C.keys.clear();
  C.ids.clear();
  C.data.clear();
}

// NewBlockSparseSorted{T,N,N2}(dims) (storage.jl:86-93)
static Tensor bs_new(const std::vector<i64>& dims, int N2) {
  int N = (int)dims.size();
  OASSERT((N == 0 && N2 == 0) || (0 < N2 && N2 <= N), "0 < N2 <= N");
  for (auto d : dims) OASSERT(d >= 1, "dims .>= 1");
  Tensor t;
  t.kind = BS;
  t.dims = dims;
  t.N2 = N2;
  t.blksize = prod(dims, N - N2, N);
  return t;
}

// getindex / setindex! (storage.jl:100-124)
static void check_index(const Tensor& A, const i64* I) {
  for (int d = 0; d < A.N(); ++d)
    if (I[d] < 1 || I[d] > A.dims[d]) fail(SB_E_BOUNDS, "BoundsError");
}
static i64 block_index(const Tensor& A, const i64* I) {
  int P = A.P();
  i64 lin = 1, stride = 1;
  for (int j = 0; j < A.N2; ++j) {
    lin += (I[P + j] - 1) * stride;
    stride *= A.dims[P + j];
  }
  return lin;
}
static double bs_getindex(const Tensor& A, const i64* I) {
  check_index(A, I);
  int P = A.P();
  if (P == 0) return A.ids.empty() ? 0.0 : block_ptr(A, A.ids[0])[block_index(A, I) - 1];
  auto [pos, found] = find_key(A, I);
  if (!found) return 0.0;
  return block_ptr(A, A.ids[pos])[block_index(A, I) - 1];
}
static void bs_setindex(Tensor& A, double v, const i64* I) {
  check_index(A, I);
  i64 id = ensure_block(A, I);
  block_ptr(A, id)[block_index(A, I) - 1] = v;
}

// _check_no_cross_perm (storage.jl:126-132)
static void check_no_cross_perm(const std::vector<i64>& perm, int N, int P) {
  OASSERT((int)perm.size() == N, "length(perm) == N");
  std::vector<i64> s(perm);
  std::sort(s.begin(), s.end());
  for (int i = 0; i < N; ++i) OASSERT(s[i] == i + 1, "perm must be a permutation of 1:N");
  for (int i = 0; i < P; ++i)
    OASSERT(perm[i] <= P, "No-cross reorder violated: output prefix axis uses old block axis");
  for (int i = P; i < N; ++i)
    OASSERT(perm[i] > P, "No-cross reorder violated: output block axis uses old prefix axis");
}

// new_lin -> old_lin map of a dense-tail permutation (storage.jl:220-244)
static std::vector<i64> tail_map(const std::vector<i64>& dimsA_tail, const std::vector<i64>& perm_block) {
  int N2 = (int)dimsA_tail.size();
  std::vector<i64> dimsB(N2), invp(N2), strides_old(N2);
  for (int i = 0; i < N2; ++i) dimsB[i] = dimsA_tail[perm_block[i] - 1];
  for (int i = 0; i < N2; ++i) invp[perm_block[i] - 1] = i + 1;
  i64 s = 1;
  for (int d = 0; d < N2; ++d) {
    strides_old[d] = s;
    s *= dimsA_tail[d];
  }
  i64 bs = prod(dimsB, 0, N2);
  std::vector<i64> map_old(bs);
  std::vector<i64> idx(N2, 1);
  for (i64 nl = 0; nl < bs; ++nl) {
    i64 ol = 1;
    for (int od = 0; od < N2; ++od) ol += (idx[invp[od] - 1] - 1) * strides_old[od];
    map_old[nl] = ol;
    for (int d = 0; d < N2; ++d) {
      if (++idx[d] <= dimsB[d]) break;
      idx[d] = 1;
    }
  }
  return map_old;
}

// stable sortperm of blocks by the column-major lin of keys (storage.jl:207)
static std::vector<i64> sortperm_lin(const std::vector<i64>& keys, i64 n, int P, const i64* pdims) {
  std::vector<i64> lin(n), p(n);
  for (i64 i = 0; i < n; ++i) lin[i] = prefix_lin(&keys[i * P], pdims, P);
  std::iota(p.begin(), p.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](i64 a, i64 b) { return lin[a] < lin[b]; });
  return p;
}

// Base.permutedims(::NewBlockSparseSorted, perm) (storage.jl:188-264)
static Tensor bs_permutedims(const Tensor& A, const std::vector<i64>& perm) {
  int N = A.N(), P = A.P(), N2 = A.N2;
  check_no_cross_perm(perm, N, P);
  Tensor B;
  B.kind = BS;
  B.N2 = N2;
  B.dims.resize(N);
  for (int j = 0; j < N; ++j) B.dims[j] = A.dims[perm[j] - 1];
  std::vector<i64> perm_block(N2);
  for (int i = 0; i < N2; ++i) perm_block[i] = perm[P + i] - P;
  std::vector<i64> tailA(A.dims.begin() + P, A.dims.end());
  B.blksize = prod(B.dims, P, N);
  i64 n = (i64)A.ids.size();
  OASSERT((i64)A.data.size() == n * A.blksize, "length(A.data) == nblocks * A.blksize");
  std::vector<i64> newkeys((size_t)n * P);
  for (i64 i = 0; i < n; ++i)
    for (int j = 0; j < P; ++j) newkeys[i * P + j] = A.keys[i * P + perm[j] - 1];
  std::vector<i64> p = sortperm_lin(newkeys, n, P, B.dims.data());
  B.keys.resize((size_t)n * P);
  B.ids.resize(n);
  B.data.resize((size_t)n * B.blksize);
  bool ident = true;
  for (int i = 0; i < N2; ++i) ident = ident && perm_block[i] == i + 1;
  bool fast = ident && A.blksize == B.blksize;
  std::vector<i64> map_old;
  if (!fast) map_old = tail_map(tailA, perm_block);
  for (i64 o = 0; o < n; ++o) {
    i64 s = p[o];
    for (int j = 0; j < P; ++j) B.keys[o * P + j] = newkeys[s * P + j];
    B.ids[o] = o + 1;
    const double* v = block_ptr(A, A.ids[s]);
    double* w = B.data.data() + o * B.blksize;
    if (fast)
      std::memcpy(w, v, sizeof(double) * B.blksize);
    else
      for (i64 t = 0; t < B.blksize; ++t) w[t] = v[map_old[t] - 1];
  }
  return B;
}

// permutedims!(::NewBlockSparseSorted, perm) (storage.jl:271-392): in place,
// needs packed ids; result equals the copying form with packed ids.
static void bs_permutedims_inplace(Tensor& A, const std::vector<i64>& perm) {
  int N = A.N(), P = A.P();
  check_no_cross_perm(perm, N, P);
  i64 n = (i64)A.ids.size();
  OASSERT((i64)A.data.size() == n * A.blksize, "length(A.data) == nblocks * A.blksize");
  for (i64 i = 0; i < n; ++i) OASSERT(A.ids[i] == i + 1, "permutedims! requires packed ids (ids[i]==i)");
  bool prefix_id = true;
  for (int i = 0; i < P; ++i) prefix_id = prefix_id && perm[i] == i + 1;
  Tensor B = bs_permutedims(A, perm);
  if (prefix_id) {  // keys untouched, no sort (storage.jl:337-341)
    A.dims = B.dims;
    A.data = std::move(B.data);
    return;
  }
  A = std::move(B);
}

// sort!(::NewBlockSparseSorted) (storage.jl:403-410): lexicographic, ids follow, data stays
static void bs_sort_lex(Tensor& C) {
  int P = C.P();
  i64 n = (i64)C.ids.size();
  std::vector<i64> p(n);
  std::iota(p.begin(), p.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](i64 a, i64 b) {
    return std::lexicographical_compare((C.keys.data() + (a * P)), (C.keys.data() + (a * P)) + P, (C.keys.data() + (b * P)), (C.keys.data() + (b * P)) + P);
  });
  std::vector<i64> k2(C.keys.size()), id2(n);
  for (i64 i = 0; i < n; ++i) {
    std::copy(&C.keys[p[i] * P], &C.keys[p[i] * P] + P, &k2[i * P]);
    id2[i] = C.ids[p[i]];
  }
  C.keys.swap(k2);
  C.ids.swap(id2);
}

// is_dense (storage.jl:413-433)
static bool bs_is_dense(const Tensor& A) {
  int P = A.P();
  if (P == 0) {
    OASSERT(A.ids.size() <= 1, "P=0 means at most one block");
    return true;
  }
  return (i64)A.ids.size() == prod(A.dims, 0, P);
}

// blocksparse_from_dense (conversions.jl:21-75)
static Tensor bs_from_dense(const Tensor& D, int N2, double atol, double rtol, bool dropzeros) {
  int N = D.N();
  OASSERT(0 <= N2 && N2 <= N, "0 <= N2 <= N");
  int P = N - N2;
  Tensor B;
  B.kind = BS;
  B.dims = D.dims;
  B.N2 = N2;
  B.blksize = prod(D.dims, P, N);
  double gmax = 0.0;
  if (rtol != 0)
    for (double v : D.data) gmax = std::max(gmax, std::fabs(v));
  double tol = std::max(atol, rtol * gmax);
  i64 np = prod(D.dims, 0, P);
  std::vector<double> buf(B.blksize);
  for (i64 pl = 0; pl < np; ++pl) {
    double bmax = 0.0;
    for (i64 sl = 0; sl < B.blksize; ++sl) {
      double v = D.data[pl + np * sl];
      buf[sl] = v;
      double av = std::fabs(v);
      if (av > bmax) bmax = av;
    }
    bool keep = !dropzeros || (bmax > tol);
    if (!keep) continue;
    i64 r = pl;
    for (int d = 0; d < P; ++d) {
      B.keys.push_back(r % D.dims[d] + 1);
      r /= D.dims[d];
    }
    B.ids.push_back((i64)B.ids.size() + 1);
    B.data.insert(B.data.end(), buf.begin(), buf.end());
  }
  return B;
}

// to_dense(::NewBlockSparseSorted; init, merged_axes) (conversions.jl:83-187)
static Tensor bs_to_dense(const Tensor& B, double init, const std::vector<std::pair<i64, i64>>& groups) {
  int N = B.N(), P = B.P();
  std::vector<std::pair<i64, i64>> plan;
  std::vector<i64> out_dims = B.dims;
  if (!groups.empty()) {
    std::vector<std::pair<i64, i64>> pairs0;
    std::vector<bool> used(N + 1, false);
    for (auto g : groups) {
      i64 a = g.first, b = g.second;
      if (!(1 <= a && a <= N) || !(1 <= b && b <= N)) fail(SB_E_ERROR, "Axis out of range in merge group");
      if (a == b) fail(SB_E_ERROR, "Duplicate axes in merge group");
      i64 keep = std::min(a, b), other = std::max(a, b);
      if (used[keep] || used[other]) fail(SB_E_ERROR, "Axis appears in multiple merge groups");
      used[keep] = used[other] = true;
      pairs0.push_back({keep, other});
    }
    std::stable_sort(pairs0.begin(), pairs0.end(), [](auto& x, auto& y) { return x.second > y.second; });
    std::vector<i64> dw = B.dims;
    for (auto [keep, other] : pairs0) {
      dw[keep - 1] *= dw[other - 1];
      dw.erase(dw.begin() + (other - 1));
      plan.push_back({keep, other});
    }
    out_dims = dw;
  }
  Tensor D;
  D.kind = DENSE;
  D.dims = out_dims;
  D.data.assign(prod(out_dims, 0, out_dims.size()), init);
  std::vector<i64> idx(N), dimw(N), full(N);
  for (i64 i = 0; i < B.nblocks(); ++i) {
    const double* blk = block_ptr(B, B.ids[i]);
    for (i64 sl = 0; sl < B.blksize; ++sl) {
      for (int d = 0; d < P; ++d) full[d] = B.keys[i * P + d];
      i64 r = sl;
      for (int d = P; d < N; ++d) {
        full[d] = r % B.dims[d] + 1;
        r /= B.dims[d];
      }
      int len = N;
      for (int d = 0; d < N; ++d) {
        idx[d] = full[d];
        dimw[d] = B.dims[d];
      }
      for (auto [keep, other] : plan) {
        i64 fused = idx[keep - 1] + (idx[other - 1] - 1) * dimw[keep - 1];
        idx[keep - 1] = fused;
        dimw[keep - 1] *= dimw[other - 1];
        for (int j = (int)other - 1; j < len - 1; ++j) {
          idx[j] = idx[j + 1];
          dimw[j] = dimw[j + 1];
        }
        --len;
      }
      i64 off = 0, st = 1;
      for (int d = 0; d < len; ++d) {
        off += (idx[d] - 1) * st;
        st *= out_dims[d];
      }
      D.data[off] = blk[sl];
    }
  }
  return D;
}

// fuse_sparse_axes (conversions.jl:228-263): keys are not re-sorted
static Tensor bs_fuse_sparse(const Tensor& A, i64 ax1, i64 ax2) {
  int N = A.N(), P = A.P();
  if (!(1 <= ax1 && ax1 <= P)) fail(SB_E_ARGUMENT, "ax1 must be in sparse head");
  if (!(1 <= ax2 && ax2 <= P)) fail(SB_E_ARGUMENT, "ax2 must be in sparse head");
  if (ax1 == ax2) fail(SB_E_ARGUMENT, "axes must be distinct");
  i64 a1 = std::min(ax1, ax2), a2 = std::max(ax1, ax2);
  i64 d1 = A.dims[a1 - 1], d2 = A.dims[a2 - 1];
  Tensor B;
  B.kind = BS;
  B.N2 = A.N2;
  B.blksize = A.blksize;
  for (int i = 1; i <= N - 1; ++i) B.dims.push_back(i < a2 ? (i == a1 ? d1 * d2 : A.dims[i - 1]) : A.dims[i]);
  i64 n = A.nblocks();
  for (i64 i = 0; i < n; ++i) {
    const i64* k = (A.keys.data() + (i * P));
    i64 fused = (k[a2 - 1] - 1) * d1 + k[a1 - 1];
    for (int j = 1; j <= P - 1; ++j) B.keys.push_back(j < a2 ? (j == a1 ? fused : k[j - 1]) : k[j]);
  }
  B.ids = A.ids;
  B.data = A.data;
  return B;
}

// fuse_dense_axes! (conversions.jl:265-297): mutates A (permutedims!) when
// the two tail axes are not adjacent, then returns the fused tensor
static Tensor bs_fuse_dense(Tensor& A, i64 ax1, i64 ax2) {
  int N = A.N(), P = A.P(), N2 = A.N2;
  if (!(P + 1 <= ax1 && ax1 <= N)) fail(SB_E_ARGUMENT, "ax1 must be in dense tail");
  if (!(P + 1 <= ax2 && ax2 <= N)) fail(SB_E_ARGUMENT, "ax2 must be in dense tail");
  if (ax1 == ax2) fail(SB_E_ARGUMENT, "axes must be distinct");
  i64 a1 = std::min(ax1, ax2), a2 = std::max(ax1, ax2);
  i64 p1 = a1 - P, p2 = a2 - P;
  if (p2 != p1 + 1) {
    std::vector<i64> pb(N2);
    std::iota(pb.begin(), pb.end(), 1);
    i64 x = pb[p2 - 1];
    pb.erase(pb.begin() + (p2 - 1));
    pb.insert(pb.begin() + p1, x);
    std::vector<i64> perm(N);
    std::iota(perm.begin(), perm.end(), 1);
    for (int i = 0; i < N2; ++i) perm[P + i] = P + pb[i];
    bs_permutedims_inplace(A, perm);
    a2 = a1 + 1;
  }
  Tensor B;
  B.kind = BS;
  B.N2 = N2 - 1;
  B.blksize = A.blksize;
  for (int i = 1; i <= N - 1; ++i)
    B.dims.push_back(i < a2 ? (i == a1 ? A.dims[a1 - 1] * A.dims[a2 - 1] : A.dims[i - 1]) : A.dims[i]);
  B.keys = A.keys;
  B.ids = A.ids;
  B.data = A.data;
  return B;
}

// ---------------------------------------------------------------- COO
static Tensor coo_new(const std::vector<i64>& dims) {
  for (auto d : dims) OASSERT(d >= 1, "dims .>= 1");
  Tensor t;
  t.kind = COO;
  t.dims = dims;
  return t;
}

// ColMajorOrder lt (coo/storage.jl:34-46)
static bool colmajor_lt(const i64* a, const i64* b, int N) {
  for (int d = N - 1; d >= 0; --d) {
    if (a[d] < b[d]) return true;
    if (a[d] > b[d]) return false;
  }
  return false;
}
static i64 coo_searchsortedfirst(const Tensor& A, const i64* key) {
  int N = A.N();
  i64 lo = 0, hi = (i64)A.data.size();
  while (lo < hi) {
    i64 mid = (lo + hi) >> 1;
    if (colmajor_lt((A.keys.data() + (mid * N)), key, N))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
// getindex (coo/storage.jl:49-60)
static double coo_getindex(const Tensor& A, const i64* I) {
  int N = A.N();
  for (int d = 0; d < N; ++d)
    if (I[d] < 1 || I[d] > A.dims[d]) fail(SB_E_BOUNDS, "BoundsError");
  i64 idx = coo_searchsortedfirst(A, I);
  if (idx < (i64)A.data.size() && std::equal(I, I + N, (A.keys.data() + (idx * N)))) return A.data[idx];
  return 0.0;
}
// setindex! (coo/storage.jl:84-134)
static void coo_setindex(Tensor& A, double vv, const i64* I, bool dirty_kw) {
  int N = A.N();
  for (int d = 0; d < N; ++d)
    if (I[d] < 1 || I[d] > A.dims[d]) fail(SB_E_BOUNDS, "BoundsError");
  auto push = [&]() {
    A.keys.insert(A.keys.end(), I, I + N);
    A.data.push_back(vv);
  };
  if (A.dirty) {
    if (vv == 0.0) {
      i64 i = 0;
      while (i < (i64)A.data.size()) {
        if (std::equal(I, I + N, (A.keys.data() + (i * N)))) {
          A.keys.erase(A.keys.begin() + i * N, A.keys.begin() + (i + 1) * N);
          A.data.erase(A.data.begin() + i);
        } else
          ++i;
      }
      return;
    }
    push();
    A.dirty = true;
    return;
  }
  i64 idx = coo_searchsortedfirst(A, I);
  if (idx < (i64)A.data.size() && std::equal(I, I + N, (A.keys.data() + (idx * N)))) {
    if (vv == 0.0) {
      A.keys.erase(A.keys.begin() + idx * N, A.keys.begin() + (idx + 1) * N);
      A.data.erase(A.data.begin() + idx);
      A.dirty = true;
    } else
      A.data[idx] = vv;
  } else if (vv != 0.0) {
    if (!dirty_kw) {
      A.keys.insert(A.keys.begin() + idx * N, I, I + N);
      A.data.insert(A.data.begin() + idx, vv);
    } else {
      push();
      A.dirty = true;
    }
  }
}
// _check_perm (coo/storage.jl:136-146)
static void coo_check_perm(const std::vector<i64>& perm, int N) {
  OASSERT((int)perm.size() == N, "length(perm) == N");
  std::vector<bool> seen(N + 1, false);
  for (int j = 0; j < N; ++j) {
    i64 p = perm[j];
    if (!(1 <= p && p <= N)) fail(SB_E_ARGUMENT, "perm out of range");
    if (seen[p]) fail(SB_E_ARGUMENT, "perm has duplicates");
    seen[p] = true;
  }
}
// permutedims (coo/storage.jl:153-166): keys remapped, order kept, dirty
static Tensor coo_permutedims(const Tensor& A, const std::vector<i64>& perm) {
  int N = A.N();
  coo_check_perm(perm, N);
  Tensor B = coo_new(std::vector<i64>(N, 1));
  for (int j = 0; j < N; ++j) B.dims[j] = A.dims[perm[j] - 1];
  i64 n = (i64)A.data.size();
  B.keys.resize(A.keys.size());
  for (i64 i = 0; i < n; ++i)
    for (int j = 0; j < N; ++j) B.keys[i * N + j] = A.keys[i * N + perm[j] - 1];
  B.data = A.data;
  B.dirty = true;
  return B;
}
// sort_coo! (coo/storage.jl:182-189): stable sortperm by reversed tuple
static void coo_sort(Tensor& A) {
  if (!A.dirty) return;
  int N = A.N();
  i64 n = (i64)A.data.size();
  std::vector<i64> p(n);
  std::iota(p.begin(), p.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](i64 a, i64 b) { return colmajor_lt((A.keys.data() + (a * N)), (A.keys.data() + (b * N)), N); });
  std::vector<i64> k2(A.keys.size());
  std::vector<double> v2(n);
  for (i64 i = 0; i < n; ++i) {
    std::copy(&A.keys[p[i] * N], &A.keys[p[i] * N] + N, &k2[i * N]);
    v2[i] = A.data[p[i]];
  }
  A.keys.swap(k2);
  A.data.swap(v2);
  A.dirty = false;
}
// to_dense (coo/conversions.jl:4-10): last write wins
static Tensor coo_to_dense(const Tensor& A) {
  Tensor D;
  D.kind = DENSE;
  D.dims = A.dims;
  D.data.assign(prod(A.dims, 0, A.dims.size()), 0.0);
  int N = A.N();
  for (i64 i = 0; i < (i64)A.data.size(); ++i) {
    i64 off = 0, st = 1;
    for (int d = 0; d < N; ++d) {
      off += (A.keys[i * N + d] - 1) * st;
      st *= A.dims[d];
    }
    D.data[off] = A.data[i];
  }
  return D;
}
// coo_from_dense (coo/conversions.jl:13-26)
static Tensor coo_from_dense(const Tensor& D, double atol, double rtol) {
  Tensor A = coo_new(D.dims);
  int N = D.N();
  for (i64 l = 0; l < (i64)D.data.size(); ++l) {
    double v = D.data[l];
    if (std::fabs(v) > atol + rtol * std::fabs(v)) {
      i64 r = l;
      for (int d = 0; d < N; ++d) {
        A.keys.push_back(r % D.dims[d] + 1);
        r /= D.dims[d];
      }
      A.data.push_back(v);
    }
  }
  return A;
}
// fuse_axes_coo (coo/conversions.jl:174-223)
static Tensor coo_fuse_axes(const Tensor& A, i64 ax1, i64 ax2) {
  int N = A.N();
  if (ax1 == ax2) fail(SB_E_ARGUMENT, "ax1 and ax2 must be distinct");
  if (!(1 <= ax1 && ax1 <= N) || !(1 <= ax2 && ax2 <= N)) fail(SB_E_BOUNDS, "BoundsError");
  OASSERT(std::llabs(ax1 - ax2) == 1, "Currently only adjacent axes can be fused");
  i64 dkeep = A.dims[ax1 - 1];
  i64 ddrop = A.dims[ax2 - 1];
  i64 ax1_out = ax1 - (ax2 < ax1 ? 1 : 0);
  auto old_axis = [&](i64 j) { return j < ax2 ? j : j + 1; };
  Tensor C = coo_new(std::vector<i64>(N - 1, 1));
  for (i64 j = 1; j <= N - 1; ++j) C.dims[j - 1] = (j == ax1_out) ? dkeep * ddrop : A.dims[old_axis(j) - 1];
  i64 n = (i64)A.data.size();
  for (i64 k = 0; k < n; ++k) {
    const i64* c = (A.keys.data() + (k * N));
    i64 fused = (c[ax2 - 1] - 1) * dkeep + c[ax1 - 1];
    for (i64 j = 1; j <= N - 1; ++j) C.keys.push_back(j == ax1_out ? fused : c[old_axis(j) - 1]);
  }
  C.data = A.data;
  C.dirty = true;
  coo_sort(C);
  return C;
}

// ---------------------------------------------------------------- dense helpers
static Tensor dense_permutedims(const Tensor& B, const std::vector<i64>& perm) {
  int N = B.N();
  Tensor R;
  R.kind = DENSE;
  R.dims.resize(N);
  for (int j = 0; j < N; ++j) R.dims[j] = B.dims[perm[j] - 1];
  i64 tot = prod(B.dims, 0, N);
  R.data.resize(tot);
  std::vector<i64> sB(N);
  i64 s = 1;
  for (int d = 0; d < N; ++d) {
    sB[d] = s;
    s *= B.dims[d];
  }
  std::vector<i64> idx(N, 0);
  for (i64 l = 0; l < tot; ++l) {
    i64 src = 0;
    for (int j = 0; j < N; ++j) src += idx[j] * sB[perm[j] - 1];
    R.data[l] = B.data[src];
    for (int j = 0; j < N; ++j) {
      if (++idx[j] < R.dims[j]) break;
      idx[j] = 0;
    }
  }
  return R;
}
static bool is_identity(const std::vector<i64>& p) {
  for (size_t i = 0; i < p.size(); ++i)
    if (p[i] != (i64)i + 1) return false;
  return true;
}
static std::vector<i64> move_axis_to_last_perm(int Nd, int axisr) {  // contract_bs_dense.jl:2-12
  std::vector<i64> v;
  for (int d = 1; d <= Nd; ++d)
    if (d != axisr) v.push_back(d);
  v.push_back(axisr);
  return v;
}

// _find_permutation_for_tensor, single label (contract_bs_bs.jl:33-45)
static std::vector<i64> find_perm_for_tensor(const Tensor& T, const LabelMap& m, const Label& rlab) {
  int P = T.P(), N = T.N(), N2 = T.N2;
  int axisr = at(m, rlab);
  std::vector<i64> perm;
  if (axisr <= P) {
    for (int a = 1; a <= P; ++a)
      if (a != axisr) perm.push_back(a);
    perm.push_back(axisr);
    for (int a = P + 1; a <= N; ++a) perm.push_back(a);
  } else {
    for (int a = 1; a <= P; ++a) perm.push_back(a);
    for (auto d : move_axis_to_last_perm(N2, axisr - P)) perm.push_back(P + d);
  }
  return perm;
}

struct Oper {  // an operand with its labels / map, possibly permuted copy
  const Tensor* t;
  Tensor own;
  Labels labels;
  LabelMap map;
  void set_own() { t = &own; }
};

// _permute_r_to_last_prefix (contract_bs_bs.jl:77-95)
static void permute_r_to_last_prefix(Oper& X, const Label& rlab) {
  int axisr = at(X.map, rlab);
  OASSERT(axisr <= X.t->P(), "rlab must be in prefix to use prefix kernel");
  if (axisr == X.t->P()) return;
  auto perm = find_perm_for_tensor(*X.t, X.map, rlab);
  X.own = bs_permutedims(*X.t, perm);
  X.set_own();
  Labels l2(perm.size());
  for (size_t i = 0; i < perm.size(); ++i) l2[i] = X.labels[perm[i] - 1];
  X.labels = l2;
  X.map = make_map(l2);
}

static Labels sub(const Labels& l, size_t lo, size_t hi) { return Labels(l.begin() + lo, l.begin() + hi); }
static Labels cat(const Labels& a, const Labels& b) {
  Labels r(a);
  r.insert(r.end(), b.begin(), b.end());
  return r;
}

// ---------------------------------------------------------------- BS x BS
// contract_prefix_outer_bb! (contract_bs_bs.jl:173-263)
static void prefix_outer_bb(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PA = A.t->P(), PB = B.t->P(), PC = C.P(), NA = A.t->N(), NB = B.t->N(), NC = C.N();
  Labels Ad = sub(A.labels, PA, NA), Bd = sub(B.labels, PB, NB), Cd = sub(labelsC, PC, NC);
  for (auto& l : Ad) OASSERT(!contains(Bd, l), "outer kernel assumes disjoint dense labels");
  int dense_order;
  if (Cd == cat(Ad, Bd))
    dense_order = 1;
  else if (Cd == cat(Bd, Ad))
    dense_order = 2;
  else
    fail(SB_E_ERROR, "C dense must be [A_dense...,B_dense...] or [B_dense..., A_dense...]");
  OASSERT(C.blksize == A.t->blksize * B.t->blksize, "C.blksize must equal A.blksize * B.blksize");
  int axisAr = at(A.map, rlab), axisBr = at(B.map, rlab);
  OASSERT(axisAr == PA, "axisAr == PA");
  OASSERT(axisBr == PB, "axisBr == PB");
  clear_bs(C);
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    if (has(A.map, labelsC[j]))
      src[j] = at(A.map, labelsC[j]);
    else {
      OASSERT(has(B.map, labelsC[j]), "Label not found in A or B");
      src[j] = -at(B.map, labelsC[j]);
    }
    if ((src[j] > 0 && src[j] > PA) || (src[j] < 0 && -src[j] > PB)) fail(SB_E_BOUNDS, "BoundsError: key tuple");
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 nA = TA.nblocks(), nB = TB.nblocks(), iA = 0, iB = 0;
  std::vector<i64> ck(PC);
  auto rA = [&](i64 i) { return TA.keys[i * PA + axisAr - 1]; };
  auto rB = [&](i64 i) { return TB.keys[i * PB + axisBr - 1]; };
  i64 ma = TA.blksize, nb = TB.blksize;
  while (iA < nA && iB < nB) {
    i64 ra = rA(iA), rb = rB(iB);
    if (ra < rb) {
      while (iA < nA && rA(iA) == ra) ++iA;
      continue;
    } else if (rb < ra) {
      while (iB < nB && rB(iB) == rb) ++iB;
      continue;
    }
    i64 blo = iB;
    while (iB < nB && rB(iB) == ra) ++iB;
    i64 bhi = iB;
    while (iA < nA && rA(iA) == ra) {
      const i64* ak = (TA.keys.data() + (iA) * PA);
      const double* Av = block_ptr(TA, TA.ids[iA]);
      for (i64 j2 = blo; j2 < bhi; ++j2) {
        const i64* bk = (TB.keys.data() + (j2) * PB);
        const double* Bv = block_ptr(TB, TB.ids[j2]);
        for (int j = 0; j < PC; ++j) ck[j] = src[j] > 0 ? ak[src[j] - 1] : bk[-src[j] - 1];
        i64 cid = ensure_block(C, ck.data());
        double* Cv = block_ptr(C, cid);
        // _outer_add! (contract_bs_bs.jl:7-19)
        if (dense_order == 1) {
          for (i64 j = 0; j < nb; ++j)
            for (i64 i = 0; i < ma; ++i) Cv[j * ma + i] += Av[i] * Bv[j];
        } else {
          for (i64 j = 0; j < ma; ++j)
            for (i64 i = 0; i < nb; ++i) Cv[j * nb + i] += Bv[i] * Av[j];
        }
      }
      ++iA;
    }
  }
}

static bool slice_any_nonzero(const double* v, i64 off, i64 len) {  // contract_bs_bs.jl:280-285
  for (i64 i = 0; i < len; ++i)
    if (v[off + i] != 0.0) return true;
  return false;
}

// contract_dense_bb! (contract_bs_bs.jl:288-423)
static void dense_bb(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PA = A.t->P(), PB = B.t->P(), NA = A.t->N(), NB = B.t->N();
  int axisAr = at(A.map, rlab);
  OASSERT(axisAr > PA, "rlab must be in A dense tail");
  int axisBr = at(B.map, rlab);
  OASSERT(axisBr > PB, "rlab must be in B dense tail");
  if (axisAr != NA) {
    auto p = find_perm_for_tensor(*A.t, A.map, rlab);
    A.own = bs_permutedims(*A.t, p);
    A.set_own();
    Labels l2(p.size());
    for (size_t i = 0; i < p.size(); ++i) l2[i] = A.labels[p[i] - 1];
    A.labels = l2;
  }
  if (axisBr != NB) {
    auto p = find_perm_for_tensor(*B.t, B.map, rlab);
    B.own = bs_permutedims(*B.t, p);
    B.set_own();
    Labels l2(p.size());
    for (size_t i = 0; i < p.size(); ++i) l2[i] = B.labels[p[i] - 1];
    B.labels = l2;
  }
  OASSERT(A.labels.back() == rlab, "labelsA[end] == rlab");
  OASSERT(B.labels.back() == rlab, "labelsB[end] == rlab");
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 R = TA.dims[NA - 1];
  OASSERT(TB.dims[NB - 1] == R, "dimsB_dense[end] == R");
  i64 chunkA = TA.blksize / R, chunkB = TB.blksize / R;
  int PC = C.P(), NC = C.N();
  OASSERT(C.N2 == (TA.N2 - 1) + (TB.N2 - 1), "N2C == (N2A-1) + (N2B-1)");
  OASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(B.labels, PB, NB - 1), Cd = sub(labelsC, PC, NC);
  bool AB;
  if (Cd == cat(Awo, Bwo))
    AB = true;
  else if (Cd == cat(Bwo, Awo))
    AB = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense..., B_dense...] or [B_dense..., A_dense...]");
  clear_bs(C);
  // prefix sourcing (contract_bs_bs.jl:343-354); the map of the unpermuted
  // operands is used, which is valid: prefix axes do not move here
  std::vector<int> srcA(PC, 0), srcB(PC, 0);
  for (int j = 0; j < PC; ++j) {
    if (has(A.map, labelsC[j])) {
      int ax = at(A.map, labelsC[j]);
      OASSERT(ax <= PA, "ax <= PA");
      srcA[j] = ax;
    } else {
      int ax = at(B.map, labelsC[j]);
      OASSERT(ax <= PB, "ax <= PB");
      srcB[j] = ax;
    }
  }
  std::vector<i64> ck(PC);
  for (i64 ai = 0; ai < TA.nblocks(); ++ai) {
    const i64* ak = (TA.keys.data() + (ai) * PA);
    const double* Af = block_ptr(TA, TA.ids[ai]);
    for (i64 bi = 0; bi < TB.nblocks(); ++bi) {
      const i64* bk = (TB.keys.data() + (bi) * PB);
      const double* Bf = block_ptr(TB, TB.ids[bi]);
      for (int j = 0; j < PC; ++j) ck[j] = srcA[j] != 0 ? ak[srcA[j] - 1] : bk[srcB[j] - 1];
      double* Cv = nullptr;
      for (i64 r = 0; r < R; ++r) {
        i64 aoff = r * chunkA, boff = r * chunkB;
        if (!slice_any_nonzero(Af, aoff, chunkA)) continue;
        if (!slice_any_nonzero(Bf, boff, chunkB)) continue;
        if (!Cv) Cv = block_ptr(C, ensure_block(C, ck.data()));  // created lazily, once
        if (AB)
          ger_acc(chunkA, chunkB, Af + aoff, Bf + boff, Cv);
        else
          ger_acc(chunkB, chunkA, Bf + boff, Af + aoff, Cv);
      }
    }
  }
}

// contract! BS x BS single label (contract_bs_bs.jl:126-151)
static void contract_bs_bs(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  int axisBr = at(B.map, rlab);
  if (axisBr <= B.t->P()) {
    permute_r_to_last_prefix(A, rlab);
    permute_r_to_last_prefix(B, rlab);
    prefix_outer_bb(C, labelsC, A, B, rlab);
  } else {
    dense_bb(C, labelsC, A, B, rlab);
  }
}

// _cdense_grouping_and_orders (contract_bs_bs.jl:470-505)
enum Mode { AthenB, BthenA, Interleaved };
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
static bool same_set(const Labels& a, const Labels& b) {
  std::set<Label> x(a.begin(), a.end()), y(b.begin(), b.end());
  return x == y;
}

// contract_shared! BS x BS (contract_bs_bs.jl:572-730).  join_exact selects
// the merge-join comparison (reference order when false).
static void shared_bs_bs(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared,
                         bool join_exact) {
  for (auto& l : shared) OASSERT(!contains(labelsC, l), "shared label must be reduced (must not appear in labelsC)");
  int PA = A.t->P(), PB = B.t->P(), NA = A.t->N(), NB = B.t->N();
  Labels sp, sd;
  for (auto& l : shared) {
    OASSERT(has(A.map, l) && has(B.map, l), "shared label must exist in both A and B");
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
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims");
  OASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA");
  OASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB");
  // permutations (contract_bs_bs.jl:513-567)
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
    OASSERT((int)(ka.size() + ra.size()) == X.t->N2, "perm covers dense axes");
    for (auto a : (red_first ? ra : ka)) perm.push_back(P + a);
    for (auto a : (red_first ? ka : ra)) perm.push_back(P + a);
    return perm;
  };
  auto pA = build_perm(A, false, dA);
  auto pB = build_perm(B, true, dB);
  auto apply = [&](Oper& X, const std::vector<i64>& p) {
    X.own = bs_permutedims(*X.t, p);
    X.set_own();
    Labels l2(p.size());
    for (size_t i = 0; i < p.size(); ++i) l2[i] = X.labels[p[i] - 1];
    X.labels = l2;
    X.map = make_map(l2);
  };
  apply(A, pA);
  apply(B, pB);
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
  for (auto& l : sp) {
    jA.push_back(at(A.map, l));
    jB.push_back(at(B.map, l));
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 M = prod(TA.dims, PA, PA + dA.size());
  i64 K = prod(TA.dims, PA + dA.size(), NA);
  i64 Nn = prod(TB.dims, PB + red.size(), NB);
  if (!red.empty()) OASSERT(K == prod(TB.dims, PB, PB + red.size()), "Reduced dense extents mismatch between A and B");
  OASSERT(C.blksize == M * Nn, "C.blksize must equal M*N");
  clear_bs(C);
  i64 nA = TA.nblocks(), nB = TB.nblocks();
  auto eqpos = [](const i64* k0, const i64* k1, const std::vector<int>& pos) {
    for (int p : pos)
      if (k0[p - 1] != k1[p - 1]) return false;
    return true;
  };
  auto advance = [&](const Tensor& T, int P, i64 i0, const std::vector<int>& pos) {  // _advance_run
    i64 n = T.nblocks();
    if (pos.empty()) return n;
    i64 i = i0 + 1;
    while (i < n && eqpos((T.keys.data() + (i0 * P)), (T.keys.data() + (i * P)), pos)) ++i;
    return i;
  };
  // join-tuple compare: reference order compares the first shared axis first
  // (_cmp_join_tuple, contract_bs_bs.jl:430-442); the exact order compares
  // the axis the keys are sorted by first, i.e. the last shared axis.
  auto cmp = [&](const i64* ak, const i64* bk) {
    int n = (int)jA.size();
    for (int t = 0; t < n; ++t) {
      int u = join_exact ? n - 1 - t : t;
      i64 av = ak[jA[u] - 1], bv = bk[jB[u] - 1];
      if (av < bv) return -1;
      if (av > bv) return 1;
    }
    return 0;
  };
  std::vector<i64> ck(PC);
  i64 iA = 0, iB = 0;
  while (iA < nA && iB < nB) {
    int c = jA.empty() ? 0 : cmp((TA.keys.data() + (iA) * PA), (TB.keys.data() + (iB) * PB));
    if (c < 0) {
      iA = advance(TA, PA, iA, jA);
      continue;
    } else if (c > 0) {
      iB = advance(TB, PB, iB, jB);
      continue;
    }
    i64 iA2 = advance(TA, PA, iA, jA), iB2 = advance(TB, PB, iB, jB);
    for (i64 ii = iA; ii < iA2; ++ii) {
      const i64* ak = (TA.keys.data() + (ii) * PA);
      const double* Am = block_ptr(TA, TA.ids[ii]);
      for (i64 jj = iB; jj < iB2; ++jj) {
        const i64* bk = (TB.keys.data() + (jj) * PB);
        const double* Bm = block_ptr(TB, TB.ids[jj]);
        for (int j = 0; j < PC; ++j) ck[j] = src[j] > 0 ? ak[src[j] - 1] : bk[-src[j] - 1];
        double* Cm = block_ptr(C, ensure_block(C, ck.data()));
        if (mode == AthenB)
          gemm_acc('N', 'N', M, Nn, K, Am, M, Bm, K, Cm, M);
        else
          gemm_acc('T', 'T', Nn, M, K, Bm, K, Am, M, Cm, Nn);
      }
    }
    iA = iA2;
    iB = iB2;
  }
}

// ---------------------------------------------------------------- BS x Dense
// contract_prefix_outer_bd! (contract_bs_dense.jl:34-204), Kahan loop :167-196
static void prefix_outer_bd(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PA0 = A.t->P();
  int axisAr = at(A.map, rlab);
  OASSERT(axisAr <= PA0, "rlab must be in A prefix/head");
  int NB = B.t->N();
  int axisBr = at(B.map, rlab);
  OASSERT(1 <= axisBr && axisBr <= NB, "rlab must be a dim of dense B");
  permute_r_to_last_prefix(A, rlab);
  int PA = A.t->P(), NA = A.t->N(), PC = C.P(), NC = C.N();
  OASSERT(at(A.map, rlab) == PA, "mapA[rlab] == PA");
  const Tensor* Bt = B.t;
  Tensor Bown;
  i64 R = Bt->dims[axisBr - 1];
  if (axisBr != NB) {
    Bown = dense_permutedims(*Bt, move_axis_to_last_perm(NB, axisBr));
    Bt = &Bown;
  }
  i64 chunkB = prod(Bt->dims, 0, NB) / R;
  const double* Bmat = Bt->data.data();
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
  OASSERT(C.blksize == chunkA * chunkB, "C.blksize must equal A.blksize * (prod(B dims without r))");
  clear_bs(C);
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    OASSERT(has(A.map, labelsC[j]), "C prefix label not in A");
    int ax = at(A.map, labelsC[j]);
    OASSERT(ax <= PA && labelsC[j] != rlab, "C prefix must be from A prefix excluding rlab");
    src[j] = ax;
  }
  i64 nA = TA.nblocks();
  std::vector<i64> ckeys((size_t)nA * PC);
  for (i64 i = 0; i < nA; ++i)
    for (int j = 0; j < PC; ++j) ckeys[i * PC + j] = TA.keys[i * PA + src[j] - 1];
  std::vector<i64> perm(nA);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](i64 a, i64 b) {
    return std::lexicographical_compare(&ckeys[a * PC], &ckeys[a * PC] + PC, &ckeys[b * PC], &ckeys[b * PC] + PC);
  });
  std::vector<double> Amat((size_t)chunkA * R), Bloc((size_t)chunkB * R), comp((size_t)chunkA * chunkB);
  i64 i = 0;
  while (i < nA) {
    const i64* ck = &ckeys[perm[i] * PC];
    i64 cid = ensure_block(C, ck);
    double* Cv = block_ptr(C, cid);
    i64 j = i + 1;
    while (j < nA && std::equal(ck, ck + PC, &ckeys[perm[j] * PC])) ++j;
    i64 nb = j - i;
    OASSERT(nb <= R, "group size <= R (staging buffer)");
    for (i64 k = 0; k < nb; ++k) {
      i64 ia = perm[i + k];
      i64 rv = TA.keys[ia * PA + PA - 1];
      OASSERT(1 <= rv && rv <= R, "A block key r out of bounds for dense B");
      const double* Av = block_ptr(TA, TA.ids[ia]);
      for (i64 ii = 0; ii < chunkA; ++ii) Amat[ii + k * chunkA] = Av[ii];
      for (i64 jj = 0; jj < chunkB; ++jj) Bloc[jj + k * chunkB] = Bmat[jj + (rv - 1) * chunkB];
    }
    std::fill(comp.begin(), comp.end(), 0.0);
    // Kahan-compensated outer products in the reference's exact operation order
    if (AB) {
      for (i64 k = 0; k < nb; ++k)
        for (i64 jj = 0; jj < chunkB; ++jj) {
          double bval = Bloc[jj + k * chunkB];
          for (i64 ii = 0; ii < chunkA; ++ii) {
            double p = Amat[ii + k * chunkA] * bval;
            double y = p - comp[ii + jj * chunkA];
            double t = Cv[ii + jj * chunkA] + y;
            double d = t - Cv[ii + jj * chunkA];
            comp[ii + jj * chunkA] = d - y;
            Cv[ii + jj * chunkA] = t;
          }
        }
    } else {
      for (i64 k = 0; k < nb; ++k)
        for (i64 jj = 0; jj < chunkA; ++jj) {
          double aval = Amat[jj + k * chunkA];
          for (i64 ii = 0; ii < chunkB; ++ii) {
            double p = Bloc[ii + k * chunkB] * aval;
            double y = p - comp[ii + jj * chunkB];
            double t = Cv[ii + jj * chunkB] + y;
            double d = t - Cv[ii + jj * chunkB];
            comp[ii + jj * chunkB] = d - y;
            Cv[ii + jj * chunkB] = t;
          }
        }
    }
    i = j;
  }
}

// contract_dense_bd! (contract_bs_dense.jl:248-392)
static void dense_bd(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PA = A.t->P(), NA = A.t->N(), NB = B.t->N();
  int axisAr = at(A.map, rlab);
  OASSERT(axisAr > PA, "rlab must be in A dense tail");
  int axisBr = at(B.map, rlab);
  OASSERT(1 <= axisBr && axisBr <= NB, "rlab must be an axis of B");
  if (axisAr != NA) {
    auto p = find_perm_for_tensor(*A.t, A.map, rlab);
    A.own = bs_permutedims(*A.t, p);
    A.set_own();
    Labels l2(p.size());
    for (size_t i = 0; i < p.size(); ++i) l2[i] = A.labels[p[i] - 1];
    A.labels = l2;
  }
  OASSERT(A.labels.back() == rlab, "labelsA[end] == rlab");
  const Tensor* Bt = B.t;
  Tensor Bown;
  Labels Bl = B.labels;
  if (axisBr != NB) {
    auto pB = move_axis_to_last_perm(NB, axisBr);
    Bown = dense_permutedims(*Bt, pB);
    Bt = &Bown;
    Labels l2(NB);
    for (int i = 0; i < NB; ++i) l2[i] = B.labels[pB[i] - 1];
    Bl = l2;
  }
  OASSERT(Bl.back() == rlab, "labelsB[end] == rlab");
  const Tensor& TA = *A.t;
  i64 R = TA.dims[NA - 1];
  OASSERT(Bt->dims[NB - 1] == R, "dimsB_dense[end] == R");
  i64 chunkA = TA.blksize / R;
  i64 chunkB = prod(Bt->dims, 0, NB - 1);
  int PC = C.P(), NC = C.N();
  OASSERT(C.N2 == (TA.N2 - 1) + (NB - 1), "N2C == (N2A - 1) + (NB - 1)");
  OASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(Bl, 0, NB - 1), Cd = sub(labelsC, PC, NC);
  bool AB;
  if (Cd == cat(Awo, Bwo))
    AB = true;
  else if (Cd == cat(Bwo, Awo))
    AB = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense..., B_dense...] or [B_dense..., A_dense...]");
  clear_bs(C);
  std::vector<int> srcA(PC);
  for (int j = 0; j < PC; ++j) {
    int ax = at(A.map, labelsC[j]);
    OASSERT(ax <= PA, "C sparse prefix label must come from A sparse head");
    srcA[j] = ax;
  }
  const double* Bf = Bt->data.data();
  std::vector<i64> ck(PC);
  for (i64 ai = 0; ai < TA.nblocks(); ++ai) {
    const i64* ak = (TA.keys.data() + (ai) * PA);
    const double* Af = block_ptr(TA, TA.ids[ai]);
    for (int j = 0; j < PC; ++j) ck[j] = ak[srcA[j] - 1];
    double* Cv = nullptr;
    for (i64 r = 0; r < R; ++r) {
      i64 aoff = r * chunkA, boff = r * chunkB;
      if (!slice_any_nonzero(Af, aoff, chunkA)) continue;
      if (!slice_any_nonzero(Bf, boff, chunkB)) continue;
      if (!Cv) Cv = block_ptr(C, ensure_block(C, ck.data()));  // created lazily, once
      if (AB)
        ger_acc(chunkA, chunkB, Af + aoff, Bf + boff, Cv);
      else
        ger_acc(chunkB, chunkA, Bf + boff, Af + aoff, Cv);
    }
  }
}

// contract! BS x Dense single label (contract_bs_dense.jl:207-245)
static void contract_bs_dense(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  OASSERT(has(A.map, rlab), "rlab not found in mapA");
  OASSERT(has(B.map, rlab), "rlab not found in mapB");
  int axisAr = at(A.map, rlab);
  if (axisAr <= A.t->P()) {
    permute_r_to_last_prefix(A, rlab);
    prefix_outer_bd(C, labelsC, A, B, rlab);
  } else {
    dense_bd(C, labelsC, A, B, rlab);
  }
}

// contract_shared! BS x Dense (contract_bs_dense.jl:411-619)
static void shared_bs_dense(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared) {
  for (auto& l : shared) OASSERT(!contains(labelsC, l), "shared label must be reduced");
  int PA = A.t->P(), NA = A.t->N(), NB = B.t->N();
  Labels sp, sd;
  for (auto& l : shared) {
    OASSERT(has(A.map, l), "shared label must exist in A");
    OASSERT(has(B.map, l), "shared label must exist in B");
    (at(A.map, l) <= PA ? sp : sd).push_back(l);
  }
  Labels Ad0 = sub(A.labels, PA, NA), red, keepA0, keepB0;
  for (auto& l : Ad0) (contains(sd, l) ? red : keepA0).push_back(l);
  for (auto& l : B.labels)
    if (!contains(shared, l)) keepB0.push_back(l);
  int PC = C.P(), NC = C.N();
  Labels Cd = sub(labelsC, PC, NC), dA, dB;
  Mode mode = cdense_grouping(Cd, keepA0, keepB0, dA, dB);
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims");
  OASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA");
  OASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB");
  // permA (contract_bs_bs.jl:513-536)
  std::vector<i64> spaxA, pA;
  for (auto& l : sp) spaxA.push_back(at(A.map, l));
  for (int a = 1; a <= PA; ++a)
    if (std::find(spaxA.begin(), spaxA.end(), a) == spaxA.end()) pA.push_back(a);
  for (auto a : spaxA) pA.push_back(a);
  {
    std::vector<i64> ka, ra;
    for (auto& l : dA) ka.push_back(at(A.map, l) - PA);
    for (auto& l : red) ra.push_back(at(A.map, l) - PA);
    OASSERT((int)(ka.size() + ra.size()) == A.t->N2, "perm covers dense axes");
    for (auto a : ka) pA.push_back(PA + a);
    for (auto a : ra) pA.push_back(PA + a);
  }
  std::vector<i64> pB;
  for (auto& l : sp) pB.push_back(at(B.map, l));
  for (auto& l : red) pB.push_back(at(B.map, l));
  for (auto& l : dB) pB.push_back(at(B.map, l));
  OASSERT((int)pB.size() == NB, "B perm length mismatch; labelsB must match B ndims");
  if (!is_identity(pA)) {
    A.own = bs_permutedims(*A.t, pA);
    A.set_own();
    Labels l2(pA.size());
    for (size_t i = 0; i < pA.size(); ++i) l2[i] = A.labels[pA[i] - 1];
    A.labels = l2;
    A.map = make_map(l2);
  }
  const Tensor* Bt = B.t;
  Tensor Bown;
  if (!is_identity(pB)) {
    Bown = dense_permutedims(*Bt, pB);
    Bt = &Bown;
  }
  {
    Labels l2(NB);
    for (int i = 0; i < NB; ++i) l2[i] = B.labels[pB[i] - 1];
    B.labels = l2;
    B.map = make_map(l2);
  }
  int nsp = (int)sp.size(), nred = (int)red.size(), nkA = (int)dA.size();
  std::vector<int> csrc(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    OASSERT(has(A.map, l), "C prefix label must exist in A");
    int ap = at(A.map, l);
    OASSERT(ap <= PA, "C prefix label must come from A prefix");
    OASSERT(!contains(sp, l), "C prefix label cannot be a reduced shared_prefix label");
    csrc[j] = ap;
  }
  std::vector<int> jA;
  for (auto& l : sp) jA.push_back(at(A.map, l));
  const Tensor& TA = *A.t;
  i64 M = nkA == 0 ? 1 : prod(TA.dims, PA, PA + nkA);
  i64 K = nred == 0 ? 1 : prod(TA.dims, PA + nkA, NA);
  i64 Nn = (int)dB.size() == 0 ? 1 : prod(Bt->dims, nsp + nred, NB);
  if (nred > 0) OASSERT(K == prod(Bt->dims, nsp, nsp + nred), "Reduced dense extents mismatch between A and B");
  OASSERT(C.blksize == M * Nn, "C.blksize must equal M*N");
  clear_bs(C);
  // B slice at sp_vals: Bsub[k, n] = Bp[sp_lin + S*(k + K*n)], S = prod(sp dims)
  i64 S = prod(Bt->dims, 0, nsp);
  std::vector<double> Bslice((size_t)K * Nn);
  std::vector<i64> ck(PC);
  i64 nA = TA.nblocks(), iA = 0;
  while (iA < nA) {
    i64 iA2 = iA + 1;
    if (jA.empty())
      iA2 = nA;
    else
      while (iA2 < nA) {
        bool eq = true;
        for (int p : jA) eq = eq && TA.keys[iA * PA + p - 1] == TA.keys[iA2 * PA + p - 1];
        if (!eq) break;
        ++iA2;
      }
    const double* Bm;
    if (nsp == 0)
      Bm = Bt->data.data();
    else {
      i64 sl = 0, st = 1;
      for (int t = 0; t < nsp; ++t) {
        sl += (TA.keys[iA * PA + jA[t] - 1] - 1) * st;
        st *= Bt->dims[t];
      }
      for (i64 q = 0; q < K * Nn; ++q) Bslice[q] = Bt->data[sl + S * q];
      Bm = Bslice.data();
    }
    for (i64 ii = iA; ii < iA2; ++ii) {
      const i64* ak = (TA.keys.data() + (ii) * PA);
      const double* Am = block_ptr(TA, TA.ids[ii]);
      for (int j = 0; j < PC; ++j) ck[j] = ak[csrc[j] - 1];
      double* Cm = block_ptr(C, ensure_block(C, ck.data()));
      if (mode == AthenB)
        gemm_acc('N', 'N', M, Nn, K, Am, M, Bm, K, Cm, M);
      else
        gemm_acc('T', 'T', Nn, M, K, Bm, K, Am, M, Cm, Nn);
    }
    iA = iA2;
  }
}

// ---------------------------------------------------------------- COO kernels
// ensure_rlab_last_and_sorted (contract_coo.jl:48-82)
static void coo_rlab_last_sorted(Oper& X, const Label& rlab) {
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
  X.own = coo_permutedims(*X.t, perm);
  X.own.dirty = true;
  coo_sort(X.own);
  X.set_own();
  X.labels = l2;
  X.map = make_map(l2);
}

// contract! COO x COO (contract_coo.jl:93-170); products are not coalesced
static void contract_coo_coo(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  coo_rlab_last_sorted(A, rlab);
  coo_rlab_last_sorted(B, rlab);
  int NC = C.N(), NA = A.t->N(), NB = B.t->N();
  std::vector<int> srcA(NC), pos(NC);
  for (int t = 0; t < NC; ++t) {
    if (has(A.map, labelsC[t])) {
      srcA[t] = 1;
      pos[t] = at(A.map, labelsC[t]);
    } else {
      srcA[t] = 0;
      pos[t] = at(B.map, labelsC[t]);
    }
  }
  C.keys.clear();
  C.data.clear();
  C.dirty = true;
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 nA = (i64)TA.data.size(), nB = (i64)TB.data.size(), i = 0, j = 0;
  while (i < nA && j < nB) {
    i64 ra = TA.keys[i * NA + NA - 1], rb = TB.keys[j * NB + NB - 1];
    if (ra < rb)
      ++i;
    else if (rb < ra)
      ++j;
    else {
      i64 jlo = j;
      while (j < nB && TB.keys[j * NB + NB - 1] == rb) ++j;
      while (i < nA && TA.keys[i * NA + NA - 1] == ra) {
        const i64* ac = (TA.keys.data() + (i) * NA);
        double av = TA.data[i];
        for (i64 jj = jlo; jj < j; ++jj) {
          const i64* bc = (TB.keys.data() + (jj) * NB);
          double v = av * TB.data[jj];
          if (v == 0.0) continue;  // _push_term! (contract_coo.jl:84-91)
          for (int t = 0; t < NC; ++t) C.keys.push_back(srcA[t] ? ac[pos[t] - 1] : bc[pos[t] - 1]);
          C.data.push_back(v);
          C.dirty = true;
        }
        ++i;
      }
    }
  }
}

// contract_prefix! COO x BS (contract_coo_bs.jl:51-124)
static void coo_bs_prefix(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  if (C.N2 != B.t->N2) fail(SB_E_METHOD, "MethodError: contract_prefix! needs C and B with the same N2");
  int NA = A.t->N(), PB = B.t->P(), PC = C.P();
  OASSERT(PC == NA + PB - 2, "PC == NA + PB - 2");
  clear_bs(C);
  int axisAr = at(A.map, rlab), axisBr = at(B.map, rlab);
  OASSERT(axisBr <= PB, "axisBr <= PB");
  const Tensor &TA = *A.t, &TB = *B.t;
  OASSERT(C.blksize == TB.blksize, "length(Cvec) == length(Bvec)");
  std::vector<int> srcA(PC, 0), srcB(PC, 0);
  for (int j = 0; j < PC; ++j) {
    if (has(A.map, labelsC[j]))
      srcA[j] = at(A.map, labelsC[j]);
    else {
      int ax = at(B.map, labelsC[j]);
      OASSERT(ax <= PB, "ax <= PB");
      srcB[j] = ax;
    }
  }
  i64 nA = (i64)TA.data.size(), nB = TB.nblocks(), iA = 0, iB = 0;
  auto rA = [&](i64 i) { return TA.keys[i * NA + axisAr - 1]; };
  auto rB = [&](i64 i) { return TB.keys[i * PB + axisBr - 1]; };
  std::vector<i64> ck(PC);
  while (iA < nA && iB < nB) {
    i64 ra = rA(iA), rb = rB(iB);
    if (ra < rb) {
      while (iA < nA && rA(iA) == ra) ++iA;
    } else if (rb < ra) {
      while (iB < nB && rB(iB) == rb) ++iB;
    } else {
      i64 blo = iB;
      while (iB < nB && rB(iB) == ra) ++iB;
      i64 bhi = iB;
      while (iA < nA && rA(iA) == ra) {
        const i64* ac = (TA.keys.data() + (iA) * NA);
        double al = TA.data[iA];
        for (i64 b = blo; b < bhi; ++b) {
          const i64* bk = (TB.keys.data() + (b) * PB);
          for (int j = 0; j < PC; ++j) ck[j] = srcA[j] ? ac[srcA[j] - 1] : bk[srcB[j] - 1];
          double* Cv = block_ptr(C, ensure_block(C, ck.data()));
          const double* Bv = block_ptr(TB, TB.ids[b]);
          for (i64 t = 0; t < C.blksize; ++t) Cv[t] += al * Bv[t];  // _scale_add_block!
        }
        ++iA;
      }
    }
  }
}

// contract_dense_index! COO x BS (contract_coo_bs.jl:140-241)
static void coo_bs_dense_index(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  int PB = B.t->P(), NB = B.t->N(), N2 = B.t->N2, NA = A.t->N(), PC = C.P(), NC = C.N();
  OASSERT(C.N2 == N2 - 1, "N2C == N2 - 1");
  int axisBr = at(B.map, rlab);
  OASSERT(axisBr > PB, "rlab must be in B dense tail");
  int redpos = axisBr - PB;
  int axisAr = at(A.map, rlab);
  OASSERT(axisAr == NA, "Expected A to have rlab as last axis already (axisAr == NA)");
  Labels Bd = sub(B.labels, PB, NB), Bdp = Bd;
  std::vector<i64> pd;
  for (int d = 1; d <= N2; ++d)
    if (d != redpos) pd.push_back(d);
  pd.push_back(redpos);
  if (redpos != N2) {
    std::vector<i64> pg;
    for (int a = 1; a <= PB; ++a) pg.push_back(a);
    for (auto d : pd) pg.push_back(PB + d);
    B.own = bs_permutedims(*B.t, pg);
    B.set_own();
    for (int i = 0; i < N2; ++i) Bdp[i] = Bd[pd[i] - 1];
  }
  OASSERT(Bdp.back() == rlab, "labelsB_dense_p[end] == rlab");
  const Tensor &TA = *A.t, &TB = *B.t;
  i64 dim_r = TB.dims[NB - 1];
  i64 chunk = prod(TB.dims, PB, NB - 1);
  OASSERT(sub(labelsC, PC, NC) == sub(Bdp, 0, N2 - 1), "collect(labelsC[PC+1:NC]) == expectedC_dense");
  for (int i = 0; i < C.N2; ++i) OASSERT(C.dims[PC + i] == TB.dims[PB + i], "C dense dims == B dense dims");
  OASSERT(C.blksize == chunk, "C.blksize == chunklen");
  OASSERT(TB.blksize == chunk * dim_r, "Bp.blksize == chunklen * dim_r");
  clear_bs(C);
  std::vector<int> srcA(PC, 0), srcB(PC, 0);
  for (int j = 0; j < PC; ++j) {
    if (has(A.map, labelsC[j]))
      srcA[j] = at(A.map, labelsC[j]);
    else {
      int ax = at(B.map, labelsC[j]);
      OASSERT(ax <= PB, "Label in C prefix must come from B prefix (not dense)");
      srcB[j] = ax;
    }
  }
  std::vector<double> tmp(chunk);
  std::vector<i64> ck(PC);
  for (i64 i = 0; i < (i64)TA.data.size(); ++i) {
    const i64* ac = (TA.keys.data() + (i) * NA);
    i64 rv = ac[NA - 1];
    double al = TA.data[i];
    OASSERT(1 <= rv && rv <= dim_r, "Reduction index rv out of bounds");
    for (i64 bi = 0; bi < TB.nblocks(); ++bi) {
      const i64* bk = (TB.keys.data() + (bi) * PB);
      const double* Bv = block_ptr(TB, TB.ids[bi]);
      for (int j = 0; j < PC; ++j) ck[j] = srcA[j] ? ac[srcA[j] - 1] : bk[srcB[j] - 1];
      bool anynz = false;
      for (i64 t = 0; t < chunk; ++t) {
        tmp[t] = al * Bv[(rv - 1) * chunk + t];
        anynz = anynz || tmp[t] != 0.0;
      }
      if (!anynz) continue;
      double* Cv = block_ptr(C, ensure_block(C, ck.data()));
      for (i64 t = 0; t < chunk; ++t) Cv[t] += tmp[t];
    }
  }
}

// contract! COO x BS (contract_coo_bs.jl:14-46)
static void contract_coo_bs(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  int axisBr = at(B.map, rlab);
  if (axisBr <= B.t->P()) {
    coo_rlab_last_sorted(A, rlab);
    permute_r_to_last_prefix(B, rlab);
    coo_bs_prefix(C, labelsC, A, B, rlab);
  } else {
    coo_rlab_last_sorted(A, rlab);
    coo_bs_dense_index(C, labelsC, A, B, rlab);
  }
}

// contract! COO x Dense (contract_coo_dense.jl:48-125)
static void contract_coo_dense(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab,
                               bool skip_zero) {
  int PC = C.P(), NC = C.N(), NA = A.t->N(), NB = B.t->N();
  // aligned_A_to_Cprefix (contract_coo_dense.jl:4-18)
  {
    Labels desired = sub(labelsC, 0, PC);
    desired.push_back(rlab);
    LabelMap m = make_map(A.labels);
    OASSERT((int)desired.size() == NA, "permutation length");
    std::vector<i64> perm(NA);
    for (int j = 0; j < NA; ++j) perm[j] = at(m, desired[j]);
    A.own = coo_permutedims(*A.t, perm);
    coo_sort(A.own);
    A.set_own();
    A.labels = desired;
    A.map = make_map(desired);
  }
  OASSERT(A.labels.back() == rlab, "labelsA[end] == rlab");
  Labels dnr;
  for (int j = PC; j < NC; ++j)
    if (labelsC[j] != rlab) dnr.push_back(labelsC[j]);
  int rpos = -1;
  for (int i = 0; i < NB; ++i)
    if (B.labels[i] == rlab) {
      rpos = i + 1;
      break;
    }
  OASSERT(rpos > 0, "rlab not found in labelsB");
  std::vector<i64> perm(NB, 0);
  for (size_t k = 0; k < dnr.size(); ++k) {
    int pos = -1;
    for (int i = 0; i < NB; ++i)
      if (B.labels[i] == dnr[k]) {
        pos = i + 1;
        break;
      }
    OASSERT(pos > 0, "label not found in labelsB");
    if ((int)k >= NB) fail(SB_E_BOUNDS, "BoundsError: perm");
    perm[k] = pos;
  }
  perm[NB - 1] = rpos;
  const Tensor* Bt = B.t;
  Tensor Bown;
  Labels Bl = B.labels;
  if (!is_identity(perm)) {
    for (auto p : perm)
      if (p < 1) fail(SB_E_ARGUMENT, "invalid permutation of B");
    Bown = dense_permutedims(*Bt, perm);
    Bt = &Bown;
    for (int i = 0; i < NB; ++i) Bl[i] = B.labels[perm[i] - 1];
  }
  OASSERT(Bl.back() == rlab, "labelsB[end] == rlab");
  OASSERT(!contains(labelsC, rlab), "rlab must be reduced (not in labelsC)");
  Labels ep = sub(A.labels, 0, NA - 1), ed = sub(Bl, 0, NB - 1);
  OASSERT(PC == (int)ep.size(), "PC mismatch");
  OASSERT(C.N2 == (int)ed.size(), "N2 mismatch");
  OASSERT(sub(labelsC, 0, PC) == ep, "C prefix labels must be A labels without r");
  OASSERT(sub(labelsC, PC, NC) == ed, "C dense labels must be B labels without r");
  i64 R = Bt->dims[NB - 1];
  i64 chunk = prod(Bt->dims, 0, NB - 1);
  OASSERT(C.blksize == chunk, "C.blksize must equal prod(B dims without r)");
  clear_bs(C);
  const Tensor& TA = *A.t;
  const double* Bv = Bt->data.data();
  for (i64 idx = 0; idx < (i64)TA.data.size(); ++idx) {
    const i64* ac = (TA.keys.data() + (idx) * NA);
    double al = TA.data[idx];
    i64 rv = ac[NA - 1];
    if (!(1 <= rv && rv <= R)) continue;
    if (skip_zero) {
      bool nz = false;
      for (i64 t = 0; t < chunk && !nz; ++t) nz = (al * Bv[(rv - 1) * chunk + t]) != 0.0;
      if (!nz) continue;
    }
    double* Cv = block_ptr(C, ensure_block(C, ac));
    for (i64 t = 0; t < chunk; ++t) Cv[t] += al * Bv[(rv - 1) * chunk + t];  // _axpy_chunk!
  }
}

// ================================================================ AliasedBlockSparse
// Block i of an AliasedBlockSparse is scalars[i] * templates[alias_ids[i]]
// (aliased/storage.jl:46-59).  Tensor fields: keys (n x P, sorted by the
// column-major prefix lin), ids = alias_ids (1-based template ids), scalars,
// data = templates (ntemplates x blksize).

// AliasedBlockSparse{T,N,N2}(dims) (aliased/storage.jl:122-132)
static Tensor al_new(const std::vector<i64>& dims, int N2) {
  int N = (int)dims.size();
  OASSERT(0 < N2 && N2 <= N, "0 < N2 <= N");
  for (auto d : dims) OASSERT(d >= 1, "all(dims .>= 1)");
  Tensor t;
  t.kind = ALIASED;
  t.dims = dims;
  t.N2 = N2;
  t.blksize = prod(dims, N - N2, N);
  t.ntemplates = 0;
  return t;
}

static void al_clear(Tensor& C) {  // empty!(templates / keys / alias_ids / scalars), n_templates = 0
  C.data.clear();
  C.ntemplates = 0;
  C.keys.clear();
  C.ids.clear();
  C.scalars.clear();
}

static const double* al_template(const Tensor& A, i64 t) { return A.data.data() + (t - 1) * A.blksize; }

// getindex (aliased/storage.jl:138-147) with _find_aliased_key (:97-111)
static double al_getindex(const Tensor& A, const i64* I) {
  check_index(A, I);
  int P = A.P();
  auto [pos, found] = find_key(A, I);
  if (!found) return 0.0;
  i64 tid = A.ids[pos];
  i64 bi = block_index(A, I);
  return A.scalars[pos] * A.data[(tid - 1) * A.blksize + bi - 1];
  (void)P;
}

// compression_ratio (aliased/storage.jl:163-164)
static double al_compression_ratio(const Tensor& A) {
  return A.ntemplates == 0 ? 1.0 : (double)A.nblocks() / (double)A.ntemplates;
}

// Base.permutedims(::AliasedBlockSparse, perm) (aliased/storage.jl:172-239):
// only the templates are transposed; keys are remapped and stably sorted.
static Tensor al_permutedims(const Tensor& A, const std::vector<i64>& perm) {
  int N = A.N(), P = A.P(), N2 = A.N2;
  check_no_cross_perm(perm, N, P);
  Tensor B;
  B.kind = ALIASED;
  B.N2 = N2;
  B.dims.resize(N);
  for (int j = 0; j < N; ++j) B.dims[j] = A.dims[perm[j] - 1];
  std::vector<i64> perm_block(N2);
  for (int i = 0; i < N2; ++i) perm_block[i] = perm[P + i] - P;
  B.blksize = prod(B.dims, P, N);
  OASSERT(B.blksize == A.blksize, "permutation within tail preserves block size");
  bool ident = true;
  for (int i = 0; i < N2; ++i) ident = ident && perm_block[i] == i + 1;
  B.ntemplates = A.ntemplates;
  if (ident) {
    B.data = A.data;
  } else {
    std::vector<i64> tailA(A.dims.begin() + P, A.dims.end());
    std::vector<i64> map_old = tail_map(tailA, perm_block);
    B.data.resize((size_t)A.ntemplates * B.blksize);
    for (i64 t = 0; t < A.ntemplates; ++t)
      for (i64 nl = 0; nl < B.blksize; ++nl) B.data[t * B.blksize + nl] = A.data[t * A.blksize + map_old[nl] - 1];
  }
  i64 n = A.nblocks();
  std::vector<i64> newkeys((size_t)n * P);
  for (i64 i = 0; i < n; ++i)
    for (int j = 0; j < P; ++j) newkeys[i * P + j] = A.keys[i * P + perm[j] - 1];
  std::vector<i64> p = sortperm_lin(newkeys, n, P, B.dims.data());
  B.keys.resize((size_t)n * P);
  B.ids.resize(n);
  B.scalars.resize(n);
  for (i64 o = 0; o < n; ++o) {
    for (int j = 0; j < P; ++j) B.keys[o * P + j] = newkeys[p[o] * P + j];
    B.ids[o] = A.ids[p[o]];
    B.scalars[o] = A.scalars[p[o]];
  }
  return B;
}

// to_blocksparse (aliased/conversions.jl:18-32)
static Tensor al_to_blocksparse(const Tensor& A) {
  Tensor B;
  B.kind = BS;
  B.dims = A.dims;
  B.N2 = A.N2;
  B.blksize = A.blksize;
  B.keys = A.keys;
  i64 n = A.nblocks();
  B.ids.resize(n);
  B.data.resize((size_t)n * A.blksize);
  for (i64 i = 0; i < n; ++i) {
    B.ids[i] = i + 1;
    const double* t = al_template(A, A.ids[i]);
    for (i64 j = 0; j < A.blksize; ++j) B.data[i * A.blksize + j] = A.scalars[i] * t[j];
  }
  return B;
}

// to_dense (aliased/conversions.jl:45-64)
static Tensor al_to_dense(const Tensor& A, double init) {
  Tensor D;
  D.kind = DENSE;
  D.dims = A.dims;
  int N = A.N(), P = A.P();
  i64 tot = prod(A.dims, 0, N);
  D.data.assign(tot, init);
  i64 pstride = 1;
  std::vector<i64> ps(P);
  for (int d = 0; d < P; ++d) {
    ps[d] = pstride;
    pstride *= A.dims[d];
  }
  for (i64 i = 0; i < A.nblocks(); ++i) {
    i64 base = 0;
    for (int d = 0; d < P; ++d) base += (A.keys[i * P + d] - 1) * ps[d];
    const double* t = al_template(A, A.ids[i]);
    for (i64 j = 0; j < A.blksize; ++j) D.data[base + j * pstride] = A.scalars[i] * t[j];
  }
  return D;
}

// _aliased_contribute! / _commit_aliased_dicts! (contract_aliased.jl:75-146):
// a C key fed by contributions of one combined template stays aliased (its
// scalars summed in order); a contribution of a different template demotes it
// to a concrete accumulated block.
struct AliasState {
  std::map<std::vector<i64>, std::pair<i64, double>> alias;
  std::map<std::vector<i64>, std::vector<double>> accum;
};
static void al_contribute(AliasState& S, const std::vector<double>& Ct, const std::vector<i64>& ckey, i64 tid,
                          double a, i64 bs) {
  auto it = S.accum.find(ckey);
  if (it != S.accum.end()) {
    const double* t = Ct.data() + (tid - 1) * bs;
    for (i64 j = 0; j < bs; ++j) it->second[j] += a * t[j];
    return;
  }
  auto ia = S.alias.find(ckey);
  if (ia != S.alias.end()) {
    i64 tprev = ia->second.first;
    double aprev = ia->second.second;
    if (tprev == tid) {
      ia->second.second = aprev + a;
    } else {
      std::vector<double> acc(bs);
      const double *t0 = Ct.data() + (tprev - 1) * bs, *t1 = Ct.data() + (tid - 1) * bs;
      for (i64 j = 0; j < bs; ++j) acc[j] = aprev * t0[j] + a * t1[j];
      S.alias.erase(ia);
      S.accum.emplace(ckey, std::move(acc));
    }
    return;
  }
  S.alias.emplace(ckey, std::make_pair(tid, a));
}
// Julia appends the accumulated blocks in Dict iteration order; here they are
// appended in column-major key order (their template ids are not pinned).
static void al_commit(Tensor& C, AliasState& S) {
  int PC = C.P();
  for (auto& kv : S.alias) {
    C.keys.insert(C.keys.end(), kv.first.begin(), kv.first.end());
    C.ids.push_back(kv.second.first);
    C.scalars.push_back(kv.second.second);
  }
  std::vector<const std::vector<i64>*> ak;
  for (auto& kv : S.accum) ak.push_back(&kv.first);
  std::stable_sort(ak.begin(), ak.end(), [&](const std::vector<i64>* x, const std::vector<i64>* y) {
    return prefix_lin(x->data(), C.dims.data(), PC) < prefix_lin(y->data(), C.dims.data(), PC);
  });
  for (auto* k : ak) {
    const auto& acc = S.accum[*k];
    C.ntemplates += 1;
    C.data.insert(C.data.end(), acc.begin(), acc.end());
    C.keys.insert(C.keys.end(), k->begin(), k->end());
    C.ids.push_back(C.ntemplates);
    C.scalars.push_back(1.0);
  }
  i64 n = (i64)C.ids.size();
  if (n == 0) return;
  std::vector<i64> p = sortperm_lin(C.keys, n, PC, C.dims.data());
  std::vector<i64> k2((size_t)n * PC), i2(n);
  std::vector<double> s2(n);
  for (i64 o = 0; o < n; ++o) {
    for (int j = 0; j < PC; ++j) k2[o * PC + j] = C.keys[p[o] * PC + j];
    i2[o] = C.ids[p[o]];
    s2[o] = C.scalars[p[o]];
  }
  C.keys.swap(k2);
  C.ids.swap(i2);
  C.scalars.swap(s2);
}

// combined-template registry: first-seen order of the combination key
struct Combined {
  std::map<std::pair<i64, i64>, i64> id;
  i64 get(const std::pair<i64, i64>& k) const {
    auto it = id.find(k);
    return it == id.end() ? 0 : it->second;
  }
};

static void al_apply_perm(Oper& X, const std::vector<i64>& perm) {
  X.own = al_permutedims(*X.t, perm);
  X.set_own();
  Labels l2(perm.size());
  for (size_t i = 0; i < perm.size(); ++i) l2[i] = X.labels[perm[i] - 1];
  X.labels = l2;
  X.map = make_map(l2);
}
// _aliased_r_to_last_prefix / _aliased_r_to_last_dense (contract_aliased.jl:33-65)
static void al_r_to_last_prefix(Oper& X, const Label& rlab) {
  int axisr = at(X.map, rlab), P = X.t->P(), N = X.t->N();
  OASSERT(axisr <= P, "rlab must be in sparse prefix");
  if (axisr == P) return;
  std::vector<i64> perm;
  for (int a = 1; a <= P; ++a)
    if (a != axisr) perm.push_back(a);
  perm.push_back(axisr);
  for (int a = P + 1; a <= N; ++a) perm.push_back(a);
  al_apply_perm(X, perm);
}
static void al_r_to_last_dense(Oper& X, const Label& rlab) {
  int axisr = at(X.map, rlab), P = X.t->P(), N = X.t->N(), N2 = X.t->N2;
  OASSERT(axisr > P, "rlab must be in dense tail");
  if (axisr == N) return;
  int rl = axisr - P;
  std::vector<i64> perm;
  for (int a = 1; a <= P; ++a) perm.push_back(a);
  for (int d = 1; d <= N2; ++d)
    if (d != rl) perm.push_back(P + d);
  perm.push_back(P + rl);
  al_apply_perm(X, perm);
}

// dense B with rlab moved last (contract_aliased.jl:168-177)
static void dense_r_last(Oper& B, const Label& rlab) {
  int NB = B.t->N(), ax = at(B.map, rlab);
  if (ax == NB) return;
  std::vector<i64> perm;
  for (int d = 1; d <= NB; ++d)
    if (d != ax) perm.push_back(d);
  perm.push_back(ax);
  B.own = dense_permutedims(*B.t, perm);
  B.set_own();
  Labels l2(NB);
  for (int i = 0; i < NB; ++i) l2[i] = B.labels[perm[i] - 1];
  B.labels = l2;
  B.map = make_map(l2);
}

// outer(a (m), b (n)) into a zeroed template: t[i + j*m] = 0 + a[i]*b[j] (_outer_add!, contract_bs_bs.jl:7-19)
static void outer_into(std::vector<double>& Ct, i64 off, const double* a, i64 m, const double* b, i64 n) {
  for (i64 j = 0; j < n; ++j)
    for (i64 i = 0; i < m; ++i) Ct[off + i + j * m] += a[i] * b[j];
}

// _contract_aliased_prefix_outer_ad! (contract_aliased.jl:151-233)
static void al_prefix_outer_ad(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  int PA = A.t->P(), NA = A.t->N(), PC = C.P(), NC = C.N();
  OASSERT(at(A.map, rlab) == PA, "rlab must be the last prefix axis of A");
  OASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  i64 R = B.t->dims[at(B.map, rlab) - 1];
  dense_r_last(B, rlab);
  int NB = B.t->N();
  i64 chunkB = prod(B.t->dims, 0, NB - 1);
  OASSERT(C.blksize == A.t->blksize * chunkB, "C.blksize must equal A.blksize * chunkB");
  Labels Ad = sub(A.labels, PA, NA), Bwo = sub(B.labels, 0, NB - 1), Cd = sub(labelsC, PC, NC);
  int order;
  if (Cd == cat(Ad, Bwo))
    order = 1;
  else if (Cd == cat(Bwo, Ad))
    order = 2;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense...,B_wo_r...] or [B_wo_r...,A_dense...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    int ax = at(A.map, labelsC[j]);
    OASSERT(ax <= PA && labelsC[j] != rlab, "C prefix label must come from A sparse prefix excluding rlab");
    src[j] = ax;
  }
  const Tensor& TA = *A.t;
  const double* Bv = B.t->data.data();
  Combined cm;
  AliasState S;
  std::vector<i64> ck(PC);
  for (i64 i = 0; i < TA.nblocks(); ++i) {
    const i64* ak = TA.keys.data() + i * PA;
    i64 rv = ak[PA - 1];
    if (!(1 <= rv && rv <= R)) continue;
    i64 tidA = TA.ids[i];
    double al = TA.scalars[i];
    for (int j = 0; j < PC; ++j) ck[j] = ak[src[j] - 1];
    i64 ct = cm.get({tidA, rv});
    if (ct == 0) {
      C.ntemplates += 1;
      ct = C.ntemplates;
      cm.id[{tidA, rv}] = ct;
      i64 off = (i64)C.data.size();
      C.data.resize(off + C.blksize, 0.0);
      const double* ta = al_template(TA, tidA);
      const double* bs = Bv + (rv - 1) * chunkB;
      if (order == 1)
        outer_into(C.data, off, ta, TA.blksize, bs, chunkB);
      else
        outer_into(C.data, off, bs, chunkB, ta, TA.blksize);
    }
    al_contribute(S, C.data, ck, ct, al, C.blksize);
  }
  al_commit(C, S);
}

// _contract_aliased_dense_ad! (contract_aliased.jl:238-365)
static void al_dense_ad(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  int PA = A.t->P(), NA = A.t->N(), PC = C.P(), NC = C.N();
  OASSERT(at(A.map, rlab) == NA, "rlab must be the last axis of A");
  OASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  OASSERT(PC == PA, "C sparse prefix must equal A sparse prefix for dense-tail reduction");
  dense_r_last(B, rlab);
  int NB = B.t->N();
  i64 R = B.t->dims[NB - 1];
  i64 chunkB = prod(B.t->dims, 0, NB - 1);
  OASSERT(A.t->dims[NA - 1] == R, "Reduction size R mismatch between A dense tail and B");
  i64 chunkA = A.t->blksize / R;
  OASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(B.labels, 0, NB - 1), Cd = sub(labelsC, PC, NC);
  bool AB;
  if (Cd == cat(Awo, Bwo))
    AB = true;
  else if (Cd == cat(Bwo, Awo))
    AB = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense_wo_r...,B_wo_r...] or [B_wo_r...,A_dense_wo_r...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    int ax = at(A.map, labelsC[j]);
    OASSERT(ax <= PA, "C prefix label must come from A sparse prefix");
    src[j] = ax;
  }
  const Tensor& TA = *A.t;
  const double* Bm = B.t->data.data();  // (chunkB, R)
  std::vector<i64> ctid(TA.ntemplates);
  for (i64 t = 1; t <= TA.ntemplates; ++t) {
    C.ntemplates += 1;
    ctid[t - 1] = C.ntemplates;
    i64 off = (i64)C.data.size();
    C.data.resize(off + C.blksize, 0.0);
    const double* Am = al_template(TA, t);  // (chunkA, R)
    if (AB)
      gemm_acc('N', 'T', chunkA, chunkB, R, Am, chunkA, Bm, chunkB, C.data.data() + off, chunkA);
    else
      gemm_acc('N', 'T', chunkB, chunkA, R, Bm, chunkB, Am, chunkA, C.data.data() + off, chunkB);
  }
  i64 n = TA.nblocks();
  C.keys.resize((size_t)n * PC);
  C.ids.resize(n);
  C.scalars.resize(n);
  for (i64 i = 0; i < n; ++i) {
    for (int j = 0; j < PC; ++j) C.keys[i * PC + j] = TA.keys[i * PA + src[j] - 1];
    C.ids[i] = ctid[TA.ids[i] - 1];
    C.scalars[i] = TA.scalars[i];
  }
  if (n) {
    std::vector<i64> p = sortperm_lin(C.keys, n, PC, C.dims.data());
    std::vector<i64> k2((size_t)n * PC), i2(n);
    std::vector<double> s2(n);
    for (i64 o = 0; o < n; ++o) {
      for (int j = 0; j < PC; ++j) k2[o * PC + j] = C.keys[p[o] * PC + j];
      i2[o] = C.ids[p[o]];
      s2[o] = C.scalars[p[o]];
    }
    C.keys.swap(k2);
    C.ids.swap(i2);
    C.scalars.swap(s2);
  }
}

// contract_aliased! Aliased x Dense (contract_aliased.jl:383-408)
static void contract_aliased_ad(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  al_clear(C);
  if (at(A.map, rlab) <= A.t->P()) {
    al_r_to_last_prefix(A, rlab);
    al_prefix_outer_ad(C, labelsC, A, B, rlab);
  } else {
    al_r_to_last_dense(A, rlab);
    al_dense_ad(C, labelsC, A, B, rlab);
  }
}

// _contract_aliased_prefix_outer_aa! (contract_aliased.jl:415-507)
static void al_prefix_outer_aa(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  int PA = A.t->P(), NA = A.t->N(), PB = B.t->P(), NB = B.t->N(), PC = C.P(), NC = C.N();
  OASSERT(at(A.map, rlab) == PA && at(B.map, rlab) == PB, "rlab is the last prefix axis of A and B");
  OASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  OASSERT(C.blksize == A.t->blksize * B.t->blksize, "C.blksize == A.blksize * B.blksize");
  Labels Ad = sub(A.labels, PA, NA), Bd = sub(B.labels, PB, NB), Cd = sub(labelsC, PC, NC);
  int order;
  if (Cd == cat(Ad, Bd))
    order = 1;
  else if (Cd == cat(Bd, Ad))
    order = 2;
  else
    fail(SB_E_ERROR, "C dense must be [A_dense...,B_dense...] or [B_dense...,A_dense...]");
  std::vector<int> src(PC);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    if (has(A.map, l) && at(A.map, l) != at(A.map, rlab)) {
      int ax = at(A.map, l);
      OASSERT(ax <= PA, "C prefix label from A prefix");
      src[j] = ax;
    } else {
      OASSERT(has(B.map, l), "C prefix label not found in A or B");
      int ax = at(B.map, l);
      OASSERT(ax <= PB && l != rlab, "C prefix label from B prefix");
      src[j] = -ax;
    }
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  Combined cm;
  AliasState S;
  std::vector<i64> ck(PC);
  i64 iA = 0, nA = TA.nblocks(), iB = 0, nB = TB.nblocks();
  while (iA < nA && iB < nB) {
    i64 rvA = TA.keys[iA * PA + PA - 1], rvB = TB.keys[iB * PB + PB - 1];
    if (rvA < rvB) {
      while (iA < nA && TA.keys[iA * PA + PA - 1] == rvA) ++iA;
      continue;
    } else if (rvB < rvA) {
      while (iB < nB && TB.keys[iB * PB + PB - 1] == rvB) ++iB;
      continue;
    }
    i64 blo = iB;
    while (iB < nB && TB.keys[iB * PB + PB - 1] == rvA) ++iB;
    i64 bhi = iB;
    while (iA < nA && TA.keys[iA * PA + PA - 1] == rvA) {
      const i64* ak = TA.keys.data() + iA * PA;
      i64 tidA = TA.ids[iA];
      double aA = TA.scalars[iA];
      for (i64 j2 = blo; j2 < bhi; ++j2) {
        const i64* bk = TB.keys.data() + j2 * PB;
        i64 tidB = TB.ids[j2];
        double aC = aA * TB.scalars[j2];
        i64 ct = cm.get({tidA, tidB});
        if (ct == 0) {
          C.ntemplates += 1;
          ct = C.ntemplates;
          cm.id[{tidA, tidB}] = ct;
          i64 off = (i64)C.data.size();
          C.data.resize(off + C.blksize, 0.0);
          const double *ta = al_template(TA, tidA), *tb = al_template(TB, tidB);
          if (order == 1)
            outer_into(C.data, off, ta, TA.blksize, tb, TB.blksize);
          else
            outer_into(C.data, off, tb, TB.blksize, ta, TA.blksize);
        }
        for (int j = 0; j < PC; ++j) ck[j] = src[j] > 0 ? ak[src[j] - 1] : bk[-src[j] - 1];
        al_contribute(S, C.data, ck, ct, aC, C.blksize);
      }
      ++iA;
    }
  }
  al_commit(C, S);
}

// _contract_aliased_dense_aa! (contract_aliased.jl:512-643)
static void al_dense_aa(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  int PA = A.t->P(), NA = A.t->N(), PB = B.t->P(), NB = B.t->N(), PC = C.P(), NC = C.N();
  OASSERT(at(A.map, rlab) == NA && at(B.map, rlab) == NB, "rlab is the last axis of A and B");
  OASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  i64 R = A.t->dims[NA - 1];
  OASSERT(B.t->dims[NB - 1] == R, "Reduction size mismatch");
  i64 chunkA = A.t->blksize / R, chunkB = B.t->blksize / R;
  OASSERT(C.blksize == chunkA * chunkB, "C.blksize == chunkA * chunkB");
  Labels Awo = sub(A.labels, PA, NA - 1), Bwo = sub(B.labels, PB, NB - 1), Cd = sub(labelsC, PC, NC);
  bool AB;
  if (Cd == cat(Awo, Bwo))
    AB = true;
  else if (Cd == cat(Bwo, Awo))
    AB = false;
  else
    fail(SB_E_ERROR, "C dense tail must be [A_dense_wo_r...,B_dense_wo_r...] or swapped");
  std::vector<int> srcA(PC, 0), srcB(PC, 0);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    if (has(A.map, l)) {
      int ax = at(A.map, l);
      OASSERT(ax <= PA, "C prefix label from A prefix");
      srcA[j] = ax;
    } else {
      OASSERT(has(B.map, l), "C prefix label from B");
      int ax = at(B.map, l);
      OASSERT(ax <= PB, "C prefix label from B prefix");
      srcB[j] = ax;
    }
  }
  const Tensor &TA = *A.t, &TB = *B.t;
  Combined cm;
  AliasState S;
  std::vector<i64> ck(PC);
  for (i64 ai = 0; ai < TA.nblocks(); ++ai) {
    const i64* ak = TA.keys.data() + ai * PA;
    i64 tidA = TA.ids[ai];
    double aA = TA.scalars[ai];
    for (i64 bi = 0; bi < TB.nblocks(); ++bi) {
      const i64* bk = TB.keys.data() + bi * PB;
      i64 tidB = TB.ids[bi];
      double aC = aA * TB.scalars[bi];
      i64 ct = cm.get({tidA, tidB});
      if (ct == 0) {
        C.ntemplates += 1;
        ct = C.ntemplates;
        cm.id[{tidA, tidB}] = ct;
        i64 off = (i64)C.data.size();
        C.data.resize(off + C.blksize, 0.0);
        const double *Am = al_template(TA, tidA), *Bm = al_template(TB, tidB);
        if (AB)
          gemm_acc('N', 'T', chunkA, chunkB, R, Am, chunkA, Bm, chunkB, C.data.data() + off, chunkA);
        else
          gemm_acc('N', 'T', chunkB, chunkA, R, Bm, chunkB, Am, chunkA, C.data.data() + off, chunkB);
      }
      for (int j = 0; j < PC; ++j) ck[j] = srcA[j] ? ak[srcA[j] - 1] : bk[srcB[j] - 1];
      al_contribute(S, C.data, ck, ct, aC, C.blksize);
    }
  }
  al_commit(C, S);
}

// contract_aliased! Aliased x Aliased (contract_aliased.jl:659-690)
static void contract_aliased_aa(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  OASSERT(!contains(labelsC, rlab), "rlab not in labelsC");
  al_clear(C);
  if (at(B.map, rlab) <= B.t->P()) {
    al_r_to_last_prefix(A, rlab);
    al_r_to_last_prefix(B, rlab);
    al_prefix_outer_aa(C, labelsC, A, B, rlab);
  } else {
    OASSERT(at(A.map, rlab) > A.t->P(), "rlab must be in A dense tail when it is in B dense tail");
    al_r_to_last_dense(A, rlab);
    al_r_to_last_dense(B, rlab);
    al_dense_aa(C, labelsC, A, B, rlab);
  }
}

// _advance_run (contract_bs_bs.jl:455-464): end (exclusive) of the run of rows
// equal to row i0 on the given key columns (1-based)
static i64 advance_run(const Tensor& T, i64 i0, const std::vector<int>& pos) {
  int P = T.P();
  i64 n = T.nblocks();
  if (pos.empty()) return n;
  i64 i = i0 + 1;
  while (i < n) {
    bool eq = true;
    for (int p : pos) eq = eq && T.keys[i0 * P + p - 1] == T.keys[i * P + p - 1];
    if (!eq) break;
    ++i;
  }
  return i;
}

// contract_shared! Aliased x Dense (contract_aliased_shared.jl:103-302)
static void shared_al_dense(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared) {
  for (auto& l : shared) OASSERT(!contains(labelsC, l), "shared label must be reduced (not in labelsC)");
  al_clear(C);
  int PA = A.t->P(), NA = A.t->N(), NB = B.t->N(), PC = C.P(), NC = C.N();
  Labels sp, sd;
  for (auto& l : shared) {
    OASSERT(has(A.map, l), "shared label must exist in A");
    OASSERT(has(B.map, l), "shared label must exist in B");
    (at(A.map, l) <= PA ? sp : sd).push_back(l);
  }
  Labels Ad0 = sub(A.labels, PA, NA), red, keepA0, keepB0;
  for (auto& l : Ad0) (contains(sd, l) ? red : keepA0).push_back(l);
  for (auto& l : B.labels)
    if (!contains(shared, l)) keepB0.push_back(l);
  Labels Cd = sub(labelsC, PC, NC), dA, dB;
  Mode mode = cdense_grouping(Cd, keepA0, keepB0, dA, dB);
  if (mode == Interleaved) fail(SB_E_ERROR, "Cdense interleaves A/B kept dims; unsupported");
  OASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA");
  OASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB");
  // A: prefix [keep_pref..., shared_pref...], dense [keepA..., red...]
  {
    std::vector<i64> spax, perm;
    for (auto& l : sp) spax.push_back(at(A.map, l));
    for (int a = 1; a <= PA; ++a)
      if (std::find(spax.begin(), spax.end(), a) == spax.end()) perm.push_back(a);
    for (auto a : spax) perm.push_back(a);
    std::vector<i64> ka, ra;
    for (auto& l : dA) ka.push_back(at(A.map, l) - PA);
    for (auto& l : red) ra.push_back(at(A.map, l) - PA);
    OASSERT((int)(ka.size() + ra.size()) == A.t->N2, "perm covers dense axes");
    for (auto a : ka) perm.push_back(PA + a);
    for (auto a : ra) perm.push_back(PA + a);
    if (!is_identity(perm)) al_apply_perm(A, perm);
  }
  // B: (shared_prefix..., red_dense..., keepB...)
  std::vector<i64> permB;
  for (auto& l : sp) permB.push_back(at(B.map, l));
  for (auto& l : red) permB.push_back(at(B.map, l));
  for (auto& l : dB) permB.push_back(at(B.map, l));
  OASSERT((int)permB.size() == NB, "B perm length mismatch; labelsB must match B ndims");
  if (!is_identity(permB)) {
    B.own = dense_permutedims(*B.t, permB);
    B.set_own();
  }
  int nsp = (int)sp.size(), nred = (int)red.size(), nkA = (int)dA.size(), nkB = (int)dB.size();
  const Tensor& TA = *A.t;
  const std::vector<i64>& dBp = B.t->dims;
  i64 M = nkA == 0 ? 1 : prod(TA.dims, PA, PA + nkA);
  i64 K = nred == 0 ? 1 : prod(TA.dims, PA + nkA, NA);
  i64 Nn = nkB == 0 ? 1 : prod(dBp, nsp + nred, NB);
  if (nred > 0) OASSERT(K == prod(dBp, nsp, nsp + nred), "Reduction extent mismatch between A dense tail and B");
  OASSERT(C.blksize == M * Nn, "C.blksize must equal M*N");
  std::vector<int> csrc(PC);
  LabelMap mA = make_map(A.labels);
  for (int j = 0; j < PC; ++j) {
    const Label& l = labelsC[j];
    OASSERT(has(mA, l), "C prefix label must exist in A");
    int ap = at(mA, l);
    OASSERT(ap <= PA, "C prefix label must come from A sparse prefix");
    OASSERT(!contains(sp, l), "C prefix label cannot be a reduced shared-prefix label");
    csrc[j] = ap;
  }
  std::vector<int> jA;
  for (auto& l : sp) jA.push_back(at(mA, l));
  i64 SP = prod(dBp, 0, nsp);
  std::vector<i64> spstr(nsp);
  {
    i64 s = 1;
    for (int t = 0; t < nsp; ++t) {
      spstr[t] = s;
      s *= dBp[t];
    }
  }
  const double* Bv = B.t->data.data();
  Combined cm;
  AliasState S;
  std::vector<i64> ck(PC);
  std::vector<double> Bmat((size_t)K * Nn);
  i64 iA = 0, nA = TA.nblocks();
  while (iA < nA) {
    i64 iA2 = advance_run(TA, iA, jA);
    const i64* a0 = TA.keys.data() + iA * PA;
    i64 sp_lin = 1;
    for (int t = 0; t < nsp; ++t) sp_lin += (a0[jA[t] - 1] - 1) * spstr[t];
    // Bmat(k, n) = Bp[sp..., k, n]
    for (i64 n = 0; n < Nn; ++n)
      for (i64 k = 0; k < K; ++k) Bmat[k + n * K] = Bv[(sp_lin - 1) + SP * (k + K * n)];
    for (i64 ii = iA; ii < iA2; ++ii) {
      const i64* ak = TA.keys.data() + ii * PA;
      i64 tidA = TA.ids[ii];
      i64 ct = cm.get({tidA, sp_lin});
      if (ct == 0) {
        C.ntemplates += 1;
        ct = C.ntemplates;
        cm.id[{tidA, sp_lin}] = ct;
        i64 off = (i64)C.data.size();
        C.data.resize(off + C.blksize, 0.0);
        const double* Am = al_template(TA, tidA);  // (M, K)
        if (mode == AthenB)
          gemm_acc('N', 'N', M, Nn, K, Am, M, Bmat.data(), K, C.data.data() + off, M);
        else
          gemm_acc('T', 'T', Nn, M, K, Bmat.data(), K, Am, M, C.data.data() + off, Nn);
      }
      for (int j = 0; j < PC; ++j) ck[j] = ak[csrc[j] - 1];
      al_contribute(S, C.data, ck, ct, TA.scalars[ii], C.blksize);
    }
    iA = iA2;
  }
  al_commit(C, S);
}

// contract_shared! Aliased x Aliased (contract_aliased_shared.jl:325-523)
static void shared_al_al(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Labels& shared,
                         bool join_exact) {
  for (auto& l : shared) OASSERT(!contains(labelsC, l), "shared label must be reduced (not in labelsC)");
  al_clear(C);
  int PA = A.t->P(), NA = A.t->N(), PB = B.t->P(), NB = B.t->N(), PC = C.P(), NC = C.N();
  Labels sp, sd;
  for (auto& l : shared) {
    OASSERT(has(A.map, l) && has(B.map, l), "shared label must exist in both A and B");
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
  OASSERT(dA.size() == keepA0.size() && same_set(dA, keepA0), "desired_keepA");
  OASSERT(dB.size() == keepB0.size() && same_set(dB, keepB0), "desired_keepB");
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
    OASSERT((int)(ka.size() + ra.size()) == X.t->N2, "perm covers dense axes");
    for (auto a : (red_first ? ra : ka)) perm.push_back(P + a);
    for (auto a : (red_first ? ka : ra)) perm.push_back(P + a);
    return perm;
  };
  auto pA = build_perm(A, false, dA);
  auto pB = build_perm(B, true, dB);
  al_apply_perm(A, pA);
  al_apply_perm(B, pB);
  const Tensor &TA = *A.t, &TB = *B.t;
  int nkA = (int)dA.size(), nkB = (int)dB.size(), nred = (int)red.size();
  i64 M = nkA == 0 ? 1 : prod(TA.dims, PA, PA + nkA);
  i64 K = nred == 0 ? 1 : prod(TA.dims, PA + nkA, NA);
  i64 Nn = nkB == 0 ? 1 : prod(TB.dims, PB + nred, NB);
  if (nred > 0) OASSERT(K == prod(TB.dims, PB, PB + nred), "Reduction extent mismatch between A and B dense tails");
  OASSERT(C.blksize == M * Nn, "C.blksize must equal M*N");
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
  for (auto& l : sp) {
    jA.push_back(at(A.map, l));
    jB.push_back(at(B.map, l));
  }
  auto cmp = [&](const i64* ak, const i64* bk) {  // _cmp_join_tuple (contract_bs_bs.jl:430-442)
    int n = (int)jA.size();
    for (int t = 0; t < n; ++t) {
      int u = join_exact ? n - 1 - t : t;
      i64 av = ak[jA[u] - 1], bv = bk[jB[u] - 1];
      if (av < bv) return -1;
      if (av > bv) return 1;
    }
    return 0;
  };
  Combined cm;
  AliasState S;
  std::vector<i64> ck(PC);
  i64 iA = 0, nA = TA.nblocks(), iB = 0, nB = TB.nblocks();
  while (iA < nA && iB < nB) {
    int c = jA.empty() ? 0 : cmp(TA.keys.data() + iA * PA, TB.keys.data() + iB * PB);
    if (c < 0) {
      iA = advance_run(TA, iA, jA);
      continue;
    } else if (c > 0) {
      iB = advance_run(TB, iB, jB);
      continue;
    }
    i64 iA2 = advance_run(TA, iA, jA), iB2 = advance_run(TB, iB, jB);
    for (i64 ii = iA; ii < iA2; ++ii) {
      const i64* ak = TA.keys.data() + ii * PA;
      i64 tidA = TA.ids[ii];
      double aA = TA.scalars[ii];
      const double* Am = al_template(TA, tidA);  // (M, K)
      for (i64 jj = iB; jj < iB2; ++jj) {
        const i64* bk = TB.keys.data() + jj * PB;
        i64 tidB = TB.ids[jj];
        double aC = aA * TB.scalars[jj];
        i64 ct = cm.get({tidA, tidB});
        if (ct == 0) {
          C.ntemplates += 1;
          ct = C.ntemplates;
          cm.id[{tidA, tidB}] = ct;
          i64 off = (i64)C.data.size();
          C.data.resize(off + C.blksize, 0.0);
          const double* Bm = al_template(TB, tidB);  // (K, N)
          if (mode == AthenB)
            gemm_acc('N', 'N', M, Nn, K, Am, M, Bm, K, C.data.data() + off, M);
          else
            gemm_acc('T', 'T', Nn, M, K, Bm, K, Am, M, C.data.data() + off, Nn);
        }
        for (int j = 0; j < PC; ++j) ck[j] = src[j] > 0 ? ak[src[j] - 1] : bk[-src[j] - 1];
        al_contribute(S, C.data, ck, ct, aC, C.blksize);
      }
    }
    iA = iA2;
    iB = iB2;
  }
  al_commit(C, S);
}

// contract_aliased! COO x Dense -> Aliased (contract_aliased_coo_dense.jl:34-164)
static void coo_dense_aliased(Tensor& C, const Labels& labelsC, Oper& A, Oper& B, const Label& rlab) {
  int PC = C.P(), NC = C.N(), NA = A.t->N(), NB = B.t->N();
  {  // aligned_A_to_Cprefix (contract_coo_dense.jl:4-18)
    Labels desired = sub(labelsC, 0, PC);
    desired.push_back(rlab);
    LabelMap m = make_map(A.labels);
    OASSERT((int)desired.size() == NA, "permutation length");
    std::vector<i64> perm(NA);
    for (int j = 0; j < NA; ++j) perm[j] = at(m, desired[j]);
    A.own = coo_permutedims(*A.t, perm);
    coo_sort(A.own);
    A.set_own();
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
  OASSERT(rpos > 0, "rlab not found in labelsB");
  std::vector<i64> perm(NB, 0);
  for (size_t k = 0; k < dnr.size(); ++k) {
    int pos = -1;
    for (int i = 0; i < NB; ++i)
      if (B.labels[i] == dnr[k]) {
        pos = i + 1;
        break;
      }
    OASSERT(pos > 0, "label not found in labelsB");
    if ((int)k >= NB) fail(SB_E_BOUNDS, "BoundsError: perm");
    perm[k] = pos;
  }
  perm[NB - 1] = rpos;
  if (!is_identity(perm)) {
    for (auto p : perm)
      if (p < 1) fail(SB_E_ARGUMENT, "invalid permutation of B");
    B.own = dense_permutedims(*B.t, perm);
    B.set_own();
  }
  i64 R = B.t->dims[NB - 1];
  i64 bs = prod(B.t->dims, 0, NB - 1);
  OASSERT(C.blksize == bs, "C.blksize must equal prod(B dims without r)");
  const double* Bv = B.t->data.data();
  al_clear(C);
  std::map<i64, i64> rv_tid;
  AliasState S;
  const Tensor& TA = *A.t;
  for (i64 idx = 0; idx < (i64)TA.data.size(); ++idx) {
    const i64* ac = TA.keys.data() + idx * NA;
    double al = TA.data[idx];
    i64 rv = ac[NA - 1];
    if (!(1 <= rv && rv <= R)) continue;
    std::vector<i64> ck(ac, ac + PC);
    const double* bsl = Bv + (rv - 1) * bs;
    auto it = S.accum.find(ck);
    if (it != S.accum.end()) {
      for (i64 j = 0; j < bs; ++j) it->second[j] += al * bsl[j];
      continue;
    }
    auto ia = S.alias.find(ck);
    if (ia != S.alias.end()) {  // a second contribution always demotes
      i64 tprev = ia->second.first;
      double aprev = ia->second.second;
      std::vector<double> acc(bs);
      const double* t0 = C.data.data() + (tprev - 1) * bs;
      for (i64 j = 0; j < bs; ++j) acc[j] = aprev * t0[j] + al * bsl[j];
      S.alias.erase(ia);
      S.accum.emplace(ck, std::move(acc));
      continue;
    }
    i64 tid;
    auto rt = rv_tid.find(rv);
    if (rt == rv_tid.end()) {
      C.ntemplates += 1;
      tid = C.ntemplates;
      rv_tid[rv] = tid;
      C.data.insert(C.data.end(), bsl, bsl + bs);
    } else {
      tid = rt->second;
    }
    S.alias.emplace(ck, std::make_pair(tid, al));
  }
  al_commit(C, S);
}

// ---------------------------------------------------------------- dispatcher
// contract!(C, labelsC, A, labelsA, B, labelsB) (contract.jl:37-104)
static void contract(Tensor& C, const Labels& labelsC, const Tensor& A, const Labels& labelsA, const Tensor& B,
                     const Labels& labelsB, unsigned flags) {
  OASSERT(labelsA.size() == A.dims.size(), "labelsA length must match A rank");
  OASSERT(labelsB.size() == B.dims.size(), "labelsB length must match B rank");
  OASSERT(labelsC.size() == C.dims.size(), "labelsC length must match C rank");
  Oper Ao{&A, Tensor(), labelsA, make_map(labelsA)};
  Oper Bo{&B, Tensor(), labelsB, make_map(labelsB)};
  Labels shared, reduced;
  for (auto& l : labelsA)
    if (has(Bo.map, l)) shared.push_back(l);
  for (auto& l : shared)
    if (!contains(labelsC, l)) reduced.push_back(l);
  if (reduced.size() == 1) {
    OASSERT(shared.size() == 1, "Expected exactly 1 shared label");
    const Label& r = reduced[0];
    i64 da = A.dims[at(Ao.map, r) - 1], db = B.dims[at(Bo.map, r) - 1];
    OASSERT(da == db, "Dim mismatch on shared label");
    OASSERT(!contains(labelsC, r), "Reduction label must NOT appear in output labels");
    if (C.kind == BS && A.kind == BS && B.kind == BS)
      contract_bs_bs(C, labelsC, Ao, Bo, r);
    else if (C.kind == BS && A.kind == BS && B.kind == DENSE)
      contract_bs_dense(C, labelsC, Ao, Bo, r);
    else if (C.kind == COO && A.kind == COO && B.kind == COO)
      contract_coo_coo(C, labelsC, Ao, Bo, r);
    else if (C.kind == BS && A.kind == COO && B.kind == BS)
      contract_coo_bs(C, labelsC, Ao, Bo, r);
    else if (C.kind == BS && A.kind == COO && B.kind == DENSE)
      contract_coo_dense(C, labelsC, Ao, Bo, r, (flags & SB_FLAG_SKIP_ZERO_SLICES) != 0);
    else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == DENSE)  // contract_aliased_shared.jl:535-547
      contract_aliased_ad(C, labelsC, Ao, Bo, r);
    else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == ALIASED)  // contract_aliased_shared.jl:550-562
      contract_aliased_aa(C, labelsC, Ao, Bo, r);
    else if (C.kind == ALIASED && A.kind == COO && B.kind == DENSE)  // tensor_wrappers_aliased.jl:263-275
      coo_dense_aliased(C, labelsC, Ao, Bo, r);
    else
      fail(SB_E_METHOD, "MethodError: no single-label contract! method for these storage kinds");
  } else {
    if (C.kind == BS && A.kind == BS && B.kind == BS)
      shared_bs_bs(C, labelsC, Ao, Bo, shared, (flags & SB_FLAG_EXACT_JOIN) != 0);
    else if (C.kind == BS && A.kind == BS && B.kind == DENSE)
      shared_bs_dense(C, labelsC, Ao, Bo, shared);
    else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == DENSE)
      shared_al_dense(C, labelsC, Ao, Bo, shared);
    else if (C.kind == ALIASED && A.kind == ALIASED && B.kind == ALIASED)
      shared_al_al(C, labelsC, Ao, Bo, shared, (flags & SB_FLAG_EXACT_JOIN) != 0);
    else
      fail(SB_E_METHOD, "MethodError: no contract_shared! method for these storage kinds");
  }
}

}  // namespace orc

// ================================================================ C ABI
using namespace orc;
static thread_local std::string g_err;

#define ORC_TRY(...)                      \
  try {                                   \
    __VA_ARGS__;                          \
    return SB_OK;                         \
  } catch (const SbError& e) {            \
    g_err = e.msg;                        \
    return e.code;                        \
  } catch (const std::bad_alloc&) {       \
    g_err = "out of memory";              \
    return SB_E_NOMEM;                    \
  } catch (const std::exception& e) {     \
    g_err = e.what();                     \
    return SB_E_ERROR;                    \
  }

static std::vector<i64> vec(const int64_t* p, int n) { return std::vector<i64>(p, p + n); }
static Labels labs(const uint64_t* p, int n) {
  Labels l(n);
  for (int i = 0; i < n; ++i) l[i] = Label{p[2 * i], p[2 * i + 1]};
  return l;
}
static Tensor* T(void* h) {
  if (!h) fail(SB_E_HANDLE, "null handle");
  return static_cast<Tensor*>(h);
}

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_set_blas(const char* path) {
  ORC_TRY({
    if (!path || !*path) {
      g_dgemm = nullptr;
      g_dger = nullptr;
      g_setthreads = nullptr;
      return SB_OK;
    }
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) fail(SB_E_ERROR, std::string("dlopen failed: ") + dlerror());
    g_dgemm = (dgemm_fn)dlsym(h, "scipy_dgemm_64_");
    g_dger = (dger_fn)dlsym(h, "scipy_dger_64_");
    g_setthreads = (setthreads_fn)dlsym(h, "scipy_openblas_set_num_threads64_");
    if (!g_dgemm || !g_dger) fail(SB_E_ERROR, "dgemm/dger symbols not found");
  })
}
int orc_blas_threads(int64_t n) {
  ORC_TRY({
    if (g_setthreads) g_setthreads(n);
  })
}

void orc_free(void* h) { delete static_cast<Tensor*>(h); }

int orc_kind(void* h, int* kind) { ORC_TRY(*kind = T(h)->kind) }

int orc_bs_new(int N, int N2, const int64_t* dims, void** out) {
  ORC_TRY(*out = new Tensor(bs_new(vec(dims, N), N2)))
}
int orc_bs_from_parts(int N, int N2, const int64_t* dims, int64_t nblocks, const int64_t* keys, const int64_t* ids,
                      const double* data, void** out) {
  ORC_TRY({
    Tensor t = bs_new(vec(dims, N), N2);
    int P = N - N2;
    t.keys.assign(keys, keys + nblocks * P);
    t.ids.assign(ids, ids + nblocks);
    t.data.assign(data, data + nblocks * t.blksize);
    *out = new Tensor(std::move(t));
  })
}
int orc_bs_info(void* h, int* N, int* N2, int64_t* dims, int64_t* nblocks, int64_t* blksize) {
  ORC_TRY({
    Tensor* t = T(h);
    if (t->kind != BS) fail(SB_E_HANDLE, "not block-sparse");
    *N = t->N();
    *N2 = t->N2;
    if (dims) std::copy(t->dims.begin(), t->dims.end(), dims);
    *nblocks = (int64_t)t->ids.size();
    *blksize = t->blksize;
  })
}
int orc_bs_get(void* h, int64_t* keys, int64_t* ids, double* data) {
  ORC_TRY({
    Tensor* t = T(h);
    std::copy(t->keys.begin(), t->keys.end(), keys);
    std::copy(t->ids.begin(), t->ids.end(), ids);
    std::copy(t->data.begin(), t->data.end(), data);
  })
}
int orc_bs_getindex(void* h, const int64_t* I, double* v) { ORC_TRY(*v = bs_getindex(*T(h), I)) }
int orc_bs_setindex(void* h, double v, const int64_t* I) { ORC_TRY(bs_setindex(*T(h), v, I)) }
int orc_bs_permutedims(void* h, const int64_t* perm, void** out) {
  ORC_TRY(*out = new Tensor(bs_permutedims(*T(h), vec(perm, T(h)->N()))))
}
int orc_bs_permutedims_inplace(void* h, const int64_t* perm) {
  ORC_TRY(bs_permutedims_inplace(*T(h), vec(perm, T(h)->N())))
}
int orc_bs_sort_lex(void* h) { ORC_TRY(bs_sort_lex(*T(h))) }
int orc_bs_is_dense(void* h, int* out) { ORC_TRY(*out = bs_is_dense(*T(h)) ? 1 : 0) }
int orc_bs_from_dense(void* dense, int N2, double atol, double rtol, int dropzeros, void** out) {
  ORC_TRY(*out = new Tensor(bs_from_dense(*T(dense), N2, atol, rtol, dropzeros != 0)))
}
int orc_bs_to_dense(void* h, double init, int ngroups, const int64_t* groups, void** out) {
  ORC_TRY({
    std::vector<std::pair<i64, i64>> g;
    for (int i = 0; i < ngroups; ++i) g.push_back({groups[2 * i], groups[2 * i + 1]});
    *out = new Tensor(bs_to_dense(*T(h), init, g));
  })
}
int orc_bs_fuse_two_axes(void* h, int64_t ax1, int64_t ax2, void** out) {
  ORC_TRY({
    Tensor* t = T(h);
    int N = t->N(), P = t->P();
    if (!(1 <= ax1 && ax1 <= N && 1 <= ax2 && ax2 <= N)) fail(SB_E_ARGUMENT, "axes must be in 1:N");
    if (ax1 == ax2) fail(SB_E_ARGUMENT, "axes must be distinct");
    bool s1 = ax1 <= P, s2 = ax2 <= P;
    if (s1 && s2)
      *out = new Tensor(bs_fuse_sparse(*t, ax1, ax2));
    else if (!s1 && !s2)
      *out = new Tensor(bs_fuse_dense(*t, ax1, ax2));
    else
      fail(SB_E_ARGUMENT, "Cannot fuse across sparse-head and dense-tail axes");
  })
}

int orc_dense_new(int N, const int64_t* dims, const double* data, void** out) {
  ORC_TRY({
    Tensor t;
    t.kind = DENSE;
    t.dims = vec(dims, N);
    t.data.assign(data, data + prod(t.dims, 0, N));
    *out = new Tensor(std::move(t));
  })
}
int orc_dense_info(void* h, int* N, int64_t* dims) {
  ORC_TRY({
    *N = T(h)->N();
    if (dims) std::copy(T(h)->dims.begin(), T(h)->dims.end(), dims);
  })
}
int orc_dense_get(void* h, double* data) { ORC_TRY(std::copy(T(h)->data.begin(), T(h)->data.end(), data)) }

int orc_coo_new(int N, const int64_t* dims, void** out) { ORC_TRY(*out = new Tensor(coo_new(vec(dims, N)))) }
int orc_coo_from_parts(int N, const int64_t* dims, int64_t nnz, const int64_t* keys, const double* vals, int dirty,
                       void** out) {
  ORC_TRY({
    Tensor t = coo_new(vec(dims, N));
    t.keys.assign(keys, keys + nnz * N);
    t.data.assign(vals, vals + nnz);
    t.dirty = dirty != 0;
    *out = new Tensor(std::move(t));
  })
}
int orc_coo_info(void* h, int* N, int64_t* dims, int64_t* nnz, int* dirty) {
  ORC_TRY({
    Tensor* t = T(h);
    if (t->kind != COO) fail(SB_E_HANDLE, "not COO");
    *N = t->N();
    if (dims) std::copy(t->dims.begin(), t->dims.end(), dims);
    *nnz = (int64_t)t->data.size();
    *dirty = t->dirty ? 1 : 0;
  })
}
int orc_coo_get(void* h, int64_t* keys, double* vals) {
  ORC_TRY({
    std::copy(T(h)->keys.begin(), T(h)->keys.end(), keys);
    std::copy(T(h)->data.begin(), T(h)->data.end(), vals);
  })
}
int orc_coo_getindex(void* h, const int64_t* I, double* v) { ORC_TRY(*v = coo_getindex(*T(h), I)) }
int orc_coo_setindex(void* h, double v, const int64_t* I, int dirty) { ORC_TRY(coo_setindex(*T(h), v, I, dirty != 0)) }
int orc_coo_permutedims(void* h, const int64_t* perm, void** out) {
  ORC_TRY(*out = new Tensor(coo_permutedims(*T(h), vec(perm, T(h)->N()))))
}
int orc_coo_permutedims_inplace(void* h, const int64_t* perm) {
  ORC_TRY({
    Tensor r = coo_permutedims(*T(h), vec(perm, T(h)->N()));
    *T(h) = std::move(r);
  })
}
int orc_coo_sort(void* h) { ORC_TRY(coo_sort(*T(h))) }
int orc_coo_to_dense(void* h, void** out) { ORC_TRY(*out = new Tensor(coo_to_dense(*T(h)))) }
int orc_coo_from_dense(void* dense, double atol, double rtol, void** out) {
  ORC_TRY(*out = new Tensor(coo_from_dense(*T(dense), atol, rtol)))
}
int orc_coo_fuse_axes(void* h, int64_t ax1, int64_t ax2, void** out) {
  ORC_TRY(*out = new Tensor(coo_fuse_axes(*T(h), ax1, ax2)))
}

int orc_al_new(int N, int N2, const int64_t* dims, void** out) { ORC_TRY(*out = new Tensor(al_new(vec(dims, N), N2))) }
int orc_al_from_parts(int N, int N2, const int64_t* dims, int64_t ntemplates, const double* templates, int64_t nblocks,
                      const int64_t* keys, const int64_t* alias_ids, const double* scalars, void** out) {
  ORC_TRY({
    Tensor t = al_new(vec(dims, N), N2);
    int P = t.P();
    t.ntemplates = ntemplates;
    t.data.assign(templates, templates + ntemplates * t.blksize);
    t.keys.assign(keys, keys + nblocks * P);
    t.ids.assign(alias_ids, alias_ids + nblocks);
    t.scalars.assign(scalars, scalars + nblocks);
    for (auto a : t.ids) OASSERT(1 <= a && a <= ntemplates, "alias id out of range");
    *out = new Tensor(std::move(t));
  })
}
int orc_al_info(void* h, int* N, int* N2, int64_t* dims, int64_t* nblocks, int64_t* blksize, int64_t* ntemplates) {
  ORC_TRY({
    const Tensor& t = *T(h);
    *N = t.N();
    *N2 = t.N2;
    std::copy(t.dims.begin(), t.dims.end(), dims);
    *nblocks = t.nblocks();
    *blksize = t.blksize;
    *ntemplates = t.ntemplates;
  })
}
int orc_al_get(void* h, int64_t* keys, int64_t* alias_ids, double* scalars, double* templates) {
  ORC_TRY({
    const Tensor& t = *T(h);
    std::copy(t.keys.begin(), t.keys.end(), keys);
    std::copy(t.ids.begin(), t.ids.end(), alias_ids);
    std::copy(t.scalars.begin(), t.scalars.end(), scalars);
    std::copy(t.data.begin(), t.data.end(), templates);
  })
}
int orc_al_getindex(void* h, const int64_t* I, double* v) { ORC_TRY(*v = al_getindex(*T(h), I)) }
int orc_al_permutedims(void* h, const int64_t* perm, void** out) {
  ORC_TRY(*out = new Tensor(al_permutedims(*T(h), vec(perm, T(h)->N()))))
}
int orc_al_to_blocksparse(void* h, void** out) { ORC_TRY(*out = new Tensor(al_to_blocksparse(*T(h)))) }
int orc_al_to_dense(void* h, double init, void** out) { ORC_TRY(*out = new Tensor(al_to_dense(*T(h), init))) }
int orc_al_compression_ratio(void* h, double* out) { ORC_TRY(*out = al_compression_ratio(*T(h))) }

int orc_contract(void* C, const uint64_t* labelsC, int nC, void* A, const uint64_t* labelsA, int nA, void* B,
                 const uint64_t* labelsB, int nB, uint32_t flags) {
  ORC_TRY({
    Tensor *c = T(C), *a = T(A), *b = T(B);
    contract(*c, labs(labelsC, nC), *a, labs(labelsA, nA), *b, labs(labelsB, nB), flags);
  })
}

}  // extern "C"
