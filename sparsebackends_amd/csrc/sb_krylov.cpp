// The Krylov layer above contract!: the vector operations a Krylov solver
// applies to block-sparse tensors (SparseBackends/src/tensor_wrappers.jl:792-923)
// and an FP64 Lanczos eigensolver for the lowest eigenpair of a Hermitian
// operator given as a chain of contraction plans (the eigsolve that two-site
// DMRG runs on H_eff, in the reference KrylovKit's through ITensorMPS).
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>

#include "sb_krylov.h"
#include "sb_graph.h"

namespace sb {

// ---------------------------------------------------------------- kernels
int linalg_nb();
void k_mdot(Ctx*, const double* V, i64 ldv, int k, const double* w, i64 n, bool norm, double* part, double* out);
void k_orth_update(Ctx*, const double* V, i64 ldv, int k, const double* h, double* w, i64 n, double* part,
                   double* norm_out);
void k_combine(Ctx*, const double* V, i64 ldv, int kin, const double* U, int kout, double* out, i64 ldo, i64 n);
bool k_orth_update_dot(Ctx*, const double* V, i64 ldv, int k, const double* h, double* w, i64 n, double* part,
                       double* out);
void k_axpby(Ctx*, i64 n, double a, const double* x, double b, double* y);
void k_gather_blocks_by_key(Ctx*, i64 n, i64 bs, const i32* xids, const i64* off, const double* y, double* out);
void k_match_keys(Ctx*, const i32* xk, i64 nx, const i32* yk, const i32* yids, i64 ny, int P, const i64* pdims,
                  i64 bs, i64* off);
void k_add_blocks(Ctx*, i64 n, i64 bs, const i64* soff, const double* src, const i64* doff, double* dst);
void k_match_rows(Ctx*, const i32* xk, i64 nx, const i32* yk, i64 ny, int P, const i64* pdims, i32* row);
void k_box_exceeds(Ctx*, i64 n, int N2, const i32* S, const i32* ybox, const i32* row, i32* flag);
void k_support(Ctx*, i64 n, int N2, const i32* xbox, const i32* ybox, const i32* row, i32* S);
void k_compact(Ctx*, int mode, i64 nrows, int N2, const i64* tdims, const i32* S, const i64* coff, i64 total,
               const i64* boff, const double* src, double* dst, const i32* crow);
int compact_chunk();

namespace {

template <class F>
void step(Ctx* ctx, const char* name, F&& f) {
  if (!ctx->profiled(name)) {
    f();
    return;
  }
  hipEvent_t a = ctx->event(), b = ctx->event();
  (void)hipEventRecord(a, ctx->stream);
  f();
  (void)hipEventRecord(b, ctx->stream);
  ctx->prof_push({name, a, b});
}

void need_bs(const Tensor& t, const char* what) {
  if (t.kind != BS) fail(SB_E_METHOD, std::string("MethodError: ") + what + " needs a NewBlockSparseSorted tensor");
}

// offsets (elements) in y's payload of the block with the key of each row of
// x, -1 where y has none; returns how many rows of x matched
i64 match_blocks(Ctx* ctx, const Tensor& x, const Tensor& y, DBuf<i64>& off) {
  if (x.dirty || y.dirty) fail(SB_E_ARGUMENT, "ArgumentError: block keys must be sorted and unique");
  off.alloc(ctx, x.n ? x.n : 1);
  std::vector<i64> pd(x.dims.begin(), x.dims.begin() + x.P());
  k_match_keys(ctx, x.keys.get(), x.n, y.keys.get(), y.ids.get(), y.n, x.P(), pd.data(), y.blksize, off.get());
  std::vector<i64> h(x.n);
  if (x.n) HIP_CHECK(hipMemcpyAsync(h.data(), off.get(), x.n * sizeof(i64), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  i64 m = 0;
  for (i64 v : h) m += v >= 0;
  return m;
}

double norm_of(Ctx* ctx, const double* p, i64 n) {
  DBuf<double> part(ctx, linalg_nb()), out(ctx, 1);
  k_mdot(ctx, nullptr, 0, 0, p, n, true, part.get(), out.get());
  double v = 0;
  HIP_CHECK(hipMemcpyAsync(&v, out.get(), sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return std::sqrt(v);
}

TP copy_structure(Ctx* ctx, const Tensor& A, bool copy_data) {
  TP C = bs_make(ctx, A.dims, A.N2);
  const i64 n = A.n;
  const int P = A.P();
  C->n = n;
  C->keys.alloc(ctx, (size_t)n * (P ? P : 1));
  C->ids.alloc(ctx, n ? n : 1);
  C->data.alloc(ctx, (size_t)n * A.blksize);
  if (n) {
    if (P)
      HIP_CHECK(hipMemcpyAsync(C->keys.get(), A.keys.get(), (size_t)n * P * sizeof(i32), hipMemcpyDeviceToDevice,
                               ctx->stream));
    HIP_CHECK(hipMemcpyAsync(C->ids.get(), A.ids.get(), (size_t)n * sizeof(i32), hipMemcpyDeviceToDevice,
                             ctx->stream));
    if (copy_data)
      HIP_CHECK(hipMemcpyAsync(C->data.get(), A.dptr(), (size_t)n * A.blksize * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
    else
      k_zero(ctx, C->data.get(), (size_t)n * A.blksize * sizeof(double));
  }
  return C;
}

}  // namespace

// ---------------------------------------------------------------- vector operations

// LinearAlgebra.norm(w::WrappedBlockSparse) = norm(w.blocksparse.data) (tensor_wrappers.jl:793)
double bs_norm(Ctx* ctx, const Tensor& A) {
  need_bs(A, "norm");
  return norm_of(ctx, A.dptr(), A.n * A.blksize);
}

// Base.copy(::WrappedBlockSparse) (tensor_wrappers.jl:817-823): keys, ids and data copied
TP bs_copy(Ctx* ctx, const Tensor& A) {
  need_bs(A, "copy");
  return copy_structure(ctx, A, true);
}

// _zero_similar(::WrappedBlockSparse) (tensor_wrappers.jl:898-904, via _external_similar :915-923): same keys and ids, zero data
TP bs_similar(Ctx* ctx, const Tensor& A) {
  need_bs(A, "similar");
  return copy_structure(ctx, A, false);
}

// α * wbs (tensor_wrappers.jl:801-807): a new tensor with data α .* data
TP bs_scale(Ctx* ctx, const Tensor& A, double alpha) {
  need_bs(A, "scale");
  TP C = copy_structure(ctx, A, false);
  k_axpby(ctx, A.n * A.blksize, alpha, A.dptr(), 0.0, C->data.get());
  return C;
}

// _external_fill! (tensor_wrappers.jl:850-853): fill!(data, x)
void bs_fill(Ctx* ctx, Tensor& A, double v) {
  need_bs(A, "fill!");
  k_fill_f64(ctx, A.dptr(), A.n * A.blksize, v);
  A.touch();
}

// _external_map_storage! (tensor_wrappers.jl:877-886) -> _apply_elementwise!(f, R, A) (:856-862) with
// f(r, a) = alpha*a + beta*r: position by position over the payloads, which
// must have equal length (the only check the reference makes)
void bs_axpby(Ctx* ctx, double alpha, const Tensor& A, double beta, Tensor& R) {
  need_bs(A, "map!");
  need_bs(R, "map!");
  const i64 nA = A.n * A.blksize, nR = R.n * R.blksize;
  if (nA != nR) fail(SB_E_ERROR, "Mismatched sparsity patterns in _external_map! (WrappedBlockSparse)");
  k_axpby(ctx, nR, alpha, A.dptr(), beta, R.dptr());
  R.touch();
}

// wA + wB (tensor_wrappers.jl:826-840): copy of A, then every block of B in
// B's key order added into the block of the same key, created by
// _ensure_block! (storage.jl:73-83) when A has none: new blocks take the next
// ids in that order and their keys are inserted at the sorted position
TP bs_add(Ctx* ctx, const Tensor& A, const Tensor& B) {
  need_bs(A, "+");
  need_bs(B, "+");
  if (A.N2 != B.N2 || A.dims.size() != B.dims.size())
    fail(SB_E_METHOD, "MethodError: + needs block-sparse tensors of the same N and N2");
  if (A.dims != B.dims) fail(SB_E_ARGUMENT, "ArgumentError: + of block-sparse tensors with different dims");
  const int P = A.P();
  const i64 bs = A.blksize;
  DBuf<i64> offB;  // for each B row: A's block offset or -1
  const i64 matched = match_blocks(ctx, B, A, offB);
  const i64 nnew = B.n - matched;
  const i64 nC = A.n + nnew;
  TP C = bs_make(ctx, A.dims, A.N2);
  C->n = nC;
  C->keys.alloc(ctx, (size_t)nC * (P ? P : 1));
  C->ids.alloc(ctx, nC ? nC : 1);
  C->data.alloc(ctx, (size_t)nC * bs);
  // host merge of the two sorted key lists (keys and ids are small next to the payload)
  std::vector<i32> ka((size_t)A.n * P), kb((size_t)B.n * P), ia(A.n), ib(B.n);
  std::vector<i64> ob(B.n), ofsB(B.n);
  auto down = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  };
  down(ka.data(), A.keys.get(), ka.size() * sizeof(i32));
  down(kb.data(), B.keys.get(), kb.size() * sizeof(i32));
  down(ia.data(), A.ids.get(), ia.size() * sizeof(i32));
  down(ib.data(), B.ids.get(), ib.size() * sizeof(i32));
  down(ob.data(), offB.get(), ob.size() * sizeof(i64));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::vector<i32> kc((size_t)nC * P), ic(nC);
  std::vector<i64> dstB(B.n);  // C offset each B block is added into
  {
    auto lin_less = [&](const i32* u, const i32* v) {  // column-major lin order: last axis most significant
      for (int d = P - 1; d >= 0; --d)
        if (u[d] != v[d]) return u[d] < v[d];
      return false;
    };
    i64 a = 0, c = 0, next = A.n;
    auto take_a = [&] {
      std::copy(ka.data() + a * P, ka.data() + (a + 1) * P, kc.data() + c * P);
      ic[c++] = ia[a++];
    };
    for (i64 b = 0; b < B.n; ++b) {
      const i32* kbb = kb.data() + b * P;
      while (a < A.n && lin_less(ka.data() + a * P, kbb)) take_a();
      if (ob[b] >= 0) {  // A holds this key: it is the next A row
        dstB[b] = ob[b];
        take_a();
      } else {  // _ensure_block! appends a zero block with the next id
        std::copy(kbb, kbb + P, kc.data() + c * P);
        ic[c++] = (i32)(++next);
        dstB[b] = (next - 1) * bs;
      }
    }
    while (a < A.n) take_a();
    SB_ASSERT(c == nC, "bs_add: merge size");
  }
  for (i64 b = 0; b < B.n; ++b) ofsB[b] = (i64)(ib[b] - 1) * bs;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  };
  up(C->keys.get(), kc.data(), kc.size() * sizeof(i32));
  up(C->ids.get(), ic.data(), ic.size() * sizeof(i32));
  if (A.n)
    HIP_CHECK(hipMemcpyAsync(C->data.get(), A.dptr(), (size_t)A.n * bs * sizeof(double), hipMemcpyDeviceToDevice,
                             ctx->stream));
  if (nnew) k_zero(ctx, C->data.get() + A.n * bs, (size_t)nnew * bs * sizeof(double));
  // C block += B block (every C block receives at most one B block)
  if (B.n) {
    DBuf<i64> dsrc(ctx, B.n), ddst(ctx, B.n);
    up(dsrc.get(), ofsB.data(), B.n * sizeof(i64));
    up(ddst.get(), dstB.data(), B.n * sizeof(i64));
    k_add_blocks(ctx, B.n, bs, dsrc.get(), B.dptr(), ddst.get(), C->data.get());
  }
  HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host staging vectors go out of scope
  return C;
}

// inner product <A, B> over blocks of equal key (the value ITensors' inner
// computes by contracting all indices, tensor_wrappers.jl:646-652 P_C = 0 sink)
double bs_dot(Ctx* ctx, const Tensor& A, const Tensor& B) {
  need_bs(A, "dot");
  need_bs(B, "dot");
  if (A.dims != B.dims || A.N2 != B.N2) fail(SB_E_ARGUMENT, "ArgumentError: dot of tensors with different dims");
  DBuf<i64> off;
  match_blocks(ctx, A, B, off);
  const i64 n = A.n * A.blksize;
  DBuf<double> tmp(ctx, n ? n : 1), part(ctx, 2 * linalg_nb()), out(ctx, 1);
  k_gather_blocks_by_key(ctx, A.n, A.blksize, A.ids.get(), off.get(), B.dptr(), tmp.get());
  k_mdot(ctx, A.dptr(), 0, 1, tmp.get(), n, false, part.get(), out.get());
  double v = 0;
  HIP_CHECK(hipMemcpyAsync(&v, out.get(), sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return v;
}

// B's values in A's block structure: a tensor with A's keys and ids whose
// block of key k is B's block of key k (zero where B has none).  The block
// space projection the eigensolver applies to H x (blocks of y absent from x
// are left out), for Krylov loops run outside the library.
TP bs_restrict(Ctx* ctx, const Tensor& A, const Tensor& B) {
  need_bs(A, "restrict");
  need_bs(B, "restrict");
  if (A.dims != B.dims || A.N2 != B.N2) fail(SB_E_ARGUMENT, "ArgumentError: restrict of tensors with different dims");
  DBuf<i64> off;
  match_blocks(ctx, A, B, off);
  TP C = copy_structure(ctx, A, false);
  k_gather_blocks_by_key(ctx, A.n, A.blksize, A.ids.get(), off.get(), B.dptr(), C->data.get());
  return C;
}

// ---------------------------------------------------------------- symmetric eigenproblem (host)

// Symmetric eigenproblem of a small m x m column-major matrix: Householder
// reduction to tridiagonal form, then the implicit QL iteration with
// Wilkinson's shift on the tridiagonal (the method of Bowdler, Martin, Reinsch
// and Wilkinson, 1968), every orthogonal transformation accumulated in U.
// Eigenvalues ascending in w, eigenvectors in U's columns.
void sym_eig(int m, const std::vector<double>& A_in, std::vector<double>& w, std::vector<double>& U) {
  std::vector<double> a(A_in);
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * m]; };
  U.assign((size_t)m * m, 0.0);
  auto Z = [&](int i, int j) -> double& { return U[i + (size_t)j * m]; };
  for (int i = 0; i < m; ++i) Z(i, i) = 1.0;
  std::vector<double> d(m), e(m, 0.0), v(m), p(m);
  // A <- H_k A H_k, H_k = I - 2 v v^T / |v|^2 zeroing column k below its subdiagonal
  for (int k = 0; k + 2 < m; ++k) {
    double sig = 0;
    for (int i = k + 1; i < m; ++i) sig += A(i, k) * A(i, k);
    sig = std::sqrt(sig);
    if (sig == 0.0) continue;
    const double alpha = A(k + 1, k) >= 0 ? -sig : sig;
    double vv = 0;
    for (int i = k + 1; i < m; ++i) {
      v[i] = A(i, k) - (i == k + 1 ? alpha : 0.0);
      vv += v[i] * v[i];
    }
    if (vv == 0.0) continue;
    const double beta = 2.0 / vv;
    double vp = 0;
    for (int i = k + 1; i < m; ++i) {  // p = beta B v over the trailing block B
      double s = 0;
      for (int j = k + 1; j < m; ++j) s += A(i, j) * v[j];
      p[i] = beta * s;
      vp += v[i] * p[i];
    }
    const double K = 0.5 * beta * vp;
    for (int i = k + 1; i < m; ++i) p[i] -= K * v[i];  // q = p - K v
    for (int j = k + 1; j < m; ++j)
      for (int i = k + 1; i < m; ++i) A(i, j) -= v[i] * p[j] + p[i] * v[j];
    A(k + 1, k) = A(k, k + 1) = alpha;
    for (int i = k + 2; i < m; ++i) A(i, k) = A(k, i) = 0.0;
    for (int r = 0; r < m; ++r) {  // U <- U H_k
      double t = 0;
      for (int i = k + 1; i < m; ++i) t += Z(r, i) * v[i];
      t *= beta;
      for (int i = k + 1; i < m; ++i) Z(r, i) -= t * v[i];
    }
  }
  for (int i = 0; i < m; ++i) d[i] = A(i, i);
  for (int i = 0; i + 1 < m; ++i) e[i] = A(i + 1, i);  // e[i] couples rows i and i+1
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < m; ++l) {
    for (int it = 0; it < 64; ++it) {
      int n = l;
      while (n + 1 < m && std::fabs(e[n]) > eps * (std::fabs(d[n]) + std::fabs(d[n + 1]))) ++n;
      if (n == l) break;
      // Wilkinson's shift from the leading 2 x 2 of the unreduced block l..n
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[n] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, sh = 0.0;
      int i = n - 1;
      for (; i >= l; --i) {  // chase the bulge from the bottom of the block up
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // the block splits at i: deflate and restart
          d[i + 1] -= sh;
          e[n] = 0.0;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - sh;
        r = (d[i] - g) * s + 2.0 * c * b;
        sh = s * r;
        d[i + 1] = g + sh;
        g = c * r - b;
        for (int k = 0; k < m; ++k) {
          const double zk1 = Z(k, i + 1), zk = Z(k, i);
          Z(k, i + 1) = s * zk + c * zk1;
          Z(k, i) = c * zk - s * zk1;
        }
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= sh;
      e[l] = g;
      e[n] = 0.0;
    }
  }
  std::vector<int> ord(m);
  for (int i = 0; i < m; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
  w.resize(m);
  std::vector<double> Us((size_t)m * m);
  for (int j = 0; j < m; ++j) {
    w[j] = d[ord[j]];
    std::copy(&U[(size_t)ord[j] * m], &U[(size_t)ord[j] * m] + m, &Us[(size_t)j * m]);
  }
  U.swap(Us);
}

// ---------------------------------------------------------------- Lanczos

namespace {

// One Lanczos solve as a resumable state, so that several solves can share
// each operator application (a batch of plan chains) and each host
// synchronisation.  Per iteration: [load(); the operator; applied(); store()]
// unless apply_needed() says H v is already in place, then orth(); a sync;
// redo() (when the operator's output reached beyond the compact layout:
// orth() and a sync again); update() (true when the solve is over).
class Lanczos {
 public:
  Lanczos(Ctx* ctx, Tensor& x, Tensor& y, int kdim, int maxiter, double tol)
      : ctx_(ctx), x_(x), y_(y), kdim_(kdim), maxiter_(maxiter), tol_(tol) {
    need_bs(x, "eigsolve");
    need_bs(y, "eigsolve");
    if (&x == &y) fail(SB_E_ARGUMENT, "ArgumentError: x and y must be distinct tensors");
    if (x.dims != y.dims || x.N2 != y.N2)
      fail(SB_E_ARGUMENT, "ArgumentError: the operator's output must have the vector's dims");
    if (kdim < 2 || kdim > 64) fail(SB_E_ARGUMENT, "ArgumentError: krylovdim must be in 2..64");
    if (maxiter < 1) fail(SB_E_ARGUMENT, "ArgumentError: maxiter must be positive");
    if (!(tol >= 0)) fail(SB_E_ARGUMENT, "ArgumentError: tol must be non-negative");
    nb_ = x.n;
    bs_ = x.blksize;
    nfull_ = nb_ * bs_;
    if (nfull_ == 0) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve of an empty vector");
    N2_ = x.N2;
    td_.assign(x.dims.begin() + x.P(), x.dims.end());
    ld_ = kdim + 1;
  }
  Lanczos(const Lanczos&) = delete;
  Lanczos& operator=(const Lanczos&) = delete;

  // v0 = x / |x| in x's payload, and x's own nonzero extents per block
  void begin() {
    const double b0 = norm_of(ctx_, x_.dptr(), nfull_);
    if (!(b0 > 0) || !std::isfinite(b0))
      fail(SB_E_ARGUMENT, "ArgumentError: eigsolve: initial vector has zero or non-finite norm");
    xb0_.alloc(ctx_, (size_t)(nb_ * N2_ ? nb_ * N2_ : 1));
    DBuf<i32> xf0(ctx_, (size_t)(nb_ ? nb_ : 1));
    box_scan(ctx_, x_.dptr(), x_.ids.get(), nb_, bs_, N2_, td_.data(), xb0_.get(), xf0.get());
    k_axpby(ctx_, nfull_, 1.0 / b0, x_.dptr(), 0.0, x_.dptr());
    x_.touch();
  }

  // after the operator's first application (H v0 in y): the compact layout
  // of the Krylov vectors, per block of x a box S holding every nonzero any
  // Krylov vector has: x0's extents joined with those of H v0, widened later
  // if an application reaches beyond (the padding of symmetry sectors is left
  // out).  When y's boxes are not known (an operator that does not end in a
  // box product) S is the full block.
  void first() {
    ++R_.numops;
    match_blocks(ctx_, x_, y_, off_);  // blocks of y whose key x lacks are left out: the solve stays in x's block space
    track_ = y_.box_trusted();
    yrow_.alloc(ctx_, (size_t)(nb_ ? nb_ : 1));
    S_.alloc(ctx_, (size_t)(nb_ * N2_ ? nb_ * N2_ : 1));
    if (track_) {
      std::vector<i64> pd(x_.dims.begin(), x_.dims.begin() + x_.P());
      k_match_rows(ctx_, x_.keys.get(), nb_, y_.keys.get(), y_.n, x_.P(), pd.data(), yrow_.get());
      k_support(ctx_, nb_, N2_, xb0_.get(), y_.box.get(), yrow_.get(), S_.get());
    } else {
      k_fill_rows(ctx_, S_.get(), nb_, N2_, td_.data());
    }
    std::vector<i32> hid(nb_);
    HIP_CHECK(hipMemcpyAsync(hid.data(), x_.ids.get(), nb_ * sizeof(i32), hipMemcpyDeviceToHost, ctx_->stream));
    std::vector<i64> hxoff(nb_);
    HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    for (i64 i = 0; i < nb_; ++i) hxoff[i] = (i64)(hid[i] - 1) * bs_;
    coff_.alloc(ctx_, nb_ + 1);
    xoff_.alloc(ctx_, nb_ ? nb_ : 1);
    HIP_CHECK(hipMemcpyAsync(xoff_.get(), hxoff.data(), nb_ * sizeof(i64), hipMemcpyHostToDevice, ctx_->stream));
    n_ = layout();
    if (n_ == 0) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve: the vector and its image are empty");
    const int NBP = linalg_nb();
    V_.alloc(ctx_, (size_t)(kdim_ + 1) * n_);
    part_.alloc(ctx_, (size_t)(kdim_ + 2) * NBP);
    small_.alloc(ctx_, 3 * (size_t)ld_ + 4);  // h1, h2, the norm, and the box flag's word
    dU_.alloc(ctx_, (size_t)kdim_ * kdim_);
    // v0 and H v0 in the compact layout
    k_compact(ctx_, 0, nb_, N2_, td_.data(), S_.get(), coff_.get(), n_, xoff_.get(), x_.dptr(), vec(0), crow_.get());
    k_compact(ctx_, 0, nb_, N2_, td_.data(), S_.get(), coff_.get(), n_, off_.get(), y_.dptr(), vec(1), crow_.get());
    have_first_ = true;  // H v0 is in vec(1) already
    T_.assign((size_t)ld_ * ld_, 0.0);
    hh_.assign(3 * (size_t)ld_ + 4, 0.0);
  }

  // whether this iteration applies the operator (the first one has H v0)
  bool apply_needed() {
    if (have_first_) {
      have_first_ = false;
      return false;
    }
    return true;
  }
  bool tracking() const { return track_; }

  // the basis vector into x's payload (entries outside S stay zero: x0 is zero there)
  void load() {
    step(ctx_, "lanczos_load", [&] {
      k_compact(ctx_, 1, nb_, N2_, td_.data(), S_.get(), coff_.get(), n_, xoff_.get(), vec(j_), x_.dptr(), crow_.get());
    });
    x_.touch();
  }
  // after the operator: y's boxes checked against S; the flag sits in the
  // last word of the small buffer that orth() copies to the host, and is read
  // after the step's sync by redo()
  void applied() {
    ++R_.numops;
    if (track_) {
      k_box_exceeds(ctx_, nb_, N2_, S_.get(), y_.box.get(), yrow_.get(), flag_word());
      pending_ = true;
    }
  }
  // y's blocks by key into the compact layout with the current S, queued at
  // once: whether y reached beyond S is known after the step's sync
  void store() {
    step(ctx_, "lanczos_gather", [&] {
      k_compact(ctx_, 0, nb_, N2_, td_.data(), S_.get(), coff_.get(), n_, off_.get(), y_.dptr(), vec(j_ + 1),
                crow_.get());
    });
  }
  // after the sync that follows orth(): true when this step's operator output
  // reached beyond S (rare: the support only grows while the Krylov space
  // does); S is then widened and y gathered again, and the caller
  // orthogonalises once more and syncs before update()
  bool redo() {
    if (!pending_) return false;
    pending_ = false;
    i32 hv = 0;
    std::memcpy(&hv, &hh_[3 * (size_t)ld_ + 3], sizeof(i32));
    if (!hv) return false;
    widen(j_ + 1);
    store();
    return true;
  }

  // classical Gram-Schmidt, twice, of vec(j+1) against v_0 .. v_j; the
  // coefficients reach the host after the next sync
  void orth() {
    double* w = vec(j_ + 1);
    double* h1 = small_.get();
    double* h2 = h1 + ld_ + 1;
    double* nr = h2 + ld_ + 1;
    step(ctx_, "lanczos_orth", [&] {
      k_mdot(ctx_, V_.get(), n_, j_ + 1, w, n_, true, part_.get(), h1);
      // the first update and the second pass's dot products in one read of the basis
      if (!k_orth_update_dot(ctx_, V_.get(), n_, j_ + 1, h1, w, n_, part_.get(), h2)) {
        k_orth_update(ctx_, V_.get(), n_, j_ + 1, h1, w, n_, part_.get(), nullptr);
        k_mdot(ctx_, V_.get(), n_, j_ + 1, w, n_, false, part_.get(), h2);
      }
      k_orth_update(ctx_, V_.get(), n_, j_ + 1, h2, w, n_, part_.get(), nr);
    });
    HIP_CHECK(hipMemcpyAsync(hh_.data(), small_.get(), hh_.size() * sizeof(double), hipMemcpyDeviceToHost,
                             ctx_->stream));
  }

  // the projected matrix, its lowest Ritz pair and the residual; then
  // converged (or broken down), next vector, or thick restart.  true: done
  bool update() {
    double* w = vec(j_ + 1);
    const double* H1 = hh_.data();
    const double* H2 = H1 + ld_ + 1;
    const double nAv = std::sqrt(H1[j_ + 1]);
    beta_ = std::sqrt(hh_[2 * (ld_ + 1)]);
    if (!std::isfinite(beta_) || !std::isfinite(nAv))
      fail(SB_E_ERROR, "eigsolve: the operator produced non-finite values");
    for (int i = 0; i <= j_; ++i) T_[i + (size_t)j_ * ld_] = T_[j_ + (size_t)i * ld_] = H1[i] + H2[i];
    m_ = j_ + 1;
    std::vector<double> Tm((size_t)m_ * m_);
    for (int c = 0; c < m_; ++c)
      for (int r = 0; r < m_; ++r) Tm[r + (size_t)c * m_] = T_[r + (size_t)c * ld_];
    sym_eig(m_, Tm, theta_, U_);
    const bool breakdown = beta_ <= 1e-13 * std::max(nAv, 1e-300);
    resid_ = breakdown ? 0.0 : beta_ * std::fabs(U_[(m_ - 1) + 0 * (size_t)m_]);
    if (!breakdown) k_axpby(ctx_, n_, 1.0 / beta_, w, 0.0, w);
    if (resid_ <= tol_ || breakdown) {
      converged_ = true;
      return true;
    }
    if (m_ < kdim_) {
      ++j_;
      return false;
    }
    if (cycles_ == maxiter_) return true;
    // thick restart: keep the lowest Ritz vectors, then the residual direction
    const int keep = std::max(1, std::min(kdim_ - 1, (3 * kdim_) / 5));
    step(ctx_, "lanczos_restart", [&] {
      HIP_CHECK(hipMemcpyAsync(dU_.get(), U_.data(), (size_t)m_ * keep * sizeof(double), hipMemcpyHostToDevice,
                               ctx_->stream));
      if (!tmp_.get()) tmp_.alloc(ctx_, (size_t)keep * n_);
      k_combine(ctx_, V_.get(), n_, m_, dU_.get(), keep, tmp_.get(), n_, n_);
      HIP_CHECK(hipMemcpyAsync(V_.get(), tmp_.get(), (size_t)keep * n_ * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx_->stream));
      HIP_CHECK(hipMemcpyAsync(vec(keep), vec(m_), n_ * sizeof(double), hipMemcpyDeviceToDevice, ctx_->stream));
    });
    std::fill(T_.begin(), T_.end(), 0.0);
    for (int i = 0; i < keep; ++i) {
      T_[i + (size_t)i * ld_] = theta_[i];
      T_[keep + (size_t)i * ld_] = T_[i + (size_t)keep * ld_] = beta_ * U_[(m_ - 1) + (size_t)i * m_];
    }
    HIP_CHECK(hipStreamSynchronize(ctx_->stream));  // U is reused on the host
    j_ = keep;
    ++cycles_;
    return false;
  }

  // the Ritz vector of the lowest Ritz value into x, normalised
  EigResult finish() {
    DBuf<double> ritzv(ctx_, (size_t)n_);
    step(ctx_, "lanczos_ritz", [&] {
      HIP_CHECK(
          hipMemcpyAsync(dU_.get(), U_.data(), (size_t)m_ * sizeof(double), hipMemcpyHostToDevice, ctx_->stream));
      k_combine(ctx_, V_.get(), n_, m_, dU_.get(), 1, ritzv.get(), n_, n_);
    });
    const double xn = norm_of(ctx_, ritzv.get(), n_);
    if (xn > 0) k_axpby(ctx_, n_, 1.0 / xn, ritzv.get(), 0.0, ritzv.get());
    k_compact(ctx_, 1, nb_, N2_, td_.data(), S_.get(), coff_.get(), n_, xoff_.get(), ritzv.get(), x_.dptr(),
              crow_.get());
    x_.touch();
    HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    R_.value = theta_[0];
    R_.ritz = theta_;
    R_.normres = resid_;
    R_.numiter = cycles_;
    R_.converged = converged_;
    return R_;
  }

 private:
  double* vec(int i) { return V_.get() + (i64)i * n_; }
  i32* flag_word() { return reinterpret_cast<i32*>(small_.get() + 3 * (size_t)ld_ + 3); }

  i64 layout() {  // coff from S; returns the length
    std::vector<i32> hS((size_t)nb_ * N2_);
    if (!hS.empty())
      HIP_CHECK(hipMemcpyAsync(hS.data(), S_.get(), hS.size() * sizeof(i32), hipMemcpyDeviceToHost, ctx_->stream));
    HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    std::vector<i64> hc(nb_ + 1, 0);
    for (i64 i = 0; i < nb_; ++i) {
      i64 c = 1;
      for (int a = 0; a < N2_; ++a) c *= hS[i * N2_ + a];
      hc[i + 1] = hc[i] + c;
    }
    HIP_CHECK(hipMemcpyAsync(coff_.get(), hc.data(), (nb_ + 1) * sizeof(i64), hipMemcpyHostToDevice, ctx_->stream));
    // per chunk of the compact kernel: the row holding its first entry
    const i64 ch = compact_chunk(), nch = (hc[nb_] + ch - 1) / ch;
    std::vector<i32> hr((size_t)std::max<i64>(nch, 1), 0);
    for (i64 c = 0, i = 0; c < nch; ++c) {
      while (i + 1 < nb_ && hc[i + 1] <= c * ch) ++i;
      hr[(size_t)c] = (i32)i;
    }
    crow_.alloc(ctx_, hr.size());
    HIP_CHECK(hipMemcpyAsync(crow_.get(), hr.data(), hr.size() * sizeof(i32), hipMemcpyHostToDevice, ctx_->stream));
    HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    return hc[nb_];
  }

  // S widened to y's boxes where an application reached beyond it: the
  // vectors 0..nv-1 move to the new layout (they are zero in the new entries)
  void widen(int nv) {
    DBuf<double> full(ctx_, (size_t)nfull_);
    DBuf<i32> S2(ctx_, (size_t)(nb_ * N2_ ? nb_ * N2_ : 1));
    k_support(ctx_, nb_, N2_, S_.get(), y_.box.get(), yrow_.get(), S2.get());
    DBuf<i64> coff_old(ctx_, nb_ + 1);
    HIP_CHECK(hipMemcpyAsync(coff_old.get(), coff_.get(), (nb_ + 1) * sizeof(i64), hipMemcpyDeviceToDevice,
                             ctx_->stream));
    const i64 n_old = n_;
    DBuf<i32> S_old = std::move(S_);
    S_ = std::move(S2);
    n_ = layout();
    DBuf<double> V2(ctx_, (size_t)(kdim_ + 1) * n_);
    for (int i = 0; i < nv; ++i) {
      k_zero(ctx_, full.get(), (size_t)nfull_ * sizeof(double));
      k_compact(ctx_, 1, nb_, N2_, td_.data(), S_old.get(), coff_old.get(), n_old, xoff_.get(),
                V_.get() + (i64)i * n_old, full.get(), nullptr);
      k_compact(ctx_, 0, nb_, N2_, td_.data(), S_.get(), coff_.get(), n_, xoff_.get(), full.get(),
                V2.get() + (i64)i * n_, crow_.get());
    }
    V_ = std::move(V2);
    tmp_.release();
    HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  }

  Ctx* ctx_;
  Tensor& x_;
  Tensor& y_;
  const int kdim_, maxiter_;
  const double tol_;
  i64 nb_ = 0, bs_ = 0, nfull_ = 0, n_ = 0;
  int N2_ = 0, ld_ = 0;
  std::vector<i64> td_;
  DBuf<i32> xb0_, yrow_, S_, crow_;
  DBuf<i64> off_, coff_, xoff_;
  DBuf<double> V_, part_, small_, dU_, tmp_;
  bool track_ = false, have_first_ = false, converged_ = false, pending_ = false;
  std::vector<double> T_, hh_, theta_, U_;
  int j_ = 0, cycles_ = 1, m_ = 0;
  double beta_ = 0, resid_ = 0;
  EigResult R_;
};

void check_chain(const Plan* first, const Plan* last, const Tensor& x, const Tensor& y) {
  if (!first || !last) fail(SB_E_HANDLE, "null plan");
  if (first->in_a != &x && first->in_b != &x)
    fail(SB_E_ARGUMENT, "ArgumentError: the first plan does not read the vector x");
  if (last->out != &y) fail(SB_E_ARGUMENT, "ArgumentError: the last plan does not write y");
}

}  // namespace

EigResult eigsolve_lanczos(Ctx* ctx, const std::vector<const Plan*>& plans, Tensor& x, Tensor& y, int kdim,
                           int maxiter, double tol) {
  if (plans.empty()) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve needs at least one plan");
  for (auto* p : plans)
    if (!p) fail(SB_E_HANDLE, "null plan");
  need_bs(x, "eigsolve");
  need_bs(y, "eigsolve");
  check_chain(plans.front(), plans.back(), x, y);
  Lanczos L(ctx, x, y, kdim, maxiter, tol);
  std::unique_ptr<ChainGraph> cg;
  if (eig_uses_graph(ctx, kdim, maxiter)) cg = std::make_unique<ChainGraph>(ctx, plans);
  auto apply = [&] {
    if (cg)
      cg->launch();
    else
      for (auto* p : plans) p->run();
  };
  L.begin();
  apply();  // H v0
  L.first();
  for (;;) {
    if (L.apply_needed()) {
      L.load();
      apply();
      L.applied();
      L.store();
    }
    L.orth();
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (L.redo()) {
      L.orth();
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    if (L.update()) break;
  }
  return L.finish();
}

std::vector<EigResult> eigsolve_lanczos_batch(Ctx* ctx, Batch& batch, const std::vector<Tensor*>& xs,
                                              const std::vector<Tensor*>& ys, int kdim, int maxiter, double tol) {
  const int n = batch.nchains;
  if ((int)xs.size() != n || (int)ys.size() != n)
    fail(SB_E_ARGUMENT, "ArgumentError: one vector x and one output y per chain of the batch");
  for (int c = 0; c < n; ++c) {
    if (!xs[c] || !ys[c]) fail(SB_E_HANDLE, "null tensor");
    need_bs(*xs[c], "eigsolve");
    need_bs(*ys[c], "eigsolve");
    check_chain(batch.stages.front().plans[c], batch.stages.back().plans[c], *xs[c], *ys[c]);
    for (int d = 0; d < c; ++d)
      if (xs[d] == xs[c] || ys[d] == ys[c]) fail(SB_E_ARGUMENT, "ArgumentError: the chains share a vector");
  }
  std::vector<std::unique_ptr<Lanczos>> L;
  for (int c = 0; c < n; ++c) L.emplace_back(std::make_unique<Lanczos>(ctx, *xs[c], *ys[c], kdim, maxiter, tol));
  std::vector<int> active(n);
  for (int c = 0; c < n; ++c) active[c] = c;
  // the batch's operator as one graph launch while every chain is still running
  std::unique_ptr<ChainGraph> bg;
  bool own = true;  // graphs only over this context's own plans (not over contexts on loan)
  for (auto& st : batch.stages)
    for (auto* p : st.plans) own = own && p->ctx == ctx;
  if (own && eig_uses_graph(ctx, kdim, maxiter)) bg = batch_graph(ctx, batch);
  auto apply = [&](const std::vector<int>& act) {
    if (bg && (int)act.size() == n)
      bg->launch();
    else
      batch.run(act);
  };
  for (int c : active) L[c]->begin();
  apply(active);  // H v0 of every chain
  for (int c : active) L[c]->first();
  while (!active.empty()) {
    std::vector<int> need;
    for (int c : active)
      if (L[c]->apply_needed()) need.push_back(c);
    if (!need.empty()) {
      for (int c : need) L[c]->load();
      apply(need);
      for (int c : need) {
        L[c]->applied();
        L[c]->store();
      }
    }
    for (int c : active) L[c]->orth();
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    bool again = false;
    for (int c : active)
      if (L[c]->redo()) {
        L[c]->orth();
        again = true;
      }
    if (again) HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<int> still;
    for (int c : active)
      if (!L[c]->update()) still.push_back(c);
    active.swap(still);
  }
  std::vector<EigResult> out;
  for (int c = 0; c < n; ++c) out.push_back(L[c]->finish());
  return out;
}

}  // namespace sb
