// The Krylov layer above contract!: the vector operations a Krylov solver
// applies to block-sparse tensors (SparseBackends/src/tensor_wrappers.jl:792-923)
// and an FP64 Lanczos eigensolver for the lowest eigenpair of a Hermitian
// operator given as a chain of contraction plans (the eigsolve that two-site
// DMRG runs on H_eff, in the reference KrylovKit's through ITensorMPS).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "sb_krylov.h"

namespace sb {

// ---------------------------------------------------------------- kernels
int linalg_nb();
void k_mdot(Ctx*, const double* V, i64 ldv, int k, const double* w, i64 n, bool norm, double* part, double* out);
void k_orth_update(Ctx*, const double* V, i64 ldv, int k, const double* h, double* w, i64 n, double* part,
                   double* norm_out);
void k_combine(Ctx*, const double* V, i64 ldv, int kin, const double* U, int kout, double* out, i64 ldo, i64 n);
void k_axpby(Ctx*, i64 n, double a, const double* x, double b, double* y);
void k_gather_blocks_by_key(Ctx*, i64 n, i64 bs, const i32* xids, const i64* off, const double* y, double* out);
void k_match_keys(Ctx*, const i32* xk, i64 nx, const i32* yk, const i32* yids, i64 ny, int P, const i64* pdims,
                  i64 bs, i64* off);
void k_add_blocks(Ctx*, i64 n, i64 bs, const i64* soff, const double* src, const i64* doff, double* dst);
void k_match_rows(Ctx*, const i32* xk, i64 nx, const i32* yk, i64 ny, int P, const i64* pdims, i32* row);
void k_box_exceeds(Ctx*, i64 n, int N2, const i32* S, const i32* ybox, const i32* row, i32* flag);
void k_support(Ctx*, i64 n, int N2, const i32* xbox, const i32* ybox, const i32* row, i32* S);
void k_compact(Ctx*, int mode, i64 nrows, int N2, const i64* tdims, const i32* S, const i64* coff, i64 total,
               const i64* boff, const double* src, double* dst);

namespace {

template <class F>
void step(Ctx* ctx, const char* name, F&& f) {
  if (!ctx->profile) {
    f();
    return;
  }
  hipEvent_t a = ctx->event(), b = ctx->event();
  (void)hipEventRecord(a, ctx->stream);
  f();
  (void)hipEventRecord(b, ctx->stream);
  ctx->prof_pending.push_back({name, a, b});
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

// ---------------------------------------------------------------- symmetric eigenproblem (host)

// Cyclic Jacobi (Rutishauser's rotation formulas) on an m x m symmetric
// column-major matrix; eigenvalues ascending in w, eigenvectors in U's columns.
void sym_eig(int m, const std::vector<double>& A_in, std::vector<double>& w, std::vector<double>& U) {
  std::vector<double> A(A_in);
  U.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) U[i + (size_t)i * m] = 1.0;
  auto a = [&](int i, int j) -> double& { return A[i + (size_t)j * m]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        tot += a(i, j) * a(i, j);
        if (i != j) off += a(i, j) * a(i, j);
      }
    if (off == 0.0 || off <= 1e-32 * tot) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < m; ++k) {  // A <- A J (columns p, q)
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {  // A <- J^T A (rows p, q)
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        a(p, q) = a(q, p) = 0.0;
        for (int k = 0; k < m; ++k) {
          const double ukp = U[k + (size_t)p * m], ukq = U[k + (size_t)q * m];
          U[k + (size_t)p * m] = c * ukp - s * ukq;
          U[k + (size_t)q * m] = s * ukp + c * ukq;
        }
      }
  }
  std::vector<int> ord(m);
  for (int i = 0; i < m; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  w.resize(m);
  std::vector<double> Us((size_t)m * m);
  for (int j = 0; j < m; ++j) {
    w[j] = a(ord[j], ord[j]);
    std::copy(&U[(size_t)ord[j] * m], &U[(size_t)ord[j] * m] + m, &Us[(size_t)j * m]);
  }
  U.swap(Us);
}

// ---------------------------------------------------------------- Lanczos

EigResult eigsolve_lanczos(Ctx* ctx, const std::vector<const Plan*>& plans, Tensor& x, Tensor& y, int kdim,
                           int maxiter, double tol) {
  need_bs(x, "eigsolve");
  need_bs(y, "eigsolve");
  if (plans.empty()) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve needs at least one plan");
  for (auto* p : plans)
    if (!p) fail(SB_E_HANDLE, "null plan");
  if (plans.front()->in_a != &x && plans.front()->in_b != &x)
    fail(SB_E_ARGUMENT, "ArgumentError: the first plan does not read the vector x");
  if (plans.back()->out != &y) fail(SB_E_ARGUMENT, "ArgumentError: the last plan does not write y");
  if (&x == &y) fail(SB_E_ARGUMENT, "ArgumentError: x and y must be distinct tensors");
  if (x.dims != y.dims || x.N2 != y.N2)
    fail(SB_E_ARGUMENT, "ArgumentError: the operator's output must have the vector's dims");
  if (kdim < 2 || kdim > 64) fail(SB_E_ARGUMENT, "ArgumentError: krylovdim must be in 2..64");
  if (maxiter < 1) fail(SB_E_ARGUMENT, "ArgumentError: maxiter must be positive");
  if (!(tol >= 0)) fail(SB_E_ARGUMENT, "ArgumentError: tol must be non-negative");

  const i64 nb = x.n, bs = x.blksize, nfull = nb * bs;
  if (nfull == 0) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve of an empty vector");
  const int N2 = x.N2;
  std::vector<i64> td(x.dims.begin() + x.P(), x.dims.end());
  EigResult R;

  // ---- v0 = x / |x| in x's payload
  const double b0 = norm_of(ctx, x.dptr(), nfull);
  if (!(b0 > 0) || !std::isfinite(b0)) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve: initial vector has zero or non-finite norm");
  // x's own nonzero extents per block (box)
  DBuf<i32> xb0(ctx, (size_t)(nb * N2 ? nb * N2 : 1)), xf0(ctx, (size_t)(nb ? nb : 1));
  box_scan(ctx, x.dptr(), x.ids.get(), nb, bs, N2, td.data(), xb0.get(), xf0.get());
  k_axpby(ctx, nfull, 1.0 / b0, x.dptr(), 0.0, x.dptr());
  x.touch();
  for (auto* p : plans) p->run();  // H v0
  ++R.numops;

  // ---- the compact layout of the Krylov vectors: per block of x, a box S
  // holding every nonzero any Krylov vector has: x0's extents joined with
  // those of H v0, widened later if an application reaches beyond (the
  // padding of symmetry sectors is left out).  When y's boxes are not known
  // (an operator that does not end in a box product) S is the full block.
  DBuf<i64> off;  // y's block offsets by x row (-1: y lacks the key)
  match_blocks(ctx, x, y, off);  // blocks of y whose key x lacks are left out: the solve stays in x's block space
  const bool track = y.box_trusted();
  DBuf<i32> yrow(ctx, (size_t)(nb ? nb : 1)), S(ctx, (size_t)(nb * N2 ? nb * N2 : 1));
  if (track) {
    std::vector<i64> pd(x.dims.begin(), x.dims.begin() + x.P());
    k_match_rows(ctx, x.keys.get(), nb, y.keys.get(), y.n, x.P(), pd.data(), yrow.get());
    k_support(ctx, nb, N2, xb0.get(), y.box.get(), yrow.get(), S.get());
  } else {
    k_fill_rows(ctx, S.get(), nb, N2, td.data());
  }
  std::vector<i32> hid(nb);
  HIP_CHECK(hipMemcpyAsync(hid.data(), x.ids.get(), nb * sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
  std::vector<i64> hxoff(nb);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (i64 i = 0; i < nb; ++i) hxoff[i] = (i64)(hid[i] - 1) * bs;
  DBuf<i64> coff(ctx, nb + 1), xoff(ctx, nb ? nb : 1);
  HIP_CHECK(hipMemcpyAsync(xoff.get(), hxoff.data(), nb * sizeof(i64), hipMemcpyHostToDevice, ctx->stream));
  i64 n = 0;  // Krylov vector length
  auto layout = [&]() {  // coff from S; returns the length
    std::vector<i32> hS((size_t)nb * N2);
    if (!hS.empty())
      HIP_CHECK(hipMemcpyAsync(hS.data(), S.get(), hS.size() * sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<i64> hc(nb + 1, 0);
    for (i64 i = 0; i < nb; ++i) {
      i64 c = 1;
      for (int a = 0; a < N2; ++a) c *= hS[i * N2 + a];
      hc[i + 1] = hc[i] + c;
    }
    HIP_CHECK(hipMemcpyAsync(coff.get(), hc.data(), (nb + 1) * sizeof(i64), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return hc[nb];
  };
  n = layout();
  if (n == 0) fail(SB_E_ARGUMENT, "ArgumentError: eigsolve: the vector and its image are empty");

  const int NBP = linalg_nb();
  const int ld = kdim + 1;
  DBuf<double> V(ctx, (size_t)(kdim + 1) * n), part(ctx, (size_t)(kdim + 2) * NBP);
  DBuf<double> small(ctx, 3 * (size_t)ld + 3), dU(ctx, (size_t)kdim * kdim), tmp;
  DBuf<i32> viol(ctx, 1);
  double* h1 = small.get();
  double* h2 = h1 + ld + 1;
  double* nr = h2 + ld + 1;
  auto vec = [&](int i) { return V.get() + (i64)i * n; };
  // v0 and H v0 in the compact layout
  k_compact(ctx, 0, nb, N2, td.data(), S.get(), coff.get(), n, xoff.get(), x.dptr(), vec(0));
  k_compact(ctx, 0, nb, N2, td.data(), S.get(), coff.get(), n, off.get(), y.dptr(), vec(1));

  // S widened to y's boxes where an application reached beyond it: the
  // vectors 0..nv-1 move to the new layout (they are zero in the new entries)
  auto widen = [&](int nv) {
    DBuf<double> full(ctx, (size_t)nfull);
    DBuf<i32> S2(ctx, (size_t)(nb * N2 ? nb * N2 : 1));
    k_support(ctx, nb, N2, S.get(), y.box.get(), yrow.get(), S2.get());
    DBuf<i64> coff_old(ctx, nb + 1);
    HIP_CHECK(hipMemcpyAsync(coff_old.get(), coff.get(), (nb + 1) * sizeof(i64), hipMemcpyDeviceToDevice,
                             ctx->stream));
    const i64 n_old = n;
    DBuf<i32> S_old = std::move(S);
    S = std::move(S2);
    n = layout();
    DBuf<double> V2(ctx, (size_t)(kdim + 1) * n);
    for (int i = 0; i < nv; ++i) {
      k_zero(ctx, full.get(), (size_t)nfull * sizeof(double));
      k_compact(ctx, 1, nb, N2, td.data(), S_old.get(), coff_old.get(), n_old, xoff.get(), V.get() + (i64)i * n_old,
                full.get());
      k_compact(ctx, 0, nb, N2, td.data(), S.get(), coff.get(), n, xoff.get(), full.get(), V2.get() + (i64)i * n);
    }
    V = std::move(V2);
    tmp.release();
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  };

  // later applications: the basis vector into x's payload (entries outside S
  // stay zero: x0 is zero there), the operator, y's blocks by key into the
  // compact layout; y's boxes are checked against S
  auto matvec = [&](int src, int dst) {
    step(ctx, "lanczos_load",
         [&] { k_compact(ctx, 1, nb, N2, td.data(), S.get(), coff.get(), n, xoff.get(), vec(src), x.dptr()); });
    x.touch();
    for (auto* p : plans) p->run();
    ++R.numops;
    if (track) {
      k_box_exceeds(ctx, nb, N2, S.get(), y.box.get(), yrow.get(), viol.get());
      i32 hv = 0;
      HIP_CHECK(hipMemcpyAsync(&hv, viol.get(), sizeof(i32), hipMemcpyDeviceToHost, ctx->stream));
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
      if (hv) widen(dst);
    }
    step(ctx, "lanczos_gather",
         [&] { k_compact(ctx, 0, nb, N2, td.data(), S.get(), coff.get(), n, off.get(), y.dptr(), vec(dst)); });
  };
  bool have_first = true;  // H v0 is in vec(1) already

  std::vector<double> T((size_t)ld * ld, 0.0), hh(3 * (size_t)ld + 3), theta, U;
  int j = 0;
  int cycles = 1;
  double beta = 0, resid = 0;
  bool converged = false;
  int m = 0;
  for (;;) {
    if (have_first)
      have_first = false;
    else
      matvec(j, j + 1);
    double* w = vec(j + 1);
    // classical Gram-Schmidt, twice, against v_0 .. v_j
    step(ctx, "lanczos_orth", [&] {
      k_mdot(ctx, V.get(), n, j + 1, w, n, true, part.get(), h1);
      k_orth_update(ctx, V.get(), n, j + 1, h1, w, n, part.get(), nullptr);
      k_mdot(ctx, V.get(), n, j + 1, w, n, false, part.get(), h2);
      k_orth_update(ctx, V.get(), n, j + 1, h2, w, n, part.get(), nr);
    });
    HIP_CHECK(hipMemcpyAsync(hh.data(), small.get(), hh.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const double* H1 = hh.data();
    const double* H2 = H1 + ld + 1;
    const double nAv = std::sqrt(H1[j + 1]);
    beta = std::sqrt(hh[2 * (ld + 1)]);
    if (!std::isfinite(beta) || !std::isfinite(nAv)) fail(SB_E_ERROR, "eigsolve: the operator produced non-finite values");
    for (int i = 0; i <= j; ++i) T[i + (size_t)j * ld] = T[j + (size_t)i * ld] = H1[i] + H2[i];
    m = j + 1;
    std::vector<double> Tm((size_t)m * m);
    for (int c = 0; c < m; ++c)
      for (int r = 0; r < m; ++r) Tm[r + (size_t)c * m] = T[r + (size_t)c * ld];
    sym_eig(m, Tm, theta, U);
    const bool breakdown = beta <= 1e-13 * std::max(nAv, 1e-300);
    resid = breakdown ? 0.0 : beta * std::fabs(U[(m - 1) + 0 * (size_t)m]);
    if (!breakdown) k_axpby(ctx, n, 1.0 / beta, w, 0.0, w);
    if (resid <= tol || breakdown) {
      converged = true;
      break;
    }
    if (m < kdim) {
      ++j;
      continue;
    }
    if (cycles == maxiter) break;
    // thick restart: keep the lowest Ritz vectors, then the residual direction
    const int keep = std::max(1, std::min(kdim - 1, (3 * kdim) / 5));
    step(ctx, "lanczos_restart", [&] {
      HIP_CHECK(hipMemcpyAsync(dU.get(), U.data(), (size_t)m * keep * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
      if (!tmp.get()) tmp.alloc(ctx, (size_t)keep * n);
      k_combine(ctx, V.get(), n, m, dU.get(), keep, tmp.get(), n, n);
      HIP_CHECK(hipMemcpyAsync(V.get(), tmp.get(), (size_t)keep * n * sizeof(double), hipMemcpyDeviceToDevice,
                               ctx->stream));
      HIP_CHECK(hipMemcpyAsync(vec(keep), vec(m), n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    });
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < keep; ++i) {
      T[i + (size_t)i * ld] = theta[i];
      T[keep + (size_t)i * ld] = T[i + (size_t)keep * ld] = beta * U[(m - 1) + (size_t)i * m];
    }
    HIP_CHECK(hipStreamSynchronize(ctx->stream));  // U is reused on the host
    j = keep;
    ++cycles;
  }
  // Ritz vector of the lowest Ritz value into x, normalised
  DBuf<double> ritzv(ctx, (size_t)n);
  step(ctx, "lanczos_ritz", [&] {
    HIP_CHECK(hipMemcpyAsync(dU.get(), U.data(), (size_t)m * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    k_combine(ctx, V.get(), n, m, dU.get(), 1, ritzv.get(), n, n);
  });
  const double xn = norm_of(ctx, ritzv.get(), n);
  if (xn > 0) k_axpby(ctx, n, 1.0 / xn, ritzv.get(), 0.0, ritzv.get());
  k_compact(ctx, 1, nb, N2, td.data(), S.get(), coff.get(), n, xoff.get(), ritzv.get(), x.dptr());
  x.touch();
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  R.value = theta[0];
  R.ritz = theta;
  R.normres = resid;
  R.numiter = cycles;
  R.converged = converged;
  return R;
}

}  // namespace sb
