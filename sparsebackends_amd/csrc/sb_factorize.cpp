// blocksparse_svd (SparseBackends/src/ops_factorize.jl:50-148): the two-site
// block-diagonal SVD with global truncation.  The per-group factorisations
// and the payloads run on the device (sb_svd.hip); grouping, the global
// singular-value order, truncation and the key/id bookkeeping of the two
// factors (a few values per kept singular triplet) run on the host.
#include <algorithm>
#include <map>

#include "sb_factorize.h"

namespace sb {

// the sweep cap of a one-sided Jacobi SVD was reached with rows still
// rotating: the reference's LAPACK SVD raises LAPACKException when it fails
// to converge
[[noreturn]] void no_convergence(Ctx* ctx) {
  fail(SB_E_NOCONV, "LAPACKException: the one-sided Jacobi SVD did not converge in " +
                        std::to_string(ctx->svd_sweeps) + " sweeps");
}

void k_svd_jacobi(Ctx*, const double* A, int d1, int d2, i64 ng, const i64* gstart, const i64* bsrc, const i64* woff,
                  double* W, double* sig, double* Uout, int* perm, int* noconv);
void k_svd_emit(Ctx*, i64 k, int d1, const i32* kg, const i32* kc, const i64* soff, const i64* gstart, int d2,
                const i64* woff, const double* W, const double* sig, const double* Uout, const int* perm, bool right,
                double* Ud, double* SVd);

void k_sect_gather(Ctx*, const double* psi, int d1, int d2, i64 Dl, i64 ntask, const i64* task, const i64* xoff,
                   const int* R, const int* Cc, const int* trans, double* X);
size_t sect_svd_lds(int n, int len);
void k_sect_svd(Ctx*, i64 ng, size_t lds, const double* X, const i64* xoff, const int* R, const int* Cc,
                const int* trans, const i64* uoff, const i64* soff, const i64* sgoff, double* sig, double* Uo,
                double* SVo, const int* glist, int* noconv);
bool k_sect_svd_qrcp(Ctx* ctx, const std::vector<int>& gl, const std::vector<int>& Rs, const std::vector<int>& Cs,
                     double* X, const i64* xoff, const i64* uoff, const i64* soff, const i64* sgoff, double* sig,
                     double* Uo, double* SVo);
void k_sect_emit(Ctx*, bool isA, i64 nblk, int d, i64 D1, i64 D2, i64 nside, const i64* blk, const i64* sec,
                 const i64* off, const i64* dims, const i64* kept, const int* RC, const i64* uoff_or_soff,
                 const i64* sgoff, const double* sig, const double* F, bool right, double* out);

namespace {
template <class T>
void up(Ctx* ctx, DBuf<T>& d, const std::vector<T>& h) {
  d.alloc(ctx, h.empty() ? 1 : h.size());
  if (!h.empty())
    HIP_CHECK(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}
template <class T>
void down(Ctx* ctx, std::vector<T>& h, const T* d, size_t n) {
  h.resize(n);
  if (n) HIP_CHECK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
}

// An orthonormal completion: vecs holds K vectors of length D (vector c at
// vecs[c * vs + e * es]); those with null[c] are replaced, in order, by unit
// vectors orthogonal to all the others (the kept ones and those completed
// before): the unit vector e_j of the smallest projection onto the span so far,
// orthogonalised twice (classical Gram-Schmidt) and normalised.
void complete_orthonormal(std::vector<double>& vecs, i64 vs, i64 es, i64 D, i64 K, const std::vector<char>& null) {
  std::vector<i64> basis;
  for (i64 c = 0; c < K; ++c)
    if (!null[c]) basis.push_back(c);
  auto at = [&](i64 c, i64 e) -> double& { return vecs[(size_t)(c * vs + e * es)]; };
  std::vector<double> w(D), proj(D);
  for (i64 c = 0; c < K; ++c) {
    if (!null[c]) continue;
    std::fill(proj.begin(), proj.end(), 0.0);
    for (i64 b : basis)
      for (i64 e = 0; e < D; ++e) proj[e] += at(b, e) * at(b, e);
    const i64 j = std::min_element(proj.begin(), proj.end()) - proj.begin();
    std::fill(w.begin(), w.end(), 0.0);
    w[j] = 1.0;
    for (int pass = 0; pass < 2; ++pass)
      for (i64 b : basis) {
        double dot = 0;
        for (i64 e = 0; e < D; ++e) dot += at(b, e) * w[e];
        for (i64 e = 0; e < D; ++e) w[e] -= dot * at(b, e);
      }
    double nw = 0;
    for (double x : w) nw += x * x;
    nw = std::sqrt(nw);
    for (i64 e = 0; e < D; ++e) at(c, e) = w[e] / nw;
    basis.push_back(c);
  }
}
}  // namespace

SvdResult bs_svd(Ctx* ctx, const Tensor& A, bool right, i64 maxdim, i64 mindim, double cutoff) {
  if (A.kind != BS || A.N() != 4 || A.N2 != 2)
    fail(SB_E_METHOD, "MethodError: blocksparse_svd needs a NewBlockSparseSorted{T,4,2,2}");
  const i64 nL = A.dims[0], nR = A.dims[1];
  const int d1 = (int)A.dims[2], d2 = (int)A.dims[3];
  if (d1 > 32) fail(SB_E_LIMIT, "blocksparse_svd: first dense dimension above 32");
  const i64 n = A.n, bs = A.blksize;
  // group the blocks by L, R ascending within a group (ops_factorize.jl:59-83)
  std::vector<i64> keys;
  std::vector<i32> hk, hi;
  down(ctx, hk, A.keys.get(), (size_t)n * 2);
  down(ctx, hi, A.ids.get(), (size_t)n);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::map<i32, std::vector<std::pair<i32, i32>>> groups;  // L -> (R, id)
  for (i64 i = 0; i < n; ++i) groups[hk[2 * i]].push_back({hk[2 * i + 1], hi[i]});
  const i64 ng = (i64)groups.size();
  std::vector<i64> gstart(ng + 1, 0), bsrc, woff(ng), gL(ng);
  std::vector<std::vector<i32>> gR(ng);
  bsrc.reserve(n);
  {
    i64 g = 0, w = 0;
    for (auto& kv : groups) {
      auto& v = kv.second;
      std::sort(v.begin(), v.end());
      gL[g] = kv.first;
      woff[g] = w;
      for (auto& rb : v) {
        bsrc.push_back((i64)(rb.second - 1) * bs);
        gR[g].push_back(rb.first);
      }
      gstart[g + 1] = gstart[g] + (i64)v.size();
      w += (i64)d1 * (i64)v.size() * d2;
      ++g;
    }
  }
  DBuf<i64> d_gstart, d_bsrc, d_woff;
  up(ctx, d_gstart, gstart);
  up(ctx, d_bsrc, bsrc);
  up(ctx, d_woff, woff);
  DBuf<double> W(ctx, n * bs ? n * bs : 1), sig(ctx, ng * d1 ? ng * d1 : 1), Uo(ctx, ng * d1 * d1 ? ng * d1 * d1 : 1);
  DBuf<int> perm(ctx, ng * d1 ? ng * d1 : 1), noconv(ctx, 1);
  k_zero(ctx, noconv.get(), sizeof(int));
  k_svd_jacobi(ctx, A.dptr(), d1, d2, ng, d_gstart.get(), d_bsrc.get(), d_woff.get(), W.get(), sig.get(), Uo.get(),
               perm.get(), noconv.get());
  std::vector<double> hs;
  down(ctx, hs, sig.get(), (size_t)(ng * d1));
  int hnc = 0;
  HIP_CHECK(hipMemcpyAsync(&hnc, noconv.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (hnc) no_convergence(ctx);

  // global order and truncation (ops_factorize.jl:86-104)
  struct SV {
    double s;
    i64 g;
    int c;
  };
  std::vector<SV> all;
  for (i64 g = 0; g < ng; ++g) {
    const i64 nc = (gstart[g + 1] - gstart[g]) * d2;
    const int kgs = (int)std::min<i64>(d1, nc);  // thin SVD: min(d1, ncols) values
    for (int c = 0; c < kgs; ++c) all.push_back({hs[g * d1 + c], g, c});
  }
  std::stable_sort(all.begin(), all.end(), [](const SV& a, const SV& b) { return a.s > b.s; });
  const double max_sv = all.empty() ? 1.0 : all[0].s;
  i64 keep = (maxdim < 0) ? (i64)all.size() : std::min<i64>(maxdim, (i64)all.size());
  while (keep > mindim && all[keep - 1].s < cutoff * max_sv) --keep;
  SvdResult R;
  R.truncerr = 0.0;
  for (size_t i = keep; i < all.size(); ++i) R.truncerr += all[i].s * all[i].s;
  std::vector<i64> kk(ng, 0);
  for (i64 i = 0; i < keep; ++i) ++kk[all[i].g];
  const i64 k = keep;

  // factors: U (nL, k, d1) keys (L, nb); SV (k, nR, d2) keys (nb, R); ids in
  // the creation order of the reference's loop (:118-143)
  R.U = bs_make(ctx, {nL, k, (i64)d1}, 1);
  R.SV = bs_make(ctx, {k, nR, (i64)d2}, 1);
  std::vector<i32> kg(k), kc(k), uk(2 * k), uid(k);
  std::vector<i64> soff(k + 1, 0);
  std::vector<std::array<i64, 3>> sv_blocks;  // (R, nb, id)
  R.svs.resize(k);
  {
    i64 nb = 0;
    for (i64 g = 0; g < ng; ++g)
      for (i64 c = 0; c < kk[g]; ++c, ++nb) {
        kg[nb] = (i32)g;
        kc[nb] = (i32)c;
        R.svs[nb] = hs[g * d1 + c];
        uk[2 * nb] = (i32)gL[g];
        uk[2 * nb + 1] = (i32)(nb + 1);
        uid[nb] = (i32)(nb + 1);
        for (i32 r : gR[g]) sv_blocks.push_back({(i64)r, nb + 1, (i64)sv_blocks.size() + 1});
        soff[nb + 1] = soff[nb] + (gstart[g + 1] - gstart[g]) * d2;
      }
  }
  std::stable_sort(sv_blocks.begin(), sv_blocks.end(),
                   [](const std::array<i64, 3>& a, const std::array<i64, 3>& b) {
                     return a[0] != b[0] ? a[0] < b[0] : a[1] < b[1];
                   });
  const i64 nsv = (i64)sv_blocks.size();
  std::vector<i32> sk(2 * nsv), sid(nsv);
  for (i64 i = 0; i < nsv; ++i) {
    sk[2 * i] = (i32)sv_blocks[i][1];
    sk[2 * i + 1] = (i32)sv_blocks[i][0];
    sid[i] = (i32)sv_blocks[i][2];
  }
  R.U->n = k;
  R.SV->n = nsv;
  up(ctx, R.U->keys, uk);
  up(ctx, R.U->ids, uid);
  up(ctx, R.SV->keys, sk);
  up(ctx, R.SV->ids, sid);
  R.U->data.alloc(ctx, (size_t)k * d1);
  R.SV->data.alloc(ctx, (size_t)nsv * d2);
  DBuf<i32> d_kg, d_kc;
  DBuf<i64> d_soff;
  up(ctx, d_kg, kg);
  up(ctx, d_kc, kc);
  up(ctx, d_soff, soff);
  k_svd_emit(ctx, k, d1, d_kg.get(), d_kc.get(), d_soff.get(), d_gstart.get(), d2, d_woff.get(), W.get(), sig.get(),
             Uo.get(), perm.get(), right, R.U->data.get(), R.SV->data.get());
  HIP_CHECK(hipStreamSynchronize(ctx->stream));  // host staging vectors go out of scope
  if (right && k) {
    // ortho right: the SV rows of a group are V^T rows, an isometry in LAPACK's
    // factorisation; a kept triplet of zero singular value (cutoff = 0 on a
    // group of lower rank) has a zero row there, replaced by an orthonormal
    // completion within its group
    std::vector<char> gnull(ng, 0);
    bool need = false;
    std::vector<i64> nb0(ng + 1, 0);
    for (i64 g = 0; g < ng; ++g) {
      nb0[g + 1] = nb0[g] + kk[g];
      const i64 nc = (gstart[g + 1] - gstart[g]) * d2;
      double fro2 = 0;
      for (int c = 0; c < (int)std::min<i64>(d1, nc); ++c) fro2 += hs[g * d1 + c] * hs[g * d1 + c];
      for (i64 c = 0; c < kk[g]; ++c)
        if (!(hs[g * d1 + c] > 1e-13 * std::sqrt(fro2))) gnull[g] = 1;
      need = need || gnull[g];
    }
    if (need) {
      std::vector<double> sv;
      down(ctx, sv, R.SV->data.get(), (size_t)(nsv * d2));
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (i64 g = 0; g < ng; ++g) {
        if (!gnull[g]) continue;
        const i64 nc = (gstart[g + 1] - gstart[g]) * d2, K = kk[g];
        const i64 base = soff[nb0[g]];  // the group's rows, contiguous: row c at base + c * nc
        double fro2 = 0;
        for (int c = 0; c < (int)std::min<i64>(d1, nc); ++c) fro2 += hs[g * d1 + c] * hs[g * d1 + c];
        std::vector<char> null(K);
        for (i64 c = 0; c < K; ++c) null[c] = !(hs[g * d1 + c] > 1e-13 * std::sqrt(fro2));
        std::vector<double> v(sv.begin() + base, sv.begin() + base + K * nc);
        complete_orthonormal(v, nc, 1, nc, K, null);
        std::copy(v.begin(), v.end(), sv.begin() + base);
      }
      HIP_CHECK(hipMemcpy(R.SV->data.get(), sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  return R;
}

// ================================================================ sector SVD

SectSvdResult bs_svd_sectors(Ctx* ctx, const Tensor& psi, i64 nsec, const i64* rowsec, const i64* colsec,
                             const i64* dl, const i64* dr, bool right, i64 maxdim, i64 mindim, double cutoff) {
  if (psi.kind != BS || psi.N() != 6 || psi.N2 != 4)
    fail(SB_E_METHOD, "MethodError: the sector SVD needs psi [Ql, Qr | s1, s2, l, r] (N = 6, N2 = 4)");
  const i64 nl = psi.dims[0], nr = psi.dims[1], Dl = psi.dims[4], Dr = psi.dims[5];
  const int d1 = (int)psi.dims[2], d2 = (int)psi.dims[3];
  if (nsec < 0) fail(SB_E_ARGUMENT, "ArgumentError: negative sector count");
  for (i64 i = 0; i < nl * d1; ++i)
    if (rowsec[i] < -1 || rowsec[i] >= nsec) fail(SB_E_ARGUMENT, "ArgumentError: row sector out of range");
  for (i64 i = 0; i < nr * d2; ++i)
    if (colsec[i] < -1 || colsec[i] >= nsec) fail(SB_E_ARGUMENT, "ArgumentError: column sector out of range");
  for (i64 i = 0; i < nl; ++i)
    if (dl[i] < 0 || dl[i] > Dl) fail(SB_E_ARGUMENT, "ArgumentError: left sector extent outside 0..Dl");
  for (i64 i = 0; i < nr; ++i)
    if (dr[i] < 0 || dr[i] > Dr) fail(SB_E_ARGUMENT, "ArgumentError: right sector extent outside 0..Dr");
  // rows (Ql, s1, l) and columns (Qr, s2, r) of every sector, in that order
  std::vector<i64> R(nsec, 0), Cn(nsec, 0), rowoff(nl * d1, -1), coloff(nr * d2, -1);
  for (i64 il = 0; il < nl; ++il)
    for (int s1 = 0; s1 < d1; ++s1) {
      const i64 m = rowsec[il + nl * s1];
      if (m >= 0) {
        rowoff[il + nl * s1] = R[m];
        R[m] += dl[il];
      }
    }
  for (i64 ir = 0; ir < nr; ++ir)
    for (int s2 = 0; s2 < d2; ++s2) {
      const i64 m = colsec[ir + nr * s2];
      if (m >= 0) {
        coloff[ir + nr * s2] = Cn[m];
        Cn[m] += dr[ir];
      }
    }
  // active sectors (a nonempty matrix), compact index g
  std::vector<i64> gof(nsec, -1), act;
  for (i64 m = 0; m < nsec; ++m)
    if (R[m] > 0 && Cn[m] > 0) {
      gof[m] = (i64)act.size();
      act.push_back(m);
    }
  const i64 ng = (i64)act.size();
  std::vector<int> hR(ng), hC(ng), htr(ng);
  std::vector<i64> xoff(ng + 1, 0), uoff(ng + 1, 0), soff(ng + 1, 0), sgoff(ng + 1, 0);
  // sectors with fewer than 32 rows to orthogonalise (lds_g) are factorised
  // in the LDS of one workgroup each; the others (big_g) by QR with column
  // pivoting and block Jacobi over many workgroups on the rows of R, whose
  // sequential depth per sweep is smaller from about 32 rows on
  constexpr size_t kLds = 159 * 1024;
  size_t lds = 0;
  std::vector<int> lds_g, big_g;
  for (i64 g = 0; g < ng; ++g) {
    const i64 m = act[g];
    if (R[m] > INT32_MAX || Cn[m] > INT32_MAX) fail(SB_E_LIMIT, "sector matrix dimension beyond int32");
    hR[g] = (int)R[m];
    hC[g] = (int)Cn[m];
    htr[g] = R[m] > Cn[m];
    const int n = (int)std::min(R[m], Cn[m]), len = (int)std::max(R[m], Cn[m]);
    if (n < 32 && sect_svd_lds(n, len) <= kLds) {
      lds = std::max(lds, sect_svd_lds(n, len));
      lds_g.push_back((int)g);
    } else {
      big_g.push_back((int)g);
    }
    xoff[g + 1] = xoff[g] + R[m] * Cn[m];
    uoff[g + 1] = uoff[g] + R[m] * n;
    soff[g + 1] = soff[g] + (i64)n * Cn[m];
    sgoff[g + 1] = sgoff[g] + n;
  }
  // while block Jacobi sweeps run, the small sectors ride along in their
  // launches (one round per sweep each) instead of a launch of their own that
  // the large sectors would wait for; those sectors are gathered column-major
  // (M itself) for the QR preconditioning
  std::vector<int> qp_R, qp_C;
  if (!big_g.empty()) {
    big_g.insert(big_g.end(), lds_g.begin(), lds_g.end());
    lds_g.clear();
    for (int g : big_g) {
      qp_R.push_back(hR[g]);
      qp_C.push_back(hC[g]);
    }
  }
  std::vector<int> gtr(htr);
  if (!big_g.empty()) std::fill(gtr.begin(), gtr.end(), 1);
  // gather tasks from psi's blocks
  std::vector<i32> hk, hi;
  down(ctx, hk, psi.keys.get(), (size_t)psi.n * 2);
  down(ctx, hi, psi.ids.get(), (size_t)psi.n);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::vector<i64> task;
  for (i64 b = 0; b < psi.n; ++b) {
    const i64 il = hk[2 * b] - 1, ir = hk[2 * b + 1] - 1;
    if (!dl[il] || !dr[ir]) continue;
    for (int s1 = 0; s1 < d1; ++s1) {
      const i64 m = rowsec[il + nl * s1];
      if (m < 0 || gof[m] < 0) continue;
      for (int s2 = 0; s2 < d2; ++s2) {
        if (colsec[ir + nr * s2] != m) continue;
        const i64 t[8] = {(i64)(hi[b] - 1) * psi.blksize, s1, s2, dl[il], dr[ir], gof[m], rowoff[il + nl * s1],
                          coloff[ir + nr * s2]};
        task.insert(task.end(), t, t + 8);
      }
    }
  }
  const i64 ntask = (i64)task.size() / 8;
  DBuf<i64> d_task, d_xoff, d_uoff, d_soff, d_sgoff;
  DBuf<int> d_R, d_C, d_tr;
  up(ctx, d_task, task);
  up(ctx, d_xoff, xoff);
  up(ctx, d_uoff, uoff);
  up(ctx, d_soff, soff);
  up(ctx, d_sgoff, sgoff);
  up(ctx, d_R, hR);
  up(ctx, d_C, hC);
  up(ctx, d_tr, htr);
  DBuf<int> d_gtr;
  up(ctx, d_gtr, gtr);
  DBuf<double> X(ctx, xoff[ng] ? xoff[ng] : 1), Uo(ctx, uoff[ng] ? uoff[ng] : 1), SVo(ctx, soff[ng] ? soff[ng] : 1),
      sig(ctx, sgoff[ng] ? sgoff[ng] : 1);
  k_zero(ctx, X.get(), (size_t)(xoff[ng] ? xoff[ng] : 1) * sizeof(double));
  k_sect_gather(ctx, psi.dptr(), d1, d2, Dl, ntask, d_task.get(), d_xoff.get(), d_R.get(), d_C.get(), d_gtr.get(),
                X.get());
  {
    DBuf<int> d_lg, noconv(ctx, 1);
    k_zero(ctx, noconv.get(), sizeof(int));
    if (!lds_g.empty()) {
      up(ctx, d_lg, lds_g);
      k_sect_svd(ctx, (i64)lds_g.size(), lds, X.get(), d_xoff.get(), d_R.get(), d_C.get(), d_tr.get(), d_uoff.get(),
                 d_soff.get(), d_sgoff.get(), sig.get(), Uo.get(), SVo.get(), d_lg.get(), noconv.get());
    }
    const bool nc_big = k_sect_svd_qrcp(ctx, big_g, qp_R, qp_C, X.get(), d_xoff.get(), d_uoff.get(), d_soff.get(),
                                        d_sgoff.get(), sig.get(), Uo.get(), SVo.get());
    int hnc = 0;
    HIP_CHECK(hipMemcpyAsync(&hnc, noconv.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host list goes out of scope
    if (hnc || nc_big) no_convergence(ctx);
  }
  std::vector<double> hs;
  down(ctx, hs, sig.get(), (size_t)sgoff[ng]);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));

  // global order and truncation, the rules of blocksparse_svd (ops_factorize.jl:86-104)
  struct SV {
    double s;
    i64 g;
  };
  std::vector<SV> all;
  for (i64 g = 0; g < ng; ++g)
    for (i64 c = sgoff[g]; c < sgoff[g + 1]; ++c) all.push_back({hs[c], g});
  std::stable_sort(all.begin(), all.end(), [](const SV& a, const SV& b) { return a.s > b.s; });
  const double max_sv = all.empty() ? 1.0 : all[0].s;
  i64 keep = (maxdim < 0) ? (i64)all.size() : std::min<i64>(maxdim, (i64)all.size());
  while (keep > mindim && all[keep - 1].s < cutoff * max_sv) --keep;
  SectSvdResult out;
  for (size_t i = keep; i < all.size(); ++i) out.truncerr += all[i].s * all[i].s;
  std::vector<i64> kg(ng, 0);
  for (i64 i = 0; i < keep; ++i) ++kg[all[i].g];
  out.kept.assign(nsec, 0);
  std::vector<i64> qm(ng, -1);  // new-bond sector number of each compact sector
  i64 nm = 0, Dm = 0;
  for (i64 g = 0; g < ng; ++g) {
    out.kept[act[g]] = kg[g];
    if (kg[g] > 0) {
      qm[g] = nm++;
      Dm = std::max(Dm, kg[g]);
      for (i64 c = 0; c < kg[g]; ++c) out.svs.push_back(hs[sgoff[g] + c]);
    }
  }
  // factors; blocks in key order with ids 1..n
  out.A = bs_make(ctx, {nl, nm, (i64)d1, Dl, Dm}, 3);
  out.B = bs_make(ctx, {nm, nr, (i64)d2, Dm, Dr}, 3);
  std::vector<i64> rsec_c(nl * d1, -1), csec_c(nr * d2, -1);
  for (i64 i = 0; i < nl * d1; ++i)
    if (rowsec[i] >= 0 && gof[rowsec[i]] >= 0) rsec_c[i] = gof[rowsec[i]];
  for (i64 i = 0; i < nr * d2; ++i)
    if (colsec[i] >= 0 && gof[colsec[i]] >= 0) csec_c[i] = gof[colsec[i]];
  std::vector<i64> ablk, bblk;
  std::vector<i32> akeys, bkeys;
  for (i64 g = 0; g < ng; ++g) {
    if (qm[g] < 0) continue;
    for (i64 il = 0; il < nl; ++il) {
      bool any = false;
      for (int s1 = 0; s1 < d1 && !any; ++s1) any = rsec_c[il + nl * s1] == g;
      if (!any || !dl[il]) continue;
      ablk.push_back(il);
      ablk.push_back(g);
      akeys.push_back((i32)(il + 1));
      akeys.push_back((i32)(qm[g] + 1));
    }
  }
  for (i64 ir = 0; ir < nr; ++ir)
    for (i64 g = 0; g < ng; ++g) {
      if (qm[g] < 0 || !dr[ir]) continue;
      bool any = false;
      for (int s2 = 0; s2 < d2 && !any; ++s2) any = csec_c[ir + nr * s2] == g;
      if (!any) continue;
      bblk.push_back(g);
      bblk.push_back(ir);
      bkeys.push_back((i32)(qm[g] + 1));
      bkeys.push_back((i32)(ir + 1));
    }
  const i64 nA = (i64)ablk.size() / 2, nB = (i64)bblk.size() / 2;
  std::vector<i32> aid(nA), bid(nB);
  for (i64 i = 0; i < nA; ++i) aid[i] = (i32)(i + 1);
  for (i64 i = 0; i < nB; ++i) bid[i] = (i32)(i + 1);
  out.A->n = nA;
  out.B->n = nB;
  up(ctx, out.A->keys, akeys);
  up(ctx, out.A->ids, aid);
  up(ctx, out.B->keys, bkeys);
  up(ctx, out.B->ids, bid);
  out.A->data.alloc(ctx, (size_t)(nA * out.A->blksize));
  out.B->data.alloc(ctx, (size_t)(nB * out.B->blksize));
  DBuf<i64> d_ablk, d_bblk, d_rsec, d_csec, d_roff, d_coff, d_dl, d_dr, d_kept;
  up(ctx, d_ablk, ablk);
  up(ctx, d_bblk, bblk);
  up(ctx, d_rsec, rsec_c);
  up(ctx, d_csec, csec_c);
  up(ctx, d_roff, rowoff);
  up(ctx, d_coff, coloff);
  up(ctx, d_dl, std::vector<i64>(dl, dl + nl));
  up(ctx, d_dr, std::vector<i64>(dr, dr + nr));
  up(ctx, d_kept, kg);
  // Kept triplets of (numerically) zero singular value: the Jacobi leaves null
  // rows alone, so the factor built from normalised rows of X (U for a
  // transposed sector, V^T for ortho right) would hold zero or noise there and
  // not be an isometry, as LAPACK's factors are.  Those vectors are replaced by
  // an orthonormal completion (only a state of lower rank than its kept bond,
  // with cutoff = 0, reaches this).
  DBuf<double> sigB;  // the singular values B's rows are divided by (1 for a completed V^T row)
  {
    std::vector<char> anynull(ng, 0);
    bool need = false;
    for (i64 g = 0; g < ng; ++g) {
      double fro2 = 0;
      for (i64 c = sgoff[g]; c < sgoff[g + 1]; ++c) fro2 += hs[c] * hs[c];
      const double thr = 1e-13 * std::sqrt(fro2);
      for (i64 c = 0; c < kg[g]; ++c)
        if (!(hs[sgoff[g] + c] > thr)) anynull[g] = 1;
      need = need || anynull[g];
    }
    if (need) {
      std::vector<double> hsB(hs);
      for (i64 g = 0; g < ng; ++g) {
        if (!anynull[g]) continue;
        double fro2 = 0;
        for (i64 c = sgoff[g]; c < sgoff[g + 1]; ++c) fro2 += hs[c] * hs[c];
        const double thr = 1e-13 * std::sqrt(fro2);
        const i64 K = kg[g], n = sgoff[g + 1] - sgoff[g];
        std::vector<char> null(K);
        for (i64 c = 0; c < K; ++c) null[c] = !(hs[sgoff[g] + c] > thr);
        if (!right) {  // A = U: columns c of the R x n column-major U of the sector
          const i64 Rg = hR[g];
          std::vector<double> u;
          down(ctx, u, Uo.get() + uoff[g], (size_t)(Rg * n));
          HIP_CHECK(hipStreamSynchronize(ctx->stream));
          complete_orthonormal(u, Rg, 1, Rg, K, null);
          HIP_CHECK(hipMemcpy(Uo.get() + uoff[g], u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {  // B = V^T: rows c of the n x C row-major sigma V^T, divided by sigma
          const i64 Cg = hC[g];
          std::vector<double> v;
          down(ctx, v, SVo.get() + soff[g], (size_t)(n * Cg));
          HIP_CHECK(hipStreamSynchronize(ctx->stream));
          for (i64 c = 0; c < K; ++c)
            if (!null[c])
              for (i64 j = 0; j < Cg; ++j) v[(size_t)(c * Cg + j)] /= hs[sgoff[g] + c];
          complete_orthonormal(v, Cg, 1, Cg, K, null);
          for (i64 c = 0; c < K; ++c)
            if (!null[c])
              for (i64 j = 0; j < Cg; ++j) v[(size_t)(c * Cg + j)] *= hs[sgoff[g] + c];
            else
              hsB[sgoff[g] + c] = 1.0;
          HIP_CHECK(hipMemcpy(SVo.get() + soff[g], v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        }
      }
      if (right) up(ctx, sigB, hsB);
    }
  }
  k_sect_emit(ctx, true, nA, d1, Dl, Dm, nl, d_ablk.get(), d_rsec.get(), d_roff.get(), d_dl.get(), d_kept.get(),
              d_R.get(), d_uoff.get(), d_sgoff.get(), sig.get(), Uo.get(), right, out.A->data.get());
  k_sect_emit(ctx, false, nB, d2, Dm, Dr, nr, d_bblk.get(), d_csec.get(), d_coff.get(), d_dr.get(), d_kept.get(),
              d_C.get(), d_soff.get(), d_sgoff.get(), sigB.get() ? sigB.get() : sig.get(), SVo.get(), right,
              out.B->data.get());
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return out;
}

}  // namespace sb

namespace sb {

namespace {

// one thin QR per job on the device; A is a scratch copy, Q / R / tau sized by the jobs
struct QrRun {
  DBuf<double> A, Q, R, tau;
  DBuf<QrJob> jobs;
};

void run_qr(Ctx* ctx, QrRun& q, std::vector<QrJob>& jobs, i64 na, i64 nq, i64 nr, i64 nt) {
  q.Q.alloc(ctx, (size_t)std::max<i64>(nq, 1));
  q.R.alloc(ctx, (size_t)std::max<i64>(nr, 1));
  q.tau.alloc(ctx, (size_t)std::max<i64>(nt, 1));
  (void)na;
  up(ctx, q.jobs, jobs);
  k_qr_batch(ctx, q.jobs.get(), (int)jobs.size(), q.A.get(), q.Q.get(), q.R.get(), q.tau.get());
}

void copy_map(Ctx* ctx, const std::vector<i64>& smap, const double* src, const std::vector<i64>& dmap, double* dst) {
  DBuf<i64> s, d;
  up(ctx, s, smap);
  up(ctx, d, dmap);
  k_copy_map(ctx, (i64)smap.size(), s.get(), src, d.get(), dst);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the index buffers go back to the cache
}

}  // namespace

// blocksparse_left_qr (ops_factorize.jl:8-43): A {T,3,1,2} with prefix (l, r)
// and dense s; M[(li-1) d_s + s, ri] = A[l_vals[li], r_vals[ri], s], thin QR,
// Q {T,3,2} keyed by l with dense (s, new bond), one block per l value in
// ascending order (the reference's setindex! creates them in that order), R
// the k x n_r matrix.
LeftQrResult bs_left_qr(Ctx* ctx, const Tensor& A) {
  if (A.kind != BS || A.N() != 3 || A.N2 != 1)
    fail(SB_E_METHOD, "MethodError: blocksparse_left_qr needs a NewBlockSparseSorted{T,3,1,2}");
  const i64 n = A.n, ds = A.dims[2];
  if (n == 0) fail(SB_E_ARGUMENT, "ArgumentError: blocksparse_left_qr of a tensor without blocks");
  std::vector<i32> hk, hi;
  down(ctx, hk, A.keys.get(), (size_t)n * 2);
  down(ctx, hi, A.ids.get(), (size_t)n);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  LeftQrResult res;
  for (i64 i = 0; i < n; ++i) {
    res.l_vals.push_back(hk[(size_t)(2 * i)]);
    res.r_vals.push_back(hk[(size_t)(2 * i + 1)]);
  }
  auto uniq = [](std::vector<i64>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  uniq(res.l_vals);
  uniq(res.r_vals);
  std::map<i64, i64> li, ri;
  for (size_t i = 0; i < res.l_vals.size(); ++i) li[res.l_vals[i]] = (i64)i;
  for (size_t i = 0; i < res.r_vals.size(); ++i) ri[res.r_vals[i]] = (i64)i;
  const i64 nl = (i64)res.l_vals.size(), nr = (i64)res.r_vals.size();
  const i64 m = nl * ds, k = std::min(m, nr);
  if (m >= (1ll << 31) || nr >= (1ll << 31)) fail(SB_E_LIMIT, "blocksparse_left_qr: matrix too large");
  QrRun q;
  q.A.alloc(ctx, (size_t)(m * nr));
  k_zero(ctx, q.A.get(), (size_t)(m * nr) * sizeof(double));
  std::vector<i64> sm, dm;
  for (i64 i = 0; i < n; ++i) {
    const i64 r0 = li[hk[(size_t)(2 * i)]] * ds, c = ri[hk[(size_t)(2 * i + 1)]];
    for (i64 s = 0; s < ds; ++s) {
      sm.push_back((i64)(hi[(size_t)i] - 1) * A.blksize + s);
      dm.push_back(r0 + s + m * c);
    }
  }
  copy_map(ctx, sm, A.dptr(), dm, q.A.get());
  std::vector<QrJob> jobs{QrJob{(int)m, (int)nr, 0, 0, 0, 0}};
  run_qr(ctx, q, jobs, m * nr, m * k, k * nr, k);
  // Q: block li holds Q_mat[(li) d_s + s, b] at s + d_s b
  res.Q = bs_make(ctx, {A.dims[0], ds, k}, 2);
  Tensor& Q = *res.Q;
  Q.n = nl;
  std::vector<i32> qk(res.l_vals.begin(), res.l_vals.end()), qi((size_t)nl);
  for (i64 i = 0; i < nl; ++i) qi[(size_t)i] = (i32)(i + 1);
  up(ctx, Q.keys, qk);
  up(ctx, Q.ids, qi);
  Q.data.alloc(ctx, (size_t)std::max<i64>(nl * Q.blksize, 1));
  sm.clear();
  dm.clear();
  for (i64 l = 0; l < nl; ++l)
    for (i64 b = 0; b < k; ++b)
      for (i64 s = 0; s < ds; ++s) {
        sm.push_back(l * ds + s + m * b);
        dm.push_back(l * Q.blksize + s + ds * b);
      }
  copy_map(ctx, sm, q.Q.get(), dm, Q.data.get());
  res.R = dense_new(ctx, {k, nr});
  HIP_CHECK(hipMemcpyAsync(res.R->dptr(), q.R.get(), (size_t)(k * nr) * sizeof(double), hipMemcpyDeviceToDevice,
                           ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  Q.touch();
  return res;
}

// left_factorize (tensor_wrappers.jl:954-970) below the ITensor layer: Q of
// blocksparse_left_qr and R scattered to the full k x d_r (column r_vals[ri] of
// R_full = column ri of R)
LeftQrResult bs_left_factorize(Ctx* ctx, const Tensor& A) {
  LeftQrResult res = bs_left_qr(ctx, A);
  const i64 k = res.R->dims[0], nr = res.R->dims[1], dr = A.dims[1];
  TP full = dense_new(ctx, {k, dr});
  k_zero(ctx, full->dptr(), (size_t)(k * dr) * sizeof(double));
  std::vector<i64> sm, dm;
  for (i64 c = 0; c < nr; ++c)
    for (i64 b = 0; b < k; ++b) {
      sm.push_back(b + k * c);
      dm.push_back(b + k * (res.r_vals[(size_t)c] - 1));
    }
  copy_map(ctx, sm, res.R->dptr(), dm, full->dptr());
  res.R = std::move(full);
  return res;
}

// Sector QR of a sector-format site tensor A [Ql, Qr | s, l, r]: per right
// sector ir the matrix with rows (Ql, s, l) over the blocks (il, ir) of A whose
// (il, s) has rowsec[il + nl s] == ir (ascending il, then s, then l < dl[il])
// and columns r < dr[ir]; thin QR each (one workgroup per sector).  Q has A's
// keys and ids with the columns b < kept[ir] = min(rows, dr[ir]); R [Qr, Qr | m, r]
// holds the sectors' k x dr factors on its diagonal keys, ids ascending.
SectQrResult bs_qr_sectors(Ctx* ctx, const Tensor& A, const i64* rowsec, const i64* dl, const i64* dr) {
  if (A.kind != BS || A.N() != 5 || A.N2 != 3)
    fail(SB_E_METHOD, "MethodError: the sector QR needs a site tensor [Ql, Qr | s, l, r]");
  const i64 nl = A.dims[0], nr = A.dims[1], d = A.dims[2], Dl = A.dims[3], Dr = A.dims[4];
  for (i64 i = 0; i < nl; ++i)
    if (dl[i] < 0 || dl[i] > Dl) fail(SB_E_BOUNDS, "BoundsError: sector extent beyond the padded axis");
  for (i64 i = 0; i < nr; ++i)
    if (dr[i] < 0 || dr[i] > Dr) fail(SB_E_BOUNDS, "BoundsError: sector extent beyond the padded axis");
  for (i64 i = 0; i < nl * d; ++i)
    if (rowsec[i] < -1 || rowsec[i] >= nr) fail(SB_E_BOUNDS, "BoundsError: row sector out of range");
  std::vector<i32> hk, hi;
  down(ctx, hk, A.keys.get(), (size_t)A.n * 2);
  down(ctx, hi, A.ids.get(), (size_t)A.n);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::vector<std::vector<std::pair<i64, i64>>> by_r((size_t)nr);  // ir -> (il, block row)
  for (i64 i = 0; i < A.n; ++i) by_r[(size_t)(hk[(size_t)(2 * i + 1)] - 1)].push_back({hk[(size_t)(2 * i)] - 1, i});
  SectQrResult res;
  res.kept.assign((size_t)nr, 0);
  std::vector<QrJob> jobs;
  std::vector<i64> sm, dm, jsec;
  i64 na = 0, nq = 0, nrr = 0, nt = 0;
  struct Row {
    i64 blk, s, l;
  };
  std::vector<std::vector<Row>> rows((size_t)nr);
  for (i64 ir = 0; ir < nr; ++ir) {
    auto& v = by_r[(size_t)ir];
    std::sort(v.begin(), v.end());
    for (auto& b : v)
      for (i64 s = 0; s < d; ++s)
        if (rowsec[b.first + nl * s] == ir)
          for (i64 l = 0; l < dl[b.first]; ++l) rows[(size_t)ir].push_back({b.second, s, l});
    const i64 m = (i64)rows[(size_t)ir].size(), n = dr[ir], k = std::min(m, n);
    if (k == 0) continue;
    if (m >= (1ll << 31)) fail(SB_E_LIMIT, "sector QR: sector matrix too large");
    jobs.push_back(QrJob{(int)m, (int)n, na, nq, nrr, nt});
    jsec.push_back(ir);
    for (i64 c = 0; c < n; ++c)
      for (i64 x = 0; x < m; ++x) {
        const Row& rw = rows[(size_t)ir][(size_t)x];
        sm.push_back((i64)(hi[(size_t)rw.blk] - 1) * A.blksize + rw.s + d * (rw.l + Dl * c));
        dm.push_back(na + x + m * c);
      }
    res.kept[(size_t)ir] = k;
    na += m * n;
    nq += m * k;
    nrr += k * n;
    nt += k;
  }
  QrRun q;
  q.A.alloc(ctx, (size_t)std::max<i64>(na, 1));
  if (na) k_zero(ctx, q.A.get(), (size_t)na * sizeof(double));
  copy_map(ctx, sm, A.dptr(), dm, q.A.get());
  run_qr(ctx, q, jobs, na, nq, nrr, nt);
  // Q: A's block list, zero payload, the factors' columns scattered back
  res.Q = bs_make(ctx, A.dims, 3);
  Tensor& Q = *res.Q;
  Q.n = A.n;
  up(ctx, Q.keys, hk);
  up(ctx, Q.ids, hi);
  Q.data.alloc(ctx, (size_t)std::max<i64>(A.n * A.blksize, 1));
  if (A.n) k_zero(ctx, Q.data.get(), (size_t)(A.n * A.blksize) * sizeof(double));
  sm.clear();
  dm.clear();
  for (size_t j = 0; j < jobs.size(); ++j) {
    const QrJob& J = jobs[j];
    const i64 ir = jsec[j], k = std::min(J.m, J.n);
    for (i64 b = 0; b < k; ++b)
      for (i64 x = 0; x < J.m; ++x) {
        const Row& rw = rows[(size_t)ir][(size_t)x];
        sm.push_back(J.qoff + x + (i64)J.m * b);
        dm.push_back((i64)(hi[(size_t)rw.blk] - 1) * A.blksize + rw.s + d * (rw.l + Dl * b));
      }
  }
  copy_map(ctx, sm, q.Q.get(), dm, Q.data.get());
  // R [Qr, Qr | m, r]: one diagonal block per factored sector
  res.R = bs_make(ctx, {nr, nr, Dr, Dr}, 2);
  Tensor& R = *res.R;
  const i64 nb = (i64)jobs.size();
  R.n = nb;
  std::vector<i32> rk((size_t)(2 * nb)), ri((size_t)nb);
  for (i64 j = 0; j < nb; ++j) {
    rk[(size_t)(2 * j)] = rk[(size_t)(2 * j + 1)] = (i32)(jsec[(size_t)j] + 1);
    ri[(size_t)j] = (i32)(j + 1);
  }
  up(ctx, R.keys, rk);
  up(ctx, R.ids, ri);
  R.data.alloc(ctx, (size_t)std::max<i64>(nb * R.blksize, 1));
  if (nb) k_zero(ctx, R.data.get(), (size_t)(nb * R.blksize) * sizeof(double));
  sm.clear();
  dm.clear();
  for (i64 j = 0; j < nb; ++j) {
    const QrJob& J = jobs[(size_t)j];
    const i64 k = std::min(J.m, J.n);
    for (i64 c = 0; c < J.n; ++c)
      for (i64 b = 0; b < k; ++b) {
        sm.push_back(J.roff + b + k * c);
        dm.push_back(j * R.blksize + b + Dr * c);
      }
  }
  copy_map(ctx, sm, q.R.get(), dm, R.data.get());
  Q.touch();
  R.touch();
  return res;
}

}  // namespace sb
