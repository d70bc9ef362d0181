// blocksparse_svd (SparseBackends/src/ops_factorize.jl:50-148): the two-site
// block-diagonal SVD with global truncation.  The per-group factorisations
// and the payloads run on the device (sb_svd.hip); grouping, the global
// singular-value order, truncation and the key/id bookkeeping of the two
// factors (a few values per kept singular triplet) run on the host.
#include <algorithm>
#include <map>

#include "sb_factorize.h"

namespace sb {

void k_svd_jacobi(Ctx*, const double* A, int d1, int d2, i64 ng, const i64* gstart, const i64* bsrc, const i64* woff,
                  double* W, double* sig, double* Uout, int* perm);
void k_svd_emit(Ctx*, i64 k, int d1, const i32* kg, const i32* kc, const i64* soff, const i64* gstart, int d2,
                const i64* woff, const double* W, const double* sig, const double* Uout, const int* perm, bool right,
                double* Ud, double* SVd);

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
  DBuf<int> perm(ctx, ng * d1 ? ng * d1 : 1);
  k_svd_jacobi(ctx, A.dptr(), d1, d2, ng, d_gstart.get(), d_bsrc.get(), d_woff.get(), W.get(), sig.get(), Uo.get(),
               perm.get());
  std::vector<double> hs;
  down(ctx, hs, sig.get(), (size_t)(ng * d1));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));

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
  return R;
}

}  // namespace sb
