// Measured choice of the box product's tile shape and grid density.
//
// The padded-work cost model picks a tile from the output blocks' boxes alone;
// what actually runs fastest also depends on the K extents, the operand reuse
// through L2 and how many persistent workgroups share a CU.  With
// SB_FLAG_AUTOTUNE a box plan instead times every tile shape (workgroup tiles
// and one-wave tiles) at three grid densities on its own operands at plan
// creation (one warm-up and
// three timed executions of the prepare + product pair each, the fastest
// kept), and remembers the winner under the plan's block-size histogram, so a
// later plan with the same histogram reuses it without timing.
#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>

#include "sb_contract.h"

namespace sb {

namespace {

int bitlen(i64 x) {
  int b = 0;
  while (x > 0) {
    ++b;
    x >>= 1;
  }
  return b;
}

std::mutex g_tune_mu;
std::map<std::string, BoxTuneResult>& tune_cache() {
  static std::map<std::string, BoxTuneResult> m;
  return m;
}

// the histogram of the output blocks' (row, column) box extents in log2 bins,
// counts in log2 bins too, with the precision form, the device's CU count and
// the number of reduced axes
std::string histogram_key(const Ctx* ctx, unsigned prec, const BoxGemmArgs& g, const std::vector<i32>& hb) {
  const BoxDesc& d = g.d;
  std::map<std::pair<int, int>, i64> h;
  for (i64 q = 0; q < g.ng; ++q) {
    i64 mc = 1, nc = 1;
    for (int a = 0; a < d.nm; ++a) mc *= hb[q * (d.nm + d.nn) + a];
    for (int a = 0; a < d.nn; ++a) nc *= hb[q * (d.nm + d.nn) + d.nm + a];
    if (mc && nc) ++h[{bitlen(mc), bitlen(nc)}];
  }
  std::string k = "p" + std::to_string(prec & SB_PREC_MASK) + ":cu" + std::to_string(ctx->num_cu) + ":k" +
                  std::to_string(d.nk) + ":";
  char buf[48];
  for (auto& kv : h) {
    std::snprintf(buf, sizeof(buf), "%d.%d.%d;", kv.first.first, kv.first.second, bitlen(kv.second));
    k += buf;
  }
  return k;
}

}  // namespace

BoxTuneResult autotune_box(Ctx* ctx, unsigned flags, const BoxGemmArgs& g0, const std::vector<i32>& hb, i64 M, i64 N,
                           const BoxOperands& o) {
  BoxTuneResult r;
  r.cfg = g0.cfg;
  r.wgcu = g0.wgcu;
  if (g0.ng == 0) return r;
  const std::string key = histogram_key(ctx, flags, g0, hb);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = tune_cache().find(key);
    if (it != tune_cache().end()) {  // the cached shape; nothing timed for this plan
      r.cfg = it->second.cfg;
      r.wgcu = it->second.wgcu;
      r.cached = true;
      return r;
    }
  }
  i64 maxt = 1;
  for (int c = 0; c < box_tile_configs(); ++c) {
    const BoxTile t = box_tile(c);
    maxt = std::max<i64>(maxt, g0.ng * ((M + t.BM - 1) / t.BM) * ((N + t.BN - 1) / t.BN));
  }
  if (maxt >= (1ll << 31)) return r;  // too large to try every shape: keep the cost model's choice
  DBuf<int2> trec(ctx, (size_t)maxt);
  hipEvent_t ea = ctx->event(), eb = ctx->event();
  BoxGemmArgs g = g0;
  g.trec = trec.get();
  auto run = [&] {
    box_prepare(ctx, g, o.abox, o.aflag, o.bbox, o.bflag, o.cbox, o.cflag, o.C);
    box_gemm(ctx, flags, g, o.A, o.B, o.C, o.cflag);
  };
  double best = 1e300;
  const int dens[3] = {0, 1, 2};  // resident workgroups per CU: as many as fit, one, two
  for (int c = 0; c < box_tile_configs(); ++c) {  // workgroup tiles and one-wave tiles
    for (int w : dens) {
      g.cfg = c;
      g.wgcu = w;
      run();  // warm-up
      float tmin = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        HIP_CHECK(hipEventRecord(ea, ctx->stream));
        run();
        HIP_CHECK(hipEventRecord(eb, ctx->stream));
        HIP_CHECK(hipEventSynchronize(eb));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ea, eb));
        tmin = std::min(tmin, ms);
      }
      ++r.candidates;
      if (tmin < best) {
        best = tmin;
        r.cfg = c;
        r.wgcu = w;
      }
    }
  }
  ctx->event_return(ea);
  ctx->event_return(eb);
  r.ms = best;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_cache()[key] = r;
  }
  return r;
}

void autotune_cache_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_cache().clear();
}

}  // namespace sb
