// Batches of independent plan chains: one box product launch per stage over
// the work lists of every chain.
#include "sb_batch.h"

#include <algorithm>
#include <map>

namespace sb {

namespace {

template <class F>
void timed(Ctx* ctx, const char* name, F&& f) {
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

i64 prod_ext(const int* e, int n) {
  i64 p = 1;
  for (int a = 0; a < n; ++a) p *= e[a];
  return p;
}

BoxJob job_of_plan(const BoxPlan& bp) {
  const BoxGemmArgs& g = bp.g;
  BoxJob j{};
  j.d = g.d;
  j.ng = g.ng;
  j.amf = g.amf;
  j.bnf = g.bnf;
  j.gorder = g.gorder;
  j.seg = g.seg;
  j.arow = g.arow;
  j.brow = g.brow;
  j.aoff = g.aoff;
  j.boff = g.boff;
  j.coff = g.coff;
  j.abox = bp.A->box.get();
  j.aflag = bp.A->bflag.get();
  j.bbox = bp.B->box.get();
  j.bflag = bp.B->bflag.get();
  j.gbox = g.gbox;
  j.kbox = g.kbox;
  j.kc = g.kc;
  j.cnew = g.cnew;
  j.clear = g.clear;
  j.cbox = bp.C->box.get();
  j.cflag = bp.C->bflag.get();
  j.A = bp.A->data.get();
  j.B = bp.B->data.get();
  j.C = bp.C->data.get();
  return j;
}

}  // namespace

std::unique_ptr<Batch> make_batch(Ctx* ctx, const std::vector<Plan*>& plans, int nchains, int nstages) {
  if (nchains < 1 || nstages < 1) fail(SB_E_ARGUMENT, "ArgumentError: a batch needs at least one chain and one stage");
  if ((i64)plans.size() != (i64)nchains * nstages) fail(SB_E_ARGUMENT, "ArgumentError: batch plan count");
  for (auto* p : plans) {
    if (!p) fail(SB_E_HANDLE, "null plan");
    // a plan of another context takes part while that context runs on this
    // one's stream (StreamLoan)
    if (p->ctx != ctx && (p->ctx->device != ctx->device || p->ctx->stream != ctx->stream))
      fail(SB_E_ARGUMENT, "ArgumentError: every plan of a batch must belong to its context");
  }
  auto b = std::make_unique<Batch>();
  b->ctx = ctx;
  b->nchains = nchains;
  b->nstages = nstages;
  b->stages.resize(nstages);
  for (int s = 0; s < nstages; ++s) {
    BatchStage& st = b->stages[s];
    for (int c = 0; c < nchains; ++c) {
      Plan* p = plans[(size_t)c * nstages + s];
      st.plans.push_back(p);
      b->flops += p->flops;
    }
    // fused when every plan is a single box product of the same precision form
    // with workgroup tiles; the common tile shape is the one carrying the most flops
    bool ok = true;
    std::map<int, double> vote;
    const unsigned prec0 = st.plans[0]->box ? (st.plans[0]->box->flags & SB_PREC_MASK) : 0;
    for (auto* p : st.plans) {
      if (!p->box || p->steps.size() != 3 || (p->box->flags & SB_PREC_MASK) != prec0) {
        ok = false;
        break;
      }
      // the batched launch takes workgroup tiles: plans that chose a one-wave tile do not vote
      if (!box_tile_is_wave(p->box->g.cfg)) vote[p->box->g.cfg] += p->flops + 1.0;
    }
    if (!ok) continue;
    st.fused = true;
    st.prec = prec0;
    st.cfg = vote.empty() ? 3  // 64 x 32
                          : std::max_element(vote.begin(), vote.end(), [](auto& x, auto& y) {
                              return x.second < y.second;
                            })->first;
    bool tune = false;  // a plan asked for measured tile shapes: time the stage's joint launch
    for (auto* p : st.plans) tune = tune || (p->box->flags & SB_FLAG_AUTOTUNE);
    std::vector<BoxJob> hj;
    i64 npos = 0, maxt = 0;
    for (auto* p : st.plans) {
      p->steps[0]();  // the operands' boxes exist (and are current) before their addresses are taken
      hj.push_back(job_of_plan(*p->box));
      st.ng.push_back(p->box->g.ng);
      npos += p->box->g.ng;
    }
    for (int c = 0; c < box_tile_configs(); ++c) {  // work list room for every shape the stage may take
      if (box_tile_is_wave(c)) continue;
      const BoxTile t = box_tile(c);
      i64 mt = 0;
      for (auto* p : st.plans) {
        const BoxDesc& d = p->box->g.d;
        const i64 M = prod_ext(d.m_ext, d.nm), N = prod_ext(d.n_ext, d.nn);
        mt += p->box->g.ng * ((M + t.BM - 1) / t.BM) * ((N + t.BN - 1) / t.BN);
      }
      maxt = std::max(maxt, mt);
    }
    SB_ASSERT(maxt < (1ll << 31), "batched box GEMM: work list too large");
    st.jobs.alloc(ctx, hj.size());
    HIP_CHECK(hipMemcpyAsync(st.jobs.get(), hj.data(), hj.size() * sizeof(BoxJob), hipMemcpyHostToDevice,
                             ctx->stream));
    st.act.alloc(ctx, (size_t)nchains);
    st.gbase.alloc(ctx, (size_t)nchains + 1);
    st.tiles.alloc(ctx, (size_t)std::max<i64>(1, npos));
    st.toff.alloc(ctx, (size_t)npos + 1);
    {
      const i64 zero = 0;
      k_fill_rows(ctx, st.toff.get(), 1, 1, &zero);  // toff[0]; the scan writes the rest
    }
    st.trec.alloc(ctx, (size_t)std::max<i64>(1, maxt));
    st.scan_bytes = box_scan_tmp_bytes(ctx, std::max<i64>(1, npos));
    st.scan_tmp.alloc(ctx, st.scan_bytes);
    HIP_CHECK(hipStreamSynchronize(ctx->stream));  // hj is uploaded
    if (tune) b->tune_stage(s);
  }
  return b;
}

// the fastest workgroup tile for stage s's joint launch over every chain:
// one warm-up and three timed executions of the stage per shape
void Batch::tune_stage(int s) {
  BatchStage& st = stages[s];
  std::vector<int> all(nchains);
  for (int c = 0; c < nchains; ++c) all[c] = c;
  hipEvent_t ea = ctx->event(), eb = ctx->event();
  double best = 1e300;
  int bestc = st.cfg;
  for (int c = 0; c < box_tile_configs(); ++c) {
    if (box_tile_is_wave(c)) continue;
    st.cfg = c;
    run_stage(s, all);
    float tmin = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      HIP_CHECK(hipEventRecord(ea, ctx->stream));
      run_stage(s, all);
      HIP_CHECK(hipEventRecord(eb, ctx->stream));
      HIP_CHECK(hipEventSynchronize(eb));
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, ea, eb));
      tmin = std::min(tmin, ms);
    }
    if (tmin < best) {
      best = tmin;
      bestc = c;
    }
  }
  ctx->event_return(ea);
  ctx->event_return(eb);
  st.cfg = bestc;
  st.tuned_ms = best;
}

void Batch::run_stage(int s, const std::vector<int>& active) {
  BatchStage& st = stages[s];
  if (!st.fused) {
    for (int c : active) st.plans[c]->run();
    return;
  }
  // the box scans (only where an operand changed) and the C box rows
  timed(ctx, "box_scan", [&] {
    for (int c : active) {
      Plan* p = st.plans[c];
      p->steps[0]();
      const BoxPlan& bp = *p->box;
      if (!bp.C->box_trusted()) k_fill_rows(ctx, bp.C->box.get(), bp.g.ng, bp.C->N2, bp.cfull.data());
    }
  });
  if (st.act_host != active) {  // the active set changed: new job list and position bases
    std::vector<i64> gb(active.size() + 1, 0);
    for (size_t a = 0; a < active.size(); ++a) gb[a + 1] = gb[a] + st.ng[active[a]];
    std::vector<i32> act(active.begin(), active.end());
    HIP_CHECK(hipMemcpyAsync(st.act.get(), act.data(), act.size() * sizeof(i32), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipMemcpyAsync(st.gbase.get(), gb.data(), gb.size() * sizeof(i64), hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    st.act_host = active;
  }
  BoxBatchArgs a;
  a.jobs = st.jobs.get();
  a.act = st.act.get();
  a.gbase = st.gbase.get();
  a.nact = (int)active.size();
  a.cfg = st.cfg;
  a.npos = 0;
  for (int c : active) a.npos += st.ng[c];
  a.tiles = st.tiles.get();
  a.toff = st.toff.get();
  a.trec = st.trec.get();
  a.scan_tmp = st.scan_tmp.get();
  a.scan_tmp_bytes = st.scan_bytes;
  timed(ctx, "box_prepare", [&] { box_prepare_batch(ctx, a); });
  timed(ctx, "grouped_gemm", [&] { box_gemm_batch(ctx, st.prec, a); });
  for (int c : active) st.plans[c]->finish();
}

StreamLoan::StreamLoan(Ctx* host, const std::vector<Ctx*>& guests) : host_(host) {
  for (Ctx* g : guests) {
    if (!g) fail(SB_E_HANDLE, "null context");
    if (g == host || std::find(guests_.begin(), guests_.end(), g) != guests_.end()) continue;
    if (g->device != host->device) fail(SB_E_ARGUMENT, "ArgumentError: the chains' contexts are on another device");
    guests_.push_back(g);
  }
  for (Ctx* g : guests_) {
    // the guest's queued work is done before the host's stream takes over
    hipError_t e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
      restore();
      HIP_CHECK(e);
    }
    saved_.push_back(g->stream);
    g->stream = host->stream;
  }
}

void StreamLoan::restore() {
  for (size_t i = 0; i < saved_.size(); ++i) guests_[i]->stream = saved_[i];
  saved_.clear();
}

StreamLoan::~StreamLoan() {
  // whatever the loaned contexts queued on the host's stream is done before
  // they get their own streams back (their caches then reuse blocks in order)
  (void)hipStreamSynchronize(host_->stream);
  restore();
}

void Batch::run(const std::vector<int>& active_in) {
  std::vector<int> active = active_in;
  if (active.empty())
    for (int c = 0; c < nchains; ++c) active.push_back(c);
  for (size_t i = 0; i < active.size(); ++i)
    if (active[i] < 0 || active[i] >= nchains || (i && active[i] <= active[i - 1]))
      fail(SB_E_ARGUMENT, "ArgumentError: active chains must be ascending, distinct and in range");
  for (int s = 0; s < nstages; ++s) run_stage(s, active);
}

}  // namespace sb
