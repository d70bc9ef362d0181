// Plan chains replayed as HIP graphs.
//
// The numeric phase of a box-product plan is a fixed list of kernel launches
// (box scans of operands whose values changed, the C box fill, the group
// preparation, the tile scan, the work list, the product) whose arguments are
// fixed buffers.  What varies between executions is host-side: which operands
// need a box scan (their version changed) and whether C's boxes are trusted.
// A ChainGraph therefore runs the chain eagerly once, captures the second
// execution into a graph, and from then on replays that graph whenever the
// chain's buffers are unchanged and every such host decision is the one the
// capture made, applying the same host-side bookkeeping (box versions, output
// versions) a live execution would.  Anything else runs eagerly, and a
// decision pattern seen for the first time is captured again (the executable
// graph is updated in place where HIP allows it).
#include "sb_graph.h"

#include "sb_batch.h"

#include <algorithm>
#include <cstring>

namespace sb {

namespace {

bool graphable(const Plan* p) { return p && p->box && p->steps.size() == 3; }

struct SimTrust {
  std::vector<std::pair<const Tensor*, bool>> t;
  bool get(const Tensor* x) {
    for (auto& e : t)
      if (e.first == x) return e.second;
    t.push_back({x, x->box_trusted()});
    return t.back().second;
  }
  void set(const Tensor* x) {
    get(x);
    for (auto& e : t)
      if (e.first == x) e.second = true;
  }
};

}  // namespace

ChainGraph::ChainGraph(Ctx* c, std::vector<const Plan*> ps) : ctx(c), plans(std::move(ps)) {
  if (plans.empty()) fail(SB_E_ARGUMENT, "ArgumentError: a plan chain needs at least one plan");
  for (auto* p : plans) {
    if (!p) fail(SB_E_HANDLE, "null plan");
    if (p->ctx != ctx) fail(SB_E_ARGUMENT, "ArgumentError: every plan of a chain must belong to its context");
    if (!graphable(p)) eager_only = true;  // only box-product chains are captured
  }
}

ChainGraph::ChainGraph(Ctx* c, std::vector<const Plan*> ps, std::function<void()> run,
                       std::function<uintptr_t()> extra)
    : ChainGraph(c, std::move(ps)) {
  run_ = std::move(run);
  extra_ = std::move(extra);
}

void ChainGraph::run_eager() {
  if (run_)
    run_();
  else
    for (auto* p : plans) p->run();
}

ChainGraph::~ChainGraph() {
  if (exec) (void)hipGraphExecDestroy(exec);
}

// the buffers the captured launches bake in (five words per operand and
// result of every plan)
std::vector<uintptr_t> ChainGraph::buffers() const {
  std::vector<uintptr_t> s;
  for (auto* p : plans) {
    const BoxPlan& b = *p->box;
    for (const Tensor* t : {b.A, b.B, static_cast<const Tensor*>(b.C)}) {
      s.push_back(reinterpret_cast<uintptr_t>(t->data.get()));
      s.push_back(reinterpret_cast<uintptr_t>(t->box.get()));
      s.push_back(reinterpret_cast<uintptr_t>(t->bflag.get()));
      s.push_back((uintptr_t)t->box.n);
      s.push_back((uintptr_t)t->bflag.n);
    }
  }
  return s;
}

// the buffers, then the scan / fill decisions an execution from the current
// host state takes (simulated plan by plan)
std::vector<uintptr_t> ChainGraph::state() const {
  std::vector<uintptr_t> s = buffers();
  SimTrust tr;
  for (auto* p : plans) {
    const BoxPlan& b = *p->box;
    s.push_back(tr.get(b.A) ? 1 : 0);  // box scans of the operands
    tr.set(b.A);
    s.push_back(tr.get(b.B) ? 1 : 0);
    tr.set(b.B);
    s.push_back(tr.get(b.C) ? 1 : 0);  // C box fill
    tr.set(b.C);                       // the product leaves C's boxes valid
  }
  if (extra_) s.push_back(extra_());
  return s;
}

// the host-side effects of one execution (what ensure_box and finish() do),
// after a replay
void ChainGraph::bookkeep() const {
  for (auto* p : plans) {
    const BoxPlan& b = *p->box;
    if (!b.A->box_trusted()) b.A->box_version = b.A->version;
    if (!b.B->box_trusted()) b.B->box_version = b.B->version;
    p->finish();
  }
}

namespace {

struct Saved {
  Tensor* t;
  i64 version, box_version;
};

}  // namespace

void ChainGraph::launch() {
  if (eager_only || ctx->profile) {
    run_eager();
    ++eager;
    return;
  }
  const std::vector<uintptr_t> st = state();
  if (exec && st == cap_state) {
    HIP_CHECK(hipGraphLaunch(exec, ctx->stream));
    bookkeep();
    ++replays;
    return;
  }
  // the first execution allocates what the chain needs, and one whose extra
  // state is not the replayable one (e.g. a batch whose active lists must be
  // uploaded again) synchronises: neither is captured
  if (eager == 0 || (extra_ && st.back() == 0)) {
    run_eager();
    ++eager;
    return;
  }
  // capture this execution; the host effects of the captured run are undone
  // if the graph cannot be used, and the chain then runs eagerly
  std::vector<Saved> saved;
  for (auto* p : plans)
    for (const Tensor* t : {p->box->A, p->box->B, static_cast<const Tensor*>(p->box->C)}) {
      Tensor* m = const_cast<Tensor*>(t);
      saved.push_back({m, m->version, m->box_version});
    }
  // the captured run executed nothing: undo its host effects, then run the
  // chain eagerly from now on
  auto abandon = [&](hipGraph_t g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    for (auto& s : saved) {
      s.t->version = s.version;
      s.t->box_version = s.box_version;
    }
    eager_only = true;
    run_eager();
    ++eager;
  };
  hipGraph_t g = nullptr;
  HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  try {
    run_eager();
  } catch (const Error& e) {
    (void)hipStreamEndCapture(ctx->stream, &g);
    if (e.code != SB_E_DEVICE) {  // not a call the capture refused
      if (g) (void)hipGraphDestroy(g);
      for (auto& s : saved) {
        s.t->version = s.version;
        s.t->box_version = s.box_version;
      }
      throw;
    }
    abandon(g);
    return;
  } catch (...) {
    (void)hipStreamEndCapture(ctx->stream, &g);  // never leave the stream capturing
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  if (hipStreamEndCapture(ctx->stream, &g) != hipSuccess || !g) {
    abandon(g);
    return;
  }
  bool ok = state_after_ok(st);
  // only kernel nodes: a memset node is not replayed reliably on this device
  size_t nn = 0;
  if (ok) ok = hipGraphGetNodes(g, nullptr, &nn) == hipSuccess;
  if (ok && nn) {
    std::vector<hipGraphNode_t> nodes(nn);
    ok = hipGraphGetNodes(g, nodes.data(), &nn) == hipSuccess;
    for (size_t i = 0; ok && i < nn; ++i) {
      hipGraphNodeType ty;
      ok = hipGraphNodeGetType(nodes[i], &ty) == hipSuccess &&
           (ty == hipGraphNodeTypeKernel || ty == hipGraphNodeTypeEmpty);
    }
  }
  nodes_ = nn;
  if (!ok) {
    abandon(g);
    return;
  }
  bool updated = false;
  if (exec) {
    hipGraphExecUpdateResult res;
    hipGraphNode_t err = nullptr;
    updated = hipGraphExecUpdate(exec, g, &err, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (!updated) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(exec);
      exec = nullptr;
    }
  }
  if (!exec && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) {
    exec = nullptr;
    abandon(g);
    return;
  }
  (void)hipGraphDestroy(g);
  cap_state = st;
  ++captures;
  HIP_CHECK(hipGraphLaunch(exec, ctx->stream));  // the captured execution itself
}

// the capture must not have moved a buffer (an allocation during the capture
// would be baked into the graph while the next replay's state differs)
bool ChainGraph::state_after_ok(const std::vector<uintptr_t>& before) const {
  const std::vector<uintptr_t> now = buffers();
  return before.size() >= now.size() && std::equal(now.begin(), now.end(), before.begin());
}

std::unique_ptr<ChainGraph> batch_graph(Ctx* ctx, Batch& batch) {
  std::vector<const Plan*> plans;
  for (auto& st : batch.stages)
    for (auto* p : st.plans) plans.push_back(p);
  Batch* b = &batch;
  // a replay reads the active-chain lists the stages last uploaded: they must
  // be the full ones the capture saw
  auto full = [b]() -> uintptr_t {
    for (auto& st : b->stages)
      if (st.fused && (int)st.act_host.size() != b->nchains) return 0;
    return 1;
  };
  return std::make_unique<ChainGraph>(ctx, plans, [b] { b->run({}); }, full);
}

bool eig_uses_graph(const Ctx* ctx, int kdim, int maxiter) {
  if (ctx->graphs == 0) return false;
  if (ctx->graphs > 0) return true;
  return (long)kdim * maxiter >= 8;
}

}  // namespace sb
