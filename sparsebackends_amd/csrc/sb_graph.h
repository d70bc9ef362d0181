// Plan chains replayed as HIP graphs (sb_graph.cpp).
#pragma once

#include <functional>
#include <vector>

#include "sb_contract.h"

#include <memory>

namespace sb {

// A chain of plans executed in order (e.g. the four contractions of one
// H_eff*psi).  launch() executes it: eagerly the first time, captured into a
// graph the second, replayed from then on while the chain's buffers and its
// host-side decisions match the capture.  Chains with a plan that is not one
// box product always run eagerly, and so does every launch while the context
// profiles.
class ChainGraph {
 public:
  ChainGraph(Ctx* ctx, std::vector<const Plan*> plans);
  // a batch execution: run() executes the plans (listed in the order their
  // steps run: stage by stage, the chains of a stage in order) as one
  // application, e.g. Batch::run over all chains
  // extra(): a word of further host state the captured launches depend on
  // (a replay needs the value the capture saw; 0: the execution is not
  // capturable and runs eagerly)
  ChainGraph(Ctx* ctx, std::vector<const Plan*> plans, std::function<void()> run,
             std::function<uintptr_t()> extra = nullptr);
  ~ChainGraph();
  ChainGraph(const ChainGraph&) = delete;
  ChainGraph& operator=(const ChainGraph&) = delete;
  void launch();
  long captures = 0, replays = 0, eager = 0;
  size_t nodes() const { return nodes_; }
  bool eager_only = false;

 private:
  std::vector<uintptr_t> buffers() const;
  std::vector<uintptr_t> state() const;
  bool state_after_ok(const std::vector<uintptr_t>& before) const;
  void bookkeep() const;
  void run_eager();
  Ctx* ctx;
  std::vector<const Plan*> plans;
  std::function<void()> run_;
  std::function<uintptr_t()> extra_;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> cap_state;
  size_t nodes_ = 0;
};

struct Batch;
// the graph of one execution of every chain of a batch (Batch::run over all)
std::unique_ptr<ChainGraph> batch_graph(Ctx* ctx, Batch& batch);

// whether an eigensolve uses a chain graph for its operator applications:
// ctx->graphs 1 always, 0 never, -1 (default) when it may apply the operator
// at least eight times (krylovdim * maxiter >= 8), where the capture pays off
bool eig_uses_graph(const Ctx* ctx, int kdim, int maxiter);

}  // namespace sb
