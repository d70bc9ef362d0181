// Plan chains replayed as HIP graphs (sb_graph.cpp).
#pragma once

#include <vector>

#include "sb_contract.h"

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
  Ctx* ctx;
  std::vector<const Plan*> plans;
  hipGraphExec_t exec = nullptr;
  std::vector<uintptr_t> cap_state;
  size_t nodes_ = 0;
};

// whether an eigensolve uses a chain graph for its operator applications:
// ctx->graphs 1 always, 0 never, -1 (default) when it may apply the operator
// at least eight times (krylovdim * maxiter >= 8), where the capture pays off
bool eig_uses_graph(const Ctx* ctx, int kdim, int maxiter);

}  // namespace sb
