// Contraction plans: the symbolic phase runs once at plan time; the numeric
// phase is a fixed list of kernel launches over fixed buffers, replayable
// whenever the operands' values change but their block structure does not.
#pragma once

#include <functional>
#include <memory>

#include "sb_ops.h"

namespace sb {

struct Label {
  uint64_t id, plev;
  bool operator==(const Label& o) const { return id == o.id && plev == o.plev; }
  bool operator!=(const Label& o) const { return !(*this == o); }
  bool operator<(const Label& o) const { return id < o.id || (id == o.id && plev < o.plev); }
};
using Labels = std::vector<Label>;

struct Step {
  const char* name;  // kernel family, for timing
  std::function<void()> fn;
  void operator()() const { fn(); }
};

// The block product of a plan whose numeric phase is exactly box_scan,
// box_prepare, grouped_gemm (the strided BS x BS product): what a batch needs
// to run it inside a launch shared with other plans.
struct BoxPlan {
  BoxGemmArgs g;
  const Tensor* A = nullptr;
  const Tensor* B = nullptr;
  Tensor* C = nullptr;
  unsigned flags = 0;
  std::vector<i64> cfull;  // C's full tail extents
};

// Device pointers one box product reads and writes.
struct BoxOperands {
  const i32 *abox, *aflag, *bbox, *bflag;
  i32 *cbox, *cflag;
  const double *A, *B;
  double* C;
};
// The autotuner's choice for one box product (sb_autotune.cpp): tile shape,
// workgroups per CU (0: as many as are resident), whether it came from the
// histogram cache, the best time measured and the candidates timed.
struct BoxTuneResult {
  int cfg = 0, wgcu = 0;
  bool cached = false;
  double ms = 0;
  int candidates = 0;
};
// Times the candidate tile shapes and grid densities of the box product g
// (group boxes hb on the host, full row/column extents M, N) on its operands,
// or reuses the choice cached for the same block-size histogram.  Runs the
// product several times: C ends holding its result.
BoxTuneResult autotune_box(Ctx* ctx, unsigned flags, const BoxGemmArgs& g, const std::vector<i32>& hb, i64 M, i64 N,
                           const BoxOperands& o);
void autotune_cache_clear();

struct Plan {
  Ctx* ctx = nullptr;
  std::vector<Step> steps;
  std::vector<std::shared_ptr<void>> hold;
  // accounting of the numeric phase (per execution)
  double flops = 0;        // sum of 2*M*K*N over contributing pairs (actual extents)
  double gemm_bytes = 0;   // algorithmic bytes of the block products (operands read once + C written)
  double perm_bytes = 0;   // bytes moved by operand permutes (read + write)
  i64 npairs = 0, nblocks_out = 0;
  double mfma_flops = 0;   // matrix-core FLOPs one execution issues, padding included (0: not counted)
  std::string family;

  template <class T>
  T* keep(DBuf<T>&& b) {
    auto sp = std::make_shared<DBuf<T>>(std::move(b));
    hold.push_back(sp);
    return sp->get();
  }
  void keep_tensor(TP&& t) { hold.push_back(std::shared_ptr<Tensor>(t.release())); }
  size_t done = 0;  // steps [0, done) already ran during planning
  double a_bytes = 0, b_bytes = 0, c_bytes = 0;  // distinct payload bytes of the block-product operands
  // the output tensor: every execution changes its payload; plans whose
  // last step leaves valid block boxes in it say so in out_box
  Tensor* out = nullptr;
  bool out_box = false;
  const Tensor* in_a = nullptr;  // the operands the plan reads at every execution
  const Tensor* in_b = nullptr;
  std::shared_ptr<BoxPlan> box;  // set when the steps are one box product
  BoxTuneResult tune;            // the autotuner's choice (candidates == 0: not tuned)
  void finish() const {
    if (!out) return;
    out->touch();
    if (out_box) out->box_version = out->version;
  }
  void run_step(const Step& s) const {
    if (!ctx->profiled(s.name)) {
      s();
      return;
    }
    hipEvent_t a = ctx->event(), b = ctx->event();
    (void)hipEventRecord(a, ctx->stream);
    s();
    (void)hipEventRecord(b, ctx->stream);
    ctx->prof_push({s.name, a, b});
  }
  // every step (a replay with refreshed operand values)
  void run() const {
    for (auto& s : steps) run_step(s);
    finish();
  }
  // the steps not yet run (planning needs some values mid-way)
  void run_pending() {
    for (size_t i = done; i < steps.size(); ++i) run_step(steps[i]);
    done = steps.size();
    finish();
  }
};

// permutedims(::AliasedBlockSparse, perm) (aliased/storage.jl:172-239)
TP al_permute(Ctx* ctx, const Tensor& A, const std::vector<i64>& perm);

// contract!(C, labelsC, A, labelsA, B, labelsB) (SparseBackends/src/tensoralgebra/contract.jl:37-104).
// C is overwritten: its keys/ids are set at plan time, its payload by every
// plan execution.  The plan keeps pointers to A's and B's payloads.
std::unique_ptr<Plan> contract_plan(Ctx* ctx, Tensor& C, const Labels& labelsC, const Tensor& A,
                                    const Labels& labelsA, const Tensor& B, const Labels& labelsB, unsigned flags);

// The two-site H_eff*psi (the four-call chain psi*L*W1*W2*R of two-site DMRG)
// as one plan with one fused kernel (sb_heff.hip); out gets the chain's block
// list, and every execution psi's current H_eff image.
// SB_TRACE=1 in the environment: planning prints its dispatch decisions
bool trace_on();

std::unique_ptr<Plan> heff_plan(Ctx* ctx, Tensor& out, const Tensor& psi, const Tensor& L, const Tensor& W1,
                                const Tensor& W2, const Tensor& R, unsigned flags);

}  // namespace sb
