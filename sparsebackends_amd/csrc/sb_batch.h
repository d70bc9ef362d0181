// Batches of independent plan chains (one chain per problem, e.g. the four
// contractions of one H_eff*psi): stage s of every chain runs as ONE box
// product launch over the union of their work lists, so many small problems
// fill the device together instead of one launch each.
#pragma once

#include <vector>

#include "sb_contract.h"

namespace sb {

struct BatchStage {
  std::vector<Plan*> plans;  // plan of every chain at this stage
  bool fused = false;        // every plan is one box product of one precision form
  int cfg = 0;
  double tuned_ms = 0;  // the stage's time with cfg when measured (SB_FLAG_AUTOTUNE)
  unsigned prec = 0;
  std::vector<i64> ng;       // groups per job
  DBuf<BoxJob> jobs;
  DBuf<i32> act;             // active job ids of the last upload
  DBuf<i64> gbase;
  std::vector<int> act_host;  // what act / gbase were uploaded for
  DBuf<i32> tiles, toff;
  DBuf<int4> trec;
  DBuf<unsigned char> scan_tmp;
  size_t scan_bytes = 0;
};

struct Batch {
  Ctx* ctx = nullptr;
  int nchains = 0, nstages = 0;
  std::vector<BatchStage> stages;
  double flops = 0;  // all chains, one execution
  // run every stage for the chains in active (ascending, distinct; empty: all)
  void run(const std::vector<int>& active);
  // the chains of one stage: the fused launch, or the plans one by one
  void run_stage(int s, const std::vector<int>& active);
  // time every workgroup tile for stage s over all chains and keep the fastest
  void tune_stage(int s);
};

// Lends the stream of host to the guest contexts for the lifetime of the
// object: each guest's stream is synchronised, then the guest launches on
// host's stream (work of the plans and tensors of several contexts can then
// be ordered and batched on one stream); at the end host's stream is
// synchronised and the guests get their streams back.  The guests must not be
// used by other threads meanwhile.
class StreamLoan {
 public:
  StreamLoan(Ctx* host, const std::vector<Ctx*>& guests);
  ~StreamLoan();
  StreamLoan(const StreamLoan&) = delete;
  StreamLoan& operator=(const StreamLoan&) = delete;

 private:
  void restore();
  Ctx* host_;
  std::vector<Ctx*> guests_;
  std::vector<hipStream_t> saved_;
};

// plans[c * nstages + s]: stage s of chain c.  Plans of one stage must be
// independent (no plan reads another chain's output).
std::unique_ptr<Batch> make_batch(Ctx* ctx, const std::vector<Plan*>& plans, int nchains, int nstages);

}  // namespace sb
