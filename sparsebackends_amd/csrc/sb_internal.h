// Internal runtime of libsparsebackends_amd: contexts, the caching device
// allocator, tensor handles and error mapping.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sb_status.h"

namespace sb {

using i64 = int64_t;
using i32 = int32_t;

struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg);
#define SB_ASSERT(c, m)                           \
  do {                                            \
    if (!(c)) ::sb::fail(SB_E_ASSERT, std::string(m)); \
  } while (0)
#define HIP_CHECK(x)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      ::sb::fail(e_ == hipErrorOutOfMemory ? SB_E_NOMEM : SB_E_DEVICE,                    \
                 std::string(#x) + ": " + hipGetErrorString(e_));                         \
  } while (0)

// ------------------------------------------------------------ allocator
// Size-class caching allocator: device blocks are kept after release and
// reused for later requests of the same class, so steady-state contractions
// call no hipMalloc/hipFree (both synchronise the device).  Several host
// threads may share one context: the cache is guarded by a mutex, and since a
// context has one stream, a block released by one thread and handed to
// another is reused in stream order.
class DeviceCache {
 public:
  explicit DeviceCache(int device) : device_(device) {}
  ~DeviceCache();
  void* get(size_t bytes);
  void put(void* p);
  void trim();
  size_t bytes_in_use() const {
    std::lock_guard<std::mutex> lk(mu_);
    return in_use_;
  }
  size_t bytes_cached() const {
    std::lock_guard<std::mutex> lk(mu_);
    return cached_;
  }

 private:
  static size_t cls(size_t b);
  void trim_locked();
  mutable std::mutex mu_;
  int device_;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> live_;
  size_t in_use_ = 0, cached_ = 0;
};

struct Ctx;

// RAII device buffer from a context's cache
template <class T>
struct DBuf {
  Ctx* ctx = nullptr;
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(Ctx* c, size_t count) { alloc(c, count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept { *this = std::move(o); }
  DBuf& operator=(DBuf&& o) noexcept;
  ~DBuf() { release(); }
  void alloc(Ctx* c, size_t count);
  void release();
  T* get() const { return p; }
};

// per-kernel timing of plan steps with HIP events on the context's stream
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};
struct ProfAgg {
  long count = 0;
  double ms = 0, min_ms = 1e300, max_ms = 0;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  DeviceCache cache;
  int num_cu = 256;
  // kernel timing with HIP events on the stream: 0 off, 1 every timed step,
  // 2 only the block products (grouped_gemm), the fewest events
  int profile = 0;
  int svd_sweeps = 40;  // sweep cap of the one-sided Jacobi SVDs (SB_E_NOCONV when reached)
  int graphs = -1;  // eigensolves replay their operator as a HIP graph: 1 always, 0 never, -1 for long solves
  bool profiled(const char* name) const;
  // event pool and profiling records: several host threads may share a
  // context (bind_context), so every access holds prof_mu
  std::mutex prof_mu;
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, ProfAgg> prof;
  explicit Ctx(int dev);
  ~Ctx();
  hipEvent_t event();
  void event_return(hipEvent_t e);
  void prof_push(const ProfRec& r);
  void prof_collect();  // synchronises, folds finished records into prof
};

template <class T>
DBuf<T>& DBuf<T>::operator=(DBuf&& o) noexcept {
  if (this != &o) {
    release();
    ctx = o.ctx;
    p = o.p;
    n = o.n;
    o.p = nullptr;
    o.n = 0;
  }
  return *this;
}
template <class T>
void DBuf<T>::alloc(Ctx* c, size_t count) {
  release();
  ctx = c;
  n = count;
  p = count ? static_cast<T*>(c->cache.get(count * sizeof(T))) : nullptr;
}
template <class T>
void DBuf<T>::release() {
  if (p && ctx) ctx->cache.put(p);
  p = nullptr;
  n = 0;
}

// ------------------------------------------------------------ tensors
enum Kind { BS = SB_KIND_BLOCKSPARSE, COO = SB_KIND_COO, DENSE = SB_KIND_DENSE, ALIASED = SB_KIND_ALIASED };

// A device-resident tensor.  Keys and ids are 1-based int32, row-wise per
// block (BS: n x P prefix coordinates; COO: n x N coordinates), matching the
// reference's NewBlockSparseSorted / COOTensor fields (blocksparse/storage.jl:3-9,
// coo/storage.jl:1-6).  Payload is float64, column-major within a block,
// blocks at (id-1)*blksize.
struct Tensor {
  Ctx* ctx = nullptr;
  Kind kind = BS;
  std::vector<i64> dims;
  int N2 = 0;
  i64 blksize = 1;
  i64 n = 0;            // nblocks (BS), nnz (COO); dense: 1
  DBuf<i32> keys;       // BS/COO
  DBuf<i32> ids;        // BS
  DBuf<double> data;    // payload
  double* borrowed = nullptr;  // dense tensor viewing caller memory
  bool dirty = false;   // COO
  int key_type = 0;     // declared key type (SB_KEY_*; the device keys are int32)
  int val_type = 0;     // declared payload type (SB_VAL_*; the device payload is float64)
  // aliased block-sparse (aliased/storage.jl:46-59)
  DBuf<double> scalars;
  i64 ntemplates = 0;
  // Block boxes (BS): per block and tail axis, one past the last index with a
  // nonzero entry (n x N2), and a flag per block marking non-finite values.
  // They describe the payload while box_version == version and no device
  // pointer to the payload was handed out (external).  Every change of the
  // payload or block list advances version.
  mutable DBuf<i32> box, bflag;
  mutable i64 box_version = -1;
  i64 version = 0;
  bool external = false;
  void touch() { ++version; }
  bool box_trusted() const {
    return kind == BS && !external && box_version == version && bflag.n >= (size_t)n && box.n >= (size_t)(n * N2);
  }

  int N() const { return (int)dims.size(); }
  int P() const { return (int)dims.size() - N2; }
  double* dptr() const { return borrowed ? borrowed : data.get(); }
  i64 numel() const {
    i64 p = 1;
    for (auto d : dims) p *= d;
    return p;
  }
};

Tensor* as_tensor(void* h, int kind = 0);
i64 prod(const std::vector<i64>& v, size_t lo, size_t hi);

}  // namespace sb
