// Runtime: error state, contexts (device + stream), caching device allocator.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>

#include "sb_internal.h"

namespace sb {

[[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

i64 prod(const std::vector<i64>& v, size_t lo, size_t hi) {
  i64 p = 1;
  for (size_t i = lo; i < hi; ++i) p *= v[i];
  return p;
}

// ---------------------------------------------------------------- allocator
size_t DeviceCache::cls(size_t b) {
  if (b <= 256) return 256;
  // round up to 1/8 steps of the next power of two (bounded waste)
  size_t p = 256;
  while (p < b) p <<= 1;
  size_t step = p / 8;
  return ((b + step - 1) / step) * step;
}

void* DeviceCache::get(size_t bytes) {
  size_t c = cls(bytes);
  std::lock_guard<std::mutex> lk(mu_);
  auto it = free_.lower_bound(c);
  if (it != free_.end() && it->first <= c + c / 4) {
    void* p = it->second;
    size_t sz = it->first;
    free_.erase(it);
    cached_ -= sz;
    live_[p] = sz;
    in_use_ += sz;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim_locked();
    HIP_CHECK(hipMalloc(&p, c));
  }
  live_[p] = c;
  in_use_ += c;
  return p;
}

void DeviceCache::put(void* p) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  free_.emplace(it->second, p);
  cached_ += it->second;
  in_use_ -= it->second;
  live_.erase(it);
}

void DeviceCache::trim() {
  std::lock_guard<std::mutex> lk(mu_);
  trim_locked();
}

void DeviceCache::trim_locked() {
  // caller guarantees no queued work uses cached blocks (device synchronised)
  (void)hipDeviceSynchronize();
  for (auto& kv : free_) (void)hipFree(kv.second);
  free_.clear();
  cached_ = 0;
}

DeviceCache::~DeviceCache() {
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(device_);
  (void)hipDeviceSynchronize();
  for (auto& kv : free_) (void)hipFree(kv.second);
  for (auto& kv : live_) (void)hipFree(kv.first);
  (void)hipSetDevice(cur);
}

Ctx::Ctx(int dev) : device(dev), cache(dev) {
  HIP_CHECK(hipSetDevice(dev));
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0) num_cu = n;
}

Ctx::~Ctx() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& r : prof_pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : event_pool) (void)hipEventDestroy(e);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

bool Ctx::profiled(const char* name) const {
  // level 2: the block products only (the box products and the fused H_eff kernel)
  return profile == 1 ||
         (profile == 2 && (std::strcmp(name, "grouped_gemm") == 0 || std::strcmp(name, "heff_fused") == 0));
}

hipEvent_t Ctx::event() {
  std::lock_guard<std::mutex> lk(prof_mu);
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  // timing-only events: without the system-scope fence on record, which would
  // write back and invalidate the caches between the kernels being timed
  hipEvent_t e;
  HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  return e;
}

void Ctx::event_return(hipEvent_t e) {
  std::lock_guard<std::mutex> lk(prof_mu);
  event_pool.push_back(e);
}

void Ctx::prof_push(const ProfRec& r) {
  std::lock_guard<std::mutex> lk(prof_mu);
  prof_pending.push_back(r);
}

void Ctx::prof_collect() {
  HIP_CHECK(hipStreamSynchronize(stream));
  std::lock_guard<std::mutex> lk(prof_mu);
  for (auto& r : prof_pending) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    ProfAgg& g = prof[r.name];
    g.count += 1;
    g.ms += ms;
    g.min_ms = std::min(g.min_ms, (double)ms);
    g.max_ms = std::max(g.max_ms, (double)ms);
    event_pool.push_back(r.a);
    event_pool.push_back(r.b);
  }
  prof_pending.clear();
}

Tensor* as_tensor(void* h, int kind) {
  if (!h) fail(SB_E_HANDLE, "null tensor handle");
  Tensor* t = static_cast<Tensor*>(h);
  if (kind && t->kind != kind) fail(SB_E_HANDLE, "tensor handle of the wrong storage kind");
  return t;
}

}  // namespace sb
