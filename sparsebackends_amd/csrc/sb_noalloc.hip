// Entry points that allocate no device memory, for callers that cannot call
// hipMalloc: code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation.  Their scratch is one file-scope device array per
// GPU, reached through the address hipGetSymbolAddress gives for the current
// device; the host never passes the array's name to a kernel or a copy.
//
// Every call on one GPU shares that GPU's array and leaves its last contents
// there; calls that overlap on one GPU (two streams, two threads) race on it.
#include <mutex>

#include "sb_krylov.h"

namespace sb {

int linalg_nb();
void k_mdot(Ctx*, const double* V, i64 ldv, int k, const double* w, i64 n, bool norm, double* part, double* out);

namespace {

// the block partial sums of the norm reduction (sb_linalg.hip uses 1024)
constexpr int kScratchDoubles = 1024;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

__global__ void sqrt_first_kernel(double* p) { p[0] = sqrt(p[0]); }

}  // namespace

// Looks up this device's copy of the scratch array (the device must be
// current); called once per device when a context is created on it.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) fail(SB_E_LIMIT, "device number above the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return;
  void* p = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_ptr[device] = static_cast<double*>(p);
}

static double* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) fail(SB_E_LIMIT, "device number above the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) fail(SB_E_DEVICE, "scratch of the allocation-free entry points not set up on this device");
  return g_ptr[device];
}

// norm(data) of a block-sparse payload (tensor_wrappers.jl:793) into out[0],
// device memory, on the context's stream: the same kernels and partition as
// bs_norm, the 1024 partials in the scratch array, no allocation and no
// synchronisation.  The partial count does not depend on the payload length,
// so every tensor fits.
void bs_norm_noalloc(Ctx* ctx, const Tensor& A, double* out) {
  if (A.kind != BS) fail(SB_E_METHOD, "MethodError: norm of a tensor that is not block-sparse");
  if (!out) fail(SB_E_ARGUMENT, "ArgumentError: null output pointer");
  if (linalg_nb() > kScratchDoubles) fail(SB_E_LIMIT, "reduction partials exceed the scratch array");
  double* part = scratch_ptr(ctx->device);
  k_mdot(ctx, nullptr, 0, 0, A.dptr(), A.n * A.blksize, true, part, out);
  hipLaunchKernelGGL(sqrt_first_kernel, dim3(1), dim3(1), 0, ctx->stream, out);
  HIP_CHECK(hipGetLastError());
}

}  // namespace sb
