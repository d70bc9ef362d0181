// rekey and retypeval: conversions of the key integer type and of the
// payload element type (blocksparse/conversions.jl:321-342,
// coo/conversions.jl:36-82, aliased/conversions.jl:71-93).
//
// The reference converts the stored keys to K2 (after checking that every
// prefix dimension and key coordinate fits) and the stored values to T2
// (after checking that every value survives the round trip T(T2(v)) == v).
// On the device keys are int32 and payloads float64 whatever the declared
// types, and every value those checks admit is held exactly by them; so the
// checks run on the device and the copy carries the declared type, which the
// host mirror applies when it reads the tensor back.
#include <cmath>
#include <cstdio>
#include <string>

#include "sb_convert.h"

namespace sb {

namespace {

struct KeyType {
  const char* name;
  double lo, hi;
};
// SB_KEY_* codes
const KeyType kKeys[] = {
    {"Int64", -9223372036854775808.0, 9223372036854775807.0},
    {"Int32", -2147483648.0, 2147483647.0},
    {"Int16", -32768.0, 32767.0},
    {"Int8", -128.0, 127.0},
    {"UInt64", 0.0, 18446744073709551615.0},
    {"UInt32", 0.0, 4294967295.0},
    {"UInt16", 0.0, 65535.0},
    {"UInt8", 0.0, 255.0},
};
const char* kKeyMin[] = {"-9223372036854775808", "-2147483648", "-32768", "-128", "0", "0", "0", "0"};
const char* kKeyMax[] = {"9223372036854775807", "2147483647", "32767", "127",
                         "18446744073709551615", "4294967295", "65535", "255"};
// SB_VAL_* codes
const char* kVals[] = {"Float64", "Float32", "Float16", "Int64", "Int32", "Int16", "Int8"};

__global__ void key_range_kernel(const i32* __restrict__ keys, i64 n, int* __restrict__ out) {
  int lo = INT_MAX, hi = INT_MIN;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    lo = min(lo, keys[i]);
    hi = max(hi, keys[i]);
  }
  atomicMin(&out[0], lo);
  atomicMax(&out[1], hi);
}

// value v survives T2: 0, else the reason (1 non-integer for an integer T2,
// 2 out of T2's range, 3 the round trip changes it)
__device__ __host__ inline int lossless_fail(double v, int code) {
  switch (code) {
    case SB_VAL_FLOAT64:
      return v == v ? 0 : 3;  // only NaN fails: NaN != NaN
    case SB_VAL_FLOAT32:
      return (double)(float)v == v ? 0 : 3;  // NaN fails: NaN != NaN, as in the reference's T(c) == v
    case SB_VAL_FLOAT16:
      return (double)(_Float16)v == v ? 0 : 3;
    default: {
      if (!(v == std::floor(v))) return 1;  // NaN and Inf are not integers
      double lo = -9223372036854775808.0, hi = 9223372036854775808.0;  // [lo, hi)
      if (code == SB_VAL_INT32) lo = -2147483648.0, hi = 2147483648.0;
      if (code == SB_VAL_INT16) lo = -32768.0, hi = 32768.0;
      if (code == SB_VAL_INT8) lo = -128.0, hi = 128.0;
      return (v >= lo && v < hi) ? 0 : 2;
    }
  }
}

__global__ void lossless_kernel(const double* __restrict__ v, i64 n, int code, unsigned long long* __restrict__ first) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
    if (lossless_fail(v[i], code)) atomicMin(first, (unsigned long long)i);
}

std::string fmt_num(double v) {
  if (v == std::floor(v) && std::fabs(v) < 1e15) {
    char b[64];
    std::snprintf(b, sizeof b, "%.1f", v);
    return b;
  }
  char b[64];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

// a copy of every field and buffer of T (keys, ids, payload, scalars)
TP copy_tensor(Ctx* ctx, const Tensor& A) {
  TP R(new Tensor());
  R->ctx = ctx;
  R->kind = A.kind;
  R->dims = A.dims;
  R->N2 = A.N2;
  R->blksize = A.blksize;
  R->n = A.n;
  R->dirty = A.dirty;
  R->ntemplates = A.ntemplates;
  R->key_type = A.key_type;
  R->val_type = A.val_type;
  auto cp = [&](auto& dst, const auto& src, size_t count) {
    dst.alloc(ctx, std::max<size_t>(count, 1));
    if (count)
      HIP_CHECK(hipMemcpyAsync(dst.get(), src, count * sizeof(*src), hipMemcpyDeviceToDevice, ctx->stream));
  };
  cp(R->keys, A.keys.get(), A.keys.n);
  cp(R->ids, A.ids.get(), A.ids.n);
  cp(R->scalars, A.scalars.get(), A.scalars.n);
  const size_t nd = A.borrowed ? (size_t)A.numel() : A.data.n;
  cp(R->data, A.dptr(), nd);
  R->touch();
  return R;
}

}  // namespace

TP tensor_rekey(Ctx* ctx, const Tensor& A, int key_code, bool check) {
  if (A.kind == DENSE) fail(SB_E_METHOD, "MethodError: rekey has no method for a dense array");
  if (key_code < 0 || key_code > SB_KEY_UINT8) fail(SB_E_ARGUMENT, "ArgumentError: unknown key type code");
  const KeyType& K = kKeys[key_code];
  if (check) {
    if (A.kind != COO)  // the prefix dimensions (block-sparse and aliased)
      for (int i = 0; i < A.P(); ++i)
        if ((double)A.dims[(size_t)i] > K.hi)
          fail(SB_E_ERROR, "Dimension " + std::to_string(A.dims[(size_t)i]) + " at prefix axis " +
                               std::to_string(i + 1) + " cannot be indexed by " + K.name + " (max=" +
                               kKeyMax[key_code] + ")");
    const i64 nv = A.n * (A.kind == COO ? A.N() : A.P());
    if (nv > 0) {
      DBuf<int> r(ctx, 2);
      const int init[2] = {INT_MAX, INT_MIN};
      HIP_CHECK(hipMemcpyAsync(r.get(), init, sizeof init, hipMemcpyHostToDevice, ctx->stream));
      hipLaunchKernelGGL(key_range_kernel, dim3((unsigned)std::min<i64>((nv + 255) / 256, 1024)), dim3(256), 0,
                         ctx->stream, A.keys.get(), nv, r.get());
      HIP_CHECK(hipGetLastError());
      int h[2];
      HIP_CHECK(hipMemcpyAsync(h, r.get(), sizeof h, hipMemcpyDeviceToHost, ctx->stream));
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
      for (int v : {h[0], h[1]})
        if ((double)v < K.lo || (double)v > K.hi)
          fail(SB_E_ERROR, std::string(A.kind == COO ? "Coordinate " : "Key coordinate ") + std::to_string(v) +
                               " is out of range for " + K.name + " (" + kKeyMin[key_code] + ".." +
                               kKeyMax[key_code] + ")");
    }
  }
  TP R = copy_tensor(ctx, A);
  R->key_type = key_code;
  return R;
}

TP coo_retypeval(Ctx* ctx, const Tensor& A, int val_code) {
  if (A.kind != COO) fail(SB_E_METHOD, "MethodError: retypeval is defined for COOTensor");
  if (val_code < 0 || val_code > SB_VAL_INT8) fail(SB_E_ARGUMENT, "ArgumentError: unknown value type code");
  if (A.n > 0) {
    DBuf<unsigned long long> f(ctx, 1);
    const unsigned long long none = ~0ull;
    HIP_CHECK(hipMemcpyAsync(f.get(), &none, sizeof none, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(lossless_kernel, dim3((unsigned)std::min<i64>((A.n + 255) / 256, 1024)), dim3(256), 0,
                       ctx->stream, A.data.get(), A.n, val_code, f.get());
    HIP_CHECK(hipGetLastError());
    unsigned long long first;
    HIP_CHECK(hipMemcpyAsync(&first, f.get(), sizeof first, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (first != none) {  // the first value that does not survive, in the stored order
      double v;
      HIP_CHECK(hipMemcpy(&v, A.data.get() + first, sizeof v, hipMemcpyDeviceToHost));
      const std::string T2 = kVals[val_code];
      const int why = lossless_fail(v, val_code);
      if (why == 1) fail(SB_E_ERROR, "Value " + fmt_num(v) + " is non-integer; cannot convert to " + T2 + " without loss");
      if (why == 2) fail(SB_E_ERROR, "Value " + fmt_num(v) + " cannot be converted to " + T2 + ": InexactError");
      const double c = val_code == SB_VAL_FLOAT32   ? (double)(float)v
                       : val_code == SB_VAL_FLOAT16 ? (double)(_Float16)v
                                                    : v;
      fail(SB_E_ERROR, "Value " + fmt_num(v) + " → " + fmt_num(c) + " → " + fmt_num(c) +
                           ": roundtrip failed; conversion to " + T2 + " is lossy");
    }
  }
  TP R = copy_tensor(ctx, A);
  R->val_type = val_code;
  return R;
}

}  // namespace sb
