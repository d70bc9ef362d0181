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
#include <cstdlib>
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

// the first key coordinate (in storage order: key by key, axis by axis, as
// the reference's `for key in A.keys, v in key`) outside [lo, hi]
__global__ void key_first_bad_kernel(const i32* __restrict__ keys, i64 n, double lo, double hi,
                                     unsigned long long* __restrict__ first) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    const double v = (double)keys[i];
    if (v < lo || v > hi) atomicMin(first, (unsigned long long)i);
  }
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

// Julia's shortest printing of a finite value: the fewest significant digits
// that read back as v in its type (keep(r) says whether r does), plain when
// 1e-4 <= |v| < 1e6 (with ".0" on integral values), else d.ddd and the
// exponent after `ex` ("e" for Float64, "f" for Float32); plain Float32
// values end in "f0"
std::string julia_digits(double v, bool (*keep)(double, double), const char* ex, const char* plain_suffix) {
  char b[64];
  for (int p = 1; p <= 17; ++p) {
    std::snprintf(b, sizeof b, "%.*e", p - 1, v);
    if (keep(std::strtod(b, nullptr), v)) break;
  }
  std::string s = b, sign;
  if (s[0] == '-') sign = "-", s = s.substr(1);
  const size_t e = s.find('e');
  const int x = std::atoi(s.c_str() + e + 1);
  std::string dig;
  for (size_t i = 0; i < e; ++i)
    if (s[i] != '.') dig += s[i];
  while (dig.size() > 1 && dig.back() == '0') dig.pop_back();
  std::string out;
  if (x >= -4 && x <= 5) {
    if (x >= 0) {
      std::string ip = dig.substr(0, std::min<size_t>(dig.size(), (size_t)x + 1));
      while ((int)ip.size() < x + 1) ip += '0';
      const std::string fp = dig.size() > (size_t)x + 1 ? dig.substr((size_t)x + 1) : "0";
      out = ip + "." + fp;
    } else {
      out = "0." + std::string((size_t)(-x - 1), '0') + dig;
    }
    return sign + out + plain_suffix;
  }
  out = dig.substr(0, 1) + "." + (dig.size() > 1 ? dig.substr(1) : "0");
  return sign + out + ex + std::to_string(x);
}

bool same_f64(double r, double v) { return r == v; }
bool same_f32(double r, double v) { return (double)(float)r == v; }
bool same_f16(double r, double v) { return (double)(_Float16)r == v; }

// v as Julia prints a Float64
std::string fmt_num(double v) {
  if (v != v) return "NaN";
  if (std::isinf(v)) return v > 0 ? "Inf" : "-Inf";
  if (v == 0) return std::signbit(v) ? "-0.0" : "0.0";
  return julia_digits(v, same_f64, "e", "");
}

// c, a value of the narrower float type T2 held in a double, as Julia
// prints that type: 0.1f0, 1.0f-5 (Float32); Float16(0.1) (Float16)
std::string fmt_narrow(double c, int code) {
  if (c != c || std::isinf(c)) {
    const std::string s = c != c ? "NaN" : (c > 0 ? "Inf" : "-Inf");
    return code == SB_VAL_FLOAT32 ? s + "32" : s + "16";
  }
  if (code == SB_VAL_FLOAT32)
    return c == 0 ? (std::signbit(c) ? "-0.0f0" : "0.0f0") : julia_digits(c, same_f32, "f", "f0");
  return "Float16(" + (c == 0 ? std::string(std::signbit(c) ? "-0.0" : "0.0") : julia_digits(c, same_f16, "e", "")) +
         ")";
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
      DBuf<unsigned long long> f(ctx, 1);
      const unsigned long long none = ~0ull;
      HIP_CHECK(hipMemcpyAsync(f.get(), &none, sizeof none, hipMemcpyHostToDevice, ctx->stream));
      hipLaunchKernelGGL(key_first_bad_kernel, dim3((unsigned)std::min<i64>((nv + 255) / 256, 1024)), dim3(256), 0,
                         ctx->stream, A.keys.get(), nv, K.lo, K.hi, f.get());
      HIP_CHECK(hipGetLastError());
      unsigned long long first;
      HIP_CHECK(hipMemcpyAsync(&first, f.get(), sizeof first, hipMemcpyDeviceToHost, ctx->stream));
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
      if (first != none) {
        i32 v;
        HIP_CHECK(hipMemcpy(&v, A.keys.get() + first, sizeof v, hipMemcpyDeviceToHost));
        fail(SB_E_ERROR, std::string(A.kind == COO ? "Coordinate " : "Key coordinate ") + std::to_string(v) +
                             " is out of range for " + K.name + " (" + kKeyMin[key_code] + ".." +
                             kKeyMax[key_code] + ")");
      }
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
      if (why == 2)  // the InexactError T2(v) throws, as Julia prints it
        fail(SB_E_ERROR, "Value " + fmt_num(v) + " cannot be converted to " + T2 + ": InexactError(:" + T2 + ", " +
                             T2 + ", " + fmt_num(v) + ")");
      const double c = val_code == SB_VAL_FLOAT32   ? (double)(float)v
                       : val_code == SB_VAL_FLOAT16 ? (double)(_Float16)v
                                                    : v;
      const std::string cs = val_code == SB_VAL_FLOAT64 ? fmt_num(c) : fmt_narrow(c, val_code);
      fail(SB_E_ERROR, "Value " + fmt_num(v) + " → " + cs + " → " + fmt_num(c) +
                           ": roundtrip failed; conversion to " + T2 + " is lossy");
    }
  }
  TP R = copy_tensor(ctx, A);
  R->val_type = val_code;
  return R;
}

}  // namespace sb
