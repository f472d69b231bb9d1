// Shared device/host helpers for the slab Foldy-Lax library (gfx950).
//
// Complex values are stored interleaved (re, im) as HIP double2 / float2.
// Every kernel is templated on the real type R (double -> complex128,
// float -> complex64); FP64 accumulation is used for all reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace msc {

template <typename R> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename R> using cx = typename cvec<R>::type;

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R re, R im) {
  cx<R> v;
  v.x = re;
  v.y = im;
  return v;
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) {
  return mk<R>(a.x + b.x, a.y + b.y);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) {
  return mk<R>(a.x - b.x, a.y - b.y);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R>
__host__ __device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {
  return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> cscale(cx<R> a, R s) {
  return mk<R>(a.x * s, a.y * s);
}
// acc += a * b
template <typename R>
__host__ __device__ __forceinline__ void cfma(cx<R>& acc, cx<R> a, cx<R> b) {
  acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
  acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> czero() {
  return mk<R>(R(0), R(0));
}

// Widening / narrowing conversions between storage precisions.
__host__ __device__ __forceinline__ double2 to_c128(double2 v) { return v; }
__host__ __device__ __forceinline__ double2 to_c128(float2 v) {
  return make_double2(double(v.x), double(v.y));
}
template <typename R> __host__ __device__ __forceinline__ cx<R> from_c128(double2 v);
template <> __host__ __device__ __forceinline__ double2 from_c128<double>(double2 v) {
  return v;
}
template <> __host__ __device__ __forceinline__ float2 from_c128<float>(double2 v) {
  return make_float2(float(v.x), float(v.y));
}

// Error codes returned by every C-ABI entry point.
enum Status : int {
  kOk = 0,
  kInvalid = 1,      // bad argument (maps to ValueError)
  kUnsupported = 2,  // valid request this build does not handle
  kHip = 3,          // HIP runtime error
  kNotReady = 4,     // plan not initialised for the request
};

void set_error(const std::string& msg);
const char* get_error();

}  // namespace msc

#define MSC_HIP_CHECK(expr)                                                  \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::msc::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
      return ::msc::kHip;                                                    \
    }                                                                        \
  } while (0)

#define MSC_LAUNCH_CHECK()                                                   \
  do {                                                                       \
    hipError_t _e = hipGetLastError();                                       \
    if (_e != hipSuccess) {                                                  \
      ::msc::set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
      return ::msc::kHip;                                                    \
    }                                                                        \
  } while (0)

#define MSC_REQUIRE(cond, msg)                                               \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::msc::set_error(msg);                                                 \
      return ::msc::kInvalid;                                                \
    }                                                                        \
  } while (0)
