// sdv_common.h -- shared device/host helpers for the MI355X sinc-DVR operator
// library (complex types, error reporting, launch checks).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "sincdvr.h"

namespace sdv {

constexpr double kPiD = 3.141592653589793238462643383279502884;

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T r, T i) {
  cplx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) {
  return mk(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) {
  return mk(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> scale(cplx<T> a, T s) {
  return mk(a.re * s, a.im * s);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) {
  return mk(a.re, -a.im);
}
// multiply by -i (forward) or +i (inverse)
template <typename T>
__host__ __device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) {
  return mk(a.im, -a.re);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> mul_pi(cplx<T> a) {
  return mk(-a.im, a.re);
}

using z64 = cplx<double>;
using c32 = cplx<float>;

// Last-error string, per process (the ABI is called from one host thread at a
// time, like the reference's project routines).
void set_error(const std::string& msg);
const char* last_error();

}  // namespace sdv

#define SDV_HIP_CHECK(expr)                                                                  \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      ::sdv::set_error(std::string(#expr) + " failed: " + hipGetErrorString(_e) + " at " +   \
                       __FILE__ + ":" + std::to_string(__LINE__));                            \
      return SDV_ERR_HIP;                                                                     \
    }                                                                                         \
  } while (0)

#define SDV_REQUIRE(cond, code, msg)        \
  do {                                      \
    if (!(cond)) {                          \
      ::sdv::set_error(msg);                \
      return code;                          \
    }                                       \
  } while (0)
