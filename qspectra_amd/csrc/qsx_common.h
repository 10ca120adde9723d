// Shared host/device helpers for libqsx (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/qsx.h"

namespace qsx {

// ------------------------------------------------------------ error state
void set_error(const char* fmt, ...);

#define QSX_CHECK_ARG(cond, ...)              \
  do {                                        \
    if (!(cond)) {                            \
      ::qsx::set_error(__VA_ARGS__);          \
      return QSX_EINVAL;                      \
    }                                         \
  } while (0)

#define QSX_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::qsx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                   \
      return QSX_EHIP;                                                        \
    }                                                                         \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// --------------------------------------------------------- complex algebra
template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };
template <typename T> using cplx_t = typename cplx<T>::type;

__host__ __device__ inline double2 to_d2(double2 a) { return a; }
__host__ __device__ inline double2 to_d2(float2 a) { return make_double2(a.x, a.y); }
template <typename T> __host__ __device__ inline cplx_t<T> from_d2(double2 a);
template <> __host__ __device__ inline double2 from_d2<double>(double2 a) { return a; }
template <> __host__ __device__ inline float2 from_d2<float>(double2 a) {
  return make_float2((float)a.x, (float)a.y);
}

// acc += a * b (complex)
template <typename T>
__device__ inline void cfma(cplx_t<T>& acc, cplx_t<T> a, cplx_t<T> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 caxpy(double s, double2 x, double2 y) {  // y + s*x
  return make_double2(fma(s, x.x, y.x), fma(s, x.y, y.y));
}
__device__ inline bool cfinite(double2 a) { return isfinite(a.x) && isfinite(a.y); }

}  // namespace qsx
