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

// qsx_zgemm's validation and launch, for use by other entry points of the
// library (zgemm.hip): enqueues C = alpha op(A) op(B) + beta C + gamma I.
int zgemm_enqueue(int precision, int trans, int m, int n, int k, double alpha_re,
                  double alpha_im, const void* A, int64_t lda, int64_t sA, const void* B,
                  int64_t ldb, int64_t sB, double beta_re, double beta_im, void* C, int64_t ldc,
                  int64_t sC, double gamma_re, double gamma_im, int batch, void* stream);

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

// A zero held in a VGPR that the compiler cannot see through.  Indexing a
// wave-uniform load with it makes it a vector-memory load: s_barrier is
// preceded by a wait for every outstanding scalar-memory load, so a scalar
// prefetch would stall the next barrier, while a vector load is waited for
// only at its first use.
__device__ inline int opaque_zero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}
__device__ inline int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double wave_uniform(double v) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                          __builtin_amdgcn_readfirstlane(__double2loint(v)));
}

// Wait for all outstanding global loads (s_waitcnt vmcnt(0); the expcnt and
// lgkmcnt fields at their maxima do not wait).  Placed in uniform control
// flow before a loop, it leaves the compiler no load that it must assume
// pending inside the loop.
__device__ inline void wait_global_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Save-interval table (step count, step length) of a stepping loop, one
// block of 64 intervals at a time held in registers: lane l of every
// wavefront holds interval base + l.  load() waits for the block once, so
// the intervals of the block are then read without waiting on global memory
// (on gfx9 a wait for a load also waits for every store issued before it,
// such as the states saved at the end of the previous interval).
struct IntervalBlock {
  int n_l;
  double h_l;

  // intervals base .. base + 63 of n_[0, count), h_[0, count)
  __device__ void load(const int* n_, const double* h_, int count, int base) {
    const int e = base + __lane_id();
    n_l = e < count ? n_[e] : 0;
    h_l = e < count ? h_[e] : 0.0;
    // consume the loads here, before the loop that reads the block
    asm volatile("" : : "v"(n_l), "v"(h_l));
  }
  // interval base + sl (sl wave-uniform, < 64)
  __device__ int steps(int sl) const { return __builtin_amdgcn_readlane(n_l, sl); }
  __device__ double length(int sl) const {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(h_l), sl),
                            __builtin_amdgcn_readlane(__double2loint(h_l), sl));
  }
};

// Lane move within rows of 16 lanes by a DPP control (gfx9 encoding:
// 0x00-0xFF quad permutation, 0x140 row mirror, 0x141 half-row mirror).
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false),
                          __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over aligned segments of S lanes (S = 1, 2, 4, 8 or 16); every lane of
// a segment receives the same total.  Butterfly over quad permutations
// [1,0,3,2] and [2,3,0,1], then the half-row and row mirrors (lane i with
// 7-i, then with 15-i: after the previous levels all lanes of each half
// hold that half's sum).  DPP moves stay in the VALU instead of routing
// through the LDS crossbar as wave shuffles do.
template <int S, typename F>
__device__ inline F seg_sum(F v) {
  static_assert(S == 1 || S == 2 || S == 4 || S == 8 || S == 16, "segment of 1..16 lanes");
  if (S >= 2) v += dpp_move<0xB1>(v);
  if (S >= 4) v += dpp_move<0x4E>(v);
  if (S >= 8) v += dpp_move<0x141>(v);
  if (S >= 16) v += dpp_move<0x140>(v);
  return v;
}

}  // namespace qsx
