// Register-resident fixed-step propagation of y' = L y (gfx950).
//
// One workgroup propagates one right-hand side r of one problem b.  The
// operator (L for classical RK4, or the one-step RK4 propagator) is split
// into row segments: thread t owns row i = t / S and the columns
// j = k + S*c (k = t % S, c < W) of that row, held in registers for the
// whole launch; rows beyond NT/S (only for the largest blocks) are read
// from LDS by the first threads.  A matrix-vector product is W complex
// FMAs per thread plus a DPP butterfly over the S lanes of a row (S is a
// power of two, so a row's lanes are adjacent in one wavefront), and the
// stage vector is exchanged through a double-buffered LDS vector padded
// with zeros to S*W entries (so the row segments read it without bounds
// tests): one barrier per product.  The step counts and lengths of the
// save intervals are held in registers, 64 intervals per wavefront register
// (IntervalBlock), so the loop does not wait on global memory between saves.
//
// Modes
//   RK4  (PROP = false): y <- y + h/6 (k1 + 2 k2 + 2 k3 + k4), k = L z,
//        four products per step (the reference's ODE right-hand side, applied
//        by a classical explicit Runge-Kutta stepper).
//   PROP (PROP = true): y <- P y with P = I + hL + (hL)^2/2 + (hL)^3/6 +
//        (hL)^4/24, the exact one-step map of classical RK4 for a time-
//        invariant L: one product per step, same result up to rounding.
// Precision
//   FP64: operator, stage vector and products in complex128.
//   MIXED: the off-diagonal operator in complex64 with FP32 accumulation;
//        the diagonal (largest entries: rotating-frame frequencies and
//        dephasing) in complex128 applied to the thread's FP64 stage value;
//        state and update in FP64.  In PROP mode the operator is P - I.

#include <type_traits>

#include "qsx_common.h"

namespace qsx {
namespace {

__constant__ double kRkW[4] = {1.0, 2.0, 2.0, 1.0};
__constant__ double kRkC[3] = {0.5, 0.5, 1.0};

// One propagation phase of a work-group: operator rows Ab (this problem's
// A), FP64 diagonal diagb (MIXED), state y of the thread's rows (lead lanes;
// in/out), save intervals nsub/hs (or, with nsub == nullptr, one interval of
// n_single steps), saved states to ob [nsave][M] (none when ob == nullptr).
template <typename T, int NT, int S, int W, bool MIXED, bool PROP>
__device__ __forceinline__ void step_phase(
    const cplx_t<T>* __restrict__ Ab, const double2* __restrict__ diagb, int M,
    double2 (&y)[2], const int* __restrict__ nsub, const double* __restrict__ hs, int nsave,
    int n_single, double2* __restrict__ ob, bool& bad) {
  using T2 = cplx_t<T>;
  constexpr int kRegRows = NT / S;
  constexpr int MP = S * W;                            // padded row length (>= M)
  extern __shared__ __align__(16) unsigned char smem[];
  T2* zb = reinterpret_cast<T2*>(smem);               // [2][MP] stage vector
  T2* Aov = zb + 2 * MP;                               // [M - kRegRows][MP] overflow rows

  const int tid = threadIdx.x;
  const int k = tid % S;
  const int nov = M > kRegRows ? M - kRegRows : 0;

  // this thread's rows: slot 0 = primary row (registers), slot 1 = overflow row (LDS)
  const int row0 = tid / S;
  const int row1 = kRegRows + tid / S;
  const bool has0 = row0 < M;
  const bool has1 = (tid / S) < nov;
  const bool lead = (k == 0);

  T2 a[W];
#pragma unroll
  for (int c = 0; c < W; ++c) {
    const int j = k + S * c;
    a[c] = (has0 && j < M) ? Ab[(int64_t)row0 * M + j] : T2{0, 0};
  }
  __syncthreads();                       // LDS of a previous phase is free
  for (int e = tid; e < nov * MP; e += NT) {
    const int j = e % MP;
    Aov[e] = j < M ? Ab[(int64_t)(kRegRows + e / MP) * M + j] : T2{0, 0};
  }
  for (int e = tid; e < 2 * MP; e += NT) zb[e] = T2{0, 0};

  double2 acc[2], zi[2], dd[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int row = q == 0 ? row0 : row1;
    const bool on = q == 0 ? has0 : has1;
    dd[q] = (MIXED && on) ? diagb[row] : make_double2(0.0, 0.0);
    acc[q] = zi[q] = make_double2(0.0, 0.0);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int row = q == 0 ? row0 : row1;
    const bool on = q == 0 ? has0 : has1;
    if (on && lead) zb[row] = from_d2<T>(y[q]);
  }
  __syncthreads();

  auto matvec = [&](const T2* z, int q) -> double2 {
    // FP64 segments wider than 16 columns keep a per-column test (not
    // needed for bounds: the stage vector is zero-padded): it serialises the
    // LDS loads, which otherwise would all be hoisted and exceed the 128
    // registers a 1024-thread block leaves beside the coefficients
    constexpr bool kSerial = sizeof(T) == 8 && W > 16;
    T2 s = {0, 0}, s2 = {0, 0};
    if (q == 0) {
#pragma unroll
      for (int c = 0; c < W; ++c) {
        const int j = k + S * c;
        if (!kSerial || j < M) cfma<T>((c & 1) ? s2 : s, a[c], z[j]);
      }
    } else if (has1) {
      // overflow rows use the same column split
      const T2* Ar = Aov + (int64_t)(tid / S) * MP;
#pragma unroll
      for (int c = 0; c < W; ++c) {
        const int j = k + S * c;
        if (!kSerial || j < M) cfma<T>((c & 1) ? s2 : s, Ar[j], z[j]);
      }
    }
    s.x = seg_sum<S>(s.x + s2.x);
    s.y = seg_sum<S>(s.y + s2.y);
    // keep the primary and overflow products apart: overlapping their LDS
    // loads would exceed the register budget of the largest blocks
    __builtin_amdgcn_sched_barrier(0);
    return to_d2(s);
  };

  int cur = 0;
  const int nslots = nov > 0 ? 2 : 1;
  // every load of the set-up (some waited for only under lane-divergent
  // branches) completes here, in uniform control flow: no wait on global
  // memory is then left inside the loops, where it would also wait for the
  // stores of the saved states
  wait_global_loads();
  IntervalBlock blk;
  for (int s = 0; s < nsave; ++s) {
    int n = n_single;
    double h = 0.0;
    if (nsub != nullptr) {
      if ((s & 63) == 0) blk.load(nsub, hs, nsave, s);
      n = max(blk.steps(s & 63), 0);
      h = blk.length(s & 63);
    }
    for (int step = 0; step < n; ++step) {
      if (PROP) {
        const T2* zc = zb + cur * MP;
        T2* zn = zb + (cur ^ 1) * MP;
        double2 kq[2];
#pragma unroll
        for (int q = 0; q < 2; ++q)
          kq[q] = q < nslots ? matvec(zc, q) : make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const bool on = q == 0 ? has0 : has1;
          if (!(on && lead)) continue;
          const int row = q == 0 ? row0 : row1;
          if (MIXED) {
            y[q] = cadd(y[q], cadd(kq[q], cmul(dd[q], y[q])));   // y += (P - I) y
          } else {
            y[q] = kq[q];                                         // y = P y
          }
          zn[row] = from_d2<T>(y[q]);
        }
        cur ^= 1;
        __syncthreads();
      } else {
#pragma unroll 1
        for (int st = 0; st < 4; ++st) {
          const T2* zc = zb + cur * MP;
          T2* zn = zb + (cur ^ 1) * MP;
          double2 kq[2];
#pragma unroll
          for (int q = 0; q < 2; ++q)
            kq[q] = q < nslots ? matvec(zc, q) : make_double2(0.0, 0.0);
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const bool on = q == 0 ? has0 : has1;
            if (!(on && lead)) continue;
            const int row = q == 0 ? row0 : row1;
            if (st == 0) {
              acc[q] = make_double2(0.0, 0.0);
              zi[q] = y[q];
            }
            double2 kk = kq[q];
            if (MIXED) kk = cadd(kk, cmul(dd[q], zi[q]));
            acc[q] = caxpy(kRkW[st], kk, acc[q]);
            if (st < 3) {
              zi[q] = caxpy(kRkC[st] * h, kk, y[q]);
              zn[row] = from_d2<T>(zi[q]);
            } else {
              y[q] = caxpy(h / 6.0, acc[q], y[q]);
              zn[row] = from_d2<T>(y[q]);
            }
          }
          cur ^= 1;
          __syncthreads();
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const bool on = q == 0 ? has0 : has1;
      if (!(on && lead)) continue;
      const int row = q == 0 ? row0 : row1;
      if (ob != nullptr) ob[(int64_t)s * M + row] = y[q];
      if (!cfinite(y[q])) bad = true;
    }
  }
}

template <typename T, int NT, int S, int W, bool MIXED, bool PROP>
__global__ __launch_bounds__(NT) void step_kernel(
    const cplx_t<T>* __restrict__ A, int64_t strideA, const double2* __restrict__ diag,
    int64_t strideD, int M, const double2* __restrict__ y0, int R,
    const int* __restrict__ nsub, const double* __restrict__ hs, int nsave,
    double2* __restrict__ out, double2* __restrict__ yfin, int* __restrict__ status) {
  const int tid = threadIdx.x;
  const int r = blockIdx.x, b = blockIdx.y;
  constexpr int kRegRows = NT / S;
  const int nov = M > kRegRows ? M - kRegRows : 0;
  const int row0 = tid / S, row1 = kRegRows + tid / S;
  const bool has0 = row0 < M, has1 = (tid / S) < nov, lead = (tid % S) == 0;
  const double2* yb0 = y0 + ((int64_t)b * R + r) * M;
  double2 y[2];
  y[0] = has0 ? yb0[row0] : make_double2(0.0, 0.0);
  y[1] = has1 ? yb0[row1] : make_double2(0.0, 0.0);
  bool bad = false;
  step_phase<T, NT, S, W, MIXED, PROP>(A + (int64_t)b * strideA,
                                       MIXED ? diag + (int64_t)b * strideD : nullptr, M, y, nsub,
                                       hs, nsave, 0, out + ((int64_t)b * R + r) * nsave * M, bad);
  if (bad) status[0] = 1;
  if (yfin != nullptr) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const bool on = q == 0 ? has0 : has1;
      if (!(on && lead)) continue;
      const int row = q == 0 ? row0 : row1;
      yfin[((int64_t)b * R + r) * M + row] = y[q];
    }
  }
}

// Time-chunked trajectory in one launch (PROPAGATOR mode): work-group j of
// problem b first computes its chunk start (P^c)^j y0 with the FP64 power
// Pc = P^(q c) -- the same products, in the same order, as a chain of chunk
// starts -- then steps its chunk with A (P in FP64, P - I in MIXED) over the
// save intervals nsub: out [B][nc][nsave][M].
template <int NT, int S, int W, bool MIXED>
__global__ __launch_bounds__(NT) void chunk_kernel(
    const double2* __restrict__ Pc, int64_t strideP, const void* __restrict__ A,
    int64_t strideA, const double2* __restrict__ diag, int64_t strideD, int M,
    const double2* __restrict__ y0, const int* __restrict__ nsub, const double* __restrict__ hs,
    int nsave, double2* __restrict__ out, int* __restrict__ status) {
  using T = typename std::conditional<MIXED, float, double>::type;
  const int tid = threadIdx.x;
  const int j = blockIdx.x, nc = gridDim.x, b = blockIdx.y;
  constexpr int kRegRows = NT / S;
  const int nov = M > kRegRows ? M - kRegRows : 0;
  const int row0 = tid / S, row1 = kRegRows + tid / S;
  const bool has0 = row0 < M, has1 = (tid / S) < nov;
  const double2* yb0 = y0 + (int64_t)b * M;
  double2 y[2];
  y[0] = has0 ? yb0[row0] : make_double2(0.0, 0.0);
  y[1] = has1 ? yb0[row1] : make_double2(0.0, 0.0);
  bool bad = false;
  if (j > 0)
    step_phase<double, NT, S, W, false, true>(Pc + (int64_t)b * strideP, nullptr, M, y, nullptr,
                                              nullptr, 1, j, nullptr, bad);
  step_phase<T, NT, S, W, MIXED, true>(
      reinterpret_cast<const cplx_t<T>*>(A) + (int64_t)b * strideA,
      MIXED ? diag + (int64_t)b * strideD : nullptr, M, y, nsub, hs, nsave, 0,
      out + ((int64_t)b * nc + j) * nsave * M, bad);
  if (bad) status[0] = 1;
}

struct Shape {
  int NT, S, W;
};

// Row-segment shapes by block size; the last one keeps at most 32 overflow
// rows in LDS.  Chosen so that W complex registers stay <= 128 VGPRs (FP64)
// for the occupancy the block size implies.
inline bool pick_shape(int M, Shape& sh) {
  if (M <= 8) sh = {64, 8, 1};
  else if (M <= 16) sh = {64, 4, 4};
  else if (M <= 32) sh = {128, 4, 8};
  else if (M <= 64) sh = {256, 4, 16};
  else if (M <= 128) sh = {512, 4, 32};
  else if (M <= 160) sh = {1024, 8, 20};
  else return false;
  return true;
}

template <typename T, bool MIXED, bool PROP>
int launch_step(const Shape& sh, const void* A, int64_t strideA, const void* diag, int64_t strideD,
                int M, const void* y0, int R, int B, const int* nsub, const double* h, int nsave,
                void* out, void* yfin, int* status, hipStream_t stream) {
  using T2 = cplx_t<T>;
  const int rows_reg = sh.NT / sh.S;
  const int nov = M > rows_reg ? M - rows_reg : 0;
  const int mp = sh.S * sh.W;
  const size_t shmem = (size_t)(2 * mp + nov * mp) * sizeof(T2);
  QSX_CHECK_ARG(shmem <= 160 * 1024, "qsx_propagate: block M=%d needs %zu bytes of LDS", M, shmem);
  dim3 grid(R, B);
#define QSX_STEP(NT_, S_, W_)                                                                  \
  if (sh.NT == NT_ && sh.S == S_ && sh.W == W_) {                                              \
    auto kern = step_kernel<T, NT_, S_, W_, MIXED, PROP>;                                      \
    if (shmem > 64 * 1024)                                                                     \
      QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                     \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)); \
    hipLaunchKernelGGL(kern, grid, dim3(NT_), shmem, stream, (const T2*)A, strideA,            \
                       (const double2*)diag, strideD, M, (const double2*)y0, R, nsub, h, nsave, \
                       (double2*)out, (double2*)yfin, status);                                  \
    QSX_CHECK_HIP(hipGetLastError());                                                          \
    return QSX_OK;                                                                             \
  }
  QSX_STEP(64, 8, 1)
  QSX_STEP(64, 4, 4)
  QSX_STEP(128, 4, 8)
  QSX_STEP(256, 4, 16)
  QSX_STEP(512, 4, 32)
  QSX_STEP(1024, 8, 20)
#undef QSX_STEP
  set_error("qsx_propagate: no kernel for shape (%d, %d, %d)", sh.NT, sh.S, sh.W);
  return QSX_EINVAL;
}

template <bool MIXED>
int launch_chunks(const Shape& sh, const void* Pc, int64_t strideP, const void* A,
                  int64_t strideA, const void* diag, int64_t strideD, int M, const void* y0,
                  int nc, int B, const int* nsub, const double* h, int nsave, void* out,
                  int* status, hipStream_t stream) {
  const int rows_reg = sh.NT / sh.S;
  const int nov = M > rows_reg ? M - rows_reg : 0;
  const int mp = sh.S * sh.W;
  // the FP64 phase needs the larger LDS (complex128 stage vector and rows)
  const size_t shmem = (size_t)(2 * mp + nov * mp) * sizeof(double2);
  QSX_CHECK_ARG(shmem <= 160 * 1024, "qsx_propagate_chunks: block M=%d needs %zu bytes of LDS",
                M, shmem);
  dim3 grid(nc, B);
#define QSX_CHUNK(NT_, S_, W_)                                                                 \
  if (sh.NT == NT_ && sh.S == S_ && sh.W == W_) {                                              \
    auto kern = chunk_kernel<NT_, S_, W_, MIXED>;                                              \
    if (shmem > 64 * 1024)                                                                     \
      QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                     \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)); \
    hipLaunchKernelGGL(kern, grid, dim3(NT_), shmem, stream, (const double2*)Pc, strideP, A,    \
                       strideA, (const double2*)diag, strideD, M, (const double2*)y0, nsub, h,  \
                       nsave, (double2*)out, status);                                          \
    QSX_CHECK_HIP(hipGetLastError());                                                          \
    return QSX_OK;                                                                             \
  }
  QSX_CHUNK(64, 8, 1)
  QSX_CHUNK(64, 4, 4)
  QSX_CHUNK(128, 4, 8)
  QSX_CHUNK(256, 4, 16)
  QSX_CHUNK(512, 4, 32)
  QSX_CHUNK(1024, 8, 20)
#undef QSX_CHUNK
  set_error("qsx_propagate_chunks: no kernel for shape (%d, %d, %d)", sh.NT, sh.S, sh.W);
  return QSX_EINVAL;
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_propagate(int32_t mode, int32_t precision, const void* A, int64_t strideA,
                                 const void* diag, int64_t strideD, int32_t M, const void* y0,
                                 int32_t R, int32_t B, const int32_t* nsub, const double* h,
                                 int32_t nsave, void* out, void* y_final, int32_t* status,
                                 void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(mode == QSX_MODE_RK4 || mode == QSX_MODE_PROPAGATOR, "qsx_propagate: bad mode %d",
                mode);
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED,
                "qsx_propagate: unknown precision %d", precision);
  QSX_CHECK_ARG(A && y0 && nsub && h && out && status, "qsx_propagate: null pointer");
  QSX_CHECK_ARG(M >= 1 && R >= 1 && B >= 1 && nsave >= 1 && B <= 65535 && R <= (1 << 30),
                "qsx_propagate: bad sizes M=%d R=%d B=%d nsave=%d", M, R, B, nsave);
  QSX_CHECK_ARG(strideA == 0 || strideA >= (int64_t)M * M, "qsx_propagate: bad strideA");
  Shape sh;
  QSX_CHECK_ARG(pick_shape(M, sh), "qsx_propagate: M=%d exceeds the register stepper (<= 160)", M);
  hipStream_t s = as_stream(stream);
  const bool prop = mode == QSX_MODE_PROPAGATOR;
  if (precision == QSX_MIXED) {
    QSX_CHECK_ARG(diag != nullptr, "qsx_propagate: MIXED needs the FP64 diagonal");
    QSX_CHECK_ARG(strideD == 0 || strideD >= M, "qsx_propagate: bad strideD");
    return prop ? launch_step<float, true, true>(sh, A, strideA, diag, strideD, M, y0, R, B, nsub,
                                                 h, nsave, out, y_final, status, s)
                : launch_step<float, true, false>(sh, A, strideA, diag, strideD, M, y0, R, B, nsub,
                                                  h, nsave, out, y_final, status, s);
  }
  return prop ? launch_step<double, false, true>(sh, A, strideA, nullptr, 0, M, y0, R, B, nsub, h,
                                                 nsave, out, y_final, status, s)
              : launch_step<double, false, false>(sh, A, strideA, nullptr, 0, M, y0, R, B, nsub,
                                                  h, nsave, out, y_final, status, s);
}

extern "C" int32_t qsx_propagate_chunks(int32_t precision, const void* Pc, int64_t strideP,
                                        const void* A, int64_t strideA, const void* diag,
                                        int64_t strideD, int32_t M, const void* y0, int32_t nc,
                                        int32_t B, const int32_t* nsub, const double* h,
                                        int32_t nsave, void* out, int32_t* status,
                                        void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED,
                "qsx_propagate_chunks: unknown precision %d", precision);
  QSX_CHECK_ARG(Pc && A && y0 && nsub && h && out && status, "qsx_propagate_chunks: null pointer");
  QSX_CHECK_ARG(M >= 1 && nc >= 1 && B >= 1 && nsave >= 1 && B <= 65535 && nc <= (1 << 30),
                "qsx_propagate_chunks: bad sizes M=%d nc=%d B=%d nsave=%d", M, nc, B, nsave);
  QSX_CHECK_ARG((strideP == 0 || strideP >= (int64_t)M * M) &&
                    (strideA == 0 || strideA >= (int64_t)M * M),
                "qsx_propagate_chunks: bad operator stride");
  Shape sh;
  QSX_CHECK_ARG(pick_shape(M, sh), "qsx_propagate_chunks: M=%d exceeds the register stepper "
                "(<= 160)", M);
  hipStream_t s = as_stream(stream);
  if (precision == QSX_MIXED) {
    QSX_CHECK_ARG(diag != nullptr, "qsx_propagate_chunks: MIXED needs the FP64 diagonal");
    QSX_CHECK_ARG(strideD == 0 || strideD >= M, "qsx_propagate_chunks: bad strideD");
    return launch_chunks<true>(sh, Pc, strideP, A, strideA, diag, strideD, M, y0, nc, B, nsub, h,
                               nsave, out, status, s);
  }
  return launch_chunks<false>(sh, Pc, strideP, A, strideA, nullptr, 0, M, y0, nc, B, nsub, h,
                              nsave, out, status, s);
}
