// Classical RK4 for a Liouvillian driven by classical fields (gfx950):
//     y' = L y + sum_i f_i(t) C_i y,
// C_i the commutator superoperator of the dipole operator of field term i and
// f_i(t) = -i E_i(t) (E_i conjugated for a '+' interaction), the equation of
// motion of simulate_with_fields (qspectra/simulate/eom.py:79-103).  The
// fields are sampled by the host at the three RK4 stage times of every step
// (t, t + h/2, t + h) and passed as F[b][g][3][P] for global step g < nF
// (batch stride strideF, 0 when all batch entries share one field record); a
// schedule with more steps than F holds stops with status 2.  The field
// values of step g + 1 are loaded while step g runs.
//
// The C_i are sparse (a dipole commutator links a Liouville element only to
// elements one excitation away: at most 2n of the M = (n+1)^2 columns of a
// 'gg,ge,eg,ee' row), so they come in ELL form over the union of their
// patterns (shared by the batch): K slots per row, column index col[row][k]
// and the P coefficients coef[b][row][k][i] (padding slots carry zero
// coefficients).
//
// Layout as the register stepper (one work-group per right-hand side; row i
// of L split into S-lane segments of W registers and the row's ELL slots into
// S-lane segments of KW registers; wave shuffle reductions; double-buffered
// LDS stage vector); blocks up to M = 64 and up to 4 field terms (kernels for
// at most 2 and at most 4); larger blocks (up to 1536) in driven_large_kernel.
// Each stage computes
//     sum_j a_ij z_j + sum_k (sum_i f_i(t) coef_ik) z_col(k)
// with two accumulators to halve the dependent FMA chain.
// MIXED: L off-diagonal and the ELL coefficients in complex64 (the C_i have
// no diagonal), the diagonal of L in complex128 on the FP64 stage value;
// state and update in FP64.

#include "qsx_common.h"

namespace qsx {
namespace {

constexpr int kMaxFieldTerms = 4;

__constant__ double kDrW[4] = {1.0, 2.0, 2.0, 1.0};
__constant__ double kDrC[3] = {0.5, 0.5, 1.0};

template <typename T, int NT, int S, int W, int KW, int PM, bool MIXED>
__global__ __launch_bounds__(NT) void driven_kernel(
    const cplx_t<T>* __restrict__ A, int64_t strideA, const double2* __restrict__ diag,
    int64_t strideD, int K, const int* __restrict__ Ecol, const cplx_t<T>* __restrict__ Ecoef,
    int64_t strideE, int P, const double2* __restrict__ F, int64_t strideF, int64_t nF, int M,
    const double2* __restrict__ y0, int R,
    const int* __restrict__ nsub, const double* __restrict__ hs, int nsave,
    double2* __restrict__ out, int* __restrict__ status) {
  using T2 = cplx_t<T>;
  // stage vector, zero beyond row M so that the S * W <= 64 lanes of a row
  // segment read without bounds tests; field values of the current and the
  // next step
  __shared__ T2 zb[2][64];
  __shared__ T2 fb[2][3 * PM];
  const int tid = threadIdx.x;
  const int r = blockIdx.x, b = blockIdx.y;
  const int k = tid % S;
  const int row = tid / S;
  const bool on = row < M;
  const bool lead = on && k == 0;
  const bool loader = tid < 3 * P;          // thread tid stages F value tid of a step
  const T2* Ab = A + (int64_t)b * strideA;
  const int* Cb = Ecol;                                   // one pattern for the batch
  const T2* Eb = Ecoef + (int64_t)b * strideE * P;
  const double2* Fb = F + (int64_t)b * strideF;

  T2 a[W], c[KW][PM];
  int col[KW];
#pragma unroll
  for (int q = 0; q < W; ++q) {
    const int j = k + S * q;
    a[q] = (on && j < M) ? Ab[(int64_t)row * M + j] : T2{0, 0};
  }
#pragma unroll
  for (int q = 0; q < KW; ++q) {
    const int kk = k + S * q;
    const bool in = on && kk < K;
    const int64_t slot = (int64_t)row * K + kk;
    col[q] = in ? Cb[slot] : 0;
#pragma unroll
    for (int i = 0; i < PM; ++i) c[q][i] = (in && i < P) ? Eb[slot * P + i] : T2{0, 0};
  }
  const double2* yb0 = y0 + ((int64_t)b * R + r) * M;
  double2 y = lead ? yb0[row] : make_double2(0.0, 0.0);
  const double2 dd = (MIXED && on) ? diag[(int64_t)b * strideD + row] : make_double2(0.0, 0.0);
  double2 acc = make_double2(0.0, 0.0), zi = make_double2(0.0, 0.0);
  for (int i = tid; i < 2 * 64; i += NT) zb[i / 64][i % 64] = T2{0, 0};
  if (loader && nF > 0) fb[0][tid] = from_d2<T>(Fb[tid]);
  __syncthreads();
  if (lead) zb[0][row] = from_d2<T>(y);
  __syncthreads();

  int cur = 0;
  int64_t g = 0;           // global step index (row of F)
  bool bad = false;
  double2* ob = out + ((int64_t)b * R + r) * nsave * M;
  const int vz = opaque_zero();            // schedule of the next interval: vector loads
  int n_next = nsub[vz];
  double h_next = hs[vz];
  for (int s = 0; s < nsave; ++s) {
    const int n = max(wave_uniform(n_next), 0);
    const double h = wave_uniform(h_next);
    if (s + 1 < nsave) {
      n_next = nsub[s + 1 + vz];
      h_next = hs[s + 1 + vz];
    }
    for (int step = 0; step < n; ++step, ++g) {
      if (g >= nF) {                       // fewer sampled steps than scheduled:
        if (tid == 0) status[0] = 2;       // uniform over the work-group
        return;
      }
      // next step's field values: a vector load whose wait falls at the
      // LDS store after the last stage
      double2 fnext = make_double2(0.0, 0.0);
      if (loader) fnext = Fb[(g + 1 < nF ? g + 1 : g) * 3 * P + tid];
      const T2* fc = fb[g & 1];
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int ti = st == 0 ? 0 : (st == 3 ? 2 : 1);
        const T2* zc = zb[cur];
        T2 sum = {0, 0}, sum2 = {0, 0};
#pragma unroll
        for (int q = 0; q < W; ++q) cfma<T>((q & 1) ? sum2 : sum, a[q], zc[k + S * q]);
        T2 f[PM];
#pragma unroll
        for (int i = 0; i < PM; ++i) f[i] = i < P ? fc[ti * P + i] : T2{0, 0};
#pragma unroll
        for (int q = 0; q < KW; ++q) {
          T2 e = {0, 0};
#pragma unroll
          for (int i = 0; i < PM; ++i) cfma<T>(e, f[i], c[q][i]);
          cfma<T>((q & 1) ? sum : sum2, e, zc[col[q]]);
        }
        sum.x = seg_sum<S>(sum.x + sum2.x);
        sum.y = seg_sum<S>(sum.y + sum2.y);
        if (lead) {
          if (st == 0) {
            acc = make_double2(0.0, 0.0);
            zi = y;
          }
          double2 kk = to_d2(sum);
          if (MIXED) kk = cadd(kk, cmul(dd, zi));
          acc = caxpy(kDrW[st], kk, acc);
          if (st < 3) {
            zi = caxpy(kDrC[st] * h, kk, y);
            zb[cur ^ 1][row] = from_d2<T>(zi);
          } else {
            y = caxpy(h / 6.0, acc, y);
            zb[cur ^ 1][row] = from_d2<T>(y);
          }
        }
        if (st == 3 && loader) fb[(g + 1) & 1][tid] = from_d2<T>(fnext);
        cur ^= 1;
        __syncthreads();
      }
    }
    if (lead) {
      ob[(int64_t)s * M + row] = y;
      if (!cfinite(y)) bad = true;
    }
  }
  if (bad) status[0] = 1;
}

// Blocks above 64 elements (up to kDrivenLargeMax): one work-group of four
// wavefronts per right-hand side; the stage vector, the state and the RK4
// accumulator in LDS; each wavefront takes every fourth row, its lanes
// stride over the row of L (read from memory, coalesced) and over the row's
// ELL slots, and a wavefront reduction gives the row's derivative.
constexpr int kDrivenLargeMax = 1536;   // 5 M complex128 of LDS
constexpr int kDrLargeThreads = 256;

template <typename T, bool MIXED>
__global__ __launch_bounds__(kDrLargeThreads) void driven_large_kernel(
    const cplx_t<T>* __restrict__ A, int64_t strideA, const double2* __restrict__ diag,
    int64_t strideD, int K, const int* __restrict__ Ecol, const cplx_t<T>* __restrict__ Ecoef,
    int64_t strideE, int P, const double2* __restrict__ F, int64_t strideF, int64_t nF, int M,
    const double2* __restrict__ y0, int R, const int* __restrict__ nsub,
    const double* __restrict__ hs, int nsave, double2* __restrict__ out,
    int* __restrict__ status) {
  using T2 = cplx_t<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  double2* ys = reinterpret_cast<double2*>(smem);          // [M] state
  double2* acc = ys + M;                                     // [M] RK4 accumulator
  double2* zi = acc + M;                                     // [M] FP64 stage value
  T2* zb = reinterpret_cast<T2*>(zi + M);                    // [2][M] stage vector
  __shared__ T2 fc[3 * kMaxFieldTerms];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int r = blockIdx.x, b = blockIdx.y;
  const T2* Ab = A + (int64_t)b * strideA;
  const T2* Eb = Ecoef + (int64_t)b * strideE * P;
  const double2* Fb = F + (int64_t)b * strideF;
  const double2* yb0 = y0 + ((int64_t)b * R + r) * M;
  for (int i = tid; i < M; i += kDrLargeThreads) {
    ys[i] = yb0[i];
    zb[i] = from_d2<T>(yb0[i]);
  }
  __syncthreads();
  int cur = 0;
  int64_t g = 0;
  bool bad = false;
  double2* ob = out + ((int64_t)b * R + r) * nsave * M;
  for (int s = 0; s < nsave; ++s) {
    const int n = max(nsub[s], 0);
    const double h = hs[s];
    for (int step = 0; step < n; ++step, ++g) {
      if (g >= nF) {                       // fewer sampled steps than scheduled
        if (tid == 0) status[0] = 2;
        return;
      }
      if (tid < 3 * P) fc[tid] = from_d2<T>(Fb[g * 3 * P + tid]);
      __syncthreads();
      for (int st = 0; st < 4; ++st) {
        const int ti = st == 0 ? 0 : (st == 3 ? 2 : 1);
        const T2* zc = zb + cur * M;
        T2* zn = zb + (cur ^ 1) * M;
        for (int row = wv; row < M; row += kDrLargeThreads / 64) {
          const T2* Ar = Ab + (int64_t)row * M;
          T2 sum = {0, 0}, sum2 = {0, 0};
          for (int j = lane; j < M; j += 64) cfma<T>((j & 64) ? sum2 : sum, Ar[j], zc[j]);
          for (int kk = lane; kk < K; kk += 64) {
            const int64_t slot = (int64_t)row * K + kk;
            T2 e = {0, 0};
            for (int i = 0; i < P; ++i) cfma<T>(e, fc[ti * P + i], Eb[slot * P + i]);
            cfma<T>(sum2, e, zc[Ecol[slot]]);
          }
          double sr = (double)sum.x + (double)sum2.x, si = (double)sum.y + (double)sum2.y;
          for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off);
            si += __shfl_xor(si, off);
          }
          if (lane == 0) {
            const double2 yv = ys[row];
            const double2 zv = st == 0 ? yv : zi[row];
            double2 kv = make_double2(sr, si);
            if (MIXED) kv = cadd(kv, cmul(diag[(int64_t)b * strideD + row], zv));
            const double2 a0 = st == 0 ? make_double2(0.0, 0.0) : acc[row];
            const double2 an = caxpy(kDrW[st], kv, a0);
            if (st < 3) {
              acc[row] = an;
              const double2 zz = caxpy(kDrC[st] * h, kv, yv);
              zi[row] = zz;
              zn[row] = from_d2<T>(zz);
            } else {
              const double2 yn = caxpy(h / 6.0, an, yv);
              ys[row] = yn;
              zn[row] = from_d2<T>(yn);
            }
          }
        }
        cur ^= 1;
        __syncthreads();
      }
    }
    for (int i = tid; i < M; i += kDrLargeThreads) {
      const double2 v = ys[i];
      ob[(int64_t)s * M + i] = v;
      if (!cfinite(v)) bad = true;
    }
  }
  if (bad) status[0] = 1;
}

inline size_t driven_large_lds(int M, size_t tbytes) {
  return 3 * (size_t)M * sizeof(double2) + 2 * (size_t)M * tbytes;
}

template <typename T, bool MIXED>
int launch_driven_large(const void* A, int64_t strideA, const void* diag, int64_t strideD, int K,
                        const int* Ecol, const void* Ecoef, int64_t strideE, int P,
                        const void* F, int64_t strideF, int64_t nF, int M, const void* y0,
                        int R, int B, const int* nsub, const double* h, int nsave, void* out,
                        int* status, hipStream_t stream) {
  using T2 = cplx_t<T>;
  const size_t shmem = driven_large_lds(M, sizeof(T2));
  auto kern = driven_large_kernel<T, MIXED>;
  if (shmem > 64 * 1024)
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  hipLaunchKernelGGL(kern, dim3(R, B), dim3(kDrLargeThreads), shmem, stream, (const T2*)A,
                     strideA, (const double2*)diag, strideD, K, Ecol, (const T2*)Ecoef, strideE,
                     P, (const double2*)F, strideF, nF, M, (const double2*)y0, R, nsub, h, nsave,
                     (double2*)out, status);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

struct DrShape {
  int NT, S, W, KW;
};

// S lanes per row, W registers of the L row, KW ELL slots per lane (4 when
// the K slots fit, else the whole row width W)
inline bool pick_driven_shape(int M, int K, DrShape& sh) {
  if (M <= 8) sh = {64, 8, 1, 1};
  else if (M <= 16) sh = {64, 4, 4, 4};
  else if (M <= 32) sh = {128, 4, 8, K <= 16 ? 4 : 8};
  else if (M <= 64) sh = {256, 4, 16, K <= 16 ? 4 : 16};
  else return false;
  return K <= sh.S * sh.KW;
}

template <typename T, bool MIXED>
int launch_driven(const DrShape& sh, const void* A, int64_t strideA, const void* diag,
                  int64_t strideD, int K, const int* Ecol, const void* Ecoef, int64_t strideE,
                  int P, const void* F, int64_t strideF, int64_t nF, int M,
                  const void* y0, int R, int B, const int* nsub, const double* h, int nsave,
                  void* out, int* status, hipStream_t stream) {
  using T2 = cplx_t<T>;
  dim3 grid(R, B);
#define QSX_DRV(NT_, S_, W_, KW_, PM_)                                                         \
  if (sh.NT == NT_ && sh.S == S_ && sh.W == W_ && sh.KW == KW_ && P <= PM_) {                  \
    hipLaunchKernelGGL((driven_kernel<T, NT_, S_, W_, KW_, PM_, MIXED>), grid, dim3(NT_), 0,    \
                       stream, (const T2*)A, strideA, (const double2*)diag, strideD, K, Ecol,  \
                       (const T2*)Ecoef, strideE, P, (const double2*)F, strideF, nF, M,        \
                       (const double2*)y0, R, nsub, h, nsave, (double2*)out, status);         \
    QSX_CHECK_HIP(hipGetLastError());                                                          \
    return QSX_OK;                                                                             \
  }
#define QSX_DRV_P(NT_, S_, W_, KW_) QSX_DRV(NT_, S_, W_, KW_, 2) QSX_DRV(NT_, S_, W_, KW_, 4)
  QSX_DRV_P(64, 8, 1, 1)
  QSX_DRV_P(64, 4, 4, 4)
  QSX_DRV_P(128, 4, 8, 4)
  QSX_DRV_P(128, 4, 8, 8)
  QSX_DRV_P(256, 4, 16, 4)
  QSX_DRV_P(256, 4, 16, 16)
#undef QSX_DRV_P
#undef QSX_DRV
  set_error("qsx_propagate_driven: no kernel for shape (%d, %d, %d, %d)", sh.NT, sh.S, sh.W,
            sh.KW);
  return QSX_EINVAL;
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_propagate_driven(int32_t precision, const void* A, int64_t strideA,
                                        const void* diag, int64_t strideD, int32_t K,
                                        const int32_t* col, const void* coef, int64_t strideE,
                                        int32_t P, const void* F, int64_t strideF, int64_t nF,
                                        int32_t M,
                                        const void* y0, int32_t R, int32_t B,
                                        const int32_t* nsub, const double* h, int32_t nsave,
                                        void* out, int32_t* status, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED,
                "qsx_propagate_driven: unknown precision %d", precision);
  QSX_CHECK_ARG(A && col && coef && F && y0 && nsub && h && out && status,
                "qsx_propagate_driven: null pointer");
  QSX_CHECK_ARG(P >= 1 && P <= kMaxFieldTerms, "qsx_propagate_driven: P=%d field terms (1..%d)",
                P, kMaxFieldTerms);
  QSX_CHECK_ARG(M >= 1 && R >= 1 && B >= 1 && nsave >= 1 && B <= 65535 && R <= (1 << 30),
                "qsx_propagate_driven: bad sizes M=%d R=%d B=%d nsave=%d", M, R, B, nsave);
  QSX_CHECK_ARG(strideA == 0 || strideA >= (int64_t)M * M, "qsx_propagate_driven: bad strideA");
  QSX_CHECK_ARG(K >= 1 && (strideE == 0 || strideE >= (int64_t)M * K),
                "qsx_propagate_driven: bad ELL shape (K=%d, strideE=%lld)", K, (long long)strideE);
  QSX_CHECK_ARG(nF >= 0 && (strideF == 0 || strideF >= nF * 3 * P),
                "qsx_propagate_driven: bad field record (nF=%lld, strideF=%lld)", (long long)nF,
                (long long)strideF);
  DrShape sh;
  hipStream_t s = as_stream(stream);
  if (!pick_driven_shape(M, K, sh)) {
    // beyond the register-resident shapes: the LDS-resident kernel
    QSX_CHECK_ARG(M <= kDrivenLargeMax && K <= M,
                  "qsx_propagate_driven: M=%d exceeds %d or K=%d ELL slots exceed the row", M,
                  kDrivenLargeMax, K);
    if (precision == QSX_MIXED) {
      QSX_CHECK_ARG(diag != nullptr, "qsx_propagate_driven: MIXED needs the FP64 diagonal");
      QSX_CHECK_ARG(strideD == 0 || strideD >= M, "qsx_propagate_driven: bad strideD");
      return launch_driven_large<float, true>(A, strideA, diag, strideD, K, col, coef, strideE,
                                              P, F, strideF, nF, M, y0, R, B, nsub, h, nsave,
                                              out, status, s);
    }
    return launch_driven_large<double, false>(A, strideA, nullptr, 0, K, col, coef, strideE, P,
                                              F, strideF, nF, M, y0, R, B, nsub, h, nsave, out,
                                              status, s);
  }
  if (precision == QSX_MIXED) {
    QSX_CHECK_ARG(diag != nullptr, "qsx_propagate_driven: MIXED needs the FP64 diagonal");
    QSX_CHECK_ARG(strideD == 0 || strideD >= M, "qsx_propagate_driven: bad strideD");
    return launch_driven<float, true>(sh, A, strideA, diag, strideD, K, col, coef, strideE, P, F,
                                      strideF, nF, M, y0, R, B, nsub, h, nsave, out, status, s);
  }
  return launch_driven<double, false>(sh, A, strideA, nullptr, 0, K, col, coef, strideE, P, F,
                                       strideF, nF, M, y0, R, B, nsub, h, nsave, out, status, s);
}
