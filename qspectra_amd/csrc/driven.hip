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
// Layout as the register stepper (one work-group per right-hand side; row i
// of L and of every C_i split into S-lane segments of W registers; wave
// shuffle reductions; double-buffered LDS stage vector); blocks up to M = 64
// and up to 4 field terms (kernels for at most 2 and at most 4).  Each stage
// forms the operator row on the fly, a_ij + sum_i f_i(t) c_ij, so one product
// per stage.
// MIXED: L off-diagonal and the C_i in complex64 (the C_i have no diagonal),
// the diagonal of L in complex128 on the FP64 stage value; state and update
// in FP64.

#include "qsx_common.h"

namespace qsx {
namespace {

constexpr int kMaxFieldTerms = 4;

template <typename T>
__device__ inline cplx_t<T> shfl_xor_ct(cplx_t<T> v, int m) {
  cplx_t<T> r;
  r.x = __shfl_xor(v.x, m);
  r.y = __shfl_xor(v.y, m);
  return r;
}

__constant__ double kDrW[4] = {1.0, 2.0, 2.0, 1.0};
__constant__ double kDrC[3] = {0.5, 0.5, 1.0};

template <typename T, int PM>
__device__ inline void load_fields(cplx_t<T> (&f)[3][PM], const double2* __restrict__ Fb,
                                   int64_t g, int P) {
#pragma unroll
  for (int ti = 0; ti < 3; ++ti)
#pragma unroll
    for (int i = 0; i < PM; ++i)
      f[ti][i] = i < P ? from_d2<T>(Fb[(g * 3 + ti) * P + i]) : cplx_t<T>{0, 0};
}

template <typename T, int NT, int S, int W, int PM, bool MIXED>
__global__ __launch_bounds__(NT) void driven_kernel(
    const cplx_t<T>* __restrict__ A, int64_t strideA, const double2* __restrict__ diag,
    int64_t strideD, const cplx_t<T>* __restrict__ Cop, int64_t strideC, int P,
    const double2* __restrict__ F, int64_t strideF, int64_t nF, int M,
    const double2* __restrict__ y0, int R,
    const int* __restrict__ nsub, const double* __restrict__ hs, int nsave,
    double2* __restrict__ out, int* __restrict__ status) {
  using T2 = cplx_t<T>;
  __shared__ T2 zb[2][64];
  const int tid = threadIdx.x;
  const int r = blockIdx.x, b = blockIdx.y;
  const int k = tid % S;
  const int row = tid / S;
  const bool on = row < M;
  const bool lead = on && k == 0;
  const T2* Ab = A + (int64_t)b * strideA;
  const T2* Cb = Cop + (int64_t)b * strideC;
  const double2* Fb = F + (int64_t)b * strideF;

  T2 a[W], c[PM][W];
#pragma unroll
  for (int q = 0; q < W; ++q) {
    const int j = k + S * q;
    const bool in = on && j < M;
    a[q] = in ? Ab[(int64_t)row * M + j] : T2{0, 0};
#pragma unroll
    for (int i = 0; i < PM; ++i)
      c[i][q] = (in && i < P) ? Cb[((int64_t)i * M + row) * M + j] : T2{0, 0};
  }
  const double2* yb0 = y0 + ((int64_t)b * R + r) * M;
  double2 y = lead ? yb0[row] : make_double2(0.0, 0.0);
  const double2 dd = (MIXED && on) ? diag[(int64_t)b * strideD + row] : make_double2(0.0, 0.0);
  double2 acc = make_double2(0.0, 0.0), zi = make_double2(0.0, 0.0);
  if (lead) zb[0][row] = from_d2<T>(y);
  __syncthreads();

  int cur = 0;
  int64_t g = 0;           // global step index (row of F)
  T2 fv[3][PM], fn[3][PM];
  if (nF > 0) load_fields<T, PM>(fv, Fb, 0, P);
  bool bad = false;
  double2* ob = out + ((int64_t)b * R + r) * nsave * M;
  for (int s = 0; s < nsave; ++s) {
    const int n = max(nsub[s], 0);
    const double h = hs[s];
    for (int step = 0; step < n; ++step, ++g) {
      if (g >= nF) {                       // fewer sampled steps than scheduled:
        if (tid == 0) status[0] = 2;       // uniform over the work-group
        return;
      }
      load_fields<T, PM>(fn, Fb, g + 1 < nF ? g + 1 : g, P);   // next step's fields
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int ti = st == 0 ? 0 : (st == 3 ? 2 : 1);
        const T2* zc = zb[cur];
        T2 sum = {0, 0};
#pragma unroll
        for (int q = 0; q < W; ++q) {
          const int j = k + S * q;
          if (j < M) {
            T2 e = a[q];
#pragma unroll
            for (int i = 0; i < PM; ++i) cfma<T>(e, fv[ti][i], c[i][q]);
            cfma<T>(sum, e, zc[j]);
          }
        }
#pragma unroll
        for (int m = S / 2; m > 0; m >>= 1) {
          const T2 o = shfl_xor_ct<T>(sum, m);
          sum.x += o.x;
          sum.y += o.y;
        }
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
        cur ^= 1;
        __syncthreads();
      }
#pragma unroll
      for (int ti = 0; ti < 3; ++ti)
#pragma unroll
        for (int i = 0; i < PM; ++i) fv[ti][i] = fn[ti][i];
    }
    if (lead) {
      ob[(int64_t)s * M + row] = y;
      if (!cfinite(y)) bad = true;
    }
  }
  if (bad) status[0] = 1;
}

struct DrShape {
  int NT, S, W;
};

inline bool pick_driven_shape(int M, DrShape& sh) {
  if (M <= 8) sh = {64, 8, 1};
  else if (M <= 16) sh = {64, 4, 4};
  else if (M <= 32) sh = {128, 4, 8};
  else if (M <= 64) sh = {256, 4, 16};
  else return false;
  return true;
}

template <typename T, bool MIXED>
int launch_driven(const DrShape& sh, const void* A, int64_t strideA, const void* diag,
                  int64_t strideD, const void* C, int64_t strideC, int P, const void* F,
                  int64_t strideF, int64_t nF, int M,
                  const void* y0, int R, int B, const int* nsub, const double* h, int nsave,
                  void* out, int* status, hipStream_t stream) {
  using T2 = cplx_t<T>;
  dim3 grid(R, B);
#define QSX_DRV(NT_, S_, W_, PM_)                                                              \
  if (sh.NT == NT_ && sh.S == S_ && sh.W == W_ && P <= PM_) {                                  \
    hipLaunchKernelGGL((driven_kernel<T, NT_, S_, W_, PM_, MIXED>), grid, dim3(NT_), 0, stream, \
                       (const T2*)A, strideA, (const double2*)diag, strideD, (const T2*)C,     \
                       strideC, P, (const double2*)F, strideF, nF, M, (const double2*)y0, R,   \
                       nsub, h, nsave, (double2*)out, status);                                \
    QSX_CHECK_HIP(hipGetLastError());                                                          \
    return QSX_OK;                                                                             \
  }
  QSX_DRV(64, 8, 1, 2)
  QSX_DRV(64, 4, 4, 2)
  QSX_DRV(128, 4, 8, 2)
  QSX_DRV(256, 4, 16, 2)
  QSX_DRV(64, 8, 1, 4)
  QSX_DRV(64, 4, 4, 4)
  QSX_DRV(128, 4, 8, 4)
  QSX_DRV(256, 4, 16, 4)
#undef QSX_DRV
  set_error("qsx_propagate_driven: no kernel for shape (%d, %d, %d)", sh.NT, sh.S, sh.W);
  return QSX_EINVAL;
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_propagate_driven(int32_t precision, const void* A, int64_t strideA,
                                        const void* diag, int64_t strideD, const void* C,
                                        int64_t strideC, int32_t P, const void* F,
                                        int64_t strideF, int64_t nF, int32_t M,
                                        const void* y0, int32_t R, int32_t B,
                                        const int32_t* nsub, const double* h, int32_t nsave,
                                        void* out, int32_t* status, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED,
                "qsx_propagate_driven: unknown precision %d", precision);
  QSX_CHECK_ARG(A && C && F && y0 && nsub && h && out && status,
                "qsx_propagate_driven: null pointer");
  QSX_CHECK_ARG(P >= 1 && P <= kMaxFieldTerms, "qsx_propagate_driven: P=%d field terms (1..%d)",
                P, kMaxFieldTerms);
  QSX_CHECK_ARG(M >= 1 && R >= 1 && B >= 1 && nsave >= 1 && B <= 65535 && R <= (1 << 30),
                "qsx_propagate_driven: bad sizes M=%d R=%d B=%d nsave=%d", M, R, B, nsave);
  QSX_CHECK_ARG(strideA == 0 || strideA >= (int64_t)M * M, "qsx_propagate_driven: bad strideA");
  QSX_CHECK_ARG(strideC == 0 || strideC >= (int64_t)P * M * M,
                "qsx_propagate_driven: bad strideC");
  QSX_CHECK_ARG(nF >= 0 && (strideF == 0 || strideF >= nF * 3 * P),
                "qsx_propagate_driven: bad field record (nF=%lld, strideF=%lld)", (long long)nF,
                (long long)strideF);
  DrShape sh;
  QSX_CHECK_ARG(pick_driven_shape(M, sh), "qsx_propagate_driven: M=%d exceeds 64", M);
  hipStream_t s = as_stream(stream);
  if (precision == QSX_MIXED) {
    QSX_CHECK_ARG(diag != nullptr, "qsx_propagate_driven: MIXED needs the FP64 diagonal");
    QSX_CHECK_ARG(strideD == 0 || strideD >= M, "qsx_propagate_driven: bad strideD");
    return launch_driven<float, true>(sh, A, strideA, diag, strideD, C, strideC, P, F, strideF,
                                      nF, M, y0, R, B, nsub, h, nsave, out, status, s);
  }
  return launch_driven<double, false>(sh, A, strideA, nullptr, 0, C, strideC, P, F, strideF, nF,
                                       M, y0, R, B, nsub, h, nsave, out, status, s);
}
