// Batched complex GEMM with an identity term (gfx950):
//   C[b] = alpha * A[b] * B[b] + beta * C[b] + gamma * I
// Row-major operands; 32 x 32 output tiles per 256-thread workgroup, K staged
// through LDS in chunks of 16, each thread accumulating a 2 x 2 block.
// FP64: complex128 throughout.  MIXED: complex64 A and B with FP32
// accumulation, complex128 C.  Used for the Horner construction of the RK4
// propagator, its powers, save-point dipole maps and response contractions:
// operands here are at most a few hundred wide, so the kernel favours many
// small tiles over deep register blocking.

#include "qsx_common.h"

namespace qsx {
namespace {

constexpr int kTM = 32, kTN = 32, kTK = 16;

template <typename T>
__global__ __launch_bounds__(256) void zgemm_kernel(int m, int n, int k, double2 alpha,
                                                    const cplx_t<T>* __restrict__ A, int64_t lda,
                                                    int64_t sA, const cplx_t<T>* __restrict__ B,
                                                    int64_t ldb, int64_t sB, double2 beta,
                                                    double2* __restrict__ C, int64_t ldc,
                                                    int64_t sC, double2 gamma) {
  using T2 = cplx_t<T>;
  __shared__ T2 As[kTK][kTM + 1];
  __shared__ T2 Bs[kTK][kTN + 1];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int bm = blockIdx.x * kTM, bn = blockIdx.y * kTN, bz = blockIdx.z;
  const T2* Ab = A + (int64_t)bz * sA;
  const T2* Bb = B + (int64_t)bz * sB;
  double2* Cb = C + (int64_t)bz * sC;
  T2 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = T2{0, 0};
  for (int k0 = 0; k0 < k; k0 += kTK) {
    for (int e = tid; e < kTM * kTK; e += 256) {
      const int mm = e / kTK, kk = e % kTK;     // consecutive threads along k (row-major A)
      const int gi = bm + mm, gk = k0 + kk;
      As[kk][mm] = (gi < m && gk < k) ? Ab[(int64_t)gi * lda + gk] : T2{0, 0};
    }
    for (int e = tid; e < kTK * kTN; e += 256) {
      const int kk = e / kTN, nn = e % kTN;     // consecutive threads along n (row-major B)
      const int gk = k0 + kk, gj = bn + nn;
      Bs[kk][nn] = (gk < k && gj < n) ? Bb[(int64_t)gk * ldb + gj] : T2{0, 0};
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kTK; ++kk) {
      const T2 a0 = As[kk][2 * ty], a1 = As[kk][2 * ty + 1];
      const T2 b0 = Bs[kk][2 * tx], b1 = Bs[kk][2 * tx + 1];
      cfma<T>(acc[0][0], a0, b0);
      cfma<T>(acc[0][1], a0, b1);
      cfma<T>(acc[1][0], a1, b0);
      cfma<T>(acc[1][1], a1, b1);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gi = bm + 2 * ty + i, gj = bn + 2 * tx + j;
      if (gi >= m || gj >= n) continue;
      double2 v = cmul(alpha, to_d2(acc[i][j]));
      double2* cp = Cb + (int64_t)gi * ldc + gj;
      if (beta.x != 0.0 || beta.y != 0.0) v = cadd(v, cmul(beta, *cp));
      if (gi == gj) v = cadd(v, gamma);
      *cp = v;
    }
  }
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_zgemm(int32_t precision, int32_t m, int32_t n, int32_t k, double alpha_re,
                             double alpha_im, const void* A, int64_t lda, int64_t sA,
                             const void* B, int64_t ldb, int64_t sB, double beta_re,
                             double beta_im, void* C, int64_t ldc, int64_t sC, double gamma_re,
                             double gamma_im, int32_t batch, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED, "qsx_zgemm: bad precision");
  QSX_CHECK_ARG(A && B && C, "qsx_zgemm: null pointer");
  QSX_CHECK_ARG(m >= 1 && n >= 1 && k >= 1 && batch >= 1 && batch <= 65535,
                "qsx_zgemm: bad sizes m=%d n=%d k=%d batch=%d", m, n, k, batch);
  QSX_CHECK_ARG(lda >= k && ldb >= n && ldc >= n, "qsx_zgemm: bad leading dimensions");
  QSX_CHECK_ARG(batch == 1 || sC != 0, "qsx_zgemm: a batched C needs a batch stride");
  QSX_CHECK_ARG((gamma_re == 0.0 && gamma_im == 0.0) || m == n,
                "qsx_zgemm: the identity term needs a square C");
  dim3 grid((m + kTM - 1) / kTM, (n + kTN - 1) / kTN, batch);
  const double2 al = make_double2(alpha_re, alpha_im), be = make_double2(beta_re, beta_im),
                ga = make_double2(gamma_re, gamma_im);
  if (precision == QSX_FP64)
    hipLaunchKernelGGL(zgemm_kernel<double>, grid, dim3(256), 0, as_stream(stream), m, n, k, al,
                       (const double2*)A, lda, sA, (const double2*)B, ldb, sB, be, (double2*)C,
                       ldc, sC, ga);
  else
    hipLaunchKernelGGL(zgemm_kernel<float>, grid, dim3(256), 0, as_stream(stream), m, n, k, al,
                       (const float2*)A, lda, sA, (const float2*)B, ldb, sB, be, (double2*)C, ldc,
                       sC, ga);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}
