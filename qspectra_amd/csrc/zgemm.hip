// Batched complex GEMM with an identity term (gfx950):
//   C[b] = alpha * A[b] * B[b] + beta * C[b] + gamma * I
// Row-major storage, op(A) and op(B) optionally transposed; 32 x 32 output
// tiles per 256-thread workgroup (16 x 16 for small problems), K staged
// through LDS in chunks of 16, each thread accumulating a 2 x 2 (1 x 1) block.
// FP64: complex128 throughout.  MIXED: complex64 A and B with FP32
// accumulation, complex128 C.  Used for the Horner construction of the RK4
// propagator, its powers, save-point dipole maps and response contractions:
// operands here are at most a few hundred wide, so the kernel favours many
// small tiles over deep register blocking.

#include "qsx_common.h"

namespace qsx {
namespace {

constexpr int kTK = 16;

// TM x TN output tile per 256-thread workgroup, each thread an (TM/16) x
// (TN/16) block.  op(A) = A or A^T (TA), op(B) = B or B^T (TB); the LDS
// staging loads walk the contiguous index of the stored operand.
template <typename T, int TM, int TN, bool TA, bool TB>
__global__ __launch_bounds__(256) void zgemm_kernel(int m, int n, int k, double2 alpha,
                                                    const cplx_t<T>* __restrict__ A, int64_t lda,
                                                    int64_t sA, const cplx_t<T>* __restrict__ B,
                                                    int64_t ldb, int64_t sB, double2 beta,
                                                    double2* __restrict__ C, int64_t ldc,
                                                    int64_t sC, double2 gamma) {
  using T2 = cplx_t<T>;
  constexpr int RM = TM / 16, RN = TN / 16;
  __shared__ T2 As[kTK][TM + 1];
  __shared__ T2 Bs[kTK][TN + 1];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int bm = blockIdx.x * TM, bn = blockIdx.y * TN, bz = blockIdx.z;
  const T2* Ab = A + (int64_t)bz * sA;
  const T2* Bb = B + (int64_t)bz * sB;
  double2* Cb = C + (int64_t)bz * sC;
  T2 acc[RM][RN];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j) acc[i][j] = T2{0, 0};
  for (int k0 = 0; k0 < k; k0 += kTK) {
    for (int e = tid; e < TM * kTK; e += 256) {
      int mm, kk;
      if (TA) { kk = e / TM; mm = e % TM; }       // A stored k-major: walk m
      else { mm = e / kTK; kk = e % kTK; }        // A stored m-major: walk k
      const int gi = bm + mm, gk = k0 + kk;
      const int64_t off = TA ? (int64_t)gk * lda + gi : (int64_t)gi * lda + gk;
      As[kk][mm] = (gi < m && gk < k) ? Ab[off] : T2{0, 0};
    }
    for (int e = tid; e < kTK * TN; e += 256) {
      int kk, nn;
      if (TB) { nn = e / kTK; kk = e % kTK; }     // B stored n-major: walk k
      else { kk = e / TN; nn = e % TN; }          // B stored k-major: walk n
      const int gk = k0 + kk, gj = bn + nn;
      const int64_t off = TB ? (int64_t)gj * ldb + gk : (int64_t)gk * ldb + gj;
      Bs[kk][nn] = (gk < k && gj < n) ? Bb[off] : T2{0, 0};
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kTK; ++kk) {
      T2 a[RM], b[RN];
#pragma unroll
      for (int i = 0; i < RM; ++i) a[i] = As[kk][RM * ty + i];
#pragma unroll
      for (int j = 0; j < RN; ++j) b[j] = Bs[kk][RN * tx + j];
#pragma unroll
      for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j) cfma<T>(acc[i][j], a[i], b[j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < RM; ++i) {
#pragma unroll
    for (int j = 0; j < RN; ++j) {
      const int gi = bm + RM * ty + i, gj = bn + RN * tx + j;
      if (gi >= m || gj >= n) continue;
      double2 v = cmul(alpha, to_d2(acc[i][j]));
      double2* cp = Cb + (int64_t)gi * ldc + gj;
      if (beta.x != 0.0 || beta.y != 0.0) v = cadd(v, cmul(beta, *cp));
      if (gi == gj) v = cadd(v, gamma);
      *cp = v;
    }
  }
}

template <typename T, int TM, int TN>
void launch_zgemm(int trans, dim3 grid, hipStream_t s, int m, int n, int k, double2 al,
                  const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                  double2 be, void* C, int64_t ldc, int64_t sC, double2 ga) {
  using T2 = cplx_t<T>;
#define QSX_ZG(TA_, TB_)                                                                      \
  hipLaunchKernelGGL((zgemm_kernel<T, TM, TN, TA_, TB_>), grid, dim3(256), 0, s, m, n, k, al, \
                     (const T2*)A, lda, sA, (const T2*)B, ldb, sB, be, (double2*)C, ldc, sC, ga)
  switch (trans & 3) {
    case 0: QSX_ZG(false, false); break;
    case 1: QSX_ZG(true, false); break;
    case 2: QSX_ZG(false, true); break;
    default: QSX_ZG(true, true); break;
  }
#undef QSX_ZG
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_zgemm(int32_t precision, int32_t trans, int32_t m, int32_t n, int32_t k,
                             double alpha_re, double alpha_im, const void* A, int64_t lda,
                             int64_t sA, const void* B, int64_t ldb, int64_t sB, double beta_re,
                             double beta_im, void* C, int64_t ldc, int64_t sC, double gamma_re,
                             double gamma_im, int32_t batch, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED, "qsx_zgemm: bad precision");
  QSX_CHECK_ARG(trans >= 0 && trans <= 3, "qsx_zgemm: bad transpose flags %d", trans);
  QSX_CHECK_ARG(A && B && C, "qsx_zgemm: null pointer");
  QSX_CHECK_ARG(m >= 1 && n >= 1 && k >= 1 && batch >= 1 && batch <= 65535,
                "qsx_zgemm: bad sizes m=%d n=%d k=%d batch=%d", m, n, k, batch);
  QSX_CHECK_ARG(lda >= ((trans & 1) ? m : k) && ldb >= ((trans & 2) ? k : n) && ldc >= n,
                "qsx_zgemm: bad leading dimensions");
  QSX_CHECK_ARG(batch == 1 || sC != 0, "qsx_zgemm: a batched C needs a batch stride");
  QSX_CHECK_ARG((gamma_re == 0.0 && gamma_im == 0.0) || m == n,
                "qsx_zgemm: the identity term needs a square C");
  const double2 al = make_double2(alpha_re, alpha_im), be = make_double2(beta_re, beta_im),
                ga = make_double2(gamma_re, gamma_im);
  hipStream_t s = as_stream(stream);
  // small problems: 16 x 16 tiles so that more workgroups share the work
  const int64_t tiles32 = (int64_t)((m + 31) / 32) * ((n + 31) / 32) * batch;
  if (tiles32 < 192) {
    dim3 grid((m + 15) / 16, (n + 15) / 16, batch);
    if (precision == QSX_FP64)
      launch_zgemm<double, 16, 16>(trans, grid, s, m, n, k, al, A, lda, sA, B, ldb, sB, be, C,
                                   ldc, sC, ga);
    else
      launch_zgemm<float, 16, 16>(trans, grid, s, m, n, k, al, A, lda, sA, B, ldb, sB, be, C,
                                  ldc, sC, ga);
  } else {
    dim3 grid((m + 31) / 32, (n + 31) / 32, batch);
    if (precision == QSX_FP64)
      launch_zgemm<double, 32, 32>(trans, grid, s, m, n, k, al, A, lda, sA, B, ldb, sB, be, C,
                                   ldc, sC, ga);
    else
      launch_zgemm<float, 32, 32>(trans, grid, s, m, n, k, al, A, lda, sA, B, ldb, sB, be, C,
                                  ldc, sC, ga);
  }
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}
