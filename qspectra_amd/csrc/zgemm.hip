// Batched complex GEMM with an identity term (gfx950):
//   C[b] = alpha * A[b] * B[b] + beta * C[b] + gamma * I
// Row-major storage, op(A) and op(B) optionally transposed, K staged through
// LDS.  FP64: complex128 throughout; large products on the FP64 matrix cores
// (16 x 16 tiles per wave, 32 x 32 per 4-wave work-group), small ones on the
// vector units.  MIXED: complex64 A and B with FP32 accumulation on the
// vector units.  Vector path: 32 x 32 or 16 x 16 tiles per 256-thread
// work-group, 2 x 2 or 1 x 1 block per thread, K in chunks of 16 or (for
// k >= 48) 64.  Used for the Horner construction of the RK4 propagator, its
// powers, save-point dipole maps and response contractions: operands here
// are at most a few hundred wide and the products sit on sequential chains,
// so the kernels favour many small tiles and few K chunks (each chunk's
// operand loads are issued together: one memory latency per chunk) over
// deep register blocking.

#include "qsx_common.h"

namespace qsx {
namespace {

// TM x TN output tile per 256-thread workgroup, each thread an (TM/16) x
// (TN/16) block, K in chunks of TK.  op(A) = A or A^T (TA), op(B) = B or B^T
// (TB); the LDS staging loads walk the contiguous index of the stored
// operand and are unrolled, so a chunk's loads are all in flight together,
// and the loads of chunk c + 1 are issued before the FMAs of chunk c.
template <typename T, int TM, int TN, int TK, bool TA, bool TB>
__global__ __launch_bounds__(256) void zgemm_kernel(int m, int n, int k, double2 alpha,
                                                    const cplx_t<T>* __restrict__ A, int64_t lda,
                                                    int64_t sA, const cplx_t<T>* __restrict__ B,
                                                    int64_t ldb, int64_t sB, double2 beta,
                                                    double2* __restrict__ C, int64_t ldc,
                                                    int64_t sC, double2 gamma) {
  using T2 = cplx_t<T>;
  constexpr int RM = TM / 16, RN = TN / 16;
  static_assert((TM * TK) % 256 == 0 && (TN * TK) % 256 == 0, "tile and chunk sizes");
  __shared__ T2 As[TK][TM + 1];
  __shared__ T2 Bs[TK][TN + 1];
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
  T2 av[TM * TK / 256], bv[TN * TK / 256];      // a chunk's loads, all in flight
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < TM * TK / 256; ++u) {
      const int e = tid + 256 * u;
      int mm, kk;
      if (TA) { kk = e / TM; mm = e % TM; }       // A stored k-major: walk m
      else { mm = e / TK; kk = e % TK; }          // A stored m-major: walk k
      const int gi = bm + mm, gk = k0 + kk;
      const int64_t off = TA ? (int64_t)gk * lda + gi : (int64_t)gi * lda + gk;
      av[u] = (gi < m && gk < k) ? Ab[off] : T2{0, 0};
    }
#pragma unroll
    for (int u = 0; u < TN * TK / 256; ++u) {
      const int e = tid + 256 * u;
      int kk, nn;
      if (TB) { nn = e / TK; kk = e % TK; }       // B stored n-major: walk k
      else { kk = e / TN; nn = e % TN; }          // B stored k-major: walk n
      const int gk = k0 + kk, gj = bn + nn;
      const int64_t off = TB ? (int64_t)gj * ldb + gk : (int64_t)gk * ldb + gj;
      bv[u] = (gk < k && gj < n) ? Bb[off] : T2{0, 0};
    }
  };
  load_chunk(0);
  for (int k0 = 0; k0 < k; k0 += TK) {
#pragma unroll
    for (int u = 0; u < TM * TK / 256; ++u) {
      const int e = tid + 256 * u;
      if (TA) As[e / TM][e % TM] = av[u];
      else As[e % TK][e / TK] = av[u];
    }
#pragma unroll
    for (int u = 0; u < TN * TK / 256; ++u) {
      const int e = tid + 256 * u;
      if (TB) Bs[e % TK][e / TK] = bv[u];
      else Bs[e / TN][e % TN] = bv[u];
    }
    __syncthreads();
    // the next chunk's loads are in flight while this chunk's FMAs run
    if (k0 + TK < k) load_chunk(k0 + TK);
    const int kend = min(TK, k - k0);
#pragma unroll 4
    for (int kk = 0; kk < kend; ++kk) {
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

// FP64 path on the matrix cores: each wave owns a 16 x 16 complex output
// tile held as two v_mfma_f64_16x16x4f64 accumulators (real, imaginary);
// a work-group is WM x WN waves.  Operand fragments (gfx950 layout): A lane
// l = A[l % 16][k0 + l / 16], B lane l = B[k0 + l / 16][l % 16]; result lane
// l, element i = D[l / 16 + 4 i][l % 16].  Complex product = four real MFMAs
// (re re, im im, re im, im re) into four accumulators on real/imaginary
// planes staged in LDS, K in chunks of TK (64) whose operand loads are all
// in flight together; the loads of chunk c + 1 are issued before the MFMAs
// of chunk c.
typedef double qsx_double4 __attribute__((ext_vector_type(4)));

template <bool TA, bool TB, int WM, int WN, int TK>
__global__ __launch_bounds__(64 * WM * WN) void zgemm_mfma_kernel(
    int m, int n, int k, double2 alpha, const double2* __restrict__ A, int64_t lda, int64_t sA,
    const double2* __restrict__ B, int64_t ldb, int64_t sB, double2 beta,
    double2* __restrict__ C, int64_t ldc, int64_t sC, double2 gamma) {
  constexpr int NT = 64 * WM * WN, TM = 16 * WM, TN = 16 * WN;
  __shared__ double Ar[TK][TM + 1], Ai[TK][TM + 1];
  __shared__ double Br[TK][TN + 1], Bi[TK][TN + 1];
  const int tid = threadIdx.x;
  const int lane = tid % 64, wave = tid / 64;
  const int wm = wave % WM, wn = wave / WM;
  const int bm = blockIdx.x * TM, bn = blockIdx.y * TN, bz = blockIdx.z;
  const double2* Ab = A + (int64_t)bz * sA;
  const double2* Bb = B + (int64_t)bz * sB;
  double2* Cb = C + (int64_t)bz * sC;
  // four independent accumulation chains (re re, im im, re im, im re), so
  // consecutive MFMAs never wait for each other's result
  qsx_double4 crr = {0.0, 0.0, 0.0, 0.0}, cii = crr, cri = crr, cir = crr;
  const int r16 = lane % 16, q4 = lane / 16;
  static_assert((TM * TK) % NT == 0 && (TN * TK) % NT == 0, "tile and chunk sizes");
  double2 av[TM * TK / NT], bv[TN * TK / NT];     // a chunk's loads, all in flight
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < TM * TK / NT; ++u) {
      const int e = tid + NT * u;
      int mm, kk;
      if (TA) { kk = e / TM; mm = e % TM; }
      else { mm = e / TK; kk = e % TK; }
      const int gi = bm + mm, gk = k0 + kk;
      const int64_t off = TA ? (int64_t)gk * lda + gi : (int64_t)gi * lda + gk;
      av[u] = (gi < m && gk < k) ? Ab[off] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < TN * TK / NT; ++u) {
      const int e = tid + NT * u;
      int kk, nn;
      if (TB) { nn = e / TK; kk = e % TK; }
      else { kk = e / TN; nn = e % TN; }
      const int gk = k0 + kk, gj = bn + nn;
      const int64_t off = TB ? (int64_t)gj * ldb + gk : (int64_t)gk * ldb + gj;
      bv[u] = (gk < k && gj < n) ? Bb[off] : make_double2(0.0, 0.0);
    }
  };
  load_chunk(0);
  for (int k0 = 0; k0 < k; k0 += TK) {
#pragma unroll
    for (int u = 0; u < TM * TK / NT; ++u) {
      const int e = tid + NT * u;
      const int kk = TA ? e / TM : e % TK, mm = TA ? e % TM : e / TK;
      Ar[kk][mm] = av[u].x;
      Ai[kk][mm] = av[u].y;
    }
#pragma unroll
    for (int u = 0; u < TN * TK / NT; ++u) {
      const int e = tid + NT * u;
      const int kk = TB ? e % TK : e / TN, nn = TB ? e / TK : e % TN;
      Br[kk][nn] = bv[u].x;
      Bi[kk][nn] = bv[u].y;
    }
    __syncthreads();
    // the next chunk's loads are in flight while this chunk's MFMAs run
    if (k0 + TK < k) load_chunk(k0 + TK);
    const int kend = min(TK, k - k0);
    for (int kq = 0; kq < kend; kq += 4) {
      const double ar = Ar[kq + q4][wm * 16 + r16], ai = Ai[kq + q4][wm * 16 + r16];
      const double br = Br[kq + q4][wn * 16 + r16], bi = Bi[kq + q4][wn * 16 + r16];
      crr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, crr, 0, 0, 0);
      cii = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, cii, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, cri, 0, 0, 0);
      cir = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, cir, 0, 0, 0);
    }
    __syncthreads();
  }
  const qsx_double4 cr = crr - cii, ci = cri + cir;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gi = bm + wm * 16 + q4 + 4 * i, gj = bn + wn * 16 + r16;
    if (gi >= m || gj >= n) continue;
    double2 v = cmul(alpha, make_double2(cr[i], ci[i]));
    double2* cp = Cb + (int64_t)gi * ldc + gj;
    if (beta.x != 0.0 || beta.y != 0.0) v = cadd(v, cmul(beta, *cp));
    if (gi == gj) v = cadd(v, gamma);
    *cp = v;
  }
}

template <int WM, int WN, int TK>
void launch_zgemm_mfma(int trans, hipStream_t s, int m, int n, int k, double2 al, const void* A,
                       int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                       double2 be, void* C, int64_t ldc, int64_t sC, double2 ga, int batch) {
  dim3 grid((m + 16 * WM - 1) / (16 * WM), (n + 16 * WN - 1) / (16 * WN), batch);
#define QSX_ZM(TA_, TB_)                                                                      \
  hipLaunchKernelGGL((zgemm_mfma_kernel<TA_, TB_, WM, WN, TK>), grid, dim3(64 * WM * WN), 0, s, \
                     m, n, k, al, (const double2*)A, lda, sA, (const double2*)B, ldb, sB, be,  \
                     (double2*)C, ldc, sC, ga)
  switch (trans & 3) {
    case 0: QSX_ZM(false, false); break;
    case 1: QSX_ZM(true, false); break;
    case 2: QSX_ZM(false, true); break;
    default: QSX_ZM(true, true); break;
  }
#undef QSX_ZM
}

template <typename T, int TM, int TN, int TK>
void launch_zgemm(int trans, dim3 grid, hipStream_t s, int m, int n, int k, double2 al,
                  const void* A, int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB,
                  double2 be, void* C, int64_t ldc, int64_t sC, double2 ga) {
  using T2 = cplx_t<T>;
#define QSX_ZG(TA_, TB_)                                                                      \
  hipLaunchKernelGGL((zgemm_kernel<T, TM, TN, TK, TA_, TB_>), grid, dim3(256), 0, s, m, n, k, \
                     al, (const T2*)A, lda, sA, (const T2*)B, ldb, sB, be, (double2*)C, ldc, sC, \
                     ga)
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

int qsx::zgemm_enqueue(int precision, int trans, int m, int n, int k, double alpha_re,
                       double alpha_im, const void* A, int64_t lda, int64_t sA, const void* B,
                       int64_t ldb, int64_t sB, double beta_re, double beta_im, void* C,
                       int64_t ldc, int64_t sC, double gamma_re, double gamma_im, int batch,
                       void* stream) {
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
  // FP64 with enough 32 x 32 tiles to fill the chip: matrix cores; small and
  // skinny products are latency-bound and run faster on the vector path
  // (measured on MI355X: 255 x 255 x 154 x 4 16 us vs 21 us; 255 x 147 x 128
  // 16 us vs 9 us)
  // (and output tiles at least half full: a 7-row product padded to 32 rows
  // wastes the matrix cores four times over)
  if (precision == QSX_FP64 && tiles32 >= 192 && m >= 16 && n >= 16) {
    launch_zgemm_mfma<2, 2, 64>(trans, s, m, n, k, al, A, lda, sA, B, ldb, sB, be, C, ldc, sC,
                                ga, batch);
    QSX_CHECK_HIP(hipGetLastError());
    return QSX_OK;
  }
  const bool deep = k >= 48;                 // few, deep K chunks
#define QSX_ZL(T_, TM_, TN_)                                                                  \
  (deep ? launch_zgemm<T_, TM_, TN_, 64>(trans, grid, s, m, n, k, al, A, lda, sA, B, ldb, sB, \
                                         be, C, ldc, sC, ga)                                  \
        : launch_zgemm<T_, TM_, TN_, 16>(trans, grid, s, m, n, k, al, A, lda, sA, B, ldb, sB, \
                                         be, C, ldc, sC, ga))
  if (tiles32 < 192) {
    dim3 grid((m + 15) / 16, (n + 15) / 16, batch);
    if (precision == QSX_FP64) QSX_ZL(double, 16, 16);
    else QSX_ZL(float, 16, 16);
  } else {
    dim3 grid((m + 31) / 32, (n + 31) / 32, batch);
    if (precision == QSX_FP64) QSX_ZL(double, 32, 32);
    else QSX_ZL(float, 32, 32);
  }
#undef QSX_ZL
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_zgemm(int32_t precision, int32_t trans, int32_t m, int32_t n, int32_t k,
                             double alpha_re, double alpha_im, const void* A, int64_t lda,
                             int64_t sA, const void* B, int64_t ldb, int64_t sB, double beta_re,
                             double beta_im, void* C, int64_t ldc, int64_t sC, double gamma_re,
                             double gamma_im, int32_t batch, void* stream) {
  return qsx::zgemm_enqueue(precision, trans, m, n, k, alpha_re, alpha_im, A, lda, sA, B, ldb,
                            sB, beta_re, beta_im, C, ldc, sC, gamma_re, gamma_im, batch, stream);
}
