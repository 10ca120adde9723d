// MIXED-precision construction of the one-step RK4 propagator of a large
// block on the FP32 matrix cores (gfx950 v_mfma_f32_16x16x4f32).
//
// The stepper applies a propagator P as y <- y + off y + diag * y with off
// the off-diagonal part in complex64 and diag = diag(P) - 1 in complex128
// (qsx_split_increment).  Here the split is built without an FP64 P.
// Write L = D + O (D its diagonal) and run Horner's rule for P4(h L) beside
// the one for the diagonal D alone, T_k^D (k = 4, 3, 2, 1: exact FP64
// vectors).  The differences dT_k = T_k - T_k^D obey
//     dT_4 = (h/4) O,    dT_k = c_k (L dT_{k+1} + O diag(T_{k+1}^D))
// (c_k = h/3, h/2, h), so each step is one complex GEMM L dT plus the
// column-scaled O in the epilogue; dT stays of the size of h O (far below
// the unit diagonal), so FP32 GEMMs leave its relative error at the FP32
// level while the large diagonal T^D never leaves FP64.  q RK4 steps per
// application (q a power of two) square the split: with P = diag(p) + E,
//     P^2 = diag(p^2) + [E E + (p_i + p_j) E_ij].
// The result: off = E with its diagonal zeroed (complex64), diag =
// p - 1 + E_ii (complex128).
//
// Complex products are four real MFMAs on real and imaginary planes staged
// in LDS (re re, im im, re im, im re); a work-group of four waves computes a
// 32 x 32 output tile, each wave a 16 x 16 block; K in chunks of 32, the next
// chunk's loads issued before the current chunk's MFMAs.  Measured on the
// 256-member, 147-element batch (MI355X): this tiling 0.55 ms per propagator
// launch; 80 x 80 tiles of five waves (each operand row read twice instead
// of five times, L converted once to complex64) 0.62 ms -- fewer and longer
// work-groups hide the operand loads less well.

#include "qsx_common.h"

namespace qsx {
namespace {

typedef float qsx_float4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 32, kTK = 32, kThreads = 256;

// Epilogue kinds of split_gemm_kernel
//   HORNER: out = alpha * (acc + (i != j ? A_ij : 0) * cvec_j)    (A = L)
//   SQUARE: out = acc + (cvec_i + cvec_j) * B_ij                    (A = B = E)
enum { kHorner = 0, kSquare = 1 };

__device__ inline float2 ld_c64(const double2* p) {
  const double2 v = *p;
  return make_float2((float)v.x, (float)v.y);
}
__device__ inline float2 ld_c64(const float2* p) { return *p; }

// C[b] (M x M, complex64) from A[b] (complex128 or complex64) and B[b]
// (complex64); cvec [b][M] complex128.
template <int KIND, typename TA>
__global__ __launch_bounds__(kThreads) void split_gemm_kernel(
    int M, const TA* __restrict__ A, int64_t sA, const float2* __restrict__ B,
    const double2* __restrict__ cvec, double alpha, float2* __restrict__ C) {
  __shared__ float Ar[kTK][kTile + 1], Ai[kTK][kTile + 1];
  __shared__ float Br[kTK][kTile + 1], Bi[kTK][kTile + 1];
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int wm = wave % 2, wn = wave / 2;
  const int bm = blockIdx.x * kTile, bn = blockIdx.y * kTile, b = blockIdx.z;
  const int64_t mm = (int64_t)M * M;
  const TA* Ab = A + b * sA;
  const float2* Bb = B + b * mm;
  constexpr int kPer = kTile * kTK / kThreads;          // 4 elements per operand per thread
  float2 av[kPer], bv[kPer];
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = tid + kThreads * u;
      const int mi = e / kTK, kk = e % kTK;              // A: walk k (contiguous)
      const int gi = bm + mi, gk = k0 + kk;
      av[u] = (gi < M && gk < M) ? ld_c64(Ab + (int64_t)gi * M + gk) : make_float2(0.f, 0.f);
      const int kb = e / kTile, nj = e % kTile;          // B: walk n (contiguous)
      const int gkb = k0 + kb, gj = bn + nj;
      bv[u] = (gkb < M && gj < M) ? Bb[(int64_t)gkb * M + gj] : make_float2(0.f, 0.f);
    }
  };
  qsx_float4 crr = {0.f, 0.f, 0.f, 0.f}, cii = crr, cri = crr, cir = crr;
  const int r16 = lane % 16, q4 = lane / 16;
  load_chunk(0);
  for (int k0 = 0; k0 < M; k0 += kTK) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = tid + kThreads * u;
      Ar[e % kTK][e / kTK] = av[u].x;
      Ai[e % kTK][e / kTK] = av[u].y;
      Br[e / kTile][e % kTile] = bv[u].x;
      Bi[e / kTile][e % kTile] = bv[u].y;
    }
    __syncthreads();
    if (k0 + kTK < M) load_chunk(k0 + kTK);
    const int kend = min(kTK, M - k0);
    for (int kq = 0; kq < kend; kq += 4) {
      const float ar = Ar[kq + q4][wm * 16 + r16], ai = Ai[kq + q4][wm * 16 + r16];
      const float br = Br[kq + q4][wn * 16 + r16], bi = Bi[kq + q4][wn * 16 + r16];
      crr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, crr, 0, 0, 0);
      cii = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, bi, cii, 0, 0, 0);
      cri = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cri, 0, 0, 0);
      cir = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, cir, 0, 0, 0);
    }
    __syncthreads();
  }
  const qsx_float4 cr = crr - cii, ci = cri + cir;
  const double2* cv = cvec + (int64_t)b * M;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    // v_mfma_f32_16x16x4f32 result layout: lane l, element v holds row
    // 4 (l / 16) + v, column l % 16 of the 16 x 16 block
    const int gi = bm + wm * 16 + 4 * q4 + v, gj = bn + wn * 16 + r16;
    if (gi >= M || gj >= M) continue;
    double2 acc = make_double2(cr[v], ci[v]);
    double2 out;
    if (KIND == kHorner) {
      if (gi != gj) {
        const float2 x = ld_c64(Ab + (int64_t)gi * M + gj);
        acc = cadd(acc, cmul(make_double2(x.x, x.y), cv[gj]));
      }
      out = make_double2(alpha * acc.x, alpha * acc.y);
    } else {
      const float2 x = Bb[(int64_t)gi * M + gj];
      out = cadd(acc, cmul(make_double2(x.x, x.y), cadd(cv[gi], cv[gj])));
    }
    C[b * mm + (int64_t)gi * M + gj] = make_float2((float)out.x, (float)out.y);
  }
}

// Diagonal Horner vectors t4, t3, t2, t1 = P4(h D), stored k-major as
// tvec[k][B][M] (k = 0..3 for t4..t1), and dT4 = (h/4) O (complex64)
__global__ __launch_bounds__(kThreads) void split_prep_kernel(
    const double2* __restrict__ L, int64_t sL, int M, int B, double h,
    double2* __restrict__ tvec, float2* __restrict__ dT) {
  const int64_t mm = (int64_t)M * M, total = mm * B;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kThreads) {
    const int64_t b = e / mm, r = e % mm;
    const int i = (int)(r / M), j = (int)(r % M);
    const double2 v = L[b * sL + r];
    dT[e] = i == j ? make_float2(0.f, 0.f) : make_float2((float)(h / 4 * v.x), (float)(h / 4 * v.y));
    if (i == j) {
      // Horner on the scalar d: t4 = 1 + h d / 4, t3 = 1 + h d t4 / 3, ...
      const double2 hd = make_double2(h * v.x, h * v.y);
      double2 t = make_double2(1.0 + hd.x / 4, hd.y / 4);
      const int64_t bm = (int64_t)B * M;
      tvec[b * M + i] = t;                                     // t4
      const double c[3] = {3.0, 2.0, 1.0};
      for (int s = 0; s < 3; ++s) {
        const double2 p = cmul(hd, t);
        t = make_double2(1.0 + p.x / c[s], p.y / c[s]);
        tvec[(s + 1) * bm + b * M + i] = t;                    // t3, t2, t1
      }
    }
  }
}

// p <- p^2 (diagonal of a squared split), per batch entry
__global__ __launch_bounds__(kThreads) void square_diag_kernel(const double2* __restrict__ p,
                                                                double2* __restrict__ p2,
                                                                int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * kThreads)
    p2[e] = cmul(p[e], p[e]);
}

// off = E with zero diagonal (complex64), diag = p - 1 + E_ii (complex128)
__global__ __launch_bounds__(kThreads) void split_final_kernel(const float2* __restrict__ E,
                                                                const double2* __restrict__ p,
                                                                int M, int B,
                                                                float2* __restrict__ off,
                                                                double2* __restrict__ diag) {
  const int64_t mm = (int64_t)M * M, total = mm * B;
  for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kThreads) {
    const int64_t b = e / mm, r = e % mm;
    const int i = (int)(r / M), j = (int)(r % M);
    const float2 x = E[e];
    if (i == j) {
      const double2 pi = p[b * M + i];
      off[e] = make_float2(0.f, 0.f);
      diag[b * M + i] = make_double2(pi.x - 1.0 + x.x, pi.y + x.y);
    } else {
      off[e] = x;
    }
  }
}

inline dim3 ew_grid(int64_t total) {
  int64_t g = (total + kThreads - 1) / kThreads;
  return dim3((unsigned)(g < 65535 * 16 ? g : 65535 * 16));
}

}  // namespace
}  // namespace qsx

extern "C" int64_t qsx_split_propagator_workspace_bytes(int32_t M, int32_t B) {
  if (M < 1 || B < 1) return -1;
  // two complex64 M x M buffers and five complex128 vectors of M per entry
  return (int64_t)B * ((int64_t)2 * M * M * sizeof(float2) + (int64_t)5 * M * sizeof(double2));
}

extern "C" int32_t qsx_rk4_propagator_split(const void* L, int64_t strideL, int32_t M, int32_t B,
                                            double h, int32_t q, void* off, void* diag,
                                            void* workspace, int64_t workspace_bytes,
                                            void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(L && off && diag && workspace, "qsx_rk4_propagator_split: null pointer");
  QSX_CHECK_ARG(M >= 1 && B >= 1 && B <= 65535 && q >= 1 && (q & (q - 1)) == 0,
                "qsx_rk4_propagator_split: bad sizes M=%d B=%d q=%d (q a power of two)", M, B,
                q);
  QSX_CHECK_ARG(strideL == 0 || strideL >= (int64_t)M * M,
                "qsx_rk4_propagator_split: bad strideL");
  QSX_CHECK_ARG(workspace_bytes >= qsx_split_propagator_workspace_bytes(M, B),
                "qsx_rk4_propagator_split: workspace of %lld bytes, %lld needed",
                (long long)workspace_bytes, (long long)qsx_split_propagator_workspace_bytes(M, B));
  hipStream_t s = as_stream(stream);
  const int64_t mm = (int64_t)M * M;
  float2* X0 = static_cast<float2*>(workspace);
  float2* X1 = X0 + mm * B;
  double2* tvec = reinterpret_cast<double2*>(X1 + mm * B);   // [4][B][M]: t4, t3, t2, t1
  const int64_t bm = (int64_t)B * M;
  double2* pv[2] = {tvec + 3 * bm, tvec + 4 * bm};           // t1, then the squarings
  const double2* Lp = static_cast<const double2*>(L);
  const int64_t sL = B > 1 ? strideL : 0;
  hipLaunchKernelGGL(split_prep_kernel, ew_grid(mm * B), dim3(kThreads), 0, s, Lp, sL, M, B, h,
                     tvec, X0);
  QSX_CHECK_HIP(hipGetLastError());
  const dim3 grid((M + kTile - 1) / kTile, (M + kTile - 1) / kTile, B);
  // dT3 = h/3 (L dT4 + O t4), dT2 = h/2 (L dT3 + O t3), dT1 = h (L dT2 + O t2)
  float2* cur = X0;
  float2* nxt = X1;
  const double coef[3] = {h / 3.0, h / 2.0, h};
  for (int st = 0; st < 3; ++st) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(split_gemm_kernel<kHorner, double2>), grid,
                       dim3(kThreads), 0, s, M, Lp, sL, cur, tvec + st * bm, coef[st], nxt);
    QSX_CHECK_HIP(hipGetLastError());
    float2* t = cur;
    cur = nxt;
    nxt = t;
  }
  // P4 = diag(t1) + dT1; q = 2^k: square k times
  int pp = 0;
  for (int qq = q; qq > 1; qq >>= 1) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(split_gemm_kernel<kSquare, float2>), grid, dim3(kThreads),
                       0, s, M, cur, mm, cur, pv[pp], 1.0, nxt);
    QSX_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(square_diag_kernel, ew_grid((int64_t)M * B), dim3(kThreads), 0, s,
                       pv[pp], pv[pp ^ 1], (int64_t)M * B);
    QSX_CHECK_HIP(hipGetLastError());
    pp ^= 1;
    float2* t = cur;
    cur = nxt;
    nxt = t;
  }
  hipLaunchKernelGGL(split_final_kernel, ew_grid(mm * B), dim3(kThreads), 0, s, cur, pv[pp], M,
                     B, static_cast<float2*>(off), static_cast<double2*>(diag));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}
