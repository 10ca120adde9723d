// Reverse-mode (adjoint) propagation for y_p = P y_{p-1} (gfx950).
//
// The forward map is the one a propagation with the one-step RK4 propagator
// P = P4(h L) computes: state y_0, then save interval s applies P nsub[s]
// times and saves the state.  Given the cotangents gbar_s of the saved
// states (the gradient of a real loss with respect to conj(y), as torch
// defines it), the adjoint recursion runs backwards:
//     lambda <- lambda + gbar_s                       at every save point,
//     G      <- G + lambda y_{p-1}^H,   lambda <- P^H lambda   for every step,
// and returns G = dLoss/dP (per right-hand side) and lambda = dLoss/dy_0.
//
// Checkpointed recomputation: the states y_{p-1} are not stored by the
// forward pass.  A first sweep stores every ckpt-th state in the workspace
// (global memory); the reverse sweep then recomputes one segment of ckpt
// states at a time from its checkpoint into LDS and walks it backwards.
// Memory per right-hand side: ceil(N / ckpt) checkpoints + ckpt LDS states;
// work: 3 matrix-vector products and one rank-1 update per step.
//
// Layout as the forward stepper: one workgroup per right-hand side, thread
// t owns row t / S and columns k + S c (k = t % S) of P, of P^H and of the
// gradient accumulator, all in registers; products reduce over the S lanes
// of a row with wave shuffles.  Blocks up to M = 64.
//
// Blocks of 65..160 elements (adjoint_large_kernel): work-groups of 512
// threads per right-hand side, each running the whole recursion and
// accumulating its part of the gradient (8192 entries, 16 per thread, in
// registers); P and P^H (written to the workspace by a transposition
// kernel) are read row by row from global memory (L2) by the wavefronts,
// lanes over the columns, one shuffle reduction per row; the states and the
// adjoint vector live in LDS.

#include "qsx_common.h"

namespace qsx {
namespace {

__device__ inline double2 shfl_xor_d2(double2 v, int m) {
  double2 r;
  r.x = __shfl_xor(v.x, m);
  r.y = __shfl_xor(v.y, m);
  return r;
}

__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// acc += a * conj(b)
__device__ inline void cfma_conj(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.y, b.x, acc.y);
  acc.y = fma(-a.x, b.y, acc.y);
}

template <int NT>
__device__ inline int block_sum(int v, int* red) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  const int wave = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0) red[wave] = v;
  __syncthreads();
  int s = 0;
  for (int w = 0; w < NT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

template <int NT, int S, int W>
__global__ __launch_bounds__(NT) void adjoint_kernel(
    const double2* __restrict__ A, int64_t strideA, int M, const double2* __restrict__ y0, int R,
    const int* __restrict__ nsub, int nsave, const double2* __restrict__ gout, int ckpt,
    double2* __restrict__ ck, int64_t nck_cap, double2* __restrict__ gA,
    double2* __restrict__ gy0, int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  double2* seg = reinterpret_cast<double2*>(smem);   // [ckpt][M] recomputed states
  double2* zb = seg + (int64_t)ckpt * M;              // [2][M] forward vector
  double2* lb = zb + 2 * M;                           // [2][M] adjoint vector
  int* red = reinterpret_cast<int*>(lb + 2 * M);      // [NT / 64]

  const int tid = threadIdx.x;
  const int r = blockIdx.x, b = blockIdx.y;
  const int k = tid % S;
  const int row = tid / S;
  const bool on = row < M;
  const bool lead = on && k == 0;
  const double2* Ab = A + (int64_t)b * strideA;
  const int64_t rhs = (int64_t)b * R + r;

  double2 af[W], ab[W], g[W];
#pragma unroll
  for (int c = 0; c < W; ++c) {
    const int j = k + S * c;
    const bool in = on && j < M;
    af[c] = in ? Ab[(int64_t)row * M + j] : make_double2(0.0, 0.0);
    ab[c] = in ? cconj(Ab[(int64_t)j * M + row]) : make_double2(0.0, 0.0);
    g[c] = make_double2(0.0, 0.0);
  }

  int part = 0;
  for (int s = tid; s < nsave; s += NT) part += max(nsub[s], 0);
  const int N = block_sum<NT>(part, red);
  const int nck = (N + ckpt - 1) / ckpt;
  if ((int64_t)nck > nck_cap) {            // uniform over the workgroup
    if (tid == 0) status[0] = 2;
    return;
  }

  auto matvec = [&](const double2 (&a)[W], const double2* z) -> double2 {
    double2 s = make_double2(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < W; ++c) {
      const int j = k + S * c;
      if (j < M) cfma<double>(s, a[c], z[j]);
    }
#pragma unroll
    for (int m = S / 2; m > 0; m >>= 1) {
      const double2 o = shfl_xor_d2(s, m);
      s.x += o.x;
      s.y += o.y;
    }
    return s;
  };

  // ---- forward sweep: checkpoints at positions 0, ckpt, 2 ckpt, ...
  double2* ckb = ck + rhs * nck_cap * M;
  double2 y = lead ? y0[rhs * M + row] : make_double2(0.0, 0.0);
  int cz = 0;
  if (lead) zb[row] = y;
  __syncthreads();
  for (int j = 0; j < nck; ++j) {
    if (lead) ckb[(int64_t)j * M + row] = y;
    if (j == nck - 1) break;
    for (int i = 0; i < ckpt; ++i) {
      const double2 v = matvec(af, zb + cz * M);
      if (lead) {
        y = v;
        zb[(cz ^ 1) * M + row] = y;
      }
      cz ^= 1;
      __syncthreads();
    }
  }

  // ---- reverse sweep
  const double2* gb = gout + rhs * (int64_t)nsave * M;
  double2 lam = make_double2(0.0, 0.0);
  int cl = 0;
  int p = N;                 // position of the state lambda belongs to
  int loaded = -1;           // checkpoint segment held in seg
  bool bad = false;
  for (int s = nsave - 1; s >= 0; --s) {
    if (lead) {
      lam = cadd(lam, gb[(int64_t)s * M + row]);
      lb[cl * M + row] = lam;
    }
    __syncthreads();
    const int n = max(nsub[s], 0);
    for (int t = 0; t < n; ++t) {
      const int q = p - 1;   // y_q is the input of this step
      const int jseg = q / ckpt;
      if (jseg != loaded) {
        const int m = min(ckpt, N - jseg * ckpt);
        if (lead) {
          y = ckb[(int64_t)jseg * M + row];
          seg[row] = y;
          zb[row] = y;
        }
        cz = 0;
        __syncthreads();
        for (int i = 1; i < m; ++i) {
          const double2 v = matvec(af, zb + cz * M);
          if (lead) {
            y = v;
            zb[(cz ^ 1) * M + row] = y;
            seg[(int64_t)i * M + row] = y;
          }
          cz ^= 1;
          __syncthreads();
        }
        loaded = jseg;
      }
      const double2* yq = seg + (int64_t)(q - jseg * ckpt) * M;
      const double2 lr = on ? lb[cl * M + row] : make_double2(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < W; ++c) {
        const int j = k + S * c;
        if (j < M) cfma_conj(g[c], lr, yq[j]);
      }
      const double2 v = matvec(ab, lb + cl * M);
      if (lead) {
        lam = v;
        lb[(cl ^ 1) * M + row] = lam;
      }
      cl ^= 1;
      p = q;
      __syncthreads();
    }
  }

  if (on) {
    double2* gAb = gA + rhs * (int64_t)M * M;
#pragma unroll
    for (int c = 0; c < W; ++c) {
      const int j = k + S * c;
      if (j < M) {
        gAb[(int64_t)row * M + j] = g[c];
        if (!cfinite(g[c])) bad = true;
      }
    }
  }
  if (lead) {
    if (gy0 != nullptr) gy0[rhs * M + row] = lam;
    if (!cfinite(lam)) bad = true;
  }
  if (bad) status[0] = 1;
}

constexpr int kLargeThreads = 512;
constexpr int kLargeWaves = kLargeThreads / 64;
constexpr int kLargeG = 16;            // gradient entries per thread
constexpr int kLargePart = kLargeThreads * kLargeG;   // entries per work-group
constexpr int kLargeMaxM = 160;

// work-groups per right-hand side (parts of the gradient) for a block of M
inline int large_parts(int M) { return (M * M + kLargePart - 1) / kLargePart; }

// out[i] = sum_j X[i M + j] z[j] (X row-major in global memory, z and out in
// LDS): rows over the wavefronts, columns over the lanes
__device__ inline void wave_matvec(const double2* __restrict__ X, int M, const double2* z,
                                   double2* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = w; i < M; i += kLargeWaves) {
    double2 s = make_double2(0.0, 0.0);
    for (int j = lane; j < M; j += 64) cfma<double>(s, X[(int64_t)i * M + j], z[j]);
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      const double2 o = shfl_xor_d2(s, m);
      s.x += o.x;
      s.y += o.y;
    }
    if (lane == 0) out[i] = s;
  }
}

// AH[b] = A[b]^H for the nA distinct operators
__global__ __launch_bounds__(256) void conj_transpose_kernel(const double2* __restrict__ A,
                                                             int64_t strideA, int M, int nA,
                                                             double2* __restrict__ AH) {
  const int64_t mm = (int64_t)M * M, total = mm * nA;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t b = e / mm, r = e % mm;
    const int i = (int)(r / M), j = (int)(r % M);
    AH[e] = cconj(A[b * strideA + (int64_t)j * M + i]);
  }
}

// The adjoint recursion of adjoint_kernel for 64 < M <= kLargeMaxM; work-
// group (r, b, part) accumulates gradient entries part kLargePart + tid +
// kLargeThreads c, with checkpoints of its own.
__global__ __launch_bounds__(kLargeThreads) void adjoint_large_kernel(
    const double2* __restrict__ A, int64_t strideA, const double2* __restrict__ AH, int M,
    const double2* __restrict__ y0, int R, const int* __restrict__ nsub, int nsave,
    const double2* __restrict__ gout, int ckpt, double2* __restrict__ ck, int64_t nck_cap,
    double2* __restrict__ gA, double2* __restrict__ gy0, int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  double2* seg = reinterpret_cast<double2*>(smem);   // [ckpt][M] recomputed states
  double2* zb = seg + (int64_t)ckpt * M;              // [2][M] forward vector
  double2* lb = zb + 2 * M;                           // [2][M] adjoint vector
  int* red = reinterpret_cast<int*>(lb + 2 * M);      // [kLargeWaves]

  const int tid = threadIdx.x;
  const int r = blockIdx.x, b = blockIdx.y, gpart = blockIdx.z;
  const double2* Ab = A + (int64_t)b * strideA;
  const double2* AHb = AH + (strideA ? (int64_t)b * M * M : 0);
  const int64_t rhs = (int64_t)b * R + r;
  const int mm = M * M;
  const int e0 = gpart * kLargePart + tid;         // this thread's first entry

  double2 g[kLargeG];
#pragma unroll
  for (int c = 0; c < kLargeG; ++c) g[c] = make_double2(0.0, 0.0);
  // entry e = e0 + kLargeThreads c of the accumulator is (i, j) = (e / M, e % M)
  const int di = kLargeThreads / M, dj = kLargeThreads % M;

  int part = 0;
  for (int s = tid; s < nsave; s += kLargeThreads) part += max(nsub[s], 0);
  const int N = block_sum<kLargeThreads>(part, red);
  const int nck = (N + ckpt - 1) / ckpt;
  if ((int64_t)nck > nck_cap) {            // uniform over the workgroup
    if (tid == 0) status[0] = 2;
    return;
  }

  // ---- forward sweep: checkpoints at positions 0, ckpt, 2 ckpt, ...
  double2* ckb = ck + ((int64_t)gpart * gridDim.x * gridDim.y + rhs) * nck_cap * M;
  int cz = 0;
  for (int e = tid; e < M; e += kLargeThreads) {
    zb[e] = y0[rhs * M + e];
    lb[e] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int j = 0; j < nck; ++j) {
    for (int e = tid; e < M; e += kLargeThreads) ckb[(int64_t)j * M + e] = zb[cz * M + e];
    if (j == nck - 1) break;
    for (int i = 0; i < ckpt; ++i) {
      wave_matvec(Ab, M, zb + cz * M, zb + (cz ^ 1) * M);
      cz ^= 1;
      __syncthreads();
    }
  }

  // ---- reverse sweep
  const double2* gb = gout + rhs * (int64_t)nsave * M;
  int cl = 0;
  int p = N;                 // position of the state lambda belongs to
  int loaded = -1;           // checkpoint segment held in seg
  for (int s = nsave - 1; s >= 0; --s) {
    for (int e = tid; e < M; e += kLargeThreads) {
      const double2 v = gb[(int64_t)s * M + e];
      lb[cl * M + e] = cadd(lb[cl * M + e], v);
    }
    __syncthreads();
    const int n = max(nsub[s], 0);
    for (int t = 0; t < n; ++t) {
      const int q = p - 1;   // y_q is the input of this step
      const int jseg = q / ckpt;
      if (jseg != loaded) {
        const int m = min(ckpt, N - jseg * ckpt);
        for (int e = tid; e < M; e += kLargeThreads) seg[e] = ckb[(int64_t)jseg * M + e];
        __syncthreads();
        for (int i = 1; i < m; ++i) {
          wave_matvec(Ab, M, seg + (int64_t)(i - 1) * M, seg + (int64_t)i * M);
          __syncthreads();
        }
        loaded = jseg;
      }
      const double2* yq = seg + (int64_t)(q - jseg * ckpt) * M;
      const double2* lam = lb + cl * M;
      int i = e0 / M, j = e0 % M;
#pragma unroll
      for (int c = 0; c < kLargeG; ++c) {
        if (e0 + kLargeThreads * c < mm) cfma_conj(g[c], lam[i], yq[j]);
        i += di;
        j += dj;
        if (j >= M) {
          j -= M;
          ++i;
        }
      }
      wave_matvec(AHb, M, lam, lb + (cl ^ 1) * M);
      cl ^= 1;
      p = q;
      __syncthreads();
    }
  }

  bool bad = false;
  double2* gAb = gA + rhs * (int64_t)mm;
#pragma unroll
  for (int c = 0; c < kLargeG; ++c) {
    const int e = e0 + kLargeThreads * c;
    if (e < mm) {
      gAb[e] = g[c];
      if (!cfinite(g[c])) bad = true;
    }
  }
  if (gpart == 0) {
    for (int e = tid; e < M; e += kLargeThreads) {
      const double2 v = lb[cl * M + e];
      if (gy0 != nullptr) gy0[rhs * M + e] = v;
      if (!cfinite(v)) bad = true;
    }
  }
  if (bad) status[0] = 1;
}

inline size_t adjoint_large_lds_bytes(int M, int ckpt) {
  return ((size_t)ckpt * M + 4 * (size_t)M) * sizeof(double2) + kLargeWaves * sizeof(int);
}

struct AdjShape {
  int NT, S, W;
};

inline bool pick_adjoint_shape(int M, AdjShape& sh) {
  if (M <= 8) sh = {64, 8, 1};
  else if (M <= 16) sh = {64, 4, 4};
  else if (M <= 32) sh = {128, 4, 8};
  else if (M <= 64) sh = {256, 4, 16};
  else return false;
  return true;
}

inline size_t adjoint_lds_bytes(int M, int ckpt, int NT) {
  return ((size_t)ckpt * M + 4 * (size_t)M) * sizeof(double2) + (NT / 64) * sizeof(int);
}

}  // namespace
}  // namespace qsx

extern "C" int64_t qsx_adjoint_workspace_bytes(int32_t M, int32_t R, int32_t B,
                                               int64_t total_steps, int32_t ckpt) {
  if (M < 1 || R < 1 || B < 1 || total_steps < 0 || ckpt < 1) return -1;
  const int64_t nck = total_steps > 0 ? (total_steps + ckpt - 1) / ckpt : 1;
  // above 64 elements the operators' conjugate transposes come first, and
  // every part of the gradient keeps checkpoints of its own
  const int64_t ah = M > 64 ? (int64_t)B * M * M * (int64_t)sizeof(double2) : 0;
  const int64_t parts = M > 64 ? qsx::large_parts(M) : 1;
  return ah + parts * nck * (int64_t)M * R * B * (int64_t)sizeof(double2);
}

extern "C" int32_t qsx_propagate_adjoint(const void* A, int64_t strideA, int32_t M,
                                         const void* y0, int32_t R, int32_t B,
                                         const int32_t* nsub, int32_t nsave, const void* gout,
                                         int32_t ckpt, void* workspace, int64_t workspace_bytes,
                                         void* gA, void* gy0, int32_t* status, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(A && y0 && nsub && gout && workspace && gA && status,
                "qsx_propagate_adjoint: null pointer");
  QSX_CHECK_ARG(M >= 1 && R >= 1 && B >= 1 && nsave >= 1 && B <= 65535 && R <= (1 << 30),
                "qsx_propagate_adjoint: bad sizes M=%d R=%d B=%d nsave=%d", M, R, B, nsave);
  QSX_CHECK_ARG(strideA == 0 || strideA >= (int64_t)M * M, "qsx_propagate_adjoint: bad strideA");
  QSX_CHECK_ARG(ckpt >= 1, "qsx_propagate_adjoint: ckpt must be >= 1");
  if (M > 64) {
    QSX_CHECK_ARG(M <= kLargeMaxM, "qsx_propagate_adjoint: M=%d exceeds %d", M, kLargeMaxM);
    const size_t shmem = adjoint_large_lds_bytes(M, ckpt);
    QSX_CHECK_ARG(shmem <= 160 * 1024,
                  "qsx_propagate_adjoint: ckpt=%d x M=%d states exceed the LDS (%zu bytes)",
                  ckpt, M, shmem);
    const int nA = strideA ? B : 1;
    const int64_t ah = (int64_t)B * M * M * (int64_t)sizeof(double2);
    const int parts = large_parts(M);
    const int64_t per = (int64_t)parts * M * R * B * (int64_t)sizeof(double2);
    const int64_t nck_cap = workspace_bytes > ah ? (workspace_bytes - ah) / per : 0;
    QSX_CHECK_ARG(nck_cap >= 1, "qsx_propagate_adjoint: workspace too small");
    hipStream_t s = as_stream(stream);
    double2* AH = static_cast<double2*>(workspace);
    double2* ckw = AH + (int64_t)B * M * M;
    const int64_t tot = (int64_t)M * M * nA;
    hipLaunchKernelGGL(conj_transpose_kernel, dim3((unsigned)std::min<int64_t>((tot + 255) / 256, 65535)),
                       dim3(256), 0, s, static_cast<const double2*>(A), strideA, M, nA, AH);
    QSX_CHECK_HIP(hipGetLastError());
    if (shmem > 64 * 1024)
      QSX_CHECK_HIP(hipFuncSetAttribute((const void*)adjoint_large_kernel,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL(adjoint_large_kernel, dim3(R, B, parts), dim3(kLargeThreads), shmem, s,
                       static_cast<const double2*>(A), strideA, AH, M,
                       static_cast<const double2*>(y0), R, nsub, nsave,
                       static_cast<const double2*>(gout), ckpt, ckw, nck_cap,
                       static_cast<double2*>(gA), static_cast<double2*>(gy0), status);
    QSX_CHECK_HIP(hipGetLastError());
    return QSX_OK;
  }
  AdjShape sh;
  QSX_CHECK_ARG(pick_adjoint_shape(M, sh), "qsx_propagate_adjoint: M=%d exceeds 64", M);
  const size_t shmem = adjoint_lds_bytes(M, ckpt, sh.NT);
  QSX_CHECK_ARG(shmem <= 160 * 1024,
                "qsx_propagate_adjoint: ckpt=%d x M=%d states exceed the LDS (%zu bytes)", ckpt,
                M, shmem);
  const int64_t per = (int64_t)M * R * B * (int64_t)sizeof(double2);
  const int64_t nck_cap = workspace_bytes / per;
  QSX_CHECK_ARG(nck_cap >= 1, "qsx_propagate_adjoint: workspace too small");
  hipStream_t s = as_stream(stream);
  dim3 grid(R, B);
#define QSX_ADJ(NT_, S_, W_)                                                                  \
  if (sh.NT == NT_ && sh.S == S_ && sh.W == W_) {                                             \
    auto kern = adjoint_kernel<NT_, S_, W_>;                                                  \
    if (shmem > 64 * 1024)                                                                    \
      QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                    \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)); \
    hipLaunchKernelGGL(kern, grid, dim3(NT_), shmem, s, (const double2*)A, strideA, M,        \
                       (const double2*)y0, R, nsub, nsave, (const double2*)gout, ckpt,         \
                       (double2*)workspace, nck_cap, (double2*)gA, (double2*)gy0, status);     \
    QSX_CHECK_HIP(hipGetLastError());                                                         \
    return QSX_OK;                                                                            \
  }
  QSX_ADJ(64, 8, 1)
  QSX_ADJ(64, 4, 4)
  QSX_ADJ(128, 4, 8)
  QSX_ADJ(256, 4, 16)
#undef QSX_ADJ
  set_error("qsx_propagate_adjoint: no kernel for M=%d", M);
  return QSX_EINVAL;
}
