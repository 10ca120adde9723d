// Device construction of the propagation operators of a time-invariant
// Liouvillian block (gfx950), enqueued on a stream so that a captured graph
// rebuilds them from the Liouvillian on every replay.
//
//   qsx_rk4_propagator   P = P4(h L)^q: the one-step map of classical RK4
//                        (Horner: T = I + hL/4, T = I + hL/3 T, T = I + hL/2 T,
//                        P = I + hL T; three GEMMs), raised to the q steps of
//                        one save interval by repeated squaring;
//   qsx_matrix_powers    A^e for a list of exponents, sharing one chain of
//                        squarings (binary powering);
//   qsx_split_increment  the MIXED operand of a propagator: its off-diagonal
//                        part in complex64 and its diagonal minus one in
//                        complex128 (y <- y + (P - I) y with the large
//                        diagonal applied in FP64).
// GEMMs go through the library's complex GEMM (zgemm.hip: FP64 matrix cores
// for the large batched products); the element-wise steps are the kernels
// below.  Nothing here allocates: the caller passes the workspace.

#include <mutex>
#include <vector>

#include "qsx_common.h"

// Workspace of qsx_rk4_propagator_noalloc: one copy per GPU, at its own
// address, held until the process ends.  Host code never passes this name
// to a kernel (on the host it is a stand-in, not a device address): the
// per-device address comes from hipGetSymbolAddress (scratch_address).
static __device__ __align__(256) unsigned char scratch[QSX_PROPAGATOR_SCRATCH_BYTES];

namespace qsx {
namespace {

constexpr int kEwThreads = 256;

inline dim3 ew_grid(int64_t total) {
  int64_t g = (total + kEwThreads - 1) / kEwThreads;
  return dim3((unsigned)(g < 65535 * 16 ? g : 65535 * 16));
}

// Q[b] = I/2 + (h/6) L[b] + X2[b]/24 and S[b] = I + h L[b], X2 = (h L)^2:
// with P = S + X2 Q the two-product form of P4(h L)
__global__ __launch_bounds__(kEwThreads) void p4_terms_kernel(
    const double2* __restrict__ L, int64_t strideL, const double2* __restrict__ X2, int M, int B,
    double h, double2* __restrict__ Q, double2* __restrict__ S) {
  const int64_t mm = (int64_t)M * M, total = mm * B;
  for (int64_t e = (int64_t)blockIdx.x * kEwThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kEwThreads) {
    const int64_t b = e / mm, r = e % mm;
    const double2 v = L[b * strideL + r], x2 = X2[e];
    const bool d = (r / M) == (r % M);
    Q[e] = make_double2((d ? 0.5 : 0.0) + (h / 6.0) * v.x + x2.x / 24.0,
                        (h / 6.0) * v.y + x2.y / 24.0);
    S[e] = make_double2(fma(h, v.x, d ? 1.0 : 0.0), h * v.y);
  }
}

// dst[b] = src[b] (M x M blocks with batch strides; sSrc = 0: shared source)
__global__ __launch_bounds__(kEwThreads) void copy_blocks_kernel(
    const double2* __restrict__ src, int64_t sSrc, double2* __restrict__ dst, int64_t sDst,
    int M, int B) {
  const int64_t mm = (int64_t)M * M, total = mm * B;
  for (int64_t e = (int64_t)blockIdx.x * kEwThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kEwThreads) {
    const int64_t b = e / mm, r = e % mm;
    dst[b * sDst + r] = src[b * sSrc + r];
  }
}

// dst[b] = I
__global__ __launch_bounds__(kEwThreads) void identity_blocks_kernel(double2* __restrict__ dst,
                                                                     int64_t sDst, int M, int B) {
  const int64_t mm = (int64_t)M * M, total = mm * B;
  for (int64_t e = (int64_t)blockIdx.x * kEwThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kEwThreads) {
    const int64_t b = e / mm, r = e % mm;
    dst[b * sDst + r] = make_double2((r / M) == (r % M) ? 1.0 : 0.0, 0.0);
  }
}

// off[b] = P[b] with a zero diagonal, as complex64; diag[b][i] = P[b][i][i] - 1
__global__ __launch_bounds__(kEwThreads) void split_increment_kernel(
    const double2* __restrict__ P, int M, int B, float2* __restrict__ off,
    double2* __restrict__ diag) {
  const int64_t mm = (int64_t)M * M, total = mm * B;
  for (int64_t e = (int64_t)blockIdx.x * kEwThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kEwThreads) {
    const int64_t b = e / mm, r = e % mm;
    const int i = (int)(r / M), j = (int)(r % M);
    const double2 v = P[e];
    if (i == j) {
      off[e] = make_float2(0.f, 0.f);
      diag[b * M + i] = make_double2(v.x - 1.0, v.y);
    } else {
      off[e] = make_float2((float)v.x, (float)v.y);
    }
  }
}

#define QSX_TRY(expr)              \
  do {                             \
    const int st_ = (expr);        \
    if (st_ != QSX_OK) return st_; \
  } while (0)

int enqueue_copy(const double2* src, int64_t sSrc, double2* dst, int64_t sDst, int M, int B,
                 hipStream_t s) {
  hipLaunchKernelGGL(copy_blocks_kernel, ew_grid((int64_t)M * M * B), dim3(kEwThreads), 0, s,
                     src, sSrc, dst, sDst, M, B);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

// C[b] = A[b] B[b] (complex128, M x M blocks, batch strides; sC nonzero)
int enqueue_mul(const double2* A, int64_t sA, const double2* Bm, int64_t sB, double2* C,
                int64_t sC, int M, int B, hipStream_t s) {
  return zgemm_enqueue(QSX_FP64, 0, M, M, M, 1.0, 0.0, A, M, sA, Bm, M, sB, 0.0, 0.0, C, M,
                       B > 1 ? sC : 0, 0.0, 0.0, B, s);
}

// out_s = A^{e_s} for s < ns (e_s >= 0), out_s at out + s * sOutS with batch
// stride sOutB.  Squarings ping-pong between S0 and S1; T holds a product
// before it is copied to its destination.  A power of two that no other bit
// of its exponent joins is written by its squaring directly.  None of S0,
// S1, T may alias A or the outputs.
int powers_impl(const double2* A, int64_t sA, int M, int B, const int64_t* e, int ns,
                double2* out, int64_t sOutS, int64_t sOutB, double2* S0, double2* S1, double2* T,
                hipStream_t s) {
  const int64_t mm = (int64_t)M * M;
  int64_t emax = 0;
  for (int i = 0; i < ns; ++i) emax = e[i] > emax ? e[i] : emax;
  std::vector<char> done(ns, 0);
  const double2* cur = A;
  int64_t scur = sA;
  double2* sq[2] = {S0, S1};
  int ping = 0;
  for (int k = 0; k < 63 && ((int64_t)1 << k) <= emax; ++k) {
    const int64_t bit = (int64_t)1 << k;
    for (int i = 0; i < ns; ++i) {
      if (!(e[i] & bit) || done[i] == 2) continue;
      double2* dst = out + i * sOutS;
      if (!done[i]) {
        QSX_TRY(enqueue_copy(cur, scur, dst, sOutB, M, B, s));
      } else {
        QSX_TRY(enqueue_mul(cur, scur, dst, sOutB, T, mm, M, B, s));
        QSX_TRY(enqueue_copy(T, mm, dst, sOutB, M, B, s));
      }
      done[i] = 1;
    }
    if ((bit << 1) > emax) break;
    // the next square goes straight to an output that is exactly that power
    int direct = -1;
    for (int i = 0; i < ns && direct < 0; ++i)
      if (!done[i] && e[i] == (bit << 1)) direct = i;
    if (direct >= 0) {
      double2* dst = out + direct * sOutS;
      QSX_TRY(enqueue_mul(cur, scur, cur, scur, dst, sOutB, M, B, s));
      for (int i = 0; i < ns; ++i)            // equal exponents share the result
        if (!done[i] && i != direct && e[i] == (bit << 1))
          QSX_TRY(enqueue_copy(dst, sOutB, out + i * sOutS, sOutB, M, B, s));
      for (int i = 0; i < ns; ++i)
        if (e[i] == (bit << 1)) done[i] = 2;
      cur = dst;
      scur = sOutB;
    } else {
      QSX_TRY(enqueue_mul(cur, scur, cur, scur, sq[ping], mm, M, B, s));
      cur = sq[ping];
      scur = mm;
      ping ^= 1;
    }
  }
  for (int i = 0; i < ns; ++i) {
    if (e[i] != 0) continue;
    hipLaunchKernelGGL(identity_blocks_kernel, ew_grid(mm * B), dim3(kEwThreads), 0, s,
                       out + i * sOutS, sOutB, M, B);
    QSX_CHECK_HIP(hipGetLastError());
  }
  return QSX_OK;
}

// ----------------------------------------------------------------------
// Small blocks (M <= 64): one work-group per batch entry builds the whole
// operator program in LDS -- Horner, the squarings of the powers, the split
// -- in one launch.  Two M x M complex128 matrices live in LDS (row stride
// M + 1); a product C = A B is computed by a T x T thread grid (T =
// ceil(M / 2) <= 32 of the 1024 threads), thread (tx, ty) owning rows
// ty + T r and columns tx + T c (r, c < 2), its 2 x 2 partial sums in
// registers (16 waves: four per SIMD hide the LDS latency); the result is
// written back to LDS after a barrier (so it may overwrite an operand).
constexpr int kSmallThreads = 1024;
constexpr int kSmallMaxM = 64;
constexpr int kMaxExps = 8;
constexpr int RB = 2;

struct ExpList {
  long long e[kMaxExps];
  int n;
};

struct SmallGemm {
  int M, ld, T, tx, ty;
  bool active;
  double2 acc[RB][RB];

  __device__ explicit SmallGemm(int M_) : M(M_), ld(M_ + 1), T((M_ + RB - 1) / RB) {
    tx = threadIdx.x % T;
    ty = threadIdx.x / T;
    active = ty < T;
  }

  // acc = A B (A, B in LDS, row stride ld; rows/cols beyond M read a
  // clamped index and are never stored)
  __device__ void mul(const double2* A, const double2* B) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int c = 0; c < RB; ++c) acc[r][c] = make_double2(0.0, 0.0);
    if (!active) return;
    int ia[RB], jb[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) ia[r] = min(ty + T * r, M - 1) * ld;
#pragma unroll
    for (int c = 0; c < RB; ++c) jb[c] = min(tx + T * c, M - 1);
    for (int k = 0; k < M; ++k) {
      double2 a[RB], b[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) a[r] = A[ia[r] + k];
#pragma unroll
      for (int c = 0; c < RB; ++c) b[c] = B[k * ld + jb[c]];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < RB; ++c) cfma<double>(acc[r][c], a[r], b[c]);
    }
  }
  // D = alpha acc + gamma I
  __device__ void store(double2* D, double alpha, double gamma) const {
    if (!active) return;
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int c = 0; c < RB; ++c) {
        const int i = ty + T * r, j = tx + T * c;
        if (i < M && j < M)
          D[i * ld + j] = make_double2(fma(alpha, acc[r][c].x, i == j ? gamma : 0.0),
                                       alpha * acc[r][c].y);
      }
  }
};

__device__ inline void lds_copy(const double2* src, double2* dst, int M) {
  const int ld = M + 1;
  for (int e = threadIdx.x; e < M * M; e += blockDim.x)
    dst[(e / M) * ld + e % M] = src[(e / M) * ld + e % M];
}

// P4(h L) with two dependent products instead of Horner's three: X = h L,
// X2 = X X, Q = I/2 + X/6 + X2/24, P = I + X + X2 Q (the same polynomial).
// X = L is read from Lb; P is left in Y, X is free on return.
__device__ void small_p4_two_products(const double2* __restrict__ Lb, double h, int M, double2* X,
                                      double2* Y, SmallGemm& g) {
  const int ld = M + 1;
  for (int e = threadIdx.x; e < M * M; e += blockDim.x) {
    const int i = e / M, j = e % M;
    const double2 v = Lb[e];
    X[i * ld + j] = make_double2(h * v.x, h * v.y);
  }
  __syncthreads();
  // the entries of X this thread's product block owns, for the final sum
  double2 xo[RB][RB];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < RB; ++c) {
      const int i = min(g.ty + g.T * r, M - 1), j = min(g.tx + g.T * c, M - 1);
      xo[r][c] = g.active ? X[i * ld + j] : make_double2(0.0, 0.0);
    }
  g.mul(X, X);
  __syncthreads();
  g.store(Y, 1.0, 0.0);                          // Y = X2
  __syncthreads();
  for (int e = threadIdx.x; e < M * M; e += blockDim.x) {
    const int i = e / M, j = e % M;
    const double2 x = X[i * ld + j], x2 = Y[i * ld + j];
    X[i * ld + j] = make_double2((i == j ? 0.5 : 0.0) + x.x / 6.0 + x2.x / 24.0,
                                 x.y / 6.0 + x2.y / 24.0);     // X = Q
  }
  __syncthreads();
  g.mul(Y, X);
  __syncthreads();
  if (g.active) {
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int c = 0; c < RB; ++c) {
        const int i = g.ty + g.T * r, j = g.tx + g.T * c;
        if (i < M && j < M)
          Y[i * ld + j] = make_double2((i == j ? 1.0 : 0.0) + xo[r][c].x + g.acc[r][c].x,
                                       xo[r][c].y + g.acc[r][c].y);
      }
  }
  __syncthreads();
}

// S (in LDS) raised to the power q: the result is left in S or in R (the
// returned pointer); R and S are both overwritten.  q >= 1.
__device__ double2* small_power(double2* S, double2* R, long long q, int M) {
  SmallGemm g(M);
  bool have = false;
  for (int k = 0; k < 62 && (1LL << k) <= q; ++k) {
    if (q & (1LL << k)) {
      if (!have) {
        lds_copy(S, R, M);
        have = true;
      } else {
        g.mul(R, S);
        __syncthreads();
        g.store(R, 1.0, 0.0);
      }
      __syncthreads();
    }
    if ((2LL << k) <= q) {
      g.mul(S, S);
      __syncthreads();
      g.store(S, 1.0, 0.0);
      __syncthreads();
    }
  }
  return R;
}

// P = P4(h L)^q (optionally stored), Pc = P^c (c > 0, optional), MIXED split
// of P (optional): one work-group per batch entry.
__global__ __launch_bounds__(kSmallThreads) void small_program_kernel(
    const double2* __restrict__ L, int64_t strideL, int M, double h, long long q, long long c,
    double2* __restrict__ P, double2* __restrict__ Pc, float2* __restrict__ off,
    double2* __restrict__ diag) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ld = M + 1;
  double2* X = reinterpret_cast<double2*>(smem);
  double2* Y = X + M * ld;
  const int b = blockIdx.x;
  const int64_t mm = (int64_t)M * M;
  SmallGemm g(M);
  small_p4_two_products(L + b * strideL, h, M, X, Y, g);   // Y = P4(h L)
  double2* Pq = q == 1 ? Y : small_power(Y, X, q, M);
  double2* other = Pq == Y ? X : Y;
  for (int e = threadIdx.x; e < M * M; e += blockDim.x) {
    const int i = e / M, j = e % M;
    const double2 v = Pq[i * ld + j];
    if (P != nullptr) P[b * mm + e] = v;
    if (off != nullptr) {
      off[b * mm + e] = i == j ? make_float2(0.f, 0.f) : make_float2((float)v.x, (float)v.y);
      if (i == j) diag[(int64_t)b * M + i] = make_double2(v.x - 1.0, v.y);
    }
  }
  if (Pc == nullptr || c < 1) return;
  __syncthreads();
  double2* Rc = small_power(Pq, other, c, M);
  for (int e = threadIdx.x; e < M * M; e += blockDim.x)
    Pc[b * mm + e] = Rc[(e / M) * ld + e % M];
}

// States y0 [B][R][M] propagated by P = P4(h L) to the cumulative step
// counts ex.e[s]: out[b][s][r] = P^e_s y0[b][r], one work-group per batch
// entry; the squarings P^(2^k) are shared by all save points (binary
// powering), each save point's states kept in LDS.  With `merged`, the
// state products of every save point whose count has bit k run on the
// threads the product grid leaves idle, beside the squaring that reads the
// same P^(2^k) (staging for all save points); otherwise after it, one save
// point at a time.
__global__ __launch_bounds__(kSmallThreads) void small_powers_apply_kernel(
    const double2* __restrict__ L, int64_t strideL, int M, double h, ExpList ex,
    const double2* __restrict__ y0, int R, double2* __restrict__ out, int merged) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ld = M + 1;
  double2* X = reinterpret_cast<double2*>(smem);
  double2* Y = X + M * ld;
  const int b = blockIdx.x, ns = ex.n;
  const int RM = R * M;
  double2* Z = Y + M * ld;                       // [ns][R][M] states
  double2* W = Z + ns * RM;                      // products before copy-back
  long long emax = 0;
  for (int s = 0; s < ns; ++s) emax = ex.e[s] > emax ? ex.e[s] : emax;
  for (int e = threadIdx.x; e < ns * RM; e += blockDim.x)
    Z[e] = y0[(int64_t)b * RM + e % RM];
  SmallGemm g(M);
  small_p4_two_products(L + b * strideL, h, M, X, Y, g);    // Y = P
  const int gthreads = ((g.T * g.T + 63) / 64) * 64;
  const bool side = merged && gthreads + 64 <= (int)blockDim.x;
  for (int k = 0; k < 62 && (1LL << k) <= emax; ++k) {
    const bool sq = (2LL << k) <= emax;
    if (side) {
      // the squaring's products on the grid threads; the state products
      // W_s = P^(2^k) Z_s on the others, all from the same Y
      if (sq) g.mul(Y, Y);
      const int t = (int)threadIdx.x - gthreads, nt = (int)blockDim.x - gthreads;
      if (t >= 0) {
        for (int s = 0; s < ns; ++s) {
          if (!(ex.e[s] & (1LL << k))) continue;
          const double2* Zs = Z + s * RM;
          double2* Ws = W + s * RM;
          for (int e = t; e < RM; e += nt) {
            const int r = e / M, i = e % M;
            double2 acc = make_double2(0.0, 0.0);
            for (int j = 0; j < M; ++j) cfma<double>(acc, Y[i * ld + j], Zs[r * M + j]);
            Ws[e] = acc;
          }
        }
      }
      __syncthreads();
      if (sq) g.store(Y, 1.0, 0.0);
      for (int s = 0; s < ns; ++s) {
        if (!(ex.e[s] & (1LL << k))) continue;
        for (int e = threadIdx.x; e < RM; e += blockDim.x) Z[s * RM + e] = W[s * RM + e];
      }
      __syncthreads();
      continue;
    }
    // states of every save point whose step count has bit k: v <- P^(2^k) v
    // (products staged in W, then copied back)
    for (int s = 0; s < ns; ++s) {
      if (!(ex.e[s] & (1LL << k))) continue;
      double2* Zs = Z + s * RM;
      for (int e = threadIdx.x; e < RM; e += blockDim.x) {
        const int r = e / M, i = e % M;
        double2 acc = make_double2(0.0, 0.0);
        for (int j = 0; j < M; ++j) cfma<double>(acc, Y[i * ld + j], Zs[r * M + j]);
        W[e] = acc;
      }
      __syncthreads();
      for (int e = threadIdx.x; e < RM; e += blockDim.x) Zs[e] = W[e];
      __syncthreads();
    }
    if (sq) {
      g.mul(Y, Y);
      __syncthreads();
      g.store(Y, 1.0, 0.0);
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < ns * RM; e += blockDim.x)
    out[(int64_t)b * ns * RM + e] = Z[e];
}

size_t small_program_lds(int M) { return (size_t)2 * M * (M + 1) * sizeof(double2); }

// threads of a small-block work-group: the T x T product grid (T =
// ceil(M / 2)) rounded up to whole wavefronts, at most kSmallThreads
int small_threads(int M) {
  const int T = (M + RB - 1) / RB;
  const int t = ((T * T + 63) / 64) * 64;
  return t < kSmallThreads ? t : kSmallThreads;
}
size_t small_apply_lds(int M, int R, int ns) {
  return small_program_lds(M) + (size_t)(ns + 1) * R * M * sizeof(double2);
}
// with the state products beside the squarings: staging for every save point
size_t small_apply_merged_lds(int M, int R, int ns) {
  return small_program_lds(M) + (size_t)2 * ns * R * M * sizeof(double2);
}

template <typename K>
int set_lds(K kern, size_t lds) {
  if (lds > 64 * 1024)
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return QSX_OK;
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_small_propagator_fits(int32_t M, int32_t R, int32_t ns) {
  using namespace qsx;
  if (M < 1 || M > kSmallMaxM || R < 0 || ns < 0 || ns > kMaxExps) return 0;
  const size_t lds = R > 0 ? small_apply_lds(M, R, ns) : small_program_lds(M);
  return lds <= 160 * 1024 && (R == 0 || R * M <= (1 << 20)) ? 1 : 0;
}

extern "C" int32_t qsx_small_propagator(const void* L, int64_t strideL, int32_t M, int32_t B,
                                        double h, int32_t q, int64_t c, void* P, void* Pc,
                                        void* off, void* diag, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(L, "qsx_small_propagator: null operator");
  QSX_CHECK_ARG(M >= 1 && M <= kSmallMaxM && B >= 1 && B <= (1 << 30) && q >= 1 && c >= 0,
                "qsx_small_propagator: bad sizes M=%d B=%d q=%d c=%lld", M, B, q, (long long)c);
  QSX_CHECK_ARG(strideL == 0 || strideL >= (int64_t)M * M, "qsx_small_propagator: bad strideL");
  QSX_CHECK_ARG((off == nullptr) == (diag == nullptr),
                "qsx_small_propagator: the split needs both off and diag");
  QSX_CHECK_ARG(c == 0 || Pc != nullptr, "qsx_small_propagator: power c without output");
  const size_t lds = small_program_lds(M);
  hipStream_t s = as_stream(stream);
  const int64_t sL = B > 1 ? strideL : 0;
auto kern = small_program_kernel;
  QSX_TRY(set_lds(kern, lds));
  hipLaunchKernelGGL(kern, dim3(B), dim3(small_threads(M)), lds, s,
                     static_cast<const double2*>(L), sL, M, h, (long long)q, (long long)c, static_cast<double2*>(P),
                     static_cast<double2*>(Pc), static_cast<float2*>(off),
                     static_cast<double2*>(diag));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_propagate_powers(const void* L, int64_t strideL, int32_t M, int32_t B,
                                        double h, const int64_t* exps, int32_t ns,
                                        const void* y0, int32_t R, void* out, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(L && exps && y0 && out, "qsx_propagate_powers: null pointer");
  QSX_CHECK_ARG(M >= 1 && B >= 1 && B <= (1 << 30) && R >= 1 && ns >= 1,
                "qsx_propagate_powers: bad sizes M=%d B=%d R=%d ns=%d", M, B, R, ns);
  QSX_CHECK_ARG(qsx_small_propagator_fits(M, R, ns),
                "qsx_propagate_powers: M=%d, R=%d, %d save points exceed one work-group's LDS",
                M, R, ns);
  QSX_CHECK_ARG(strideL == 0 || strideL >= (int64_t)M * M, "qsx_propagate_powers: bad strideL");
  ExpList ex{};
  ex.n = ns;
  for (int i = 0; i < ns; ++i) {
    QSX_CHECK_ARG(exps[i] >= 0 && exps[i] < ((int64_t)1 << 62),
                  "qsx_propagate_powers: bad step count %lld", (long long)exps[i]);
    ex.e[i] = exps[i];
  }
  // the state products beside the squarings (on up to six more wavefronts)
  // for blocks whose squarings are long enough to hide them (M >= 16), when
  // their staging fits and the product grid leaves a wavefront of the 1024
  // threads
  const int T = (M + RB - 1) / RB;
  const int gthreads = ((T * T + 63) / 64) * 64;
  const int merged = M >= 16 && small_apply_merged_lds(M, R, ns) <= 160 * 1024 &&
                     gthreads + 64 <= kSmallThreads;
  const int threads = merged ? std::min(kSmallThreads, gthreads + 384) : small_threads(M);
  const size_t lds = merged ? small_apply_merged_lds(M, R, ns) : small_apply_lds(M, R, ns);
  hipStream_t s = as_stream(stream);
  const int64_t sL = B > 1 ? strideL : 0;
  auto kern = small_powers_apply_kernel;
  QSX_TRY(set_lds(kern, lds));
  hipLaunchKernelGGL(kern, dim3(B), dim3(threads), lds, s,
                     static_cast<const double2*>(L), sL, M, h, ex, static_cast<const double2*>(y0),
                     R, static_cast<double2*>(out), merged);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int64_t qsx_propagator_workspace_bytes(int32_t M, int32_t B) {
  if (M < 1 || B < 1) return -1;
  return (int64_t)4 * B * M * M * (int64_t)sizeof(double2);
}

namespace qsx {
namespace {

// P = P4(h L)^q with the 4 B M^2 complex128 workspace ws (arguments checked)
int rk4_propagator_impl(const void* L, int64_t strideL, int M, int B, double h, int q, void* P,
                        void* workspace, hipStream_t s) {
  const int64_t mm = (int64_t)M * M;
  double2* ws = static_cast<double2*>(workspace);
  double2 *Ta = ws, *Tb = ws + mm * B, *Tc = ws + 2 * mm * B, *Td = ws + 3 * mm * B;
  const double2* Lp = static_cast<const double2*>(L);
  double2* Pp = static_cast<double2*>(P);
  const int64_t sL = B > 1 ? strideL : 0;
  const int64_t sT = mm;
  // two dependent products: Ta = X2 = h^2 L L;  Tc = Q = I/2 + X/6 + X2/24,
  // P' = I + X;  P' += X2 Q  (P4(hL) = I + X + X2 Q, X = h L)
  double2* P1 = q == 1 ? Pp : Tb;
  QSX_TRY(zgemm_enqueue(QSX_FP64, 0, M, M, M, h * h, 0.0, Lp, M, sL, Lp, M, sL, 0.0, 0.0, Ta, M,
                        sT, 0.0, 0.0, B, s));
  hipLaunchKernelGGL(p4_terms_kernel, ew_grid(mm * B), dim3(kEwThreads), 0, s, Lp, sL, Ta, M, B,
                     h, Tc, P1);
  QSX_CHECK_HIP(hipGetLastError());
  QSX_TRY(zgemm_enqueue(QSX_FP64, 0, M, M, M, 1.0, 0.0, Ta, M, sT, Tc, M, sT, 1.0, 0.0, P1, M,
                        sT, 0.0, 0.0, B, s));
  if (q == 1) return QSX_OK;
  const int64_t e = q;
  return powers_impl(Tb, mm, M, B, &e, 1, Pp, 0, mm, Ta, Tc, Td, s);
}

int check_rk4_args(const char* who, const void* L, int64_t strideL, int M, int B, double h,
                   int q, const void* P) {
  QSX_CHECK_ARG(L && P, "%s: null pointer", who);
  QSX_CHECK_ARG(M >= 1 && B >= 1 && B <= 65535 && q >= 1, "%s: bad sizes M=%d B=%d q=%d", who,
                M, B, q);
  QSX_CHECK_ARG(strideL == 0 || strideL >= (int64_t)M * M, "%s: bad strideL", who);
  QSX_CHECK_ARG(h == h && h > -1e300 && h < 1e300, "%s: bad step %g", who, h);
  return QSX_OK;
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_rk4_propagator(const void* L, int64_t strideL, int32_t M, int32_t B,
                                      double h, int32_t q, void* P, void* workspace,
                                      int64_t workspace_bytes, void* stream) {
  using namespace qsx;
  QSX_TRY(check_rk4_args("qsx_rk4_propagator", L, strideL, M, B, h, q, P));
  QSX_CHECK_ARG(workspace != nullptr, "qsx_rk4_propagator: null workspace");
  QSX_CHECK_ARG(workspace_bytes >= qsx_propagator_workspace_bytes(M, B),
                "qsx_rk4_propagator: workspace of %lld bytes, %lld needed",
                (long long)workspace_bytes, (long long)qsx_propagator_workspace_bytes(M, B));
  return rk4_propagator_impl(L, strideL, M, B, h, q, P, workspace, as_stream(stream));
}

extern "C" int32_t qsx_matrix_powers(const void* A, int64_t strideA, int32_t M, int32_t B,
                                     const int64_t* exps_host, int32_t ns, void* out,
                                     void* workspace, int64_t workspace_bytes, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(A && exps_host && out && workspace, "qsx_matrix_powers: null pointer");
  QSX_CHECK_ARG(M >= 1 && B >= 1 && B <= 65535 && ns >= 1,
                "qsx_matrix_powers: bad sizes M=%d B=%d ns=%d", M, B, ns);
  QSX_CHECK_ARG(strideA == 0 || strideA >= (int64_t)M * M, "qsx_matrix_powers: bad strideA");
  for (int i = 0; i < ns; ++i)
    QSX_CHECK_ARG(exps_host[i] >= 0 && exps_host[i] < ((int64_t)1 << 62),
                  "qsx_matrix_powers: bad exponent %lld", (long long)exps_host[i]);
  QSX_CHECK_ARG(workspace_bytes >= qsx_propagator_workspace_bytes(M, B),
                "qsx_matrix_powers: workspace of %lld bytes, %lld needed",
                (long long)workspace_bytes, (long long)qsx_propagator_workspace_bytes(M, B));
  const int64_t mm = (int64_t)M * M;
  double2* ws = static_cast<double2*>(workspace);
  return powers_impl(static_cast<const double2*>(A), B > 1 ? strideA : 0, M, B, exps_host, ns,
                     static_cast<double2*>(out), mm, (int64_t)ns * mm, ws, ws + mm * B,
                     ws + 2 * mm * B, as_stream(stream));
}

extern "C" int32_t qsx_split_increment(const void* P, int32_t M, int32_t B, void* off,
                                       void* diag, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(P && off && diag, "qsx_split_increment: null pointer");
  QSX_CHECK_ARG(M >= 1 && B >= 1, "qsx_split_increment: bad sizes M=%d B=%d", M, B);
  hipLaunchKernelGGL(split_increment_kernel, ew_grid((int64_t)M * M * B), dim3(kEwThreads), 0,
                     as_stream(stream), static_cast<const double2*>(P), M, B,
                     static_cast<float2*>(off), static_cast<double2*>(diag));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

// ----------------------------------------------------------------------
// Allocation-free qsx_rk4_propagator for callers that cannot allocate
// (code under graph capture, serving loops): the workspace is `scratch`.
namespace qsx {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device
// (with that device current) and kept.
int scratch_address(void** out) {
  int dev = -1;
  QSX_CHECK_HIP(hipGetDevice(&dev));
  QSX_CHECK_ARG(dev >= 0 && dev < kMaxDevices,
                "qsx_rk4_propagator_noalloc: device ordinal %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_address[dev] == nullptr) {
    void* p = nullptr;
    QSX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QSX_CHECK_ARG(p != nullptr, "qsx_rk4_propagator_noalloc: no device address for the scratch "
                  "array on device %d", dev);
    g_scratch_address[dev] = p;
  }
  *out = g_scratch_address[dev];
  return QSX_OK;
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_rk4_propagator_noalloc(const void* L, int64_t strideL, int32_t M,
                                              int32_t B, double h, int32_t q, void* P,
                                              void* stream) {
  using namespace qsx;
  QSX_TRY(check_rk4_args("qsx_rk4_propagator_noalloc", L, strideL, M, B, h, q, P));
  const int64_t need = qsx_propagator_workspace_bytes(M, B);
  QSX_CHECK_ARG(need <= (int64_t)QSX_PROPAGATOR_SCRATCH_BYTES,
                "qsx_rk4_propagator_noalloc: M=%d, B=%d needs %lld bytes of workspace; the "
                "scratch array holds %lld (use qsx_rk4_propagator with a workspace)",
                M, B, (long long)need, (long long)QSX_PROPAGATOR_SCRATCH_BYTES);
  void* ws = nullptr;
  QSX_TRY(scratch_address(&ws));
  // every workspace element is written before it is read: nothing to clear
  return rk4_propagator_impl(L, strideL, M, B, h, q, P, ws, as_stream(stream));
}
