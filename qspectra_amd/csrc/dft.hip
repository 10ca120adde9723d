// The Fourier transform of qspectra/simulate/utils.py:128-219 on the GPU, as
// a direct sum (gfx950).
//
// The reference zero-pads a signal of n samples to a grid of N points
// symmetric about t = 0 (the samples at positions lo .. lo+n-1), scales by
// dt, and applies ifftshift, FFT, fftshift (and a reversal for sign = +1).
// Written out, output r is
//
//   X[r] = dt * sum_j x[j] * exp(-2 pi i s_r (lo + j - c) / N),   c = N / 2,
//   s_r  = r - c            (sign = -1)
//   s_r  = N - 1 - c - r    (sign = +1).
//
// Only the n nonzero samples enter, so the direct sum costs N*n complex
// multiply-adds, with no plan and nothing compiled at run time (a library FFT
// of a new length is compiled on first use).  The exponent is reduced exactly
// in integers, m = s_r * (lo + j - c) mod N; every 16th twiddle is read from
// a table exp(-2 pi i m / N) built once per call, the ones in between are
// advanced by one complex multiplication with the step twiddle.  The adjoint (the backward of the
// transform, x -> conj(F)^T X) is the same sum over r with conjugated
// twiddles.
//
// Layout: one thread per output index, 256 threads per work-group, the input
// row staged through LDS in tiles (every thread reads the same sample:
// broadcast); grid.y walks the batch rows.

#include "qsx_common.h"

#include <algorithm>

namespace qsx {
namespace {

constexpr int kDftThreads = 256;
constexpr int kDftTile = 1024;
constexpr int kAnchor = 16;   // samples between exact twiddles (divides kDftTile)

__global__ __launch_bounds__(kDftThreads) void dft_twiddle_kernel(int N, double2* __restrict__ T) {
  const int m = blockIdx.x * kDftThreads + threadIdx.x;
  if (m >= N) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)N, &s, &c);
  T[m] = make_double2(c, -s);
}

__device__ inline int64_t mod_n(int64_t a, int64_t N) {
  const int64_t r = a % N;
  return r < 0 ? r + N : r;
}

// out[b, u] = scale * sum_v in[b, v] * T[(a_u * d_v) mod N] (conj if adj),
// a_u = a0 + da * u, d_v = d0 + dd * v
__global__ __launch_bounds__(kDftThreads) void dft_kernel(
    const double2* __restrict__ in, int64_t batch, int n_in, int n_out, int N, int64_t a0,
    int da, int64_t d0, int dd, int adj, double scale, const double2* __restrict__ T,
    double2* __restrict__ out) {
  __shared__ double2 xs[kDftTile];
  const int u = blockIdx.x * kDftThreads + threadIdx.x;
  const int64_t a = mod_n(a0 + (int64_t)da * u, N);
  const int64_t step = mod_n(a * dd, N);
  const int64_t anchor_step = mod_n(step * kAnchor, N);
  for (int64_t b = blockIdx.y; b < batch; b += gridDim.y) {
    const double2* row = in + b * n_in;
    double2 acc = make_double2(0.0, 0.0);
    for (int v0 = 0; v0 < n_in; v0 += kDftTile) {
      const int len = min(kDftTile, n_in - v0);
      __syncthreads();
      for (int k = threadIdx.x; k < len; k += kDftThreads) xs[k] = row[v0 + k];
      __syncthreads();
      if (u < n_out) {
        // exact twiddle from the table every kAnchor samples, in between
        // multiplied by the step twiddle (error <= kAnchor rounding steps)
        int64_t m = mod_n(a * mod_n(d0 + (int64_t)dd * v0, N), N);
        const double sg = adj ? -1.0 : 1.0;
        const double2 w = make_double2(T[step].x, sg * T[step].y);
        for (int k0 = 0; k0 < len; k0 += kAnchor) {
          double2 t = make_double2(T[m].x, sg * T[m].y);
          const int kend = min(k0 + kAnchor, len);
          for (int k = k0; k < kend; ++k) {
            const double2 x = xs[k];
            acc.x = fma(x.x, t.x, acc.x);
            acc.x = fma(-x.y, t.y, acc.x);
            acc.y = fma(x.x, t.y, acc.y);
            acc.y = fma(x.y, t.x, acc.y);
            t = make_double2(fma(t.x, w.x, -t.y * w.y), fma(t.x, w.y, t.y * w.x));
          }
          m += anchor_step;
          if (m >= N) m -= N;
        }
      }
    }
    if (u < n_out) out[b * n_out + u] = make_double2(scale * acc.x, scale * acc.y);
  }
}

}  // namespace
}  // namespace qsx

using namespace qsx;

extern "C" int64_t qsx_fourier_workspace_bytes(int32_t N) {
  if (N < 1) return -1;
  return (int64_t)N * (int64_t)sizeof(double2);
}

extern "C" int32_t qsx_fourier_transform(const void* x, int64_t batch, int32_t n, int32_t N,
                                         int32_t lo, int32_t sign, double dt, int32_t adjoint,
                                         void* workspace, void* out, void* stream) {
  QSX_CHECK_ARG(x && out && workspace, "qsx_fourier_transform: null pointer");
  QSX_CHECK_ARG(n >= 1 && N >= n && lo >= 0 && lo + n <= N && batch >= 1,
                "qsx_fourier_transform: bad sizes n=%d N=%d lo=%d batch=%lld", n, N, lo,
                (long long)batch);
  QSX_CHECK_ARG(sign == 1 || sign == -1, "qsx_fourier_transform: sign must be +1 or -1");
  hipStream_t s = as_stream(stream);
  double2* T = static_cast<double2*>(workspace);
  hipLaunchKernelGGL(dft_twiddle_kernel, dim3((N + kDftThreads - 1) / kDftThreads), dim3(kDftThreads),
                     0, s, N, T);
  const int64_t c = N / 2;
  // output index r -> s_r = a0 + da * r; sample index j -> lo + j - c
  const int64_t sa0 = sign == -1 ? -c : (int64_t)N - 1 - c;
  const int sda = sign == -1 ? 1 : -1;
  const int64_t ja0 = (int64_t)lo - c;
  const int n_out = adjoint ? n : N, n_in = adjoint ? N : n;
  const int64_t a0 = adjoint ? ja0 : sa0, d0 = adjoint ? sa0 : ja0;
  const int da = adjoint ? 1 : sda, dd = adjoint ? sda : 1;
  const unsigned gy = (unsigned)std::min<int64_t>(batch, 65535);
  hipLaunchKernelGGL(dft_kernel, dim3((n_out + kDftThreads - 1) / kDftThreads, gy),
                     dim3(kDftThreads), 0, s, static_cast<const double2*>(x), batch, n_in,
                     n_out, N, a0, da, d0, dd, (int)(adjoint != 0), dt, T,
                     static_cast<double2*>(out));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}
