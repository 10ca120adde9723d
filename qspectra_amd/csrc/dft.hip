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
//
// Long transforms (both lengths >= kChirpMin) take Bluestein's chirp-z form
// instead, O(N log N) per row.  Both directions are one sum
//
//   out[u] = scale * sum_v in[v] * exp(-2 pi i sigma (u + A)(v + D) / N)
//
// (sigma = +-1), and with pq = (p^2 + q^2 - (q - p)^2) / 2
//
//   out[u] = scale * c(u + A) * sum_v [in[v] c(v + D)] * conj(c(u - v + A - D)),
//   c(k)   = exp(-pi i sigma (k^2 mod 2N) / N)    (exact integer reduction),
//
// a linear convolution.  It is cut into input pieces of P samples and output
// chunks of H points with P + H - 1 <= L = 8192, so that every piece/chunk
// pair is one cyclic convolution of length L: FFT of the chirped piece (in
// LDS), times the FFT of the pair's chirp kernel (a table built once per call,
// it depends on the pair only), inverse FFT by conjugation.  A row of the
// absorption workload (4096 samples onto 8191 points) is one piece and two
// chunks: four FFTs of 8192 points on one row held in LDS, the row read twice
// (the second time from the cache) and its transform written once.  The FFT is a Stockham autosort FFT
// (one radix-2 pass, three radix-16 passes; Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC 2008)
// on 512 threads, 16 points per thread, the LDS rows padded by one entry in
// 16 against bank conflicts.

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


// ---------------------------------------------------------------------------
// chirp-z transform for long rows

constexpr int kFftL = 8192;                    // cyclic convolution length
constexpr int kFftThreads = 512;
constexpr int kFftPer = kFftL / kFftThreads;   // 16 points per thread
constexpr int kFftPadded = kFftL + kFftL / 16; // one pad entry per 16
constexpr int kTw1 = 128, kTw2 = kFftL / kTw1; // two-level twiddle table
constexpr size_t kFftLds = (size_t)(kFftPadded + kTw1 + kTw2) * sizeof(double2);
constexpr int kChirpMin = 2048;                // shorter transforms: direct sum
constexpr int kChirpMaxTables = 64;

struct ChirpGeom {
  int n_in, n_out;
  int P, H;              // input piece and output chunk lengths, P + H - 1 <= L
  int pieces, chunks;
  int sigma;
  int64_t N, A, D;
};

__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// a * (-i) and a * (+i)
__device__ __forceinline__ double2 mul_mi(double2 a) { return make_double2(a.y, -a.x); }
__device__ __forceinline__ double2 mul_pi(double2 a) { return make_double2(-a.y, a.x); }

// c(k) = exp(-pi i sigma (k^2 mod 2N) / N), the exponent reduced exactly
// in integers (64-bit divisions: evaluated once per call, into tables)
__device__ __forceinline__ double2 chirp_exact(int64_t k, int64_t N) {
  const int64_t N2 = 2 * N;
  int64_t m = k % N2;
  if (m < 0) m += N2;
  double s, c;
  sincospi((double)((m * m) % N2) / (double)N, &s, &c);
  return make_double2(c, -s);
}

// exp(-2 pi i k / L) = tw1[k mod 128] * tw2[k / 128]
__device__ __forceinline__ void build_fft_twiddles(double2* tw1, double2* tw2) {
  const int t = threadIdx.x;
  double s, c;
  if (t < kTw1) {
    sincospi(2.0 * (double)t / (double)kFftL, &s, &c);
    tw1[t] = make_double2(c, -s);
  } else if (t < kTw1 + kTw2) {
    sincospi(2.0 * (double)((t - kTw1) * kTw1) / (double)kFftL, &s, &c);
    tw2[t - kTw1] = make_double2(c, -s);
  }
}

// 4-point forward DFT in place: y_k = sum_m x_m (-i)^(mk)
__device__ __forceinline__ void dft4(double2& x0, double2& x1, double2& x2, double2& x3) {
  const double2 s02 = cadd(x0, x2), d02 = csub(x0, x2);
  const double2 s13 = cadd(x1, x3), d13 = mul_mi(csub(x1, x3));
  x0 = cadd(s02, s13);
  x2 = csub(s02, s13);
  x1 = cadd(d02, d13);
  x3 = csub(d02, d13);
}

// 16-point forward DFT of v (natural order in and out) as 4 x 4: DFTs over
// the stride-4 subsequences, twiddles exp(-2 pi i q k1 / 16), DFTs over q
__device__ __forceinline__ void dft16(double2 (&v)[16]) {
  constexpr double c1 = 0.92387953251128674;   // cos(pi / 8)
  constexpr double s1 = 0.38268343236508978;   // sin(pi / 8)
  constexpr double h = 0.70710678118654752;    // sqrt(1 / 2)
#pragma unroll
  for (int q = 0; q < 4; ++q) dft4(v[q], v[q + 4], v[q + 8], v[q + 12]);
  // v[q + 4 k1] now holds b[q][k1]; twiddle by w16^(q k1)
  v[5] = cmul(v[5], make_double2(c1, -s1));     // q 1, k1 1: e 1
  v[9] = cmul(v[9], make_double2(h, -h));       // q 1, k1 2: e 2
  v[13] = cmul(v[13], make_double2(s1, -c1));   // q 1, k1 3: e 3
  v[6] = cmul(v[6], make_double2(h, -h));       // q 2, k1 1: e 2
  v[10] = mul_mi(v[10]);                        // q 2, k1 2: e 4
  v[14] = cmul(v[14], make_double2(-h, -h));    // q 2, k1 3: e 6
  v[7] = cmul(v[7], make_double2(s1, -c1));     // q 3, k1 1: e 3
  v[11] = cmul(v[11], make_double2(-h, -h));    // q 3, k1 2: e 6
  v[15] = cmul(v[15], make_double2(-c1, s1));   // q 3, k1 3: e 9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
    dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
  // v[4 k1 + k2] holds X[k1 + 4 k2]: transpose the 4 x 4 index
  double2 t;
  t = v[1]; v[1] = v[4]; v[4] = t;
  t = v[2]; v[2] = v[8]; v[8] = t;
  t = v[3]; v[3] = v[12]; v[12] = t;
  t = v[6]; v[6] = v[9]; v[9] = t;
  t = v[7]; v[7] = v[13]; v[13] = t;
  t = v[11]; v[11] = v[14]; v[14] = t;
}

// The butterflies of one radix-16 Stockham pass over the padded LDS row:
// thread j reads entries j + r L/16 and leaves in v[r] the pass's outputs
// for entries (j - j mod Ns) 16 + j mod Ns + r Ns.
__device__ __forceinline__ void radix16_butterflies(const double2* buf, int Ns,
                                                    const double2* tw1, const double2* tw2,
                                                    double2 (&v)[16]) {
  const int j = threadIdx.x;   // butterfly index, kFftL / 16 of them
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = buf[fpad(j + r * (kFftL / 16))];
  const int jm = j & (Ns - 1);
  const int k1 = jm * (kFftL / (Ns * 16));
  const double2 w1 = cmul(tw1[k1 & (kTw1 - 1)], tw2[k1 / kTw1]);
  double2 w = w1;
#pragma unroll
  for (int r = 1; r < 16; ++r) {
    v[r] = cmul(v[r], w);
    w = cmul(w, w1);
  }
  dft16(v);
}

// One radix-16 Stockham pass, in place: every thread's reads finish before
// the barrier, its writes follow it.
__device__ __forceinline__ void radix16_pass(double2* buf, int Ns, const double2* tw1,
                                             const double2* tw2) {
  double2 v[16];
  radix16_butterflies(buf, Ns, tw1, tw2, v);
  const int j = threadIdx.x, jm = j & (Ns - 1);
  const int idx = (j - jm) * 16 + jm;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) buf[fpad(idx + r * Ns)] = v[r];
  __syncthreads();
}

// The last three passes of an FFT whose radix-2 pass is in LDS: the final
// pass (Ns = L / 16) leaves entry t + r L/16 of the transform in v[r] of
// thread t, the entries that thread read, so it needs no write-back.
__device__ __forceinline__ void fft_tail_to_registers(double2* buf, const double2* tw1,
                                                      const double2* tw2, double2 (&v)[16]) {
  radix16_pass(buf, 2, tw1, tw2);
  radix16_pass(buf, 32, tw1, tw2);
  radix16_butterflies(buf, kFftL / 16, tw1, tw2, v);
}

// Forward FFT of the L-point row in LDS (natural order in and out).  Expects
// the row's writes to have been made visible by a barrier.
__device__ void fft_row(double2* buf, const double2* tw1, const double2* tw2) {
  const int t = threadIdx.x;
  double2 v[16];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = t + q * kFftThreads;
    v[2 * q] = buf[fpad(j)];
    v[2 * q + 1] = buf[fpad(j + kFftL / 2)];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int j = t + q * kFftThreads;
    buf[fpad(2 * j)] = cadd(v[2 * q], v[2 * q + 1]);
    buf[fpad(2 * j + 1)] = csub(v[2 * q], v[2 * q + 1]);
  }
  __syncthreads();
  radix16_pass(buf, 2, tw1, tw2);
  radix16_pass(buf, 32, tw1, tw2);
  radix16_pass(buf, 512, tw1, tw2);
}

// The chirps of the input samples, cin[v] = c(v + D), and of the output
// points, cout[u] = c(u + A) (conjugated for sigma = -1): the same for
// every row.
__global__ __launch_bounds__(256) void chirp_vectors_kernel(ChirpGeom g, double2* __restrict__ cin,
                                                           double2* __restrict__ cout) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < g.n_in) {
    const double2 c = chirp_exact(j + g.D, g.N);
    cin[j] = make_double2(c.x, g.sigma > 0 ? c.y : -c.y);
  }
  if (j < g.n_out) {
    const double2 c = chirp_exact(j + g.A, g.N);
    cout[j] = make_double2(c.x, g.sigma > 0 ? c.y : -c.y);
  }
}

// The FFT of each piece/chunk pair's cyclic chirp kernel: G[k] =
// conj(c(delta + k + A - D)) for k < H and conj(c(delta - (L - k) + A - D))
// for k > L - P, delta = j0 - i0.  One work-group per pair.
__global__ __launch_bounds__(kFftThreads) void chirp_tables_kernel(ChirpGeom g,
                                                                  double2* __restrict__ Gt) {
  extern __shared__ double2 lds[];
  double2* buf = lds;
  double2* tw1 = lds + kFftPadded;
  double2* tw2 = tw1 + kTw1;
  build_fft_twiddles(tw1, tw2);
  const int id = blockIdx.x;
  const int j = id / g.pieces, i = id - j * g.pieces;
  const int64_t delta = (int64_t)j * g.H - (int64_t)i * g.P, dl = g.A - g.D;
  for (int r = 0; r < kFftPer; ++r) {
    const int k = threadIdx.x + r * kFftThreads;
    const int64_t e = k < g.H ? delta + k : delta - (kFftL - k);
    const double2 c = chirp_exact(e + dl, g.N);   // conj(c) for sigma = +1
    buf[fpad(k)] = make_double2(c.x, g.sigma > 0 ? -c.y : c.y);
  }
  __syncthreads();
  fft_row(buf, tw1, tw2);
  double2* G = Gt + (size_t)id * kFftL;
#pragma unroll
  for (int r = 0; r < kFftPer; ++r) {
    const int k = threadIdx.x + r * kFftThreads;
    G[k] = buf[fpad(k)];
  }
}

// Piece i of a row, chirped and zero-padded to L points, into LDS.
__device__ __forceinline__ void load_piece(double2* buf, const double2* __restrict__ row,
                                           const ChirpGeom& g, int i,
                                           const double2* __restrict__ cin) {
  const int i0 = i * g.P, pe = min(g.P, g.n_in - i0);
  // four points at a time: their sample and chirp loads issue together
#pragma unroll
  for (int r0 = 0; r0 < kFftPer; r0 += 4) {
    double2 x[4], c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int v = threadIdx.x + (r0 + q) * kFftThreads;
      x[q] = make_double2(0.0, 0.0);
      c[q] = make_double2(0.0, 0.0);
      if (v < pe) {
        x[q] = row[i0 + v];
        c[q] = cin[i0 + v];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      buf[fpad(threadIdx.x + (r0 + q) * kFftThreads)] = cmul(x[q], c[q]);
  }
  __syncthreads();
}

// One work-group per row (grid-stride over the batch): the chirp-z form of
// the transform.  Every piece/chunk pair is two FFTs in LDS with nothing
// held in registers across them (16 points per thread of the in-place passes
// already take most of the 256 registers a thread has here): the chirped
// piece, its product with the pair's kernel spectrum, the inverse FFT.  A
// chunk of several pieces adds their contributions in the output row.
__global__ __launch_bounds__(kFftThreads) void chirp_rows_kernel(
    const double2* __restrict__ in, int64_t batch, ChirpGeom g, const double2* __restrict__ cin,
    const double2* __restrict__ cout, const double2* __restrict__ Gt, double scale,
    double2* __restrict__ out) {
  extern __shared__ double2 lds[];
  double2* buf = lds;
  double2* tw1 = lds + kFftPadded;
  double2* tw2 = tw1 + kTw1;
  build_fft_twiddles(tw1, tw2);
  __syncthreads();
  const int t = threadIdx.x;
  const double oscale = scale / (double)kFftL;   // the inverse FFT's 1 / L
  for (int64_t b = blockIdx.x; b < batch; b += gridDim.x) {
    const double2* row = in + b * g.n_in;
    double2* orow = out + b * g.n_out;
    if (g.pieces == 1) {
      // one piece (P = L/2 samples, zero-padded to L): its spectrum, kept in
      // registers (entry t + r L/16 in S[r] of thread t), serves every
      // chunk.  The radix-2 pass of its FFT only duplicates each sample (the
      // upper half is zero), so the samples go straight to both outputs.
      __syncthreads();                 // the previous row's reads are done
#pragma unroll
      for (int r0 = 0; r0 < kFftPer / 2; r0 += 4) {
        double2 x[4], c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int v = t + (r0 + q) * kFftThreads;
          x[q] = make_double2(0.0, 0.0);
          c[q] = make_double2(0.0, 0.0);
          if (v < g.n_in) {
            x[q] = row[v];
            c[q] = cin[v];
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int v = t + (r0 + q) * kFftThreads;
          const double2 y = cmul(x[q], c[q]);
          buf[fpad(2 * v)] = y;
          buf[fpad(2 * v + 1)] = y;
        }
      }
      __syncthreads();
      double2 S[kFftPer];
      fft_tail_to_registers(buf, tw1, tw2, S);
      for (int j = 0; j < g.chunks; ++j) {
        const int j0 = j * g.H, he = min(g.H, g.n_out - j0);
        const double2* G = Gt + (size_t)j * kFftL;
        // the inverse FFT (conj(FFT(conj(.))) / L) of S times the kernel
        // spectrum, its radix-2 pass from registers: entries t + q L/16 and
        // t + (q + 8) L/16 are this thread's S[q], S[q + 8]
        __syncthreads();               // every thread's reads of buf are done
#pragma unroll
        for (int q = 0; q < kFftPer / 2; ++q) {
          const int jj = t + q * kFftThreads;
          const double2 a = cconj(cmul(S[q], G[jj]));
          const double2 b = cconj(cmul(S[q + kFftPer / 2], G[jj + kFftL / 2]));
          buf[fpad(2 * jj)] = cadd(a, b);
          buf[fpad(2 * jj + 1)] = csub(a, b);
        }
        __syncthreads();
        // the last pass reads and writes only this thread's entries t + r
        // L/16 (no barrier), which the output then reads back
        {
          double2 y[kFftPer];
          fft_tail_to_registers(buf, tw1, tw2, y);
#pragma unroll
          for (int r = 0; r < kFftPer; ++r) buf[fpad(t + r * kFftThreads)] = y[r];
        }
#pragma unroll
        for (int r = 0; r < kFftPer; ++r) {
          const int u = t + r * kFftThreads;
          if (u < he) {
            const double2 z = cmul(cout[j0 + u], cconj(buf[fpad(u)]));
            orow[j0 + u] = make_double2(oscale * z.x, oscale * z.y);
          }
        }
      }
      continue;
    }
    for (int j = 0; j < g.chunks; ++j) {
      const int j0 = j * g.H, he = min(g.H, g.n_out - j0);
      for (int i = 0; i < g.pieces; ++i) {
        load_piece(buf, row, g, i, cin);
        fft_row(buf, tw1, tw2);
        // times the kernel spectrum, conjugated: the inverse FFT is
        // conj(FFT(conj(.))) / L
        const double2* G = Gt + (size_t)(j * g.pieces + i) * kFftL;
#pragma unroll
        for (int r = 0; r < kFftPer; ++r) {
          const int k = t + r * kFftThreads;
          buf[fpad(k)] = cconj(cmul(buf[fpad(k)], G[k]));
        }
        __syncthreads();
        fft_row(buf, tw1, tw2);
#pragma unroll
        for (int r = 0; r < kFftPer; ++r) {
          const int u = t + r * kFftThreads;
          if (u < he) {
            const double2 y = cconj(buf[fpad(u)]);
            double2 z = cmul(cout[j0 + u], y);
            z = make_double2(oscale * z.x, oscale * z.y);
            // the same thread wrote this entry for the previous piece
            orow[j0 + u] = i == 0 ? z : cadd(orow[j0 + u], z);
          }
        }
        __syncthreads();
      }
    }
  }
}

// Piece/chunk geometry of a transform, or false for the direct sum.
bool chirp_geometry(int n_in, int n_out, int N, int64_t a0, int da, int64_t d0, int dd, int adj,
                    ChirpGeom* g) {
  if (n_in < kChirpMin || n_out < kChirpMin) return false;
  g->n_in = n_in;
  g->n_out = n_out;
  g->P = std::min(n_in, kFftL / 2);
  g->H = kFftL - g->P + 1;
  g->pieces = (n_in + g->P - 1) / g->P;
  g->chunks = (n_out + g->H - 1) / g->H;
  if (g->pieces * g->chunks > kChirpMaxTables) return false;
  // (a0 + da u)(d0 + dd v) = da dd (u + da a0)(v + dd d0), da, dd = +-1
  g->sigma = da * dd * (adj ? -1 : 1);
  g->N = N;
  g->A = da * a0;
  g->D = dd * d0;
  return true;
}

// Entries (complex) of the chirp tables' workspace beyond the direct sum's
// N-entry twiddle table: the input and output chirps (n_in + n_out <= 2N)
// and the pair spectra.
int64_t chirp_workspace_entries(int64_t N) {
  if (N < kChirpMin) return 0;
  const int64_t per = (N + kFftL / 2 - 1) / (kFftL / 2);   // pieces or chunks, at most
  return 2 * N + std::min<int64_t>(kChirpMaxTables, per * per) * kFftL;
}

}  // namespace
}  // namespace qsx

using namespace qsx;

extern "C" int64_t qsx_fourier_workspace_bytes(int32_t N) {
  if (N < 1) return -1;
  return ((int64_t)N + chirp_workspace_entries(N)) * (int64_t)sizeof(double2);
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
  const int64_t c = N / 2;
  // output index r -> s_r = a0 + da * r; sample index j -> lo + j - c
  const int64_t sa0 = sign == -1 ? -c : (int64_t)N - 1 - c;
  const int sda = sign == -1 ? 1 : -1;
  const int64_t ja0 = (int64_t)lo - c;
  const int n_out = adjoint ? n : N, n_in = adjoint ? N : n;
  const int64_t a0 = adjoint ? ja0 : sa0, d0 = adjoint ? sa0 : ja0;
  const int da = adjoint ? 1 : sda, dd = adjoint ? sda : 1;
  ChirpGeom g;
  if (chirp_geometry(n_in, n_out, N, a0, da, d0, dd, adjoint != 0, &g)) {
    double2* cin = T + N;
    double2* cout = cin + n_in;
    double2* Gt = cin + 2 * (int64_t)N;
    hipLaunchKernelGGL(chirp_vectors_kernel, dim3((unsigned)((std::max(n_in, n_out) + 255) / 256)),
                       dim3(256), 0, s, g, cin, cout);
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)chirp_tables_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFftLds));
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)chirp_rows_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFftLds));
    hipLaunchKernelGGL(chirp_tables_kernel, dim3(g.pieces * g.chunks), dim3(kFftThreads), kFftLds,
                       s, g, Gt);
    const unsigned gx = (unsigned)std::min<int64_t>(batch, 1 << 20);
    hipLaunchKernelGGL(chirp_rows_kernel, dim3(gx), dim3(kFftThreads), kFftLds, s,
                       static_cast<const double2*>(x), batch, g, (const double2*)cin,
                       (const double2*)cout, (const double2*)Gt, dt, static_cast<double2*>(out));
    QSX_CHECK_HIP(hipGetLastError());
    return QSX_OK;
  }
  hipLaunchKernelGGL(dft_twiddle_kernel, dim3((N + kDftThreads - 1) / kDftThreads), dim3(kDftThreads),
                     0, s, N, T);
  const unsigned gy = (unsigned)std::min<int64_t>(batch, 65535);
  hipLaunchKernelGGL(dft_kernel, dim3((n_out + kDftThreads - 1) / kDftThreads, gy),
                     dim3(kDftThreads), 0, s, static_cast<const double2*>(x), batch, n_in,
                     n_out, N, a0, da, d0, dd, (int)(adjoint != 0), dt, T,
                     static_cast<double2*>(out));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}
