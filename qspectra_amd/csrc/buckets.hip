// Linear-response signals of many small problems of unequal size in one
// launch (gfx950).
//
// The ensemble absorption workload (static-disorder members and aggregates
// of different site counts, qspectra/simulate/response.py:15-43 and
// :103-154 per member) propagates, for every problem b, one Heisenberg
// state y_{k+1} = P_b y_k with its one-step propagator P_b (M_b x M_b,
// M_b <= 31) and records s_k = v_b . y_k for k = 0 .. nt-1.
//
// J steps per round trip: the state advances by P_b^J, and the J signal
// values s_{Jk+j} = v_b . P_b^j y_{Jk} = w_j . y_{Jk}, w_j = (P_b^j)^T v_b,
// come out of the same pass.  Each problem occupies a group of G lanes of
// one wavefront (G the power of two at or above M_b + J): lane i < M_b keeps
// row i of P_b^J in registers and produces y_{J(k+1)}[i]; lane M_b + j keeps
// w_j and produces s_{Jk+j}, so no reduction is needed.  The state of the
// wavefront's groups is exchanged through LDS (double-buffered, one barrier
// per round trip), and the signal values are staged in LDS and stored kFlush
// steps at a time.  A round trip costs what a single step cost (the same M
// multiply-adds per lane, one barrier), so J = 4 takes a quarter of the
// round trips.
//
// A round trip's product is 4 M FP64 FMAs per lane in four independent
// chains after the M LDS reads (straight-line code per M); with packed FP32
// disabled, FP32 FMAs issue at the FP64 rate on the vector ALUs, so the
// products stay FP64.
//
// Work description (device, built by the host once per plan):
//   waves   int32 [nwaves][4]: first problem, problems in the wave, M, G
//           (every problem of a wave has the same M; G * count <= 64,
//           G >= M + J)
//   offs    int64 [B][3]: offset of P_b^J in P (complex elements, M_b^2 of
//           them, row-major), of y0_b in y0 (M_b; the projection rows of
//           problem b sit at J times it in W, [J][M_b]), and the output row
//           r_b of problem b
//   out     complex128 [B][nt], row r_b; status int32 [B]: 1 at r_b when a
//           signal value was not finite.

#include "qsx_common.h"

namespace qsx {
namespace {

constexpr int kWave = 64;
constexpr int kMaxStates = 31;  // rows M plus J projection rows in 64 lanes
constexpr int kFlush = 32;      // signal values staged in LDS between stores
constexpr int kMaxJ = 8;

// One wavefront's problems of M states through all nt steps, J per round
// trip.  M is a template parameter so that the row product is straight-line
// code: the M LDS reads of a round trip issue together before its
// multiply-adds.  S lanes share a row (S = 1 or 2): lane s of row i takes
// the columns j = s, s + S, ... and the S partial sums are added across
// neighbouring lanes, which halves a large problem's chain per round trip.
template <int M, int J, int S>
__device__ void bucket_wave(const int4 w, const int64_t* __restrict__ offs,
                            const double2* __restrict__ P, const double2* __restrict__ y0,
                            const double2* __restrict__ W, int nt, double2* __restrict__ out,
                            int32_t* __restrict__ status, double2 (*ys)[kWave],
                            double2 (*sbuf)[kWave / 2], int64_t* rows, int (*sbad)[kMaxJ]) {
  constexpr int kCols = (M + S - 1) / S;                 // columns per lane
  const int lane = threadIdx.x, G = w.w;
  const int g = lane / G, li = lane - g * G;
  const int i = li / S, s = li - i * S;                  // row, column phase
  const bool live = g < w.y;
  const int64_t mem = (int64_t)w.x + g;
  const int jj = i - M;                                  // projection index
  const bool srow = live && i < M, prow = live && jj >= 0 && jj < J;
  const bool lead = s == 0;                              // writes the row's result
  int64_t offP = 0, offY = 0, orow = 0;
  if (live) {
    offP = offs[3 * mem];
    offY = offs[3 * mem + 1];
    orow = offs[3 * mem + 2];
  }
  // row i's columns s, s + S, ...; a column past M (odd M, s = 1) reads
  // the last state entry with a zero coefficient
  double2 row[kCols];
#pragma unroll
  for (int q = 0; q < kCols; ++q) {
    const int j = q * S + s;
    double2 r = make_double2(0.0, 0.0);
    if (j < M) {
      if (srow) r = P[offP + (int64_t)i * M + j];
      else if (prow) r = W[J * offY + (int64_t)jj * M + j];
    }
    row[q] = r;
  }
  const int base = g * G;
  if (lead) ys[0][base + i] = srow ? y0[offY + i] : make_double2(0.0, 0.0);
  if (live && li == 0) rows[g] = orow;   // for the coalesced flushes
  __syncthreads();
  const int nr = (nt + J - 1) / J;
  bool bad = false;
  int cur = 0;
  for (int k = 0; k < nr; ++k) {
    double2 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
    // the state in chunks of up to 8 entries: a chunk's reads issue together
#pragma unroll
    for (int q0 = 0; q0 < kCols; q0 += 8) {
      constexpr int kChunk = 8;
      double2 yv[kChunk];
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
        if (q0 + q < kCols) yv[q] = ys[cur][base + min((q0 + q) * S + s, M - 1)];
#pragma unroll
      for (int q = 0; q < kChunk; ++q) {
        if (q0 + q < kCols) {
          double2& a = acc[q & 3];
          const double2 r = row[q0 + q];
          a.x = fma(r.x, yv[q].x, a.x);
          a.x = fma(-r.y, yv[q].y, a.x);
          a.y = fma(r.x, yv[q].y, a.y);
          a.y = fma(r.y, yv[q].x, a.y);
        }
      }
    }
    double2 sum = make_double2((acc[0].x + acc[1].x) + (acc[2].x + acc[3].x),
                               (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y));
    if (S == 2) {                                        // add the neighbour lane's half
      sum.x += __shfl_xor(sum.x, 1);
      sum.y += __shfl_xor(sum.y, 1);
    }
    const int step = k * J + jj;
    if (prow && lead && step < nt) {
      sbuf[step % kFlush][g] = sum;
      bad |= !isfinite(sum.x) || !isfinite(sum.y);
    }
    if (lead) ys[cur ^ 1][base + i] = sum;
    __syncthreads();
    cur ^= 1;
    const int last = min((k + 1) * J, nt) - 1;           // last step of this round trip
    if (last % kFlush == kFlush - 1 || k == nr - 1) {    // uniform: flush the staged values
      const int k0 = last - last % kFlush, cnt = last - k0 + 1;
      for (int e = lane; e < w.y * cnt; e += kWave) {
        const int gi = e / cnt, t = e - gi * cnt;
        out[rows[gi] * nt + k0 + t] = sbuf[t][gi];
      }
      __syncthreads();
    }
  }
  if (prow && lead) sbad[g][jj] = bad ? 1 : 0;
  __syncthreads();
  if (prow && lead && jj == 0) {
    int any = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) any |= sbad[g][j];
    status[orow] = any;
  }
}

// Entry N of a 16-lane row of v, broadcast to the row (DPP row_newbcast;
// two 32-bit moves per double).
template <int N>
__device__ __forceinline__ double row_bcast(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x150 + N, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x150 + N, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// acc[j & 3] += row[j] y_j for j = Q .. M - 1, y_j held by lane j of the row
template <int Q, int M>
struct RowDot {
  static __device__ __forceinline__ void run(const double2* row, double2 y, double2* acc) {
    const double2 yj = make_double2(row_bcast<Q>(y.x), row_bcast<Q>(y.y));
    double2& a = acc[Q & 3];
    a.x = fma(row[Q].x, yj.x, a.x);
    a.x = fma(-row[Q].y, yj.y, a.x);
    a.y = fma(row[Q].x, yj.y, a.y);
    a.y = fma(row[Q].y, yj.x, a.y);
    RowDot<Q + 1, M>::run(row, y, acc);
  }
};
template <int M>
struct RowDot<M, M> {
  static __device__ __forceinline__ void run(const double2*, double2, double2*) {}
};

// Problems of M + J <= 16 rows, one per 16-lane row of the wavefront (G =
// 16): lane i keeps y[i] in a register and the row product takes y_j from
// lane j by a cross-lane broadcast, so a round trip has no LDS exchange and
// no barrier.  Signals are staged and flushed as in bucket_wave.
template <int M, int J>
__device__ void bucket_wave_row(const int4 w, const int64_t* __restrict__ offs,
                                const double2* __restrict__ P, const double2* __restrict__ y0,
                                const double2* __restrict__ W, int nt, double2* __restrict__ out,
                                int32_t* __restrict__ status, double2 (*sbuf)[kWave / 2],
                                int64_t* rows, int (*sbad)[kMaxJ]) {
  static_assert(M + J <= 16, "one problem per 16-lane row");
  const int lane = threadIdx.x;
  const int g = lane >> 4, i = lane & 15;
  const bool live = g < w.y;
  const int64_t mem = (int64_t)w.x + g;
  const int jj = i - M;
  const bool srow = live && i < M, prow = live && jj >= 0 && jj < J;
  int64_t offP = 0, offY = 0, orow = 0;
  if (live) {
    offP = offs[3 * mem];
    offY = offs[3 * mem + 1];
    orow = offs[3 * mem + 2];
  }
  double2 row[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    double2 r = make_double2(0.0, 0.0);
    if (srow) r = P[offP + (int64_t)i * M + j];
    else if (prow) r = W[J * offY + (int64_t)jj * M + j];
    row[j] = r;
  }
  double2 y = srow ? y0[offY + i] : make_double2(0.0, 0.0);
  if (live && i == 0) rows[g] = orow;
  __syncthreads();
  const int nr = (nt + J - 1) / J;
  bool bad = false;
  for (int k = 0; k < nr; ++k) {
    double2 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
    RowDot<0, M>::run(row, y, acc);
    const double2 sum = make_double2((acc[0].x + acc[1].x) + (acc[2].x + acc[3].x),
                                     (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y));
    const int step = k * J + jj;
    if (prow && step < nt) {
      sbuf[step % kFlush][g] = sum;
      bad |= !isfinite(sum.x) || !isfinite(sum.y);
    }
    y = sum;            // projection lanes' values are never broadcast
    const int last = min((k + 1) * J, nt) - 1;
    if (last % kFlush == kFlush - 1 || k == nr - 1) {    // uniform: flush the staged values
      __syncthreads();
      const int k0 = last - last % kFlush, cnt = last - k0 + 1;
      for (int e = lane; e < w.y * cnt; e += kWave) {
        const int gi = e / cnt, t = e - gi * cnt;
        out[rows[gi] * nt + k0 + t] = sbuf[t][gi];
      }
      __syncthreads();
    }
  }
  if (prow) sbad[g][jj] = bad ? 1 : 0;
  __syncthreads();
  if (prow && jj == 0) {
    int any = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) any |= sbad[g][j];
    status[orow] = any;
  }
}

// Lanes per problem for S lanes per row: S times the power of two at or
// above M + J.
__host__ __device__ constexpr int bucket_lanes(int M, int J, int S) {
  int p = 1;
  while (p < M + J) p *= 2;
  return S * p;
}

__global__ __launch_bounds__(kWave) void response_buckets_kernel(
    const int4* __restrict__ waves, const int64_t* __restrict__ offs,
    const double2* __restrict__ P, const double2* __restrict__ y0,
    const double2* __restrict__ W, int J, int nt, double2* __restrict__ out,
    int32_t* __restrict__ status) {
  __shared__ double2 ys[2][kWave];
  // s_k of every group, kFlush steps at a time: the stores leave the step
  // loop, so its barriers do not wait on them, and go out coalesced
  __shared__ double2 sbuf[kFlush][kWave / 2];
  __shared__ int64_t rows[kWave / 2];   // output row of each group
  __shared__ int sbad[kWave / 2][kMaxJ];
  const int4 w = waves[blockIdx.x];
  // M is the same for every problem of the wavefront; G picks the form:
  // 16 with M + J <= 16 the register-resident one (a problem per 16-lane
  // row), twice the power of two at or above M + J two lanes per row
  if (J == 8 || J == 4) {
#define QSX_BUCKET_M(m_)                                                                      \
  case m_:                                                                                    \
    if (m_ + kJ <= 16 && w.w == 16)                                                           \
      bucket_wave_row<(m_ + kJ <= 16 ? m_ : 1), kJ>(w, offs, P, y0, W, nt, out, status, sbuf, \
                                                    rows, sbad);                              \
    else if (bucket_lanes(m_, kJ, 2) <= kWave && w.w == bucket_lanes(m_, kJ, 2))              \
      bucket_wave<m_, kJ, 2>(w, offs, P, y0, W, nt, out, status, ys, sbuf, rows, sbad);        \
    else                                                                                      \
      bucket_wave<m_, kJ, 1>(w, offs, P, y0, W, nt, out, status, ys, sbuf, rows, sbad);        \
    break;
#define QSX_BUCKET_SWITCH                                                                     \
  switch (w.z) {                                                                              \
    QSX_BUCKET_M(1) QSX_BUCKET_M(2) QSX_BUCKET_M(3) QSX_BUCKET_M(4) QSX_BUCKET_M(5)           \
    QSX_BUCKET_M(6) QSX_BUCKET_M(7) QSX_BUCKET_M(8) QSX_BUCKET_M(9) QSX_BUCKET_M(10)          \
    QSX_BUCKET_M(11) QSX_BUCKET_M(12) QSX_BUCKET_M(13) QSX_BUCKET_M(14) QSX_BUCKET_M(15)      \
    QSX_BUCKET_M(16) QSX_BUCKET_M(17) QSX_BUCKET_M(18) QSX_BUCKET_M(19) QSX_BUCKET_M(20)      \
    QSX_BUCKET_M(21) QSX_BUCKET_M(22) QSX_BUCKET_M(23) QSX_BUCKET_M(24) QSX_BUCKET_M(25)      \
    QSX_BUCKET_M(26) QSX_BUCKET_M(27) QSX_BUCKET_M(28) QSX_BUCKET_M(29) QSX_BUCKET_M(30)      \
    QSX_BUCKET_M(31)                                                                          \
    default: break;                                                                           \
  }
    if (J == 8) {
      constexpr int kJ = 8;
      QSX_BUCKET_SWITCH
    } else {
      constexpr int kJ = 4;
      QSX_BUCKET_SWITCH
    }
#undef QSX_BUCKET_SWITCH
#undef QSX_BUCKET_M
  } else {
    switch (w.z) {
#define QSX_BUCKET_M(m_) \
  case m_: bucket_wave<m_, 1, 1>(w, offs, P, y0, W, nt, out, status, ys, sbuf, rows, sbad); break;
      QSX_BUCKET_M(1) QSX_BUCKET_M(2) QSX_BUCKET_M(3) QSX_BUCKET_M(4) QSX_BUCKET_M(5)
      QSX_BUCKET_M(6) QSX_BUCKET_M(7) QSX_BUCKET_M(8) QSX_BUCKET_M(9) QSX_BUCKET_M(10)
      QSX_BUCKET_M(11) QSX_BUCKET_M(12) QSX_BUCKET_M(13) QSX_BUCKET_M(14) QSX_BUCKET_M(15)
      QSX_BUCKET_M(16) QSX_BUCKET_M(17) QSX_BUCKET_M(18) QSX_BUCKET_M(19) QSX_BUCKET_M(20)
      QSX_BUCKET_M(21) QSX_BUCKET_M(22) QSX_BUCKET_M(23) QSX_BUCKET_M(24) QSX_BUCKET_M(25)
      QSX_BUCKET_M(26) QSX_BUCKET_M(27) QSX_BUCKET_M(28) QSX_BUCKET_M(29) QSX_BUCKET_M(30)
      QSX_BUCKET_M(31)
#undef QSX_BUCKET_M
      default: break;
    }
  }
}

// The projection rows W[J offY_b + j M_b + c] = ((P_b^j)^T v_b)[c], j < J,
// from the one-step propagators P (offsets as the bucket kernel's, P_b^1
// here) and v: one wavefront per entry of the wave table, lane c of group g
// holds entry c of the group's current vector, one barrier per power.
__global__ __launch_bounds__(kWave) void bucket_weights_kernel(
    const int4* __restrict__ waves, const int64_t* __restrict__ offs,
    const double2* __restrict__ P, const double2* __restrict__ v, int J,
    double2* __restrict__ W) {
  __shared__ double2 wv[kWave];
  const int4 w = waves[blockIdx.x];
  const int lane = threadIdx.x, G = w.w, M = w.z;
  const int g = lane / G, c = lane - g * G;
  const bool live = g < w.y && c < M;
  const int64_t mem = (int64_t)w.x + g;
  int64_t offP = 0, offY = 0;
  if (live) {
    offP = offs[3 * mem];
    offY = offs[3 * mem + 1];
  }
  double2 cur = live ? v[offY + c] : make_double2(0.0, 0.0);
  if (live) W[J * offY + c] = cur;
  const int base = g * G;
  for (int j = 1; j < J; ++j) {
    wv[lane] = cur;
    __syncthreads();
    double2 acc = make_double2(0.0, 0.0);
    if (live)
      for (int r = 0; r < M; ++r) cfma<double>(acc, P[offP + (int64_t)r * M + c], wv[base + r]);
    __syncthreads();
    cur = acc;
    if (live) W[J * offY + (int64_t)j * M + c] = cur;
  }
}

}  // namespace
}  // namespace qsx

using namespace qsx;

extern "C" int32_t qsx_response_buckets(const int32_t* waves, int32_t nwaves,
                                        const int64_t* offsets, const void* P, const void* y0,
                                        const void* W, int32_t J, int32_t nt, void* out,
                                        int32_t* status, void* stream) {
  QSX_CHECK_ARG(waves && offsets && P && y0 && W && out && status,
                "qsx_response_buckets: null pointer");
  QSX_CHECK_ARG(nwaves >= 1 && nt >= 1, "qsx_response_buckets: bad sizes nwaves=%d nt=%d",
                nwaves, nt);
  QSX_CHECK_ARG(J == 1 || J == 4 || J == 8,
                "qsx_response_buckets: J=%d (1, 4 or 8 steps per round trip)", J);
  hipLaunchKernelGGL(response_buckets_kernel, dim3(nwaves), dim3(kWave), 0, as_stream(stream),
                     reinterpret_cast<const int4*>(waves), offsets,
                     static_cast<const double2*>(P), static_cast<const double2*>(y0),
                     static_cast<const double2*>(W), J, nt, static_cast<double2*>(out), status);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_response_bucket_weights(const int32_t* waves, int32_t nwaves,
                                               const int64_t* offsets, const void* P,
                                               const void* v, int32_t J, void* W, void* stream) {
  QSX_CHECK_ARG(waves && offsets && P && v && W, "qsx_response_bucket_weights: null pointer");
  QSX_CHECK_ARG(nwaves >= 1 && J >= 1 && J <= kMaxJ,
                "qsx_response_bucket_weights: bad sizes nwaves=%d J=%d", nwaves, J);
  hipLaunchKernelGGL(bucket_weights_kernel, dim3(nwaves), dim3(kWave), 0, as_stream(stream),
                     reinterpret_cast<const int4*>(waves), offsets,
                     static_cast<const double2*>(P), static_cast<const double2*>(v), J,
                     static_cast<double2*>(W));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_response_buckets_max_states(void) { return kMaxStates; }
