// Linear-response signals of many small problems of unequal size in one
// launch (gfx950).
//
// The ensemble absorption workload (static-disorder members and aggregates
// of different site counts, qspectra/simulate/response.py:15-43 and
// :103-154 per member) propagates, for every problem b, one Heisenberg
// state y_{k+1} = P_b y_k with its one-step propagator P_b (M_b x M_b,
// M_b <= 31) and records s_k = v_b . y_k for k = 0 .. nt-1.  Each problem
// occupies a group of G lanes of one wavefront (G the power of two above
// M_b): lane i < M_b keeps row i of P_b in registers and produces y_{k+1}[i];
// lane M_b keeps v_b and produces s_k from the same y_k in the same pass,
// so no reduction is needed.  The state of the wavefront's groups is
// exchanged through LDS (double-buffered, one barrier per step), and the
// signal values are staged in LDS and stored kFlush steps at a time.
//
// A one-step product is 4 M FP64 FMAs per lane in four independent chains
// after the step's M LDS reads (straight-line code per M); with packed FP32 disabled, FP32 FMAs issue at the FP64 rate
// on the vector ALUs, so the products stay FP64.
//
// Work description (device, built by the host once per plan):
//   waves   int32 [nwaves][4]: first problem, problems in the wave, M, G
//           (every problem of a wave has the same M; G * count <= 64)
//   offs    int64 [B][3]: offset of P_b in P (complex elements, M_b^2 of
//           them, row-major), of y0_b / v_b in y0 / v (M_b each), and the
//           output row r_b of problem b
//   out     complex128 [B][nt], row r_b; status int32 [B]: 1 at r_b when a
//           signal value was not finite.

#include "qsx_common.h"

namespace qsx {
namespace {

constexpr int kWave = 64;
constexpr int kMaxStates = 31;  // rows M plus the projection row in 32 lanes
constexpr int kFlush = 32;      // signal values staged in LDS between stores

// One wavefront's problems of M states through all nt steps.  M is a
// template parameter so that the row product is straight-line code: the M
// LDS reads of a step issue together before its multiply-adds.
template <int M>
__device__ void bucket_wave(const int4 w, const int64_t* __restrict__ offs,
                            const double2* __restrict__ P, const double2* __restrict__ y0,
                            const double2* __restrict__ v, int nt, double2* __restrict__ out,
                            int32_t* __restrict__ status, double2 (*ys)[kWave],
                            double2 (*sbuf)[kWave / 2], int64_t* rows) {
  const int lane = threadIdx.x, G = w.w;
  const int g = lane / G, i = lane - g * G;
  const bool live = g < w.y;
  const int64_t mem = (int64_t)w.x + g;
  const bool srow = live && i < M, prow = live && i == M;
  int64_t offP = 0, offY = 0, orow = 0;
  if (live) {
    offP = offs[3 * mem];
    offY = offs[3 * mem + 1];
    orow = offs[3 * mem + 2];
  }
  if (live && i == 0) rows[g] = orow;   // for the coalesced flushes
  double2 row[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    double2 r = make_double2(0.0, 0.0);
    if (srow) r = P[offP + (int64_t)i * M + j];
    else if (prow) r = v[offY + j];
    row[j] = r;
  }
  ys[0][lane] = srow ? y0[offY + i] : make_double2(0.0, 0.0);
  __syncthreads();
  const int base = g * G;
  bool bad = false;
  int cur = 0;
  for (int k = 0; k < nt; ++k) {
    double2 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = make_double2(0.0, 0.0);
    // the state in chunks of up to 8 entries: a chunk's reads issue together
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += 8) {
      constexpr int kChunk = 8;
      double2 yv[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j)
        if (j0 + j < M) yv[j] = ys[cur][base + j0 + j];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        if (j0 + j < M) {
          double2& a = acc[j & 3];
          const double2 r = row[j0 + j];
          a.x = fma(r.x, yv[j].x, a.x);
          a.x = fma(-r.y, yv[j].y, a.x);
          a.y = fma(r.x, yv[j].y, a.y);
          a.y = fma(r.y, yv[j].x, a.y);
        }
      }
    }
    const double2 sum = make_double2((acc[0].x + acc[1].x) + (acc[2].x + acc[3].x),
                                     (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y));
    const int kk = k % kFlush;
    if (prow) {
      sbuf[kk][g] = sum;
      bad |= !isfinite(sum.x) || !isfinite(sum.y);
    }
    ys[cur ^ 1][lane] = sum;
    __syncthreads();
    cur ^= 1;
    if (kk == kFlush - 1 || k == nt - 1) {   // uniform: flush the staged values
      const int k0 = k - kk, cnt = kk + 1;
      for (int e = lane; e < w.y * cnt; e += kWave) {
        const int gi = e / cnt, s = e - gi * cnt;
        out[rows[gi] * nt + k0 + s] = sbuf[s][gi];
      }
      __syncthreads();
    }
  }
  if (prow) status[orow] = bad ? 1 : 0;
}

__global__ __launch_bounds__(kWave) void response_buckets_kernel(
    const int4* __restrict__ waves, const int64_t* __restrict__ offs,
    const double2* __restrict__ P, const double2* __restrict__ y0,
    const double2* __restrict__ v, int nt, double2* __restrict__ out,
    int32_t* __restrict__ status) {
  __shared__ double2 ys[2][kWave];
  // s_k of every group, kFlush steps at a time: the stores leave the step
  // loop, so its barriers do not wait on them, and go out coalesced
  __shared__ double2 sbuf[kFlush][kWave / 2];
  __shared__ int64_t rows[kWave / 2];   // output row of each group
  const int4 w = waves[blockIdx.x];
  switch (w.z) {   // M, the same for every problem of the wavefront
#define QSX_BUCKET_M(m_) \
  case m_: bucket_wave<m_>(w, offs, P, y0, v, nt, out, status, ys, sbuf, rows); break;
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

}  // namespace
}  // namespace qsx

using namespace qsx;

extern "C" int32_t qsx_response_buckets(const int32_t* waves, int32_t nwaves,
                                        const int64_t* offsets, const void* P, const void* y0,
                                        const void* v, int32_t nt, void* out, int32_t* status,
                                        void* stream) {
  QSX_CHECK_ARG(waves && offsets && P && y0 && v && out && status,
                "qsx_response_buckets: null pointer");
  QSX_CHECK_ARG(nwaves >= 1 && nt >= 1, "qsx_response_buckets: bad sizes nwaves=%d nt=%d",
                nwaves, nt);
  hipLaunchKernelGGL(response_buckets_kernel, dim3(nwaves), dim3(kWave), 0, as_stream(stream),
                     reinterpret_cast<const int4*>(waves), offsets,
                     static_cast<const double2*>(P), static_cast<const double2*>(y0),
                     static_cast<const double2*>(v), nt, static_cast<double2*>(out), status);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_response_buckets_max_states(void) { return kMaxStates; }
