// Dense Liouvillian application and the fused RK4 stepper (gfx950).
//
// One workgroup owns one problem b and a tile of RT right-hand sides.  The
// Liouvillian L_b is staged once into LDS (when it fits), the stage vectors
// z live in a double-buffered LDS tile [2][RT][M], and every thread keeps
// the FP64 state y, the RK accumulator and its own stage value for the
// outputs (r, i) it owns in registers, so one barrier separates RK stages.
//
// FP64 mode: L, z and products in complex128.
// MIXED mode: L holds the off-diagonal part in complex64, z is staged in
// complex64 and products accumulate in FP32; the diagonal of L (the
// rotating-frame frequencies and pure dephasing, the largest entries) is
// applied in FP64 to the thread's own FP64 stage value, and y / the RK
// update stay FP64.

#include "qsx_common.h"

#include <algorithm>

namespace qsx {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxEPT = 8;                  // outputs per thread
constexpr size_t kLdsBudget = 64 * 1024;    // bytes per workgroup

__global__ __launch_bounds__(kThreads) void apply_dense_c128_kernel(
    const double2* __restrict__ L, int64_t strideL, int M, const double2* __restrict__ y,
    int R, int B, double2* __restrict__ out) {
  const int64_t o = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t total = (int64_t)B * R * M;
  if (o >= total) return;
  const int i = (int)(o % M);
  const int64_t br = o / M;          // b * R + r
  const int b = (int)(br / R);
  const double2* Lrow = L + b * strideL + (int64_t)i * M;
  const double2* yv = y + br * M;
  double2 acc = make_double2(0.0, 0.0);
  for (int j = 0; j < M; ++j) cfma<double>(acc, Lrow[j], yv[j]);
  out[o] = acc;
}

// RK4 stage weights: k_st enters the update with kW[st]/6 and the next stage
// argument is y + kC[st] * h * k_st.
__constant__ double kW[4] = {1.0, 2.0, 2.0, 1.0};
__constant__ double kC[3] = {0.5, 0.5, 1.0};

template <typename LT, bool MIXED, int EPT>
__global__ __launch_bounds__(kThreads) void rk4_dense_kernel(
    const cplx_t<LT>* __restrict__ L, int64_t strideL, const double2* __restrict__ diag,
    int64_t strideD, int M, const double2* __restrict__ y0, int R, int RT,
    const int* __restrict__ nsub, const double* __restrict__ hs, int nsave,
    const cplx_t<LT>* __restrict__ proj, int64_t strideP, int K, double2* __restrict__ out,
    double2* __restrict__ yfin, int* __restrict__ status, int l_in_lds) {
  using LT2 = cplx_t<LT>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int r0 = blockIdx.x * RT;
  const int rt = min(RT, R - r0);
  const int nout = rt * M;
  const int MM = M * M;

  const LT2* Lb = L + (int64_t)b * strideL;
  LT2* Ls = reinterpret_cast<LT2*>(smem);
  LT2* zs = reinterpret_cast<LT2*>(smem + (l_in_lds ? (size_t)MM * sizeof(LT2) : 0));
  const int zstride = RT * M;  // one buffer of the double buffer
  if (l_in_lds) {
    for (int e = tid; e < MM; e += kThreads) Ls[e] = Lb[e];
  }
  const LT2* Lu = l_in_lds ? Ls : Lb;

  double2 y[EPT], acc[EPT], zi[EPT], dd[EPT];
  int row[EPT], rhs[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int o = tid + e * kThreads;
    const bool on = o < nout;
    rhs[e] = on ? o / M : 0;
    row[e] = on ? o % M : 0;
    y[e] = on ? y0[((int64_t)b * R + r0 + rhs[e]) * M + row[e]] : make_double2(0.0, 0.0);
    dd[e] = (MIXED && on) ? diag[(int64_t)b * strideD + row[e]] : make_double2(0.0, 0.0);
    if (on) zs[o] = from_d2<LT>(y[e]);
  }
  __syncthreads();

  int cur = 0;
  bool bad = false;
  for (int s = 0; s < nsave; ++s) {
    const int n = nsub[s];
    const double h = hs[s];
    for (int step = 0; step < n; ++step) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        acc[e] = make_double2(0.0, 0.0);
        zi[e] = y[e];
      }
#pragma unroll 1
      for (int st = 0; st < 4; ++st) {
        const LT2* zc = zs + cur * zstride;
        LT2* zn = zs + (cur ^ 1) * zstride;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int o = tid + e * kThreads;
          if (o < nout) {
            const LT2* Lrow = Lu + row[e] * M;
            const LT2* zr = zc + rhs[e] * M;
            LT2 ks = {0, 0};
            for (int j = 0; j < M; ++j) cfma<LT>(ks, Lrow[j], zr[j]);
            double2 k = to_d2(ks);
            if (MIXED) k = cadd(k, cmul(dd[e], zi[e]));
            acc[e] = caxpy(kW[st], k, acc[e]);
            if (st < 3) {
              zi[e] = caxpy(kC[st] * h, k, y[e]);
              zn[o] = from_d2<LT>(zi[e]);
            } else {
              y[e] = caxpy(h / 6.0, acc[e], y[e]);
              zn[o] = from_d2<LT>(y[e]);
            }
          }
        }
        cur ^= 1;
        __syncthreads();
      }
    }
    // ---- save point s
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int o = tid + e * kThreads;
      if (o < nout && !cfinite(y[e])) bad = true;
    }
    if (proj == nullptr) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int o = tid + e * kThreads;
        if (o < nout)
          out[(((int64_t)b * R + r0 + rhs[e]) * nsave + s) * K + row[e]] = y[e];
      }
    } else {
      // The projection reads the whole state of an RHS: in FP64 mode the
      // current z buffer already holds y; in MIXED mode stage an FP64 copy
      // over the (now idle) z double buffer, which is exactly RT*M*16 bytes.
      const double2* ysh;
      if (MIXED) {
        double2* yb = reinterpret_cast<double2*>(zs);
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int o = tid + e * kThreads;
          if (o < nout) yb[o] = y[e];
        }
        __syncthreads();
        ysh = yb;
      } else {
        ysh = reinterpret_cast<const double2*>(zs + cur * zstride);
      }
      const LT2* Pb = proj + (int64_t)b * strideP;
      for (int p = tid; p < rt * K; p += kThreads) {
        const int r = p / K, k = p % K;
        const LT2* Prow = Pb + (int64_t)k * M;
        const double2* yr = ysh + r * M;
        double2 a = make_double2(0.0, 0.0);
        for (int j = 0; j < M; ++j) cfma<double>(a, to_d2(Prow[j]), yr[j]);
        out[(((int64_t)b * R + r0 + r) * nsave + s) * K + k] = a;
      }
      if (MIXED) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int o = tid + e * kThreads;
          if (o < nout) zs[cur * zstride + o] = from_d2<LT>(y[e]);
        }
        __syncthreads();
      }
    }
  }
  if (bad) status[0] = 1;
  if (yfin != nullptr) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int o = tid + e * kThreads;
      if (o < nout) yfin[((int64_t)b * R + r0 + rhs[e]) * M + row[e]] = y[e];
    }
  }
}

template <typename LT, bool MIXED>
int launch_rk4(const void* L, int64_t strideL, const void* diag, int64_t strideD, int M,
               const void* y0, int R, int B, const int* nsub, const double* h, int nsave,
               const void* proj, int64_t strideP, int K, void* out, void* yfin, int* status,
               hipStream_t stream) {
  using LT2 = cplx_t<LT>;
  const size_t lbytes = (size_t)M * M * sizeof(LT2);
  // z double buffer per RHS: 2*M*sizeof(LT2); MIXED needs >= M*16 bytes per
  // RHS for the FP64 save copy, which 2*M*8 provides.
  const size_t per_rhs = 2 * (size_t)M * sizeof(LT2);
  const bool l_in_lds = lbytes + per_rhs <= kLdsBudget;
  const size_t avail = kLdsBudget - (l_in_lds ? lbytes : 0);
  int RT = (int)std::min<size_t>((size_t)R, avail / per_rhs);
  RT = std::min(RT, kThreads * kMaxEPT / M);
  if (RT >= 1) RT = (R + (R + RT - 1) / RT - 1) / ((R + RT - 1) / RT);  // balance tiles
  QSX_CHECK_ARG(RT >= 1, "qsx_rk4_dense: M=%d too large for one workgroup tile", M);
  const int ept = (RT * M + kThreads - 1) / kThreads;
  const size_t shmem = (l_in_lds ? lbytes : 0) + (size_t)RT * per_rhs;
  dim3 grid((R + RT - 1) / RT, B);
  auto args = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(kThreads), shmem, stream,
                       (const LT2*)L, strideL, (const double2*)diag, strideD, M,
                       (const double2*)y0, R, RT, nsub, h, nsave, (const LT2*)proj, strideP, K,
                       (double2*)out, (double2*)yfin, status, (int)l_in_lds);
  };
  if (ept <= 1) args(rk4_dense_kernel<LT, MIXED, 1>);
  else if (ept <= 2) args(rk4_dense_kernel<LT, MIXED, 2>);
  else if (ept <= 4) args(rk4_dense_kernel<LT, MIXED, 4>);
  else args(rk4_dense_kernel<LT, MIXED, 8>);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

}  // namespace
}  // namespace qsx

extern "C" int32_t qsx_apply_dense_c128(const void* L, int64_t strideL, int32_t M,
                                        const void* y, int32_t R, int32_t B, void* out,
                                        void* stream) {
  QSX_CHECK_ARG(L && y && out, "qsx_apply_dense_c128: null pointer");
  QSX_CHECK_ARG(M >= 1 && R >= 1 && B >= 1, "qsx_apply_dense_c128: bad sizes M=%d R=%d B=%d",
                M, R, B);
  QSX_CHECK_ARG(strideL == 0 || strideL >= (int64_t)M * M, "qsx_apply_dense_c128: bad strideL");
  const int64_t total = (int64_t)B * R * M;
  const int64_t blocks = (total + qsx::kThreads - 1) / qsx::kThreads;
  QSX_CHECK_ARG(blocks < (1LL << 31), "qsx_apply_dense_c128: problem too large");
  hipLaunchKernelGGL(qsx::apply_dense_c128_kernel, dim3((unsigned)blocks), dim3(qsx::kThreads),
                     0, qsx::as_stream(stream), (const double2*)L, strideL, M,
                     (const double2*)y, R, B, (double2*)out);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_rk4_dense(int32_t precision, const void* L, int64_t strideL,
                                 const void* diag, int64_t strideD, int32_t M, const void* y0,
                                 int32_t R, int32_t B, const int32_t* nsub, const double* h,
                                 int32_t nsave, const void* proj, int64_t strideP, int32_t K,
                                 void* out, void* y_final, int32_t* status, void* stream) {
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED,
                "qsx_rk4_dense: unknown precision %d", precision);
  QSX_CHECK_ARG(L && y0 && nsub && h && out && status, "qsx_rk4_dense: null pointer");
  QSX_CHECK_ARG(M >= 1 && R >= 1 && B >= 1 && nsave >= 1 && B <= 65535,
                "qsx_rk4_dense: bad sizes M=%d R=%d B=%d nsave=%d", M, R, B, nsave);
  QSX_CHECK_ARG(strideL == 0 || strideL >= (int64_t)M * M, "qsx_rk4_dense: bad strideL");
  QSX_CHECK_ARG(proj != nullptr || K == M, "qsx_rk4_dense: K must equal M without proj");
  QSX_CHECK_ARG(K >= 1, "qsx_rk4_dense: bad K=%d", K);
  QSX_CHECK_ARG(proj == nullptr || strideP == 0 || strideP >= (int64_t)K * M,
                "qsx_rk4_dense: bad strideP");
  if (precision == QSX_MIXED) {
    QSX_CHECK_ARG(diag != nullptr, "qsx_rk4_dense: MIXED needs the FP64 diagonal");
    QSX_CHECK_ARG(strideD == 0 || strideD >= M, "qsx_rk4_dense: bad strideD");
    return qsx::launch_rk4<float, true>(L, strideL, diag, strideD, M, y0, R, B, nsub, h, nsave,
                                        proj, strideP, K, out, y_final, status,
                                        qsx::as_stream(stream));
  }
  return qsx::launch_rk4<double, false>(L, strideL, nullptr, 0, M, y0, R, B, nsub, h, nsave, proj,
                                        strideP, K, out, y_final, status, qsx::as_stream(stream));
}
