// Classical RK4 for the ZOFE master equation (gfx950), optionally driven by
// classical fields.  The equation of qspectra/dynamics/zofe.py:121-225 for
// the density matrix rho (N x N) and the auxiliary operators O[p, s]
// (pseudomode p, site s), with L_s the coupling operators (the negated
// system-bath couplings), Osum_s = sum_p O[p, s]:
//     a      = sum_s L_s^+ Osum_s,           b = -i H - a
//     c      = sum_s L_s rho Osum_s^+,       d = b rho + c
//     f      = rho (i H - a^+) + sum_s Osum_s rho L_s^+          (default)
//              rho b^+ + sum_s Osum_s rho L_s^+                   (ham_hermit)
//              rho (i H - a^+) + c^+                             (rho_hermit)
//              d^+                                               (both)
//     rho'   = unit (d + f)
//     O'[p,s]= unit (Gamma[p,s] L_s - w[p,s] O[p,s] + b O[p,s] - O[p,s] b)
// plus, for field terms, sum_i f_i(t) [V_i, X] on rho and on every O[p, s]
// (the ZOFE commutator of qspectra/dynamics/zofe.py:24-34; f_i = -i E_i, E
// conjugated for '+', qspectra/simulate/eom.py:86-93).
//
// One work-group per right-hand side.  The stage vector (rho and all O) and
// the per-site intermediates (Osum_s, L_s rho, Osum_s rho, L_s^+ Osum_s,
// L_s rho Osum_s^+, Osum_s rho L_s^+, b, i H - a^+ and be = unit b +
// sum_i f_i V_i) live in LDS (the operators are read through the cache).
// Per stage, in phases separated by barriers and split into per-site items
// so that no item loops longer than max(N, S, P): (1) Osum; (2) the three
// per-site products; (3) the two per-site triple products, and a summed
// over sites (stored as b, i H - a^+ and be); (4) the derivative of every
// owned element and its RK4 update into the workspace; (5) the next stage
// vector copied into LDS.  The O derivative is unit (Gamma L - w O) +
// be O - O be (the field commutators folded into be).  Thread t owns the
// elements t, t + NT, ...; their state y, RK4 accumulator and next stage
// value sit in a workspace slice of the right-hand side, read and written by
// the owner only (L2 resident), the next element's entries loaded while the
// current one is computed.  The host state layout is the reference's
// (vec(rho) then vec(O) of shape (P, S, N, N), both column-major); in LDS
// rho and each O[p, s] are row-major N x N blocks, O ordered p-major.

#include "qsx_common.h"

namespace qsx {
namespace {

constexpr int kZofeMaxFields = 4;

struct ZofeArgs {
  int N, S, P, size, flags, nfield, K, R, apply;
  double unit;
  const double2* H;       // [B|1][N][N] rotating-frame H (not multiplied by unit)
  int64_t strideH;
  const double2* L;       // [B|1][S][N][N] coupling operators L_s
  int64_t strideL;
  const double2* G;       // [B|1][P][S] Gamma = Omega^2 huang
  const double2* W;       // [B|1][P][S] w = i Omega + gamma
  int64_t strideG;
  const double2* V;       // [B|1][nfield][N][N] dipole operators (fields)
  int64_t strideV;
  const double2* F;       // [B|1][nF][3][nfield] field values at the stage times
  int64_t strideF, nF;
  const double2* proj;    // [B|1][K][N*N] rows on vec(rho) (column-major); null: full state
  int64_t strideP;
};

__constant__ double kZW[4] = {1.0, 2.0, 2.0, 1.0};
__constant__ double kZC[3] = {0.5, 0.5, 1.0};

__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline void cfmad(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// reference index (column-major vec of rho, then of O[p][s][i][j] with p
// fastest) of internal element e (rho row-major, then O[p*S+s] row-major)
__device__ inline int zofe_ref_index(int e, int N, int S, int P) {
  const int N2 = N * N;
  if (e < N2) return (e / N) + N * (e % N);
  const int r = e - N2, ps = r / N2, ij = r % N2;
  const int i = ij / N, j = ij % N, p = ps / S, s = ps % S;
  return N2 + p + P * (s + S * (i + N * j));
}

// LDS of a stage work-group (row-major N x N blocks)
struct ZofeLds {
  double2* z;    // [size] stage vector: rho, then O[p * S + s]
  double2* Os;   // [S][N2] sum over pseudomodes
  double2* T1;   // [S][N2] L_s rho
  double2* T2;   // [S][N2] Osum_s rho
  double2* U;    // [S][N2] L_s^+ Osum_s
  double2* X1;   // [S][N2] L_s rho Osum_s^+
  double2* X2;   // [S][N2] Osum_s rho L_s^+
  double2* bm;   // [N2]    b = -i H - a
  double2* gm;   // [N2]    i H - a^+
  double2* be;   // [N2]    unit b + sum_i f_i V_i (the O blocks' commutator operator)
  double2* fb;   // [2][3 * kZofeMaxFields] field values
};

__device__ inline ZofeLds zofe_lds(unsigned char* smem, int N, int S, int size) {
  const int N2 = N * N, SN2 = S * N2;
  ZofeLds l;
  l.z = reinterpret_cast<double2*>(smem);
  l.Os = l.z + size;
  l.T1 = l.Os + SN2;
  l.T2 = l.T1 + SN2;
  l.U = l.T2 + SN2;
  l.X1 = l.U + SN2;
  l.X2 = l.X1 + SN2;
  l.bm = l.X2 + SN2;
  l.gm = l.bm + N2;
  l.be = l.gm + N2;
  l.fb = l.be + N2;
  return l;
}

// Phases (1)-(3) of a stage at stage input l.z: Osum; the per-site products
// L rho, Osum rho, L^+ Osum; the triple products, a = sum_s L_s^+ Osum_s and
// b, i H - a^+ and be = unit b + sum_i f_i V_i.  Every sum over sites or pseudomodes is split
// into per-site items, so no item loops longer than max(N, S, P).  Ends
// after a barrier.
template <int NT>
__device__ __forceinline__ void zofe_prepare(const ZofeArgs& a, const ZofeLds& l,
                                             const double2* Hb, const double2* Lb,
                                             const double2* Vb, const double2* fs, int tid) {
  const int N = a.N, S = a.S, P = a.P, N2 = N * N, SN2 = S * N2;
  const double2* rho = l.z;
  for (int it = tid; it < SN2; it += NT) {             // (1) Osum
    const int s = it / N2, ij = it % N2;
    double2 v = make_double2(0.0, 0.0);
    for (int p = 0; p < P; ++p) v = cadd(v, l.z[N2 + (p * S + s) * N2 + ij]);
    l.Os[it] = v;
  }
  __syncthreads();
  for (int it = tid; it < 3 * SN2; it += NT) {         // (2) L rho, Osum rho, L^+ Osum
    const int which = it / SN2, w = it % SN2;
    const int s = w / N2, i = (w % N2) / N, k = w % N;
    const double2* Ls = Lb + s * N2;
    const double2* Oss = l.Os + s * N2;
    double2 v = make_double2(0.0, 0.0);
    if (which == 0) {
      for (int j = 0; j < N; ++j) cfmad(v, Ls[i * N + j], rho[j * N + k]);
      l.T1[w] = v;
    } else if (which == 1) {
      for (int j = 0; j < N; ++j) cfmad(v, Oss[i * N + j], rho[j * N + k]);
      l.T2[w] = v;
    } else {
      for (int j = 0; j < N; ++j) cfmad(v, cconj(Ls[j * N + i]), Oss[j * N + k]);
      l.U[w] = v;
    }
  }
  __syncthreads();
  for (int it = tid; it < 2 * SN2 + N2; it += NT) {    // (3) X1, X2, b, i H - a^+, be
    if (it < 2 * SN2) {
      const int which = it / SN2, w = it % SN2;
      const int s = w / N2, i = (w % N2) / N, j = w % N;
      double2 v = make_double2(0.0, 0.0);
      if (which == 0) {                                   // (L rho)(Osum)^+
        for (int k = 0; k < N; ++k)
          cfmad(v, l.T1[s * N2 + i * N + k], cconj(l.Os[s * N2 + j * N + k]));
        l.X1[w] = v;
      } else {                                            // (Osum rho) L^+
        for (int k = 0; k < N; ++k)
          cfmad(v, l.T2[s * N2 + i * N + k], cconj(Lb[(s * N + j) * N + k]));
        l.X2[w] = v;
      }
    } else {
      const int ik = it - 2 * SN2, i = ik / N, k = ik % N;
      double2 v = make_double2(0.0, 0.0);
      for (int s = 0; s < S; ++s) v = cadd(v, l.U[s * N2 + ik]);     // a_ik
      const double2 hh = Hb[ik], hk = Hb[k * N + i];
      const double2 bik = make_double2(hh.y - v.x, -hh.x - v.y);     // -i H_ik - a_ik
      l.bm[ik] = bik;
      l.gm[k * N + i] = make_double2(-hk.y - v.x, hk.x + v.y);      // i H_ki - conj(a_ik)
      double2 e = make_double2(a.unit * bik.x, a.unit * bik.y);
      for (int fi = 0; fi < a.nfield; ++fi) e = cadd(e, cmul(fs[fi], Vb[fi * N2 + ik]));
      l.be[ik] = e;
    }
  }
  __syncthreads();
}

// d rho_ij / dt (element e = i N + j of the row-major rho), with the field
// terms; needs the intermediates of zofe_prepare.
__device__ __forceinline__ double2 zofe_rho_derivative(const ZofeArgs& a, const ZofeLds& l,
                                                       const double2* Vb, const double2* fs,
                                                       int e) {
  const int N = a.N, S = a.S, N2 = N * N;
  const bool ham_h = a.flags & 1, rho_h = a.flags & 2;
  const double2* rho = l.z;
  const int i = e / N, j = e % N;
  double2 d = make_double2(0.0, 0.0), f = make_double2(0.0, 0.0);
  if (rho_h && ham_h) {                                 // f = d^+
    double2 dt = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) cfmad(dt, l.bm[j * N + k], rho[k * N + i]);
    for (int s = 0; s < S; ++s) dt = cadd(dt, l.X1[s * N2 + j * N + i]);
    f = cconj(dt);
  } else if (ham_h) {                                   // rho b^+
    for (int k = 0; k < N; ++k) cfmad(f, rho[i * N + k], cconj(l.bm[j * N + k]));
  } else {                                              // rho (i H - a^+)
    for (int k = 0; k < N; ++k) cfmad(f, rho[i * N + k], l.gm[k * N + j]);
  }
  if (!rho_h) {                                         // + sum_s Osum rho L^+
    for (int s = 0; s < S; ++s) f = cadd(f, l.X2[s * N2 + e]);
  } else if (!ham_h) {                                  // + c^+_ij = conj(c_ji)
    double2 cc = make_double2(0.0, 0.0);
    for (int s = 0; s < S; ++s) cc = cadd(cc, l.X1[s * N2 + j * N + i]);
    f = cadd(f, cconj(cc));
  }
  for (int k = 0; k < N; ++k) cfmad(d, l.bm[i * N + k], rho[k * N + j]);
  for (int s = 0; s < S; ++s) d = cadd(d, l.X1[s * N2 + e]);
  double2 v = cadd(d, f);
  v = make_double2(a.unit * v.x, a.unit * v.y);
  for (int fi = 0; fi < a.nfield; ++fi) {               // f_i [V_i, rho]
    const double2* Vi = Vb + fi * N2;
    double2 cm = make_double2(0.0, 0.0);
    for (int k = 0; k < N; ++k) {
      cfmad(cm, Vi[i * N + k], rho[k * N + j]);
      const double2 t = cmul(rho[i * N + k], Vi[k * N + j]);
      cm.x -= t.x;
      cm.y -= t.y;
    }
    v = cadd(v, cmul(fs[fi], cm));
  }
  return v;
}

// RK4 bookkeeping of one element: returns the next stage value, updates
// the accumulator an (stages 0-2) or the state yv (stage 3).
__device__ __forceinline__ double2 zofe_rk4(double2 v, double2& yv, double2& an, int st,
                                            double h) {
  const double2 acc = caxpy(kZW[st] * h / 6.0, v, st == 0 ? yv : an);
  if (st < 3) {
    an = acc;
    return caxpy(kZC[st] * h, v, yv);
  }
  yv = acc;
  return acc;
}

// Saves of the state in l.z at save point s: the projections of vec(rho) or
// the whole state (reference layout).
template <int NT>
__device__ __forceinline__ void zofe_save(const ZofeArgs& a, const ZofeLds& l,
                                          const double2* Pb, double2* ob, int s, int tid,
                                          bool& bad) {
  const int N = a.N, N2 = N * N;
  if (Pb) {
    for (int k = tid; k < a.K; k += NT) {
      double2 v = make_double2(0.0, 0.0);
      for (int e = 0; e < N2; ++e) cfmad(v, Pb[(int64_t)k * N2 + (e % N) * N + e / N], l.z[e]);
      ob[(int64_t)s * a.K + k] = v;
      if (!cfinite(v)) bad = true;
    }
  } else {
    for (int e = tid; e < a.size; e += NT) {
      const double2 v = l.z[e];
      ob[(int64_t)s * a.size + zofe_ref_index(e, N, a.S, a.P)] = v;
      if (!cfinite(v)) bad = true;
    }
  }
}

// Per-element kernel (any N): thread t owns the elements t, t + NT, ...;
// their state y, RK4 accumulator and next stage value sit in a workspace
// slice of the right-hand side, read and written by the owner only.
template <int NT>
__global__ __launch_bounds__(NT) void zofe_kernel(ZofeArgs a, const double2* __restrict__ y0,
                                                  const int* __restrict__ nsub,
                                                  const double* __restrict__ hs, int nsave,
                                                  double2* __restrict__ out,
                                                  double2* __restrict__ work,
                                                  int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = a.N, S = a.S, P = a.P, N2 = N * N, size = a.size;
  const ZofeLds l = zofe_lds(smem, N, S, size);
  const int tid = threadIdx.x;
  const int r = blockIdx.x, b = blockIdx.y;
  const double2* Hb = a.H + (int64_t)b * a.strideH;
  const double2* Lb = a.L + (int64_t)b * a.strideL;
  const double2* Gb = a.G + (int64_t)b * a.strideG;
  const double2* Wb = a.W + (int64_t)b * a.strideG;
  const double2* Vb = a.V ? a.V + (int64_t)b * a.strideV : nullptr;
  const double2* Fb = a.F ? a.F + (int64_t)b * a.strideF : nullptr;
  const int nfld = a.nfield;
  const bool loader = tid < 3 * nfld;

  const double2* yb0 = y0 + ((int64_t)b * a.R + r) * size;
  double2* yw = work + ((int64_t)b * a.R + r) * 3 * size;   // state y
  double2* aw = yw + size;                                  // RK4 accumulator
  double2* zw = aw + size;                                  // next stage value
  for (int e = tid; e < size; e += NT) {
    const double2 v = yb0[zofe_ref_index(e, N, S, P)];
    yw[e] = v;
    l.z[e] = v;
  }
  if (loader && a.nF > 0) l.fb[tid] = Fb[tid];
  __syncthreads();

  auto stage = [&](const double2* fs, int st, double h) {
    zofe_prepare<NT>(a, l, Hb, Lb, Vb, fs, tid);
    // the workspace entries of the next owned element are loaded while the
    // current one is computed
    double2 y_next = make_double2(0.0, 0.0), a_next = make_double2(0.0, 0.0);
    if (!a.apply && tid < size) {
      y_next = yw[tid];
      a_next = st == 0 ? y_next : aw[tid];
    }
    for (int e = tid; e < size; e += NT) {                 // (4) owned elements
      double2 yv = y_next, an = a_next;
      if (!a.apply && e + NT < size) {
        y_next = yw[e + NT];
        a_next = st == 0 ? y_next : aw[e + NT];
      }
      double2 v;
      if (e < N2) {
        v = zofe_rho_derivative(a, l, Vb, fs, e);
      } else {
        const int rr = e - N2, ps = rr / N2, ij = rr % N2, s = ps % S;
        const int i = ij / N, j = ij % N;
        const double2* X = l.z + N2 + ps * N2;
        const double2 g = cmul(Gb[ps], Lb[s * N2 + ij]), wx = cmul(Wb[ps], X[ij]);
        v = make_double2(a.unit * (g.x - wx.x), a.unit * (g.y - wx.y));
        for (int k = 0; k < N; ++k) {                       // be X - X be
          cfmad(v, l.be[i * N + k], X[k * N + j]);
          const double2 t = cmul(X[i * N + k], l.be[k * N + j]);
          v.x -= t.x;
          v.y -= t.y;
        }
      }
      if (a.apply) {                                        // derivative only
        zw[e] = v;
        continue;
      }
      zw[e] = zofe_rk4(v, yv, an, st, h);
      if (st < 3) aw[e] = an;
      else yw[e] = yv;
    }
    __syncthreads();                                       // all reads of z done
  };

  int64_t g = 0;
  bool bad = false;
  if (a.apply) {                                           // out = F(y0), no fields
    stage(l.fb, 0, 0.0);
    double2* ob = out + ((int64_t)b * a.R + r) * size;
    for (int e = tid; e < size; e += NT) ob[zofe_ref_index(e, N, S, P)] = zw[e];
    return;
  }
  double2* ob = out + ((int64_t)b * a.R + r) * nsave * (a.proj ? a.K : size);
  const double2* Pb = a.proj ? a.proj + (int64_t)b * a.strideP : nullptr;
  const int vz = opaque_zero();
  int n_next = nsub[vz];
  double h_next = hs[vz];
  for (int s = 0; s < nsave; ++s) {
    const int n = max(wave_uniform(n_next), 0);
    const double h = wave_uniform(h_next);
    if (s + 1 < nsave) {
      n_next = nsub[s + 1 + vz];
      h_next = hs[s + 1 + vz];
    }
    for (int step = 0; step < n; ++step, ++g) {
      if (nfld > 0 && g >= a.nF) {
        if (tid == 0) status[0] = 2;
        return;
      }
      double2 fnext = make_double2(0.0, 0.0);
      if (loader) fnext = Fb[(g + 1 < a.nF ? g + 1 : g) * 3 * nfld + tid];
      const double2* fc = l.fb + (g & 1) * 3 * kZofeMaxFields;
      for (int st = 0; st < 4; ++st) {
        const int ti = st == 0 ? 0 : (st == 3 ? 2 : 1);
        stage(fc + ti * nfld, st, h);                      // field values in LDS
        for (int e = tid; e < size; e += NT) l.z[e] = zw[e];
        if (st == 3 && loader) l.fb[((g + 1) & 1) * 3 * kZofeMaxFields + tid] = fnext;
        __syncthreads();
      }
    }
    zofe_save<NT>(a, l, Pb, ob, s, tid, bad);
  }
  if (bad) status[0] = 1;
}

inline size_t zofe_lds_bytes(int N, int S, int size) {
  const size_t N2 = (size_t)N * N;
  return ((size_t)size + 6 * (size_t)S * N2 + 3 * N2 + 2 * 3 * kZofeMaxFields) *
         sizeof(double2);
}

}  // namespace
}  // namespace qsx

static int32_t zofe_launch(int apply,
    int32_t B, int32_t R, int32_t N, int32_t S, int32_t P, const void* H, int64_t strideH,
    const void* L, int64_t strideL, const void* Gamma, const void* w, int64_t strideG,
    int32_t flags, double unit_convert, int32_t nfield, const void* V, int64_t strideV,
    const void* F, int64_t strideF, int64_t nF, const void* y0, const int32_t* nsub,
    const double* h, int32_t nsave, int32_t K, const void* proj, int64_t strideP, void* out,
    void* workspace, int32_t* status, void* stream);


extern "C" int64_t qsx_zofe_workspace_bytes(int32_t B, int32_t R, int32_t N, int32_t S,
                                            int32_t P) {
  if (B < 1 || R < 1 || N < 1 || S < 1 || P < 1) return -1;
  return 3 * (int64_t)B * R * N * N * (1 + (int64_t)P * S) * 16;
}

extern "C" int32_t qsx_zofe_apply(int32_t B, int32_t R, int32_t N, int32_t S, int32_t P,
                                  const void* H, int64_t strideH, const void* L,
                                  int64_t strideL, const void* Gamma, const void* w,
                                  int64_t strideG, int32_t flags, double unit_convert,
                                  const void* y, void* out, void* workspace, void* stream) {
  // the apply launch returns before it reads a schedule or writes a status
  return zofe_launch(1, B, R, N, S, P, H, strideH, L, strideL, Gamma, w, strideG, flags,
                     unit_convert, 0, nullptr, 0, nullptr, 0, 0, y, nullptr, nullptr, 1, 0,
                     nullptr, 0, out, workspace, nullptr, stream);
}

extern "C" int32_t qsx_zofe_propagate(
    int32_t B, int32_t R, int32_t N, int32_t S, int32_t P, const void* H, int64_t strideH,
    const void* L, int64_t strideL, const void* Gamma, const void* w, int64_t strideG,
    int32_t flags, double unit_convert, int32_t nfield, const void* V, int64_t strideV,
    const void* F, int64_t strideF, int64_t nF, const void* y0, const int32_t* nsub,
    const double* h, int32_t nsave, int32_t K, const void* proj, int64_t strideP, void* out,
    void* workspace, int32_t* status, void* stream) {
  return zofe_launch(0, B, R, N, S, P, H, strideH, L, strideL, Gamma, w, strideG, flags,
                     unit_convert, nfield, V, strideV, F, strideF, nF, y0, nsub, h, nsave, K,
                     proj, strideP, out, workspace, status, stream);
}

static int32_t zofe_launch(int apply,
    int32_t B, int32_t R, int32_t N, int32_t S, int32_t P, const void* H, int64_t strideH,
    const void* L, int64_t strideL, const void* Gamma, const void* w, int64_t strideG,
    int32_t flags, double unit_convert, int32_t nfield, const void* V, int64_t strideV,
    const void* F, int64_t strideF, int64_t nF, const void* y0, const int32_t* nsub,
    const double* h, int32_t nsave, int32_t K, const void* proj, int64_t strideP, void* out,
    void* workspace, int32_t* status, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(H && L && Gamma && w && y0 && out && workspace &&
                    (apply || (nsub && h && status)),
                "qsx_zofe_propagate: null pointer");
  QSX_CHECK_ARG(B >= 1 && R >= 1 && N >= 1 && S >= 1 && P >= 1 && nsave >= 1 && B <= 65535,
                "qsx_zofe_propagate: bad sizes B=%d R=%d N=%d S=%d P=%d", B, R, N, S, P);
  QSX_CHECK_ARG(nfield >= 0 && nfield <= kZofeMaxFields,
                "qsx_zofe_propagate: %d field terms (0..%d)", nfield, kZofeMaxFields);
  QSX_CHECK_ARG(nfield == 0 || (V && F && nF >= 0 && (strideF == 0 || strideF >= nF * 3 * nfield)),
                "qsx_zofe_propagate: field terms need V and F");
  QSX_CHECK_ARG(K == 0 || (proj && K >= 1), "qsx_zofe_propagate: projection rows need proj");
  const int64_t size64 = (int64_t)N * N * (1 + (int64_t)P * S);
  QSX_CHECK_ARG(size64 <= 8192, "qsx_zofe_propagate: state of %lld elements exceeds 8192",
                (long long)size64);
  const int size = (int)size64;
  const size_t shmem = zofe_lds_bytes(N, S, size);
  QSX_CHECK_ARG(shmem <= 160 * 1024, "qsx_zofe_propagate: %zu bytes of LDS", shmem);
  ZofeArgs a;
  a.N = N;
  a.S = S;
  a.P = P;
  a.size = size;
  a.flags = flags;
  a.nfield = nfield;
  a.K = proj ? K : 0;
  a.R = R;
  a.apply = apply;
  a.unit = unit_convert;
  a.H = (const double2*)H;
  a.strideH = strideH;
  a.L = (const double2*)L;
  a.strideL = strideL;
  a.G = (const double2*)Gamma;
  a.W = (const double2*)w;
  a.strideG = strideG;
  a.V = nfield ? (const double2*)V : nullptr;
  a.strideV = strideV;
  a.F = nfield ? (const double2*)F : nullptr;
  a.strideF = strideF;
  a.nF = nfield ? nF : 0;
  a.proj = (const double2*)proj;
  a.strideP = strideP;
  hipStream_t st = as_stream(stream);
  dim3 grid(R, B);
#define QSX_ZOFE(NT_, LIMIT_)                                                                  \
  if (size <= LIMIT_) {                                                                        \
    auto kern = zofe_kernel<NT_>;                                                              \
    if (shmem > 64 * 1024)                                                                     \
      QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                     \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)); \
    hipLaunchKernelGGL(kern, grid, dim3(NT_), shmem, st, a, (const double2*)y0, nsub, h,       \
                       nsave, (double2*)out, (double2*)workspace, status);                     \
    QSX_CHECK_HIP(hipGetLastError());                                                          \
    return QSX_OK;                                                                             \
  }
  QSX_ZOFE(256, 1024)
  QSX_ZOFE(1024, 8192)
#undef QSX_ZOFE
  set_error("qsx_zofe_propagate: no kernel for %d elements", size);
  return QSX_EINVAL;
}
