// Batched FP64 construction of Redfield / unitary Liouvillian blocks (gfx950).
//
// Inputs per problem b: the non-rotating one-exciton Hamiltonian H1_b
// (n x n, real symmetric) and the rotating-frame frequency rw_b.  Hilbert
// blocks: g (1 state), e (n), f (n(n-1)/2 pairs i<j, hard-core bosons).
// The Liouvillian is block-diagonal over Liouville blocks XY (rho_ab with
// a in X, b in Y); one call builds one block, M = nX * nY, with the local
// index p = a + nX * b (column stacking, the reference's vec order).
//
// Pipeline (one kernel each, intermediates in the caller's workspace):
//   1. block_eigh:  Jacobi eigensystem of H_X - x*rw (x = excitation number)
//   2. block_bath:  C_X[d][c] = C(E_d - E_c) (Debye bath), K_i = U^T P_i U
//   3. block_gs:    Gs_X[a][c] = sum_{i,m} K_i[a][m] K_i[m][c] C_X[c][m]
//   4. liouv_cols:  column q of L_eig = -i dE - R, then W = T L_eig (pass 1)
//   5. liouv_rows:  L_site = W T^T (pass 2), T = kron(U_Y, U_X)
// following qspectra/dynamics/redfield.py:9-104 (May & Kuhn 3.319/3.322)
// and qspectra/bath.py:17-31, 84-102.
//
// Dense mode (qsx_redfield_block_dense, qsx_unitary_block_dense,
// qsx_dense_eigh): the Hilbert blocks are given as dense rotating-frame
// matrices and the system-bath couplings as diagonals over the block's basis
// (the vibronic Hamiltonian of qspectra/hamiltonian.py:633-807: electronic
// number operators extended by the identity on the vibrations); the
// pipeline is the same from the eigensystem on.

#include "qsx_common.h"

#include <algorithm>

namespace qsx {
namespace {

constexpr int kMaxBlock = 64;  // Jacobi in LDS
constexpr int kThreads = 256;

__host__ __device__ inline int block_size(int X, int n) {
  return X == 0 ? 1 : (X == 1 ? n : n * (n - 1) / 2);
}

// per-(problem, Hilbert block) record, in doubles:
//   E[nb] | U[nb*nb] | C[2*nb*nb] | Gs[2*nb*nb] | K[n*nb*nb]
// rounded up to an even count so that every record (and its complex C, Gs,
// at the even offset nb*(nb+1)) stays 16-byte aligned
__host__ __device__ inline size_t rec_doubles(int nb, int n) {
  return ((size_t)nb + 5 * (size_t)nb * nb + (size_t)n * nb * nb + 1) & ~(size_t)1;
}
struct Rec {
  double* E;
  double* U;
  double2* C;
  double2* Gs;
  double* K;
};
__host__ __device__ inline Rec rec_at(double* base, int nb, int n) {
  Rec r;
  r.E = base;
  r.U = base + nb;
  r.C = reinterpret_cast<double2*>(base + nb + (size_t)nb * nb);
  r.Gs = r.C + (size_t)nb * nb;
  r.K = reinterpret_cast<double*>(r.Gs + (size_t)nb * nb);
  return r;
}

// f-block state s -> site pair (i, j), i < j, lexicographic (operator_tools.py:122-152)
__device__ inline void pair_of(int s, int n, int& i, int& j) {
  int k = s;
  i = 0;
  while (k >= n - 1 - i) {
    k -= n - 1 - i;
    ++i;
  }
  j = i + 1 + k;
}

// ------------------------------------------------------------------ 1. eigh
// Cyclic two-sided Jacobi on the symmetric m x m matrix A in LDS (m even),
// accumulating the rotations into V: m/2 disjoint rotations per round
// (circle method), sweeps until the off-diagonal norm is below 1e-30 of the
// total.  On return A is diagonal (the eigenvalues) and the columns of V are
// the eigenvectors.  Called by every thread of an NT-thread work-group; red
// holds 2 NT / 64 doubles.
template <int NT>
__device__ void jacobi_sweeps(double* A, double* V, int m, double (*rot)[2], int (*pqs)[2],
                              double* red) {
  const int tid = threadIdx.x, h = m / 2;
  // q / h and q / m for the small q here, exact: the quotient's fraction
  // stays at least 1 / (2 m) from an integer
  const float rh = 1.0f / (float)h, rm = 1.0f / (float)m;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < m * m; e += NT) {
      const double a = A[e];
      const int r = __float2int_rz(((float)e + 0.5f) * rm);
      tot += a * a;
      if (e != r * (m + 1)) off += a * a;
    }
    for (int o = 32; o > 0; o >>= 1) {                    // within each wavefront
      off += __shfl_xor(off, o);
      tot += __shfl_xor(tot, o);
    }
    if ((tid & 63) == 0) {
      red[tid >> 6] = off;
      red[NT / 64 + (tid >> 6)] = tot;
    }
    __syncthreads();
    off = 0.0;
    tot = 0.0;
    for (int q = 0; q < NT / 64; ++q) {
      off += red[q];
      tot += red[NT / 64 + q];
    }
    __syncthreads();
    if (!(off > 1e-30 * tot)) break;
    for (int round = 0; round < m - 1; ++round) {
      if (tid < h) {
        auto idx = [&](int j) { return ((j + round) % (m - 1)) + 1; };
        int p = tid == 0 ? 0 : idx(tid);
        int q = tid == 0 ? idx(0) : idx(m - 1 - tid);
        if (p > q) { const int t = p; p = q; q = t; }
        const double apq = A[p * m + q];
        double c = 1.0, s = 0.0;
        if (apq != 0.0) {
          // t = sgn(theta) / (|theta| + sqrt(theta^2 + 1)), theta = d / e,
          // with numerator and denominator times |e|: one division
          const double d = A[q * m + q] - A[p * m + p], e = 2.0 * apq;
          const double num = d > 0.0 ? e : (d < 0.0 ? -e : fabs(e));
          const double t = num / (fabs(d) + sqrt(d * d + e * e));
          c = rsqrt(t * t + 1.0);
          s = t * c;
        }
        pqs[tid][0] = p;
        pqs[tid][1] = q;
        rot[tid][0] = c;
        rot[tid][1] = s;
      }
      __syncthreads();
      // A <- J^T A J by 2 x 2 blocks (rows of rotation k1, columns of
      // rotation k2: each block read and written by one thread; the pivot
      // blocks k1 = k2 lose their off-diagonal pair), V <- V J by row pairs
      for (int w = tid; w < h * h + m * h; w += NT) {
        if (w < h * h) {
          const int k1 = __float2int_rz(((float)w + 0.5f) * rh), k2 = w - k1 * h;
          const int p1 = pqs[k1][0], q1 = pqs[k1][1], p2 = pqs[k2][0], q2 = pqs[k2][1];
          const double c1 = rot[k1][0], s1 = rot[k1][1], c2 = rot[k2][0], s2 = rot[k2][1];
          const double app = A[p1 * m + p2], apq = A[p1 * m + q2];
          const double aqp = A[q1 * m + p2], aqq = A[q1 * m + q2];
          const double bpp = c2 * app - s2 * apq, bpq = s2 * app + c2 * apq;
          const double bqp = c2 * aqp - s2 * aqq, bqq = s2 * aqp + c2 * aqq;
          A[p1 * m + p2] = c1 * bpp - s1 * bqp;
          A[q1 * m + q2] = s1 * bpq + c1 * bqq;
          A[p1 * m + q2] = k1 == k2 ? 0.0 : c1 * bpq - s1 * bqq;
          A[q1 * m + p2] = k1 == k2 ? 0.0 : s1 * bpp + c1 * bqp;
        } else {
          const int v = w - h * h;
          const int r = __float2int_rz(((float)v + 0.5f) * rh), k = v - r * h;
          const int p = pqs[k][0], q = pqs[k][1];
          const double c = rot[k][0], s = rot[k][1];
          const double vp = V[r * m + p], vq = V[r * m + q];
          V[r * m + p] = c * vp - s * vq;
          V[r * m + q] = s * vp + c * vq;
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kThreads) void block_eigh_kernel(const double* __restrict__ H1, int n,
                                                              int X, const double* __restrict__ rw,
                                                              const double* __restrict__ Hd,
                                                              int nbd, double* __restrict__ ws,
                                                              size_t rstride) {
  __shared__ double A[kMaxBlock * kMaxBlock];
  __shared__ double V[kMaxBlock * kMaxBlock];
  __shared__ double rot[kMaxBlock / 2][2];
  __shared__ int pqs[kMaxBlock / 2][2];
  __shared__ int pi_[kMaxBlock], pj_[kMaxBlock];
  __shared__ double red[2][kThreads];
  __shared__ int perm[kMaxBlock];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int nb = Hd ? nbd : block_size(X, n);
  const int m = nb + (nb & 1);
  const double* H = Hd ? Hd + (size_t)b * nb * nb : H1 + (size_t)b * n * n;
  const double shift = Hd ? 0.0 : X * (rw ? rw[b] : 0.0);
  if (!Hd && X == 2)
    for (int s = tid; s < nb; s += kThreads) pair_of(s, n, pi_[s], pj_[s]);
  __syncthreads();
  for (int e = tid; e < m * m; e += kThreads) {
    const int r = e / m, c = e % m;
    double v = 0.0;
    if (r < nb && c < nb) {
      if (Hd) {
        v = H[r * nb + c];
      } else if (X == 1) {
        v = H[r * n + c];
      } else if (X == 2) {  // operator_1_to_2 (operator_tools.py:155-191)
        const int i = pi_[r], j = pj_[r], k = pi_[c], l = pj_[c];
        v = (i == k ? H[j * n + l] : 0.0) + (i == l ? H[j * n + k] : 0.0) +
            (j == k ? H[i * n + l] : 0.0) + (j == l ? H[i * n + k] : 0.0);
      }
      if (r == c) v -= shift;
    }
    A[e] = v;
    V[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  jacobi_sweeps<kThreads>(A, V, m, rot, pqs, &red[0][0]);
  // ascending eigenvalues (stable ranking), canonical sign: the largest-
  // magnitude component of each eigenvector is made positive
  for (int k = tid; k < nb; k += kThreads) {
    const double vk = A[k * m + k];
    int rank = 0;
    for (int l = 0; l < nb; ++l) {
      const double vl = A[l * m + l];
      rank += (vl < vk) || (vl == vk && l < k);
    }
    perm[rank] = k;
  }
  __syncthreads();
  Rec rec = rec_at(ws + (size_t)b * rstride, nb, n);
  for (int k = tid; k < nb; k += kThreads) {
    const int c = perm[k];
    int best = 0;
    double bv = -1.0;
    for (int r = 0; r < nb; ++r)
      if (fabs(V[r * m + c]) > bv) {
        bv = fabs(V[r * m + c]);
        best = r;
      }
    const double sg = V[best * m + c] < 0.0 ? -1.0 : 1.0;
    rec.E[k] = A[c * m + c];
    for (int r = 0; r < nb; ++r) rec.U[r * nb + k] = sg * V[r * m + c];
  }
}

// ------------------------------------------------------------ 2. C(w) and K
// Debye bath correlation function, qspectra/bath.py:84-102 (complex, 1000-term
// Matsubara sum, k = 0 term zero) and bath.py:17-31 (real, detailed balance).
// 1 / d from the hardware reciprocal and two Newton steps: no shared
// condition register, unlike the IEEE division sequence, so several of them
// overlap; for the well-scaled denominators of the Matsubara terms (no
// overflow, underflow or zero) it is within an ulp of the quotient
__device__ __forceinline__ double recip_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

// The terms k = k0, k0 + stride, ... (< 1000) of the Matsubara sum: four
// interleaved Kahan-compensated sums (their reciprocals independent, so
// they overlap), added at the end
__device__ double2 matsubara_partial(double x, double T, double gam, int k0, int stride) {
  double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
  double cr[4] = {0.0, 0.0, 0.0, 0.0}, ci[4] = {0.0, 0.0, 0.0, 0.0};
  const double twopiT = 2.0 * M_PI * T;
  for (int k = k0; k < 1000; k += 4 * stride) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int ku = k + u * stride;
      if (ku < 1000) {
        const double nu = twopiT * ku;
        const double inv = recip_nr((nu * nu - gam * gam) * (nu * nu + x * x));
        const double tr = nu * nu * inv - cr[u];
        const double ti = nu * x * inv - ci[u];
        const double ur = sr[u] + tr, ui = si[u] + ti;
        cr[u] = (ur - sr[u]) - tr;
        ci[u] = (ui - si[u]) - ti;
        sr[u] = ur;
        si[u] = ui;
      }
    }
  }
  return make_double2((sr[0] + sr[1]) + (sr[2] + sr[3]), (si[0] + si[1]) + (si[2] + si[3]));
}

// sum_{k=1}^{999} nu_k / ((nu_k^2 - g^2)(nu_k - i x)), nu_k = 2 pi k T,
// compensated as above; its value at -x is the complex conjugate (each
// term's imaginary part is odd in x, its real part even), exactly so in
// floating point since only signs change
__device__ double2 matsubara_sum(double x, double T, double gam) {
  return matsubara_partial(x, T, gam, 1, 1);
}

// digamma of a complex argument away from the poles 0, -1, -2, ...: the
// recurrence psi(z) = psi(z + 1) - 1/z until Re z >= 20 (at most 4096
// steps), then the asymptotic series ln z - 1/(2z) - sum_k B_2k / (2k z^2k)
// through k = 6 (the first omitted term is below 1e-19 there)
__device__ double2 digamma_c(double zr, double zi) {
  double sr = 0.0, si = 0.0;
  for (int i = 0; i < 4096 && zr < 20.0; ++i) {
    const double d = 1.0 / (zr * zr + zi * zi);
    sr -= zr * d;
    si += zi * d;
    zr += 1.0;
  }
  const double d = 1.0 / (zr * zr + zi * zi);
  const double wr = zr * d, wi = -zi * d;                     // w = 1/z
  const double w2r = wr * wr - wi * wi, w2i = 2.0 * wr * wi;  // w^2
  double tr = 691.0 / 32760.0, ti = 0.0;
  const double cf[5] = {1.0 / 132.0, 1.0 / 240.0, 1.0 / 252.0, 1.0 / 120.0, 1.0 / 12.0};
#pragma unroll
  for (int k = 0; k < 5; ++k) {                                // t <- c_k - w^2 t
    const double ur = cf[k] - (w2r * tr - w2i * ti);
    const double ui = -(w2r * ti + w2i * tr);
    tr = ur;
    ti = ui;
  }
  const double fr = w2r * tr - w2i * ti, fi = w2r * ti + w2i * tr;   // w^2 t
  return make_double2(sr + 0.5 * log(zr * zr + zi * zi) - 0.5 * wr - fr,
                      si + atan2(zi, zr) - 0.5 * wi - fi);
}

// p = gam / (2 pi T) inside which matsubara_closed is used (relative error
// below 1e-13 of the summed form there; outside, the sum)
__device__ __forceinline__ bool matsubara_closed_ok(double T, double gam) {
  const double p = gam / (2.0 * M_PI * T);
  return T > 0.0 && p >= 0.01 && p <= 32.0;
}

// The same sum in closed form: with nu_k = a k, a = 2 pi T, p = gam / a and
// q = i x / a, each term is (k / a^2) / ((k - p)(k + p)(k - q)); in partial
// fractions A/(k - p) + B/(k + p) + C/(k - q) with A = 1/(2(p - q)),
// B = -1/(2(p + q)), C = q/(q^2 - p^2), and sum_{k=1}^{999} 1/(k - z) =
// psi(1000 - z) - psi(1 - z).  dm and dp are the brackets at z = p and
// z = -p (real, per problem).
__device__ double2 matsubara_closed(double x, double T, double dm, double dp, double gam) {
  const double a = 2.0 * M_PI * T, p = gam / a, y = x / a;
  const double r2 = p * p + y * y;
  const double Ar = 0.5 * p / r2, Ai = 0.5 * y / r2;          // 1/(2(p - iy))
  const double Br = -0.5 * p / r2, Bi = 0.5 * y / r2;         // -1/(2(p + iy))
  const double Ci = -y / r2;                                  // iy / (-y^2 - p^2)
  const double2 hi = digamma_c(1000.0, -y), lo = digamma_c(1.0, -y);
  const double Dr = hi.x - lo.x, Di = hi.y - lo.y;
  const double inv = 1.0 / (a * a);
  return make_double2((Ar * dm + Br * dp - Ci * Di) * inv, (Ai * dm + Bi * dp + Ci * Dr) * inv);
}

// C(x) given the Matsubara sum S(x) (unused for the real correlation
// function and at x = 0)
__device__ double2 debye_corr_with(double x, double2 S, double T, double lam, double gam,
                                   int real_corr) {
  if (real_corr) {
    if (x == 0.0) return make_double2(T * 2.0 * lam / gam, 0.0);
    const double J = 2.0 * lam * gam * x / (gam * gam + x * x);  // odd: J_anti = J
    return make_double2((1.0 / expm1(x / T) + 1.0) * J, 0.0);
  }
  if (x == 0.0) return make_double2(lam * 2.0 * T / gam, -lam);
  // (cot(g/2T) - i) / (g - i x) + 4T S(x)
  const double ct = 1.0 / tan(gam / (2.0 * T));
  const double den = gam * gam + x * x;
  const double re = (ct * gam + x) / den + 4.0 * T * S.x;
  const double im = (ct * x - gam) / den + 4.0 * T * S.y;
  return make_double2(lam * gam * re, lam * gam * im);
}

__global__ __launch_bounds__(kThreads) void block_bath_kernel(int n, int X, int nb,
                                                              const double* __restrict__ occ,
                                                              double T, double lam,
                                                              double gam, int real_corr,
                                                              double* __restrict__ ws,
                                                              size_t rstride,
                                                              const double2* __restrict__ ctab) {
  const int b = blockIdx.x, tid = threadIdx.x;
  Rec rec = rec_at(ws + (size_t)b * rstride, nb, n);
  // a tabulated bath: C[d][c] = C(E_d - E_c) given per problem
  if (ctab != nullptr)
    for (int e = tid; e < nb * nb; e += kThreads) rec.C[e] = ctab[(size_t)b * nb * nb + e];
  // pairs (d, c) with d <= c: one Matsubara sum serves C[d][c] and C[c][d]
  const int npairs = ctab != nullptr ? 0 : nb * (nb + 1) / 2;
  for (int e = tid; e < npairs; e += kThreads) {
    int d = 0, r = e;
    while (r >= nb - d) {
      r -= nb - d;
      ++d;
    }
    const int c = d + r;
    const double x = rec.E[d] - rec.E[c];
    const bool sum = !real_corr && x != 0.0;
    const double2 S = sum ? matsubara_sum(x, T, gam) : make_double2(0.0, 0.0);
    rec.C[d * nb + c] = debye_corr_with(x, S, T, lam, gam, real_corr);
    if (c != d) rec.C[c * nb + d] = debye_corr_with(-x, make_double2(S.x, -S.y), T, lam, gam,
                                                    real_corr);
  }
  // K_i[a][m] = sum over states s of X that contain site i of U[s][a] U[s][m]
  const int nk = n * nb * nb;
  for (int e = tid; e < nk; e += kThreads) {
    const int i = e / (nb * nb), a = (e / nb) % nb, m = e % nb;
    double v = 0.0;
    if (occ) {                       // diagonal coupling: occ[i][s] over the block basis
      for (int s2 = 0; s2 < nb; ++s2) {
        const double o = occ[(size_t)i * nb + s2];
        if (o != 0.0) v += o * rec.U[s2 * nb + a] * rec.U[s2 * nb + m];
      }
    } else if (X == 1) {
      v = rec.U[i * nb + a] * rec.U[i * nb + m];
    } else if (X == 2) {
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        const int lo = min(i, j), hi = max(i, j);
        const int s = lo * n - lo * (lo + 1) / 2 + (hi - lo - 1);
        v += rec.U[s * nb + a] * rec.U[s * nb + m];
      }
    }
    rec.K[e] = v;
  }
}

// --------------------------------------------------------------------- 3. Gs
__global__ __launch_bounds__(kThreads) void block_gs_kernel(int n, int nb, double* __restrict__ ws,
                                                            size_t rstride) {
  const int b = blockIdx.x, tid = threadIdx.x;
  Rec rec = rec_at(ws + (size_t)b * rstride, nb, n);
  for (int e = tid; e < nb * nb; e += kThreads) {
    const int a = e / nb, c = e % nb;
    double2 g = make_double2(0.0, 0.0);
    for (int i = 0; i < n; ++i) {
      const double* Ki = rec.K + (size_t)i * nb * nb;
      for (int m = 0; m < nb; ++m) {
        const double kk = Ki[a * nb + m] * Ki[m * nb + c];
        const double2 cc = rec.C[c * nb + m];
        g.x = fma(kk, cc.x, g.x);
        g.y = fma(kk, cc.y, g.y);
      }
    }
    rec.Gs[e] = g;
  }
}

// sandwich: out[i + nX j] = sum_{a,b} UX[i][a] Z[a + nX b] UY[j][b]  (LDS)
__device__ void sandwich(const double2* Z, double2* tmp, double2* outp, const double* UX,
                         const double* UY, int nX, int nY, int tid) {
  for (int e = tid; e < nX * nY; e += kThreads) {  // tmp[a][j] = sum_b Z[a][b] UY[j][b]
    const int a = e % nX, j = e / nX;
    double2 t = make_double2(0.0, 0.0);
    for (int bb = 0; bb < nY; ++bb) {
      const double u = UY[j * nY + bb];
      const double2 z = Z[a + nX * bb];
      t.x = fma(u, z.x, t.x);
      t.y = fma(u, z.y, t.y);
    }
    tmp[a + nX * j] = t;
  }
  __syncthreads();
  for (int e = tid; e < nX * nY; e += kThreads) {  // out[i][j] = sum_a UX[i][a] tmp[a][j]
    const int i = e % nX, j = e / nX;
    double2 t = make_double2(0.0, 0.0);
    for (int a = 0; a < nX; ++a) {
      const double u = UX[i * nX + a];
      const double2 z = tmp[a + nX * j];
      t.x = fma(u, z.x, t.x);
      t.y = fma(u, z.y, t.y);
    }
    outp[e] = t;
  }
  __syncthreads();
}

// ----------------------------------------------- 4. L_eig columns, pass 1
// L_eig[(a,b),(c,d)] = unit * (-i (E_a - E_b) d_ac d_bd - R[(a,b),(c,d)]),
// R = d_bd Gs_X[a][c] + d_ac conj(Gs_Y[b][d])
//     - sum_i K^X_i[a][c] K^Y_i[b][d] (conj(C_Y[d][b]) + C_X[c][a])
// (redfield.py:61-69 with K real symmetric), secular mask redfield.py:77-83.
// The Heisenberg block is the transpose; since T = kron(U_Y, U_X) is real,
// T L^T T^T = (T L T^T)^T, so both passes work on A = L_eig or L_eig^T and
// every global write is a contiguous row: pass 1 writes column q of W = T A
// as row q of W^T, pass 2 reads row p of W (a strided read of W^T, L2
// resident) and writes row p of T A T^T.
enum { kSecular = 1, kRealCorr = 2, kEigenBasis = 4, kHeisenberg = 8 };

__device__ inline double2 leig_entry(const Rec& rx, const Rec& ry, int n, int nX, int nY,
                                     int same, int flags, double unit, int p, int q) {
  const int a = p % nX, bb = p / nX, c = q % nX, d = q / nX;
  double2 r = make_double2(0.0, 0.0);
  if (bb == d) r = cadd(r, rx.Gs[a * nX + c]);
  if (a == c) {
    const double2 g = ry.Gs[bb * nY + d];
    r = cadd(r, make_double2(g.x, -g.y));
  }
  const double2 cy = ry.C[d * nY + bb], cx = rx.C[c * nX + a];
  const double2 cs = make_double2(cy.x + cx.x, -cy.y + cx.y);
  double kk = 0.0;
  for (int i = 0; i < n; ++i)
    kk = fma(rx.K[((size_t)i * nX + a) * nX + c], ry.K[((size_t)i * nY + bb) * nY + d], kk);
  r.x -= kk * cs.x;
  r.y -= kk * cs.y;
  const bool keep = (a == c && bb == d) || (same && a == bb && c == d);
  if ((flags & kSecular) && !keep) r = make_double2(0.0, 0.0);
  double2 v = make_double2(-r.x, -r.y);
  if (a == c && bb == d) v.y -= rx.E[a] - ry.E[bb];
  return make_double2(unit * v.x, unit * v.y);
}

// block q of problem b: eigen basis -> row q of A into L; site basis ->
// column q of A, times T, into row q of W^T
__global__ __launch_bounds__(kThreads) void liouv_cols_kernel(
    int n, int nX, int nY, int same, const double* __restrict__ wsX,
    const double* __restrict__ wsY, size_t rstride, int flags, double unit,
    double2* __restrict__ WT, double2* __restrict__ L) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int q = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const int M = nX * nY;
  Rec rx = rec_at(const_cast<double*>(wsX) + (size_t)b * rstride, nX, n);
  Rec ry = rec_at(const_cast<double*>(wsY) + (size_t)b * rstride, nY, n);
  double2* z = reinterpret_cast<double2*>(smem);
  double2* tmp = z + M;
  double2* w = tmp + M;
  double* UX = reinterpret_cast<double*>(w + M);
  double* UY = UX + nX * nX;
  const bool heis = flags & kHeisenberg;
  if (flags & kEigenBasis) {
    double2* Lrow = L + (size_t)b * M * M + (size_t)q * M;
    for (int p = tid; p < M; p += kThreads)       // A[q][p]
      Lrow[p] = heis ? leig_entry(rx, ry, n, nX, nY, same, flags, unit, p, q)
                     : leig_entry(rx, ry, n, nX, nY, same, flags, unit, q, p);
    return;
  }
  for (int p = tid; p < M; p += kThreads)         // A[p][q]
    z[p] = heis ? leig_entry(rx, ry, n, nX, nY, same, flags, unit, q, p)
                : leig_entry(rx, ry, n, nX, nY, same, flags, unit, p, q);
  for (int e = tid; e < nX * nX; e += kThreads) UX[e] = rx.U[e];
  for (int e = tid; e < nY * nY; e += kThreads) UY[e] = ry.U[e];
  __syncthreads();
  sandwich(z, tmp, w, UX, UY, nX, nY, tid);
  double2* WTrow = WT + (size_t)b * M * M + (size_t)q * M;
  for (int p = tid; p < M; p += kThreads) WTrow[p] = w[p];
}

// ----------------------------------------------------- 5. site rows, pass 2
__global__ __launch_bounds__(kThreads) void liouv_rows_kernel(
    int n, int nX, int nY, const double* __restrict__ wsX, const double* __restrict__ wsY,
    size_t rstride, int flags, const double2* __restrict__ WT, double2* __restrict__ L) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int p = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const int M = nX * nY;
  Rec rx = rec_at(const_cast<double*>(wsX) + (size_t)b * rstride, nX, n);
  Rec ry = rec_at(const_cast<double*>(wsY) + (size_t)b * rstride, nY, n);
  double2* z = reinterpret_cast<double2*>(smem);
  double2* tmp = z + M;
  double2* w = tmp + M;
  double* UX = reinterpret_cast<double*>(w + M);
  double* UY = UX + nX * nX;
  const double2* WTb = WT + (size_t)b * M * M;
  for (int e = tid; e < M; e += kThreads) z[e] = WTb[(size_t)e * M + p];   // row p of W
  for (int e = tid; e < nX * nX; e += kThreads) UX[e] = rx.U[e];
  for (int e = tid; e < nY * nY; e += kThreads) UY[e] = ry.U[e];
  __syncthreads();
  sandwich(z, tmp, w, UX, UY, nX, nY, tid);
  double2* Lrow = L + (size_t)b * M * M + (size_t)p * M;
  for (int q = tid; q < M; q += kThreads) Lrow[q] = w[q];
}

// ---------------------------------------- 4'. non-secular block, one pass
// Without the secular mask every term of L_eig factorises over (a,c) x (b,d):
//   -i(E_a - E_b) d_ac d_bd - R = sum_t A_t[a][c] B_t[b][d],
//   t = 0:      A = -i E_X - Gs_X,                  B = I
//   t = 1:      A = I,                              B = i E_Y - conj(Gs_Y)
//   t = 2+i:    A = K^X_i,                          B[b][d] = K^Y_i[b][d] conj(C_Y[d][b])
//   t = 2+n+i:  A[a][c] = K^X_i[a][c] C_X[c][a],    B = K^Y_i
// and the basis change of a Kronecker product factorises too:
// T (A (x) B) T^T = (U_X A U_X^T) (x) (U_Y B U_Y^T).  So the site-basis block
// is L[(i,j),(k,l)] = unit sum_t Ah_t[i][k] Bh_t[j][l] with the 2 + 2n small
// transformed factors held in LDS: one work-group per problem, one pass, no
// intermediate in global memory.  Thread r owns index pairs (i, k) (or
// (k, i) for the Heisenberg transpose, so that consecutive threads write
// consecutive addresses), keeps its 2 + 2n coefficients Ah_t[i][k] in
// registers and reads Bh_t[j][l] as LDS broadcasts.
constexpr int kKronMaxSites = 9;

__device__ inline void small_mul(const double* U, const double2* A, double2* out, double2* tmp,
                                 int nb, int tid) {
  // out = U A U^T (U real nb x nb, A complex nb x nb; out may alias A)
  for (int e = tid; e < nb * nb; e += kThreads) {   // tmp[a][k] = sum_c A[a][c] U[k][c]
    const int a = e / nb, k = e % nb;
    double2 t = make_double2(0.0, 0.0);
    for (int c = 0; c < nb; ++c) {
      const double u = U[k * nb + c];
      const double2 v = A[a * nb + c];
      t.x = fma(u, v.x, t.x);
      t.y = fma(u, v.y, t.y);
    }
    tmp[e] = t;
  }
  __syncthreads();
  for (int e = tid; e < nb * nb; e += kThreads) {   // out[i][k] = sum_a U[i][a] tmp[a][k]
    const int i = e / nb, k = e % nb;
    double2 t = make_double2(0.0, 0.0);
    for (int a = 0; a < nb; ++a) {
      const double u = U[i * nb + a];
      const double2 v = tmp[a * nb + k];
      t.x = fma(u, v.x, t.x);
      t.y = fma(u, v.y, t.y);
    }
    out[e] = t;
  }
  __syncthreads();
}

__host__ __device__ inline size_t kron_lds_bytes(int n, int nX, int nY) {
  const int nt = 2 + 2 * n, big = nX > nY ? nX : nY;
  return (size_t)nt * (nX * nX + nY * nY) * sizeof(double2) + (size_t)big * big * sizeof(double2) +
         (size_t)(nX * nX + nY * nY) * sizeof(double);
}

template <int NS>   // sites: 2 + 2 NS terms
__global__ __launch_bounds__(kThreads) void liouv_kron_kernel(
    int nX, int nY, const double* __restrict__ wsX, const double* __restrict__ wsY,
    size_t rstride, int flags, double unit, double2* __restrict__ L) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int n = NS, nt = 2 + 2 * NS;
  const int b = blockIdx.x, tid = threadIdx.x;
  const int M = nX * nY, nX2 = nX * nX, nY2 = nY * nY;
  Rec rx = rec_at(const_cast<double*>(wsX) + (size_t)b * rstride, nX, n);
  Rec ry = rec_at(const_cast<double*>(wsY) + (size_t)b * rstride, nY, n);
  double2* Ah = reinterpret_cast<double2*>(smem);       // [nt][nX2]
  double2* Bh = Ah + (size_t)nt * nX2;                   // [nt][nY2]
  double2* tmp = Bh + (size_t)nt * nY2;                  // [max(nX,nY)^2]
  double* UX = reinterpret_cast<double*>(tmp + (nX > nY ? nX2 : nY2));
  double* UY = UX + nX2;
  for (int e = tid; e < nX2; e += kThreads) UX[e] = rx.U[e];
  for (int e = tid; e < nY2; e += kThreads) UY[e] = ry.U[e];
  for (int e = tid; e < nt * nX2; e += kThreads) {
    const int t = e / nX2, r = e % nX2, a = r / nX, c = r % nX;
    double2 v;
    if (t == 0) {
      const double2 g = rx.Gs[a * nX + c];
      v = make_double2(-g.x, (a == c ? -rx.E[a] : 0.0) - g.y);
    } else if (t == 1) {
      v = make_double2(a == c ? 1.0 : 0.0, 0.0);
    } else if (t < 2 + n) {
      v = make_double2(rx.K[((size_t)(t - 2) * nX + a) * nX + c], 0.0);
    } else {
      const double k = rx.K[((size_t)(t - 2 - n) * nX + a) * nX + c];
      const double2 cx = rx.C[c * nX + a];
      v = make_double2(k * cx.x, k * cx.y);
    }
    Ah[e] = v;
  }
  for (int e = tid; e < nt * nY2; e += kThreads) {
    const int t = e / nY2, r = e % nY2, bb = r / nY, d = r % nY;
    double2 v;
    if (t == 0) {
      v = make_double2(bb == d ? 1.0 : 0.0, 0.0);
    } else if (t == 1) {
      const double2 g = ry.Gs[bb * nY + d];
      v = make_double2(-g.x, (bb == d ? ry.E[bb] : 0.0) + g.y);
    } else if (t < 2 + n) {
      const double k = ry.K[((size_t)(t - 2) * nY + bb) * nY + d];
      const double2 cy = ry.C[d * nY + bb];
      v = make_double2(k * cy.x, -k * cy.y);
    } else {
      v = make_double2(ry.K[((size_t)(t - 2 - n) * nY + bb) * nY + d], 0.0);
    }
    Bh[e] = v;
  }
  __syncthreads();
  if (!(flags & kEigenBasis)) {
    // the identity factors (Bh_0, Ah_1) are invariant under U . U^T
    for (int t = 0; t < nt; ++t) {
      if (t != 1) small_mul(UX, Ah + (size_t)t * nX2, Ah + (size_t)t * nX2, tmp, nX, tid);
      if (t != 0) small_mul(UY, Bh + (size_t)t * nY2, Bh + (size_t)t * nY2, tmp, nY, tid);
    }
  }
  const bool heis = flags & kHeisenberg;
  double2* Lb = L + (size_t)b * M * M;
  for (int pr = tid; pr < nX2; pr += kThreads) {
    // Heisenberg: consecutive threads take consecutive i (consecutive
    // columns of the transposed output row); otherwise consecutive k
    const int i = heis ? pr % nX : pr / nX, k = heis ? pr / nX : pr % nX;
    double2 a[nt];
#pragma unroll
    for (int t = 0; t < nt; ++t) a[t] = Ah[(size_t)t * nX2 + i * nX + k];
    for (int j = 0; j < nY; ++j) {
      for (int l = 0; l < nY; ++l) {
        double2 v = make_double2(0.0, 0.0);
#pragma unroll
        for (int t = 0; t < nt; ++t) cfma<double>(v, a[t], Bh[(size_t)t * nY2 + j * nY + l]);
        const int p = i + nX * j, q = k + nX * l;
        const double2 out = make_double2(unit * v.x, unit * v.y);
        if (heis) Lb[(size_t)q * M + p] = out;
        else Lb[(size_t)p * M + q] = out;
      }
    }
  }
}

// ------------------------------------------------------------------ unitary
// L[(i,j),(k,l)] = -i unit (H_X[i][k] d_jl - d_ik H_Y[l][j]) in the rotating
// frame (qspectra/dynamics/unitary.py:5-9, liouville_space.py:93-131).
__device__ double block_h(const double* H, int n, int X, double rw, int r, int c) {
  if (X == 0) return 0.0;
  double v;
  if (X == 1) {
    v = H[r * n + c];
  } else {
    int i, j, k, l;
    pair_of(r, n, i, j);
    pair_of(c, n, k, l);
    v = (i == k ? H[j * n + l] : 0.0) + (i == l ? H[j * n + k] : 0.0) +
        (j == k ? H[i * n + l] : 0.0) + (j == l ? H[i * n + k] : 0.0);
  }
  if (r == c) v -= X * rw;
  return v;
}

__global__ __launch_bounds__(kThreads) void unitary_block_kernel(const double* __restrict__ H1,
                                                                 int n, const double* __restrict__ rw,
                                                                 double unit, int X, int Y,
                                                                 int heisenberg,
                                                                 double2* __restrict__ L) {
  const int b = blockIdx.y;
  const int nX = block_size(X, n), nY = block_size(Y, n), M = nX * nY;
  const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= (int64_t)M * M) return;
  const int p = (int)(e / M), q = (int)(e % M);
  const int i = p % nX, j = p / nX, k = q % nX, l = q / nX;
  const double* H = H1 + (size_t)b * n * n;
  const double w = rw ? rw[b] : 0.0;
  double v = 0.0;
  if (j == l) v += block_h(H, n, X, w, i, k);
  if (i == k) v -= block_h(H, n, Y, w, l, j);
  double2* Lb = L + (size_t)b * M * M;
  const double2 val = make_double2(0.0, -unit * v);
  if (heisenberg) Lb[(size_t)q * M + p] = val;
  else Lb[e] = val;
}

// dense blocks: L[(i,j),(k,l)] = -i unit (HX[i][k] d_jl - d_ik HY[l][j])
__global__ __launch_bounds__(kThreads) void unitary_dense_kernel(const double* __restrict__ HX,
                                                                 const double* __restrict__ HY,
                                                                 int nX, int nY, double unit,
                                                                 int heisenberg,
                                                                 double2* __restrict__ L) {
  const int b = blockIdx.y;
  const int M = nX * nY;
  const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= (int64_t)M * M) return;
  const int p = (int)(e / M), q = (int)(e % M);
  const int i = p % nX, j = p / nX, k = q % nX, l = q / nX;
  const double* hx = HX + (size_t)b * nX * nX;
  const double* hy = HY + (size_t)b * nY * nY;
  double v = 0.0;
  if (j == l) v += hx[i * nX + k];
  if (i == k) v -= hy[l * nY + j];
  double2* Lb = L + (size_t)b * M * M;
  const double2 val = make_double2(0.0, -unit * v);
  if (heisenberg) Lb[(size_t)q * M + p] = val;
  else Lb[e] = val;
}

// ------------------------------------------------- absorption operators
// The operators of linear absorption for a ragged batch of aggregates of up
// to kAbsMaxSites sites, one work-group per problem, everything in LDS: the
// Heisenberg 'eg' Redfield block of the one-exciton Hamiltonian (Debye bath)
// and its RK4 propagator P = P4(h L) with P^J.  For the eg block (Y = g: one
// state of energy 0 that no bath couples to, so K^Y = 0 and Gs_Y = 0) the
// pipeline of the kernels above reduces to
//   L_eig[a][c] = unit (-Gs[a][c] - i E_a d_ac)    (secular: Gs diagonal only)
//   Gs[a][c]    = sum_i U[i][a] U[i][c] sum_m U[i][m]^2 C(E_c - E_m)
//   L           = U L_eig U^T  (its transpose for the Heisenberg picture)
// (redfield.py:9-104, with the e block's couplings K_i = U^T |i><i| U).
// The LDS layout follows each problem's n (m x m doubles A and V, m = n
// rounded up to even, then four n x n complex matrices), the launch's LDS
// size its largest n, so launches of small aggregates keep several
// work-groups per CU.  Each Matsubara sum is split over the threads left
// per pair (a power of two, at most 64, reduced by wavefront shuffles).
// table int64 [B][3]: n, offset of H1 (doubles, n x n), offset of P and P^J
// (complex, n x n); bath double [B][3]: T, lam, gam.
constexpr int kAbsMaxSites = 31;

__host__ __device__ inline size_t abs_lds_bytes(int n) {
  const int m = n + (n & 1);
  return 2 * (size_t)m * m * sizeof(double) + 4 * (size_t)n * n * sizeof(double2);
}

// out = alpha A B + gamma I (n x n row-major, LDS); the caller synchronises
// before and after
template <int NT>
__device__ inline void abs_gemm(const double2* A, const double2* B, double2* out, int n,
                                double alpha, double gamma) {
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const int i = e / n, j = e % n;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < n; ++k) cfma<double>(acc, A[i * n + k], B[k * n + j]);
    out[i * n + j] = make_double2(fma(alpha, acc.x, i == j ? gamma : 0.0), alpha * acc.y);
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void absorption_ops_kernel(
    const int64_t* __restrict__ table, int max_n, const double* __restrict__ H1,
    const double* __restrict__ rw, const double* __restrict__ bath, int flags, double unit,
    double h, int J, double2* __restrict__ P1, double2* __restrict__ PJ) {
  extern __shared__ __align__(16) unsigned char absmem[];
  __shared__ double rot[kAbsMaxSites / 2 + 1][2];
  __shared__ int pqs[kAbsMaxSites / 2 + 1][2];
  __shared__ double red[2 * NT / 64];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int n = (int)table[3 * b];
  if (n < 1 || n > max_n) return;                          // uniform per work-group
  const int64_t oh = table[3 * b + 1], op = table[3 * b + 2];
  const int m = n + (n & 1), nn = n * n;
  double* A = reinterpret_cast<double*>(absmem);           // [m x m] H_e, then diag(E)
  double* V = A + m * m;                                    // [m x m] eigenvectors
  double2* C = reinterpret_cast<double2*>(V + m * m);       // C, then L_eig
  double2* G = C + nn;                                      // Gs, then L
  double2* X = G + nn;                                      // scratch
  double2* Y = X + nn;                                      // P, its powers
  const double w = rw[b];
  for (int e = tid; e < m * m; e += NT) {
    const int r = e / m, c = e % m;
    double v = (r < n && c < n) ? H1[oh + r * n + c] : 0.0;
    if (r == c && r < n) v -= w;
    A[e] = v;
    V[e] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();
  jacobi_sweeps<NT>(A, V, m, rot, pqs, red);
  // C[d][c] = C(E_d - E_c) over the pairs d <= c, one Matsubara sum each,
  // split over tpp threads
  const double T = bath[3 * b], lam = bath[3 * b + 1], gam = bath[3 * b + 2];
  const int real_corr = (flags & kRealCorr) != 0;
  const int npairs = n * (n + 1) / 2;
  // the closed form where it is accurate (one thread per pair); else threads
  // per pair: the power of two (at most 64) with the fewest terms per thread
  // over the rounds of NT threads
  const bool closed = !real_corr && matsubara_closed_ok(T, gam);
  double dm = 0.0, dp = 0.0;
  if (closed) {
    const double pg = gam / (2.0 * M_PI * T);
    dm = digamma_c(1000.0 - pg, 0.0).x - digamma_c(1.0 - pg, 0.0).x;
    dp = digamma_c(1000.0 + pg, 0.0).x - digamma_c(1.0 + pg, 0.0).x;
  }
  int tpp = 1, cost = 1 << 30;
  for (int t = 1; t <= (closed ? 1 : 64); t *= 2) {
    const int c = ((npairs * t + NT - 1) / NT) * ((999 + t - 1) / t);
    if (c < cost) {
      cost = c;
      tpp = t;
    }
  }
  for (int e0 = 0; e0 < npairs * tpp; e0 += NT) {     // uniform trip count
    const int e = e0 + tid, pair = e / tpp, part = e % tpp;
    int d = 0, c = 0;
    double x = 0.0;
    double2 S = make_double2(0.0, 0.0);
    if (pair < npairs) {
      int r = pair;
      while (r >= n - d) {
        r -= n - d;
        ++d;
      }
      c = d + r;
      x = A[d * m + d] - A[c * m + c];
      if (!real_corr && x != 0.0)
        S = closed ? matsubara_closed(x, T, dm, dp, gam) : matsubara_partial(x, T, gam, 1 + part, tpp);
    }
    for (int o = tpp / 2; o > 0; o >>= 1) {                  // within aligned groups of tpp lanes
      S.x += __shfl_xor(S.x, o);
      S.y += __shfl_xor(S.y, o);
    }
    if (pair < npairs && part == 0) {
      C[d * n + c] = debye_corr_with(x, S, T, lam, gam, real_corr);
      if (c != d) C[c * n + d] = debye_corr_with(-x, make_double2(S.x, -S.y), T, lam, gam, real_corr);
    }
  }
  __syncthreads();
  // X[i][c] = sum_m U[i][m]^2 C[c][m], then Gs[a][c] = sum_i U[i][a] U[i][c] X[i][c]
  for (int e = tid; e < nn; e += NT) {
    const int i = e / n, c = e % n;
    double2 acc = make_double2(0.0, 0.0);
    for (int q = 0; q < n; ++q) {
      const double u = V[i * m + q];
      const double2 cc = C[c * n + q];
      acc.x = fma(u * u, cc.x, acc.x);
      acc.y = fma(u * u, cc.y, acc.y);
    }
    X[i * n + c] = acc;
  }
  __syncthreads();
  for (int e = tid; e < nn; e += NT) {
    const int a = e / n, c = e % n;
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < n; ++i) {
      const double uu = V[i * m + a] * V[i * m + c];
      const double2 xx = X[i * n + c];
      acc.x = fma(uu, xx.x, acc.x);
      acc.y = fma(uu, xx.y, acc.y);
    }
    G[a * n + c] = acc;
  }
  __syncthreads();
  // L_eig (transposed for the Heisenberg picture) into C
  const bool secular = (flags & kSecular) != 0, heis = (flags & kHeisenberg) != 0;
  for (int e = tid; e < nn; e += NT) {
    const int a = e / n, c = e % n;
    double2 v = make_double2(0.0, 0.0);
    if (!secular || a == c) v = make_double2(-G[a * n + c].x, -G[a * n + c].y);
    if (a == c) v.y -= A[a * m + a];
    v = make_double2(unit * v.x, unit * v.y);
    if (heis) C[c * n + a] = v;
    else C[a * n + c] = v;
  }
  __syncthreads();
  // L = U L_eig U^T: X = L_eig U^T, G = U X
  for (int e = tid; e < nn; e += NT) {
    const int a = e / n, j = e % n;
    double2 acc = make_double2(0.0, 0.0);
    for (int c = 0; c < n; ++c) {
      const double u = V[j * m + c];
      const double2 l = C[a * n + c];
      acc.x = fma(u, l.x, acc.x);
      acc.y = fma(u, l.y, acc.y);
    }
    X[a * n + j] = acc;
  }
  __syncthreads();
  for (int e = tid; e < nn; e += NT) {
    const int i = e / n, j = e % n;
    double2 acc = make_double2(0.0, 0.0);
    for (int a = 0; a < n; ++a) {
      const double u = V[i * m + a];
      const double2 t = X[a * n + j];
      acc.x = fma(u, t.x, acc.x);
      acc.y = fma(u, t.y, acc.y);
    }
    G[i * n + j] = acc;
  }
  __syncthreads();
  // P = P4(h L) by Horner: Y = I + h/4 L, Y = I + c L Y for c = h/3, h/2, h
  for (int e = tid; e < nn; e += NT) {
    const int i = e / n, j = e % n;
    const double2 l = G[e];
    Y[e] = make_double2(fma(0.25 * h, l.x, i == j ? 1.0 : 0.0), 0.25 * h * l.y);
  }
  __syncthreads();
  const double coef[3] = {h / 3.0, h / 2.0, h};
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    abs_gemm<NT>(G, Y, X, n, coef[t], 1.0);
    __syncthreads();
    for (int e = tid; e < nn; e += NT) Y[e] = X[e];
    __syncthreads();
  }
  for (int e = tid; e < nn; e += NT) P1[op + e] = Y[e];
  // P^J by squaring (J a power of two)
  for (int q = 1; q < J; q *= 2) {
    abs_gemm<NT>(Y, Y, X, n, 1.0, 0.0);
    __syncthreads();
    for (int e = tid; e < nn; e += NT) Y[e] = X[e];
    __syncthreads();
  }
  for (int e = tid; e < nn; e += NT) PJ[op + e] = Y[e];
}

// ------------------------------------------------ vector-Jacobian product
// Reverse mode of a Redfield block with respect to the one-exciton
// Hamiltonian H1 and the reorganization energy (Debye bath).  With
// Lu = leig_entry (unit * mask o L_eig, eigen basis) and the site block
// L = T Lu T^T, T = kron(U_Y, U_X), a cotangent G of L (dl = Re sum conj(G) o dL,
// torch's convention) pulls back as
//   W   = T^T G T                        cotangent of Lu (G itself in the eigen basis)
//   Lam = unit * mask o W                cotangent of -i w - R in leig_entry
//   U_X^T Ubar_X[a'][a] = Re sum_b (W Lu^H + W^H Lu)[(a',b),(a,b)]  (and Y alike)
// then through -i(E_a - E_b), Gs_X[a][c] = sum_{i,m} K_i[a][m] K_i[m][c] C[c][m],
// the Kronecker terms K^X_i[a][c] K^Y_i[b][d] (conj(C_Y[d][b]) + C_X[c][a]),
// K_i = U^T P_i U (U^T Ubar += K_i (Kbar_i + Kbar_i^T)) and C[d][c] = C(E_d - E_c)
// (Ebar_d += Re(Cbar[d][c] conj(C'(x))), Ebar_c -= the same), and reaches the
// block Hamiltonian through the symmetric eigenproblem's VJP
//   Hbar = U (diag(Ebar) + F o (Om - Om^T) / 2) U^T,  F[a][c] = 1 / (E_c - E_a),
// Om = U^T Ubar (F = 0 where E_c = E_a); the f block's Hbar is gathered back
// onto H1 through operator_1_to_2's four terms.

// C(x)/lam and dC/dx of the Debye bath (the x = 0 branch is a constant)
__device__ void debye_corr_grad(double x, double T, double lam, double gam, int real_corr,
                                double2& c1, double2& dc) {
  if (real_corr) {
    if (x == 0.0) {
      c1 = make_double2(T * 2.0 / gam, 0.0);
      dc = make_double2(0.0, 0.0);
      return;
    }
    const double den = gam * gam + x * x;
    const double J1 = 2.0 * gam * x / den;                       // J / lam
    const double dJ1 = 2.0 * gam * (gam * gam - x * x) / (den * den);
    const double nb = 1.0 / expm1(x / T);
    c1 = make_double2((nb + 1.0) * J1, 0.0);
    dc = make_double2(lam * ((nb + 1.0) * dJ1 - nb * (nb + 1.0) * J1 / T), 0.0);
    return;
  }
  if (x == 0.0) {
    c1 = make_double2(2.0 * T / gam, -1.0);
    dc = make_double2(0.0, 0.0);
    return;
  }
  const double2 S = matsubara_sum(x, T, gam);
  const double ct = 1.0 / tan(gam / (2.0 * T));
  const double den = gam * gam + x * x;
  c1 = make_double2(gam * ((ct * gam + x) / den + 4.0 * T * S.x),
                    gam * ((ct * x - gam) / den + 4.0 * T * S.y));
  // d/dx (ct - i)/(g - i x) = (1 + i ct)(g^2 - x^2 + 2 i g x) / (g^2 + x^2)^2
  const double pr = gam * gam - x * x, pi = 2.0 * gam * x, d2 = den * den;
  double dr = (pr - ct * pi) / d2, di = (ct * pr + pi) / d2;
  // dS/dx = sum_k nu/(nu^2 - g^2) (-2 nu x + i (nu^2 - x^2)) / (nu^2 + x^2)^2
  double sr = 0.0, si = 0.0;
  const double twopiT = 2.0 * M_PI * T;
  for (int k = 1; k < 1000; ++k) {
    const double nu = twopiT * k;
    const double q = nu * nu + x * x;
    const double f = nu / ((nu * nu - gam * gam) * q * q);
    sr = fma(f, -2.0 * nu * x, sr);
    si = fma(f, nu * nu - x * x, si);
  }
  dr += 4.0 * T * sr;
  di += 4.0 * T * si;
  dc = make_double2(lam * gam * dr, lam * gam * di);
}

// cotangent of L_eig's bracket: Lam(p, q) = unit * mask * W(p, q)
__device__ inline double2 vjp_w(const double2* Wb, const double2* Gb, int M, int flags, int p,
                                int q) {
  if (flags & kEigenBasis)
    return (flags & kHeisenberg) ? Gb[(size_t)q * M + p] : Gb[(size_t)p * M + q];
  return Wb[(size_t)p * M + q];
}
__device__ inline double2 vjp_lam(const double2* Wb, const double2* Gb, int M, int nX,
                                  int same, int flags, double unit, int a, int bb, int c,
                                  int d) {
  if ((flags & kSecular) && !((a == c && bb == d) || (same && a == bb && c == d)))
    return make_double2(0.0, 0.0);
  const double2 w = vjp_w(Wb, Gb, M, flags, a + nX * bb, c + nX * d);
  return make_double2(unit * w.x, unit * w.y);
}

// T^T applied to column q of G (the caller's block transposed back for the
// Heisenberg picture): row q of W1^T, W1 = T^T G
__global__ __launch_bounds__(kThreads) void vjp_cols_kernel(
    int n, int nX, int nY, const double* __restrict__ wsX, const double* __restrict__ wsY,
    size_t rstride, int flags, const double2* __restrict__ G, double2* __restrict__ W1T) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int q = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const int M = nX * nY;
  Rec rx = rec_at(const_cast<double*>(wsX) + (size_t)b * rstride, nX, n);
  Rec ry = rec_at(const_cast<double*>(wsY) + (size_t)b * rstride, nY, n);
  double2* z = reinterpret_cast<double2*>(smem);
  double2* tmp = z + M;
  double2* w = tmp + M;
  double* UXt = reinterpret_cast<double*>(w + M);
  double* UYt = UXt + nX * nX;
  const double2* Gb = G + (size_t)b * M * M;
  const bool heis = flags & kHeisenberg;
  for (int p = tid; p < M; p += kThreads)
    z[p] = heis ? Gb[(size_t)q * M + p] : Gb[(size_t)p * M + q];
  for (int e = tid; e < nX * nX; e += kThreads) UXt[e] = rx.U[(e % nX) * nX + e / nX];
  for (int e = tid; e < nY * nY; e += kThreads) UYt[e] = ry.U[(e % nY) * nY + e / nY];
  __syncthreads();
  sandwich(z, tmp, w, UXt, UYt, nX, nY, tid);
  double2* row = W1T + (size_t)b * M * M + (size_t)q * M;
  for (int p = tid; p < M; p += kThreads) row[p] = w[p];
}

// row p of W = W1 T: T^T applied to row p of W1 (column p of W1^T)
__global__ __launch_bounds__(kThreads) void vjp_rows_kernel(
    int n, int nX, int nY, const double* __restrict__ wsX, const double* __restrict__ wsY,
    size_t rstride, const double2* __restrict__ W1T, double2* __restrict__ W) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int p = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const int M = nX * nY;
  Rec rx = rec_at(const_cast<double*>(wsX) + (size_t)b * rstride, nX, n);
  Rec ry = rec_at(const_cast<double*>(wsY) + (size_t)b * rstride, nY, n);
  double2* z = reinterpret_cast<double2*>(smem);
  double2* tmp = z + M;
  double2* w = tmp + M;
  double* UXt = reinterpret_cast<double*>(w + M);
  double* UYt = UXt + nX * nX;
  const double2* src = W1T + (size_t)b * M * M;
  for (int q = tid; q < M; q += kThreads) z[q] = src[(size_t)q * M + p];
  for (int e = tid; e < nX * nX; e += kThreads) UXt[e] = rx.U[(e % nX) * nX + e / nX];
  for (int e = tid; e < nY * nY; e += kThreads) UYt[e] = ry.U[(e % nY) * nY + e / nY];
  __syncthreads();
  sandwich(z, tmp, w, UXt, UYt, nX, nY, tid);
  double2* row = W + (size_t)b * M * M + (size_t)p * M;
  for (int q = tid; q < M; q += kThreads) row[q] = w[q];
}

// per-problem reductions of Lam (complex):
//   dX[nX]  dY[nY]                 sum over the other index of Lam's diagonal
//   GX[nX2] GY[nY2]                cotangents of Gs_X, Gs_Y
//   OX[nX2] OY[nY2]                U^T Ubar from the basis change (real part)
//   V[n][nX2] V2[n][nX2]           sum_{b,d} Lam K^Y_i[b][d] (x C_Y[d][b])
//   Z[n][nY2] Z2[n][nY2]           sum_{a,c} Lam K^X_i[a][c] (x conj(C_X[c][a]))
__host__ __device__ inline size_t vjp_red_count(int n, int nX, int nY) {
  const size_t s2 = (size_t)nX * nX + (size_t)nY * nY;
  return (size_t)nX + nY + 2 * s2 + 2 * (size_t)n * s2;
}

__global__ __launch_bounds__(kThreads) void vjp_reduce_kernel(
    int n, int nX, int nY, int same, int flags, double unit, const double* __restrict__ wsX,
    const double* __restrict__ wsY, size_t rstride, const double2* __restrict__ G,
    const double2* __restrict__ W, const double2* __restrict__ Lu, double2* __restrict__ red,
    size_t rcount) {
  const int b = blockIdx.y;
  const size_t e0 = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (e0 >= rcount) return;
  const int M = nX * nY, nX2 = nX * nX, nY2 = nY * nY;
  Rec rx = rec_at(const_cast<double*>(wsX) + (size_t)b * rstride, nX, n);
  Rec ry = rec_at(const_cast<double*>(wsY) + (size_t)b * rstride, nY, n);
  const double2* Wb = W + (size_t)b * M * M;
  const double2* Gb = G + (size_t)b * M * M;
  const double2* Lb = Lu + (size_t)b * M * M;
  auto lam_at = [&](int a, int bb, int c, int d) {
    return vjp_lam(Wb, Gb, M, nX, same, flags, unit, a, bb, c, d);
  };
  size_t e = e0;
  double2 acc = make_double2(0.0, 0.0);
  if (e < (size_t)nX) {
    const int a = (int)e;
    for (int bb = 0; bb < nY; ++bb) acc = cadd(acc, lam_at(a, bb, a, bb));
  } else if ((e -= nX) < (size_t)nY) {
    const int bb = (int)e;
    for (int a = 0; a < nX; ++a) acc = cadd(acc, lam_at(a, bb, a, bb));
  } else if ((e -= nY) < (size_t)nX2) {
    const int a = (int)e / nX, c = (int)e % nX;
    for (int bb = 0; bb < nY; ++bb) acc = cadd(acc, lam_at(a, bb, c, bb));
    acc = make_double2(-acc.x, -acc.y);
  } else if ((e -= nX2) < (size_t)nY2) {
    const int bb = (int)e / nY, d = (int)e % nY;
    for (int a = 0; a < nX; ++a) acc = cadd(acc, lam_at(a, bb, a, d));
    acc = make_double2(-acc.x, acc.y);
  } else if ((e -= nY2) < (size_t)(nX2 + nY2)) {
    // Re sum (W(p', q) conj(Lu(p, q)) + conj(W(q, p')) Lu(q, p)) over the
    // other block's index and q (site basis only)
    if (!(flags & kEigenBasis)) {
      const bool xs = e < (size_t)nX2;
      const int r = xs ? (int)e / nX : (int)(e - nX2) / nY;
      const int s = xs ? (int)e % nX : (int)(e - nX2) % nY;
      const int nothers = xs ? nY : nX;
      double v = 0.0;
      for (int o = 0; o < nothers; ++o) {
        const int pp = xs ? r + nX * o : o + nX * r;
        const int p = xs ? s + nX * o : o + nX * s;
        for (int q = 0; q < M; ++q) {
          const double2 w1 = Wb[(size_t)pp * M + q], l1 = Lb[(size_t)p * M + q];
          const double2 w2 = Wb[(size_t)q * M + pp], l2 = Lb[(size_t)q * M + p];
          v += w1.x * l1.x + w1.y * l1.y + w2.x * l2.x + w2.y * l2.y;
        }
      }
      acc = make_double2(v, 0.0);
    }
  } else if ((e -= nX2 + nY2) < (size_t)2 * n * nX2) {
    const bool withc = e >= (size_t)n * nX2;
    const size_t f = withc ? e - (size_t)n * nX2 : e;
    const int i = (int)(f / nX2), a = (int)(f % nX2) / nX, c = (int)(f % nX2) % nX;
    const double* Ky = ry.K + (size_t)i * nY2;
    for (int bb = 0; bb < nY; ++bb)
      for (int d = 0; d < nY; ++d) {
        const double k = Ky[bb * nY + d];
        if (k == 0.0) continue;
        double2 l = lam_at(a, bb, c, d);
        l = make_double2(k * l.x, k * l.y);
        acc = cadd(acc, withc ? cmul(l, ry.C[d * nY + bb]) : l);
      }
  } else {
    e -= (size_t)2 * n * nX2;
    const bool withc = e >= (size_t)n * nY2;
    const size_t f = withc ? e - (size_t)n * nY2 : e;
    const int i = (int)(f / nY2), bb = (int)(f % nY2) / nY, d = (int)(f % nY2) % nY;
    const double* Kx = rx.K + (size_t)i * nX2;
    for (int a = 0; a < nX; ++a)
      for (int c = 0; c < nX; ++c) {
        const double k = Kx[a * nX + c];
        if (k == 0.0) continue;
        double2 l = lam_at(a, bb, c, d);
        l = make_double2(k * l.x, k * l.y);
        if (withc) {
          const double2 cx = rx.C[c * nX + a];
          l = cmul(l, make_double2(cx.x, -cx.y));
        }
        acc = cadd(acc, l);
      }
  }
  red[(size_t)b * rcount + e0] = acc;
}

// per-(problem, Hilbert block) slab of the chain (doubles):
//   Cbar dC C1 Gsbar (complex nb^2 each) | Om Kbar Mid Tmp Hb (nb^2) | Ebar (nb) | glam
__host__ __device__ inline size_t vjp_slab_doubles(int nb) {
  return (13 * (size_t)nb * nb + nb + 1 + 1) & ~(size_t)1;
}

// one work-group per (problem, block k): the block's Hbar and its share of dl/dlam
__global__ __launch_bounds__(kThreads) void vjp_chain_kernel(
    int n, int nX, int nY, int same, double T, double lam, double gam, int flags,
    const double* __restrict__ wsX, const double* __restrict__ wsY, size_t rstride,
    const double2* __restrict__ red, size_t rcount, double* __restrict__ chain,
    size_t cstride) {
  __shared__ double part[kThreads];
  const int b = blockIdx.x, k = blockIdx.y, tid = threadIdx.x;
  const int nb = k == 0 ? nX : nY, nb2 = nb * nb;
  const int nX2 = nX * nX, nY2 = nY * nY;
  const bool isx = k == 0 || same, isy = k == 1 || same;
  Rec r = rec_at(const_cast<double*>(k == 0 ? wsX : wsY) + (size_t)b * rstride, nb, n);
  double* slab = chain + ((size_t)b * 2 + k) * cstride;
  double2* Cbar = reinterpret_cast<double2*>(slab);
  double2* dC = Cbar + nb2;
  double2* C1 = dC + nb2;
  double2* Gsb = C1 + nb2;
  double* Om = reinterpret_cast<double*>(Gsb + nb2);
  double* Kbar = Om + nb2;
  double* Mid = Kbar + nb2;
  double* Tmp = Mid + nb2;
  double* Hb = Tmp + nb2;
  double* Ebar = Hb + nb2;
  double* glam = Ebar + nb;
  const double2* rb = red + (size_t)b * rcount;
  const double2* dX = rb;
  const double2* dY = dX + nX;
  const double2* GX = dY + nY;
  const double2* GY = GX + nX2;
  const double2* OX = GY + nY2;
  const double2* OY = OX + nX2;
  const double2* V = OY + nY2;
  const double2* V2 = V + (size_t)n * nX2;
  const double2* Z = V2 + (size_t)n * nX2;
  const double2* Z2 = Z + (size_t)n * nY2;
  const int real_corr = (flags & kRealCorr) != 0;
  // phase 0: eigenvalue cotangents of -i w, Gs cotangent, basis-change Om,
  // C(x)/lam and C'(x) (pairs d <= c: S(-x) = conj S(x), S'(-x) = -conj S'(x)
  // hold for the Debye sums, but the closed-form part is not symmetric, so
  // each ordered pair is evaluated)
  for (int a = tid; a < nb; a += kThreads)
    Ebar[a] = (isx ? -dX[a].y : 0.0) + (isy ? dY[a].y : 0.0);
  for (int e = tid; e < nb2; e += kThreads) {
    double2 g = make_double2(0.0, 0.0);
    if (isx) g = cadd(g, GX[e]);
    if (isy) g = cadd(g, GY[e]);
    Gsb[e] = g;
    Om[e] = (isx ? OX[e].x : 0.0) + (isy ? OY[e].x : 0.0);
    const int d = e / nb, c = e % nb;
    double2 c1, dc;
    debye_corr_grad(r.E[d] - r.E[c], T, lam, gam, real_corr, c1, dc);
    C1[e] = c1;
    dC[e] = dc;
  }
  __syncthreads();
  // phase 1: Cbar[d][c] (index [r][s]) from the Kronecker terms and Gs
  for (int e = tid; e < nb2; e += kThreads) {
    const int rr = e / nb, ss = e % nb;
    double2 cb = make_double2(0.0, 0.0);
    for (int i = 0; i < n; ++i) {
      const double* Ki = r.K + (size_t)i * nb2;
      const double kk = Ki[ss * nb + rr];
      if (isx) {
        const double2 v = V[(size_t)i * nX2 + ss * nb + rr];
        cb.x += kk * v.x;
        cb.y += kk * v.y;
      }
      if (isy) {
        const double2 z = Z[(size_t)i * nY2 + ss * nb + rr];
        cb.x += kk * z.x;
        cb.y -= kk * z.y;
      }
      // Gs: Cbar[c][m] += K_i[m][c] sum_a K_i[m][a] Gsbar[a][c]
      double2 t = make_double2(0.0, 0.0);
      for (int a = 0; a < nb; ++a) {
        const double km = Ki[ss * nb + a];
        const double2 g = Gsb[a * nb + rr];
        t.x = fma(km, g.x, t.x);
        t.y = fma(km, g.y, t.y);
      }
      cb.x += kk * t.x;
      cb.y += kk * t.y;
    }
    Cbar[e] = cb;
  }
  __syncthreads();
  // phase 2: per site, Kbar_i, then Om += K_i (Kbar_i + Kbar_i^T)
  for (int i = 0; i < n; ++i) {
    const double* Ki = r.K + (size_t)i * nb2;
    for (int e = tid; e < nb2; e += kThreads) {
      const int rr = e / nb, ss = e % nb;
      const double2 csr = r.C[ss * nb + rr];
      double kb = 0.0;
      if (isx) {
        const double2 v = V[(size_t)i * nX2 + e], v2 = V2[(size_t)i * nX2 + e];
        kb += csr.x * v.x + csr.y * v.y + v2.x;            // Re(conj(C[s][r]) V + V2)
      }
      if (isy) {
        const double2 z = Z[(size_t)i * nY2 + e], z2 = Z2[(size_t)i * nY2 + e];
        kb += csr.x * z.x - csr.y * z.y + z2.x;            // Re(C[s][r] Z + Z2)
      }
      // Gs: Re sum_c Gsbar[r][c] K_i[s][c] conj(C[c][s])
      //   + Re conj(C[s][r]) sum_a Gsbar[a][s] K_i[a][r]
      double t1 = 0.0;
      double2 t2 = make_double2(0.0, 0.0);
      for (int c = 0; c < nb; ++c) {
        const double2 g = Gsb[rr * nb + c];
        const double2 cc = r.C[c * nb + ss];
        const double kc = Ki[ss * nb + c];
        t1 += kc * (g.x * cc.x + g.y * cc.y);
        const double2 g2 = Gsb[c * nb + ss];
        const double ka = Ki[c * nb + rr];
        t2.x = fma(ka, g2.x, t2.x);
        t2.y = fma(ka, g2.y, t2.y);
      }
      kb += t1 + csr.x * t2.x + csr.y * t2.y;
      Kbar[e] = kb;
    }
    __syncthreads();
    for (int e = tid; e < nb2; e += kThreads) {
      const int rr = e / nb, ss = e % nb;
      double o = 0.0;
      for (int t = 0; t < nb; ++t) o += Ki[rr * nb + t] * (Kbar[t * nb + ss] + Kbar[ss * nb + t]);
      Om[e] += o;
    }
    __syncthreads();
  }
  // phase 3: eigenvalue cotangents through C(E_d - E_c), and dl/dlam
  double gl = 0.0;
  for (int e = tid; e < nb2; e += kThreads) {
    const double2 cb = Cbar[e], c1 = C1[e];
    gl += cb.x * c1.x + cb.y * c1.y;
  }
  part[tid] = gl;
  for (int d = tid; d < nb; d += kThreads) {
    double eb = 0.0;
    for (int c = 0; c < nb; ++c) {
      if (c == d) continue;
      const double2 a1 = Cbar[d * nb + c], g1 = dC[d * nb + c];
      const double2 a2 = Cbar[c * nb + d], g2 = dC[c * nb + d];
      eb += (a1.x * g1.x + a1.y * g1.y) - (a2.x * g2.x + a2.y * g2.y);
    }
    Ebar[d] += eb;
  }
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (tid < w) part[tid] += part[tid + w];
    __syncthreads();
  }
  if (tid == 0) glam[0] = part[0];
  // phase 4: Hbar = U (diag(Ebar) + F o (Om - Om^T)/2) U^T
  for (int e = tid; e < nb2; e += kThreads) {
    const int a = e / nb, c = e % nb;
    double v;
    if (a == c) {
      v = Ebar[a];
    } else {
      const double gap = r.E[c] - r.E[a];
      v = gap != 0.0 ? 0.5 * (Om[e] - Om[c * nb + a]) / gap : 0.0;
    }
    Mid[e] = v;
  }
  __syncthreads();
  for (int e = tid; e < nb2; e += kThreads) {
    const int p = e / nb, c = e % nb;
    double v = 0.0;
    for (int a = 0; a < nb; ++a) v += r.U[p * nb + a] * Mid[a * nb + c];
    Tmp[e] = v;
  }
  __syncthreads();
  for (int e = tid; e < nb2; e += kThreads) {
    const int p = e / nb, q = e % nb;
    double v = 0.0;
    for (int c = 0; c < nb; ++c) v += Tmp[p * nb + c] * r.U[q * nb + c];
    Hb[e] = v;
  }
}

// Hbar of the blocks onto H1: e block directly, f block through the four
// terms of operator_1_to_2 (H_f[(i,j),(k,l)] = d_ik H[j][l] + d_il H[j][k] +
// d_jk H[i][l] + d_jl H[i][k]); g has no H1 dependence
__device__ double block_hbar_to_h1(const double* Hb, int n, int X, int p, int q) {
  if (X == 1) return Hb[p * n + q];
  if (X != 2) return 0.0;
  const int nf = n * (n - 1) / 2;
  auto pair_index = [n](int lo, int hi) { return lo * n - lo * (lo + 1) / 2 + (hi - lo - 1); };
  double v = 0.0;
  // H[p][q] appears in H_f[(p,m),(q,m)] for every site m other than p and q
  // (and, for p = q, on the diagonal H_f[(p,m),(p,m)])
  for (int m = 0; m < n; ++m) {
    if (m == p || m == q) continue;
    const int s = pair_index(min(p, m), max(p, m)), t = pair_index(min(q, m), max(q, m));
    v += Hb[(size_t)s * nf + t];
  }
  return v;
}

__global__ __launch_bounds__(kThreads) void vjp_map_kernel(
    int n, int X, int Y, int nX, int nY, const double* __restrict__ chain, size_t cstride,
    double* __restrict__ gH1, double* __restrict__ glam) {
  const int b = blockIdx.y;
  const int e = blockIdx.x * kThreads + threadIdx.x;
  const int nk = X == Y ? 1 : 2;
  if (e < n * n) {
    const int p = e / n, q = e % n;
    double v = 0.0;
    for (int k = 0; k < nk; ++k) {
      const int nb = k == 0 ? nX : nY;
      const double* Hb = chain + ((size_t)b * 2 + k) * cstride + 12 * (size_t)nb * nb;
      v += block_hbar_to_h1(Hb, n, k == 0 ? X : Y, p, q);
    }
    gH1[(size_t)b * n * n + e] = v;
  }
  if (e == 0 && glam != nullptr) {
    double g = 0.0;
    for (int k = 0; k < nk; ++k) {
      const int nb = k == 0 ? nX : nY;
      g += chain[((size_t)b * 2 + k) * cstride + 13 * (size_t)nb * nb + nb];
    }
    glam[b] = g;
  }
}

size_t ws_doubles_dense(int n, int nX, int nY, int B) {
  const size_t r = std::max(rec_doubles(nX, n), rec_doubles(nY, n));
  const size_t M = (size_t)nX * nY;
  return 2 * r * B + 2 * M * M * B;
}

size_t ws_doubles(int n, int X, int Y, int B) {
  const int nX = block_size(X, n), nY = block_size(Y, n);
  const size_t r = std::max(rec_doubles(nX, n), rec_doubles(nY, n));
  const size_t M = (size_t)nX * nY;
  return 2 * r * B + 2 * M * M * B;
}

}  // namespace
}  // namespace qsx

using namespace qsx;

extern "C" int64_t qsx_redfield_workspace_bytes(int32_t n, int32_t X, int32_t Y, int32_t B) {
  if (n < 1 || B < 1 || X < 0 || X > 2 || Y < 0 || Y > 2) return -1;
  return (int64_t)(ws_doubles(n, X, Y, B) * sizeof(double));
}

static int check_block_args(int n, int X, int Y, int B) {
  QSX_CHECK_ARG(n >= 1 && B >= 1 && B <= 65535, "bad sizes n=%d B=%d", n, B);
  QSX_CHECK_ARG(X >= 0 && X <= 2 && Y >= 0 && Y <= 2, "bad Hilbert blocks X=%d Y=%d", X, Y);
  QSX_CHECK_ARG(block_size(X, n) >= 1 && block_size(Y, n) >= 1,
                "empty Hilbert block (f needs n >= 2)");
  QSX_CHECK_ARG(block_size(X, n) <= kMaxBlock && block_size(Y, n) <= kMaxBlock,
                "Hilbert block larger than %d states", kMaxBlock);
  return QSX_OK;
}

extern "C" int32_t qsx_block_eigh(const double* H1, int32_t n, int32_t B, int32_t X,
                                  const double* rw, double* E, double* U, void* workspace,
                                  void* stream) {
  QSX_CHECK_ARG(H1 && E && U && workspace, "qsx_block_eigh: null pointer");
  int st = check_block_args(n, X, X, B);
  if (st) return st;
  const int nb = block_size(X, n);
  const size_t rs = rec_doubles(nb, n);
  double* ws = static_cast<double*>(workspace);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(block_eigh_kernel, dim3(B), dim3(kThreads), 0, s, H1, n, X, rw,
                     (const double*)nullptr, 0, ws, rs);
  QSX_CHECK_HIP(hipGetLastError());
  QSX_CHECK_HIP(hipMemcpy2DAsync(E, nb * sizeof(double), ws, rs * sizeof(double),
                                 nb * sizeof(double), B, hipMemcpyDeviceToDevice, s));
  QSX_CHECK_HIP(hipMemcpy2DAsync(U, (size_t)nb * nb * sizeof(double), ws + nb, rs * sizeof(double),
                                 (size_t)nb * nb * sizeof(double), B, hipMemcpyDeviceToDevice, s));
  return QSX_OK;
}

static int redfield_block_impl(const double* H1, int n, int B, const double* rw, double T,
                               double lam, double gam, int flags, double unit_convert, int X,
                               int Y, const void* ctabX, const void* ctabY, void* L,
                               void* workspace, hipStream_t s);

extern "C" int32_t qsx_redfield_block(const double* H1, int32_t n, int32_t B, const double* rw,
                                      double T, double lam, double gam, int32_t flags,
                                      double unit_convert, int32_t X, int32_t Y, void* L,
                                      void* workspace, void* stream) {
  QSX_CHECK_ARG(H1 && L && workspace, "qsx_redfield_block: null pointer");
  int st = check_block_args(n, X, Y, B);
  if (st) return st;
  QSX_CHECK_ARG((flags & kRealCorr) ? T > 0 : (T > 0 && gam > 0),
                "qsx_redfield_block: bath needs T > 0 and cutoff > 0");
  QSX_CHECK_ARG(gam > 0, "qsx_redfield_block: cutoff frequency must be > 0");
  return redfield_block_impl(H1, n, B, rw, T, lam, gam, flags, unit_convert, X, Y, nullptr,
                             nullptr, L, workspace, as_stream(stream));
}

extern "C" int32_t qsx_redfield_block_tabulated(const double* H1, int32_t n, int32_t B,
                                                const double* rw, const void* ctabX,
                                                const void* ctabY, int32_t flags,
                                                double unit_convert, int32_t X, int32_t Y,
                                                void* L, void* workspace, void* stream) {
  QSX_CHECK_ARG(H1 && L && workspace && ctabX && ctabY,
                "qsx_redfield_block_tabulated: null pointer");
  int st = check_block_args(n, X, Y, B);
  if (st) return st;
  return redfield_block_impl(H1, n, B, rw, 1.0, 0.0, 1.0, flags, unit_convert, X, Y, ctabX,
                             ctabY, L, workspace, as_stream(stream));
}

static int redfield_block_impl(const double* H1, int n, int B, const double* rw, double T,
                               double lam, double gam, int flags, double unit_convert, int X,
                               int Y, const void* ctabX, const void* ctabY, void* L,
                               void* workspace, hipStream_t s) {
  const int nX = block_size(X, n), nY = block_size(Y, n), M = nX * nY;
  const size_t rs = std::max(rec_doubles(nX, n), rec_doubles(nY, n));
  double* wsX = static_cast<double*>(workspace);
  double* wsY = (X == Y) ? wsX : wsX + rs * B;
  double2* W = reinterpret_cast<double2*>(wsX + 2 * rs * B);
  const int blocks[2] = {X, Y};
  double* recs[2] = {wsX, wsY};
  const double2* ctabs[2] = {static_cast<const double2*>(ctabX),
                             static_cast<const double2*>(ctabY)};
  for (int k = 0; k < (X == Y ? 1 : 2); ++k) {
    const int nb = block_size(blocks[k], n);
    hipLaunchKernelGGL(block_eigh_kernel, dim3(B), dim3(kThreads), 0, s, H1, n, blocks[k], rw,
                       (const double*)nullptr, 0, recs[k], rs);
    hipLaunchKernelGGL(block_bath_kernel, dim3(B), dim3(kThreads), 0, s, n, blocks[k], nb,
                       (const double*)nullptr, T, lam, gam, (int)((flags & kRealCorr) != 0),
                       recs[k], rs, ctabs[k]);
    hipLaunchKernelGGL(block_gs_kernel, dim3(B), dim3(kThreads), 0, s, n, nb, recs[k], rs);
  }
  // non-secular blocks whose transformed factors fit the LDS: one pass
  const size_t kron = kron_lds_bytes(n, nX, nY);
  if (!(flags & kSecular) && n <= kKronMaxSites && kron <= 160 * 1024) {
#define QSX_KRON(NS_)                                                                          \
  case NS_: {                                                                                  \
    auto kern = liouv_kron_kernel<NS_>;                                                        \
    if (kron > 64 * 1024)                                                                      \
      QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                     \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)kron)); \
    hipLaunchKernelGGL(kern, dim3(B), dim3(kThreads), kron, s, nX, nY, wsX, wsY, rs, flags,    \
                       unit_convert, (double2*)L);                                             \
    break;                                                                                     \
  }
    switch (n) {
      QSX_KRON(1) QSX_KRON(2) QSX_KRON(3) QSX_KRON(4) QSX_KRON(5) QSX_KRON(6) QSX_KRON(7)
      QSX_KRON(8) QSX_KRON(9)
      default: break;
    }
#undef QSX_KRON
    QSX_CHECK_HIP(hipGetLastError());
    return QSX_OK;
  }
  const size_t shmem = 3 * (size_t)M * sizeof(double2) + (size_t)(nX * nX + nY * nY) * sizeof(double);
  QSX_CHECK_ARG(shmem <= 64 * 1024, "qsx_redfield_block: block too large (M=%d)", M);
  hipLaunchKernelGGL(liouv_cols_kernel, dim3(M, B), dim3(kThreads), shmem, s, n, nX, nY,
                     (int)(X == Y), wsX, wsY, rs, flags, unit_convert, W, (double2*)L);
  if (!(flags & kEigenBasis))
    hipLaunchKernelGGL(liouv_rows_kernel, dim3(M, B), dim3(kThreads), shmem, s, n, nX, nY, wsX,
                       wsY, rs, flags, (const double2*)W, (double2*)L);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_unitary_block(const double* H1, int32_t n, int32_t B, const double* rw,
                                     double unit_convert, int32_t X, int32_t Y,
                                     int32_t heisenberg, void* L, void* stream) {
  QSX_CHECK_ARG(H1 && L, "qsx_unitary_block: null pointer");
  int st = check_block_args(n, X, Y, B);
  if (st) return st;
  const int M = block_size(X, n) * block_size(Y, n);
  const int64_t blocks = ((int64_t)M * M + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(unitary_block_kernel, dim3((unsigned)blocks, B), dim3(kThreads), 0,
                     as_stream(stream), H1, n, rw, unit_convert, X, Y, heisenberg, (double2*)L);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int64_t qsx_redfield_dense_workspace_bytes(int32_t n, int32_t nX, int32_t nY,
                                                      int32_t B) {
  if (n < 1 || nX < 1 || nY < 1 || B < 1) return -1;
  return (int64_t)(ws_doubles_dense(n, nX, nY, B) * sizeof(double));
}

extern "C" int32_t qsx_dense_eigh(const double* H, int32_t nb, int32_t B, double* E, double* U,
                                  void* workspace, void* stream) {
  QSX_CHECK_ARG(H && E && U && workspace, "qsx_dense_eigh: null pointer");
  QSX_CHECK_ARG(nb >= 1 && nb <= kMaxBlock && B >= 1 && B <= 65535,
                "qsx_dense_eigh: bad sizes nb=%d B=%d", nb, B);
  const size_t rs = rec_doubles(nb, 1);
  double* ws = static_cast<double*>(workspace);
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(block_eigh_kernel, dim3(B), dim3(kThreads), 0, s, (const double*)nullptr, 1,
                     0, (const double*)nullptr, H, nb, ws, rs);
  QSX_CHECK_HIP(hipGetLastError());
  QSX_CHECK_HIP(hipMemcpy2DAsync(E, nb * sizeof(double), ws, rs * sizeof(double),
                                 nb * sizeof(double), B, hipMemcpyDeviceToDevice, s));
  QSX_CHECK_HIP(hipMemcpy2DAsync(U, (size_t)nb * nb * sizeof(double), ws + nb, rs * sizeof(double),
                                 (size_t)nb * nb * sizeof(double), B, hipMemcpyDeviceToDevice, s));
  return QSX_OK;
}

extern "C" int32_t qsx_redfield_block_dense(const double* HX, const double* HY, int32_t nX,
                                            int32_t nY, const double* occX, const double* occY,
                                            int32_t n, int32_t B, double T, double lam,
                                            double gam, int32_t flags, double unit_convert,
                                            void* L, void* workspace, void* stream) {
  QSX_CHECK_ARG(HX && HY && occX && occY && L && workspace,
                "qsx_redfield_block_dense: null pointer");
  QSX_CHECK_ARG(n >= 1 && B >= 1 && B <= 65535 && nX >= 1 && nY >= 1 && nX <= kMaxBlock &&
                    nY <= kMaxBlock,
                "qsx_redfield_block_dense: bad sizes n=%d nX=%d nY=%d B=%d", n, nX, nY, B);
  QSX_CHECK_ARG((flags & kRealCorr) ? T > 0 : (T > 0 && gam > 0),
                "qsx_redfield_block_dense: bath needs T > 0 and cutoff > 0");
  QSX_CHECK_ARG(gam > 0, "qsx_redfield_block_dense: cutoff frequency must be > 0");
  const int M = nX * nY;
  const bool same = HX == HY && occX == occY && nX == nY;
  const size_t rs = std::max(rec_doubles(nX, n), rec_doubles(nY, n));
  double* wsX = static_cast<double*>(workspace);
  double* wsY = same ? wsX : wsX + rs * B;
  double2* W = reinterpret_cast<double2*>(wsX + 2 * rs * B);
  hipStream_t s = as_stream(stream);
  const double* Hs[2] = {HX, HY};
  const double* occs[2] = {occX, occY};
  const int nbs[2] = {nX, nY};
  double* recs[2] = {wsX, wsY};
  for (int k = 0; k < (same ? 1 : 2); ++k) {
    hipLaunchKernelGGL(block_eigh_kernel, dim3(B), dim3(kThreads), 0, s, (const double*)nullptr,
                       n, 0, (const double*)nullptr, Hs[k], nbs[k], recs[k], rs);
    hipLaunchKernelGGL(block_bath_kernel, dim3(B), dim3(kThreads), 0, s, n, 0, nbs[k], occs[k],
                       T, lam, gam, (int)((flags & kRealCorr) != 0), recs[k], rs,
                       (const double2*)nullptr);
    hipLaunchKernelGGL(block_gs_kernel, dim3(B), dim3(kThreads), 0, s, n, nbs[k], recs[k], rs);
  }
  const size_t shmem = 3 * (size_t)M * sizeof(double2) + (size_t)(nX * nX + nY * nY) * sizeof(double);
  QSX_CHECK_ARG(shmem <= 64 * 1024, "qsx_redfield_block_dense: block too large (M=%d)", M);
  hipLaunchKernelGGL(liouv_cols_kernel, dim3(M, B), dim3(kThreads), shmem, s, n, nX, nY,
                     (int)same, wsX, wsY, rs, flags, unit_convert, W, (double2*)L);
  if (!(flags & kEigenBasis))
    hipLaunchKernelGGL(liouv_rows_kernel, dim3(M, B), dim3(kThreads), shmem, s, n, nX, nY, wsX,
                       wsY, rs, flags, (const double2*)W, (double2*)L);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_unitary_block_dense(const double* HX, const double* HY, int32_t nX,
                                           int32_t nY, int32_t B, double unit_convert,
                                           int32_t heisenberg, void* L, void* stream) {
  QSX_CHECK_ARG(HX && HY && L, "qsx_unitary_block_dense: null pointer");
  QSX_CHECK_ARG(nX >= 1 && nY >= 1 && B >= 1 && B <= 65535,
                "qsx_unitary_block_dense: bad sizes nX=%d nY=%d B=%d", nX, nY, B);
  const int M = nX * nY;
  const int64_t blocks = ((int64_t)M * M + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(unitary_dense_kernel, dim3((unsigned)blocks, B), dim3(kThreads), 0,
                     as_stream(stream), HX, HY, nX, nY, unit_convert, heisenberg, (double2*)L);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_absorption_operators(const int64_t* table, int32_t B, int32_t max_n,
                                            const double* H1, const double* rw,
                                            const double* bath, int32_t flags,
                                            double unit_convert, double h, int32_t J, void* P1,
                                            void* PJ, void* stream) {
  QSX_CHECK_ARG(table && H1 && rw && bath && P1 && PJ, "qsx_absorption_operators: null pointer");
  QSX_CHECK_ARG(B >= 1 && (J == 1 || J == 2 || J == 4 || J == 8) && max_n >= 1 && max_n <= kAbsMaxSites,
                "qsx_absorption_operators: bad sizes B=%d J=%d max_n=%d", B, J, max_n);
  QSX_CHECK_ARG((flags & ~(kSecular | kRealCorr | kHeisenberg)) == 0,
                "qsx_absorption_operators: flags %d (secular, real correlation, Heisenberg only)",
                flags);
  const size_t lds = abs_lds_bytes(max_n);
  // 256 threads per problem at every size: wider work-groups shorten a
  // problem's Jacobi rounds but leave fewer compute units to the other
  // size classes and the propagation running beside them
  constexpr int kAbsThreads = 256;
  if (lds > 64 * 1024)
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)absorption_ops_kernel<kAbsThreads>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(absorption_ops_kernel<kAbsThreads>, dim3(B), dim3(kAbsThreads), lds,
                     as_stream(stream), table, max_n, H1, rw, bath, flags, unit_convert, h, J,
                     static_cast<double2*>(P1), static_cast<double2*>(PJ));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_absorption_operators_max_sites(void) { return kAbsMaxSites; }

// -------------------------------------------------------------------- VJP
static size_t vjp_ws_doubles(int n, int X, int Y, int B) {
  const int nX = block_size(X, n), nY = block_size(Y, n);
  const size_t rs = std::max(rec_doubles(nX, n), rec_doubles(nY, n));
  const size_t M = (size_t)nX * nY;
  return 2 * rs * B + 3 * 2 * M * M * B + 2 * vjp_red_count(n, nX, nY) * B +
         2 * vjp_slab_doubles(std::max(nX, nY)) * B;
}

extern "C" int64_t qsx_redfield_vjp_workspace_bytes(int32_t n, int32_t X, int32_t Y,
                                                    int32_t B) {
  if (n < 1 || B < 1 || X < 0 || X > 2 || Y < 0 || Y > 2) return -1;
  return (int64_t)(vjp_ws_doubles(n, X, Y, B) * sizeof(double));
}

extern "C" int32_t qsx_redfield_block_vjp(const double* H1, int32_t n, int32_t B,
                                          const double* rw, double T, double lam, double gam,
                                          int32_t flags, double unit_convert, int32_t X,
                                          int32_t Y, const void* gL, double* gH1, double* glam,
                                          void* workspace, void* stream) {
  QSX_CHECK_ARG(H1 && gL && gH1 && workspace, "qsx_redfield_block_vjp: null pointer");
  int st = check_block_args(n, X, Y, B);
  if (st) return st;
  QSX_CHECK_ARG(T > 0 && gam > 0, "qsx_redfield_block_vjp: bath needs T > 0 and cutoff > 0");
  const int nX = block_size(X, n), nY = block_size(Y, n), M = nX * nY;
  const int same = X == Y;
  const size_t rs = std::max(rec_doubles(nX, n), rec_doubles(nY, n));
  const size_t shmem = 3 * (size_t)M * sizeof(double2) + (size_t)(nX * nX + nY * nY) * sizeof(double);
  QSX_CHECK_ARG(shmem <= 64 * 1024, "qsx_redfield_block_vjp: block too large (M=%d)", M);
  hipStream_t s = as_stream(stream);
  double* wsX = static_cast<double*>(workspace);
  double* wsY = same ? wsX : wsX + rs * B;
  double2* W1T = reinterpret_cast<double2*>(wsX + 2 * rs * B);
  double2* W = W1T + (size_t)M * M * B;
  double2* Lu = W + (size_t)M * M * B;
  double2* red = Lu + (size_t)M * M * B;
  const size_t rcount = vjp_red_count(n, nX, nY);
  double* chain = reinterpret_cast<double*>(red + rcount * B);
  const size_t cstride = vjp_slab_doubles(std::max(nX, nY));
  const int blocks[2] = {X, Y};
  double* recs[2] = {wsX, wsY};
  for (int k = 0; k < (same ? 1 : 2); ++k) {
    const int nb = block_size(blocks[k], n);
    hipLaunchKernelGGL(block_eigh_kernel, dim3(B), dim3(kThreads), 0, s, H1, n, blocks[k], rw,
                       (const double*)nullptr, 0, recs[k], rs);
    hipLaunchKernelGGL(block_bath_kernel, dim3(B), dim3(kThreads), 0, s, n, blocks[k], nb,
                       (const double*)nullptr, T, lam, gam, (int)((flags & kRealCorr) != 0),
                       recs[k], rs, (const double2*)nullptr);
    hipLaunchKernelGGL(block_gs_kernel, dim3(B), dim3(kThreads), 0, s, n, nb, recs[k], rs);
  }
  const double2* G = static_cast<const double2*>(gL);
  if (!(flags & kEigenBasis)) {
    hipLaunchKernelGGL(vjp_cols_kernel, dim3(M, B), dim3(kThreads), shmem, s, n, nX, nY, wsX,
                       wsY, rs, flags, G, W1T);
    hipLaunchKernelGGL(vjp_rows_kernel, dim3(M, B), dim3(kThreads), shmem, s, n, nX, nY, wsX,
                       wsY, rs, (const double2*)W1T, W);
    // the eigen-basis block itself (Schroedinger picture) for the basis-change term
    hipLaunchKernelGGL(liouv_cols_kernel, dim3(M, B), dim3(kThreads), shmem, s, n, nX, nY, same,
                       wsX, wsY, rs, (flags | kEigenBasis) & ~kHeisenberg, unit_convert,
                       (double2*)nullptr, Lu);
  }
  const unsigned rblocks = (unsigned)((rcount + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(vjp_reduce_kernel, dim3(rblocks, B), dim3(kThreads), 0, s, n, nX, nY, same,
                     flags, unit_convert, wsX, wsY, rs, G, (const double2*)W, (const double2*)Lu,
                     red, rcount);
  hipLaunchKernelGGL(vjp_chain_kernel, dim3(B, same ? 1 : 2), dim3(kThreads), 0, s, n, nX, nY,
                     same, T, lam, gam, flags, wsX, wsY, rs, (const double2*)red, rcount, chain,
                     cstride);
  hipLaunchKernelGGL(vjp_map_kernel, dim3((n * n + kThreads - 1) / kThreads, B), dim3(kThreads),
                     0, s, n, X, Y, nX, nY, (const double*)chain, cstride, gH1, glam);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}
