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
  // cyclic two-sided Jacobi, m/2 disjoint rotations per round (circle method)
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < m * m; e += kThreads) {
      const double a = A[e];
      tot += a * a;
      if (e / m != e % m) off += a * a;
    }
    red[0][tid] = off;
    red[1][tid] = tot;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
      if (tid < w) {
        red[0][tid] += red[0][tid + w];
        red[1][tid] += red[1][tid + w];
      }
      __syncthreads();
    }
    off = red[0][0];
    tot = red[1][0];
    __syncthreads();
    if (!(off > 1e-30 * tot)) break;
    for (int round = 0; round < m - 1; ++round) {
      if (tid < m / 2) {
        auto idx = [&](int j) { return ((j + round) % (m - 1)) + 1; };
        int p = tid == 0 ? 0 : idx(tid);
        int q = tid == 0 ? idx(0) : idx(m - 1 - tid);
        if (p > q) { const int t = p; p = q; q = t; }
        const double apq = A[p * m + q];
        double c = 1.0, s = 0.0;
        if (apq != 0.0) {
          const double theta = (A[q * m + q] - A[p * m + p]) / (2.0 * apq);
          const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          c = 1.0 / sqrt(t * t + 1.0);
          s = t * c;
        }
        pqs[tid][0] = p;
        pqs[tid][1] = q;
        rot[tid][0] = c;
        rot[tid][1] = s;
      }
      __syncthreads();
      for (int w = tid; w < (m / 2) * m; w += kThreads) {  // A <- A J, V <- V J
        const int k = w / m, r = w % m;
        const int p = pqs[k][0], q = pqs[k][1];
        const double c = rot[k][0], s = rot[k][1];
        const double ap = A[r * m + p], aq = A[r * m + q];
        A[r * m + p] = c * ap - s * aq;
        A[r * m + q] = s * ap + c * aq;
        const double vp = V[r * m + p], vq = V[r * m + q];
        V[r * m + p] = c * vp - s * vq;
        V[r * m + q] = s * vp + c * vq;
      }
      __syncthreads();
      for (int w = tid; w < (m / 2) * m; w += kThreads) {  // A <- J^T A
        const int k = w / m, r = w % m;
        const int p = pqs[k][0], q = pqs[k][1];
        const double c = rot[k][0], s = rot[k][1];
        const double ap = A[p * m + r], aq = A[q * m + r];
        A[p * m + r] = c * ap - s * aq;
        A[q * m + r] = s * ap + c * aq;
      }
      __syncthreads();
      if (tid < m / 2) {
        A[pqs[tid][0] * m + pqs[tid][1]] = 0.0;
        A[pqs[tid][1] * m + pqs[tid][0]] = 0.0;
      }
      __syncthreads();
    }
  }
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
// sum_{k=1}^{999} nu_k / ((nu_k^2 - g^2)(nu_k - i x)), nu_k = 2 pi k T,
// Kahan-compensated; its value at -x is the complex conjugate (each term's
// imaginary part is odd in x, its real part even), exactly so in floating
// point since only signs change
__device__ double2 matsubara_sum(double x, double T, double gam) {
  double sr = 0.0, si = 0.0, cr = 0.0, ci = 0.0;
  const double twopiT = 2.0 * M_PI * T;
  for (int k = 1; k < 1000; ++k) {
    const double nu = twopiT * k;
    const double inv = 1.0 / ((nu * nu - gam * gam) * (nu * nu + x * x));
    const double tr = nu * nu * inv - cr;
    const double ti = nu * x * inv - ci;
    const double ur = sr + tr, ui = si + ti;
    cr = (ur - sr) - tr;
    ci = (ui - si) - ti;
    sr = ur;
    si = ui;
  }
  return make_double2(sr, si);
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
                                                              size_t rstride) {
  const int b = blockIdx.x, tid = threadIdx.x;
  Rec rec = rec_at(ws + (size_t)b * rstride, nb, n);
  // pairs (d, c) with d <= c: one Matsubara sum serves C[d][c] and C[c][d]
  const int npairs = nb * (nb + 1) / 2;
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
  const int nX = block_size(X, n), nY = block_size(Y, n), M = nX * nY;
  const size_t rs = std::max(rec_doubles(nX, n), rec_doubles(nY, n));
  double* wsX = static_cast<double*>(workspace);
  double* wsY = (X == Y) ? wsX : wsX + rs * B;
  double2* W = reinterpret_cast<double2*>(wsX + 2 * rs * B);
  hipStream_t s = as_stream(stream);
  const int blocks[2] = {X, Y};
  double* recs[2] = {wsX, wsY};
  for (int k = 0; k < (X == Y ? 1 : 2); ++k) {
    const int nb = block_size(blocks[k], n);
    hipLaunchKernelGGL(block_eigh_kernel, dim3(B), dim3(kThreads), 0, s, H1, n, blocks[k], rw,
                       (const double*)nullptr, 0, recs[k], rs);
    hipLaunchKernelGGL(block_bath_kernel, dim3(B), dim3(kThreads), 0, s, n, blocks[k], nb,
                       (const double*)nullptr, T, lam, gam, (int)((flags & kRealCorr) != 0),
                       recs[k], rs);
    hipLaunchKernelGGL(block_gs_kernel, dim3(B), dim3(kThreads), 0, s, n, nb, recs[k], rs);
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
                       T, lam, gam, (int)((flags & kRealCorr) != 0), recs[k], rs);
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
