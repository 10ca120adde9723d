// Complex-Hermitian Hamiltonians (gfx950): batched eigensystems by complex
// Jacobi rotations in LDS, and unitary Liouvillian blocks -i [H, .].
//
// The reference accepts any Hermitian Hamiltonian (check_hermitian,
// qspectra/hamiltonian.py:18-22) and diagonalises it with eigh
// (hamiltonian.py:310-328); its unitary model is -i [H, .]
// (qspectra/dynamics/unitary.py:5-9).  Hilbert blocks of a one-exciton matrix
// H1 (complex, n x n): g (1 state), e (n), f (n(n-1)/2 pairs i<j, hard-core
// bosons), as for real H1.
//
// Complex Jacobi: for the pivot (p, q), b = A[p][q] = |b| e^{i phi}, the
// phase D = diag(1, e^{-i phi}) makes the pivot real, D^H A D has |b| at
// (p, q), and the real rotation J of the symmetric case zeroes it; the
// unitary G = D J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] updates
// A <- G^H A G and V <- V G.  Rounds of m/2 disjoint pivots (circle method),
// sweeps until the off-diagonal norm is below 1e-30 of the total.  The
// eigenvalues come out ascending, each eigenvector with its largest-
// magnitude component real and positive (for a real H1 the eigensystem of
// the real solver).

#include "qsx_common.h"

#include <algorithm>

namespace qsx {
namespace {

constexpr int kMaxBlock = 64;
constexpr int kThreads = 256;

__host__ __device__ inline int hblock_size(int X, int n) {
  return X == 0 ? 1 : (X == 1 ? n : n * (n - 1) / 2);
}

// f-block state s -> site pair (i, j), i < j, lexicographic (operator_tools.py:122-152)
__device__ inline void hpair_of(int s, int n, int& i, int& j) {
  int k = s;
  i = 0;
  while (k >= n - 1 - i) {
    k -= n - 1 - i;
    ++i;
  }
  j = i + 1 + k;
}

__device__ inline double2 cconj2(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ inline double2 csub2(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// entry (r, c) of Hilbert block X of the one-exciton matrix H (n x n),
// shifted by -X rw on the diagonal (operator_1_to_2, operator_tools.py:155-191)
__device__ inline double2 hblock_entry(const double2* H, int n, int X, double rw, int r, int c) {
  if (X == 0) return make_double2(0.0, 0.0);
  double2 v;
  if (X == 1) {
    v = H[r * n + c];
  } else {
    int i, j, k, l;
    hpair_of(r, n, i, j);
    hpair_of(c, n, k, l);
    v = make_double2(0.0, 0.0);
    if (i == k) v = cadd(v, H[j * n + l]);
    if (i == l) v = cadd(v, H[j * n + k]);
    if (j == k) v = cadd(v, H[i * n + l]);
    if (j == l) v = cadd(v, H[i * n + k]);
  }
  if (r == c) v.x -= X * rw;
  return v;
}

// One problem per work-group: H1 mode (block X of H1[b]) or dense mode
// (Hd[b], nbd x nbd).  LDS: A and V, m x m complex each (m = nb rounded up to
// even), dynamic.
__global__ __launch_bounds__(kThreads) void herm_eigh_kernel(const double2* __restrict__ H1, int n,
                                                             int X, const double* __restrict__ rw,
                                                             const double2* __restrict__ Hd,
                                                             int nbd, double* __restrict__ Eout,
                                                             double2* __restrict__ Uout) {
  extern __shared__ double2 hlds[];
  __shared__ double rot[kMaxBlock / 2][4];   // c, s, e^{-i phi}
  __shared__ int pqs[kMaxBlock / 2][2];
  __shared__ double red[2][kThreads];
  __shared__ int perm[kMaxBlock];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int nb = Hd ? nbd : hblock_size(X, n);
  const int m = nb + (nb & 1);
  double2* A = hlds;
  double2* V = hlds + m * m;
  const double2* H = Hd ? Hd + (size_t)b * nb * nb : H1 + (size_t)b * n * n;
  const double w = (Hd || !rw) ? 0.0 : rw[b];
  for (int e = tid; e < m * m; e += kThreads) {
    const int r = e / m, c = e % m;
    double2 v = make_double2(0.0, 0.0);
    if (r < nb && c < nb) v = Hd ? H[r * nb + c] : hblock_entry(H, n, X, w, r, c);
    if (r == c) v.y = 0.0;   // Hermitian: real diagonal
    A[e] = v;
    V[e] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < m * m; e += kThreads) {
      const double2 a = A[e];
      const double a2 = a.x * a.x + a.y * a.y;
      tot += a2;
      if (e / m != e % m) off += a2;
    }
    red[0][tid] = off;
    red[1][tid] = tot;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
      if (tid < s) {
        red[0][tid] += red[0][tid + s];
        red[1][tid] += red[1][tid + s];
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
        const double2 apq = A[p * m + q];
        const double mag = hypot(apq.x, apq.y);
        double c = 1.0, s = 0.0, phr = 1.0, phi = 0.0;
        if (mag != 0.0) {
          phr = apq.x / mag;    // e^{-i phi} = conj(apq) / |apq|
          phi = -apq.y / mag;
          const double theta = (A[q * m + q].x - A[p * m + p].x) / (2.0 * mag);
          const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          c = 1.0 / sqrt(t * t + 1.0);
          s = t * c;
        }
        pqs[tid][0] = p;
        pqs[tid][1] = q;
        rot[tid][0] = c;
        rot[tid][1] = s;
        rot[tid][2] = phr;
        rot[tid][3] = phi;
      }
      __syncthreads();
      for (int e = tid; e < (m / 2) * m; e += kThreads) {   // A <- A G, V <- V G
        const int k = e / m, r = e % m;
        const int p = pqs[k][0], q = pqs[k][1];
        const double c = rot[k][0], s = rot[k][1];
        const double2 ph = make_double2(rot[k][2], rot[k][3]);
        const double2 ap = A[r * m + p], aq = cmul(ph, A[r * m + q]);
        A[r * m + p] = csub2(cscale(c, ap), cscale(s, aq));
        A[r * m + q] = cadd(cscale(s, ap), cscale(c, aq));
        const double2 vp = V[r * m + p], vq = cmul(ph, V[r * m + q]);
        V[r * m + p] = csub2(cscale(c, vp), cscale(s, vq));
        V[r * m + q] = cadd(cscale(s, vp), cscale(c, vq));
      }
      __syncthreads();
      for (int e = tid; e < (m / 2) * m; e += kThreads) {   // A <- G^H A
        const int k = e / m, r = e % m;
        const int p = pqs[k][0], q = pqs[k][1];
        const double c = rot[k][0], s = rot[k][1];
        const double2 phc = make_double2(rot[k][2], -rot[k][3]);
        const double2 ap = A[p * m + r], aq = cmul(phc, A[q * m + r]);
        A[p * m + r] = csub2(cscale(c, ap), cscale(s, aq));
        A[q * m + r] = cadd(cscale(s, ap), cscale(c, aq));
      }
      __syncthreads();
      if (tid < m / 2) {
        const int p = pqs[tid][0], q = pqs[tid][1];
        A[p * m + q] = make_double2(0.0, 0.0);
        A[q * m + p] = make_double2(0.0, 0.0);
        A[p * m + p].y = 0.0;
        A[q * m + q].y = 0.0;
      }
      __syncthreads();
    }
  }
  // ascending eigenvalues (stable ranking over the nb real states)
  for (int k = tid; k < nb; k += kThreads) {
    const double vk = A[k * m + k].x;
    int rank = 0;
    for (int l = 0; l < nb; ++l) {
      const double vl = A[l * m + l].x;
      rank += (vl < vk) || (vl == vk && l < k);
    }
    perm[rank] = k;
  }
  __syncthreads();
  double* E = Eout + (size_t)b * nb;
  double2* U = Uout + (size_t)b * nb * nb;
  for (int k = tid; k < nb; k += kThreads) {
    const int c = perm[k];
    int best = 0;
    double bv = -1.0;
    for (int r = 0; r < nb; ++r) {
      const double2 v = V[r * m + c];
      const double a = v.x * v.x + v.y * v.y;
      if (a > bv) {
        bv = a;
        best = r;
      }
    }
    // phase making the largest component real positive
    const double2 vb = V[best * m + c];
    const double mag = hypot(vb.x, vb.y);
    const double2 ph = mag > 0.0 ? make_double2(vb.x / mag, -vb.y / mag) : make_double2(1.0, 0.0);
    E[k] = A[c * m + c].x;
    for (int r = 0; r < nb; ++r) {
      double2 u = cmul(V[r * m + c], ph);
      if (r == best) u.y = 0.0;
      U[r * nb + k] = u;
    }
  }
}

// L[(i,j),(k,l)] = -i unit (H_X[i][k] d_jl - d_ik H_Y[l][j]), p = i + nX j,
// q = k + nX l (column stacking); H_X, H_Y blocks of a complex H1 (H1 mode) or
// given dense (HX, HY non-null)
__global__ __launch_bounds__(kThreads) void herm_unitary_kernel(
    const double2* __restrict__ H1, int n, const double* __restrict__ rw, int X, int Y,
    const double2* __restrict__ HX, const double2* __restrict__ HY, int nXd, int nYd, double unit,
    int heisenberg, double2* __restrict__ L) {
  const int b = blockIdx.y;
  const bool dense = HX != nullptr;
  const int nX = dense ? nXd : hblock_size(X, n), nY = dense ? nYd : hblock_size(Y, n);
  const int M = nX * nY;
  const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= (int64_t)M * M) return;
  const int p = (int)(e / M), q = (int)(e % M);
  const int i = p % nX, j = p / nX, k = q % nX, l = q / nX;
  double2 v = make_double2(0.0, 0.0);
  if (dense) {
    if (j == l) v = cadd(v, HX[(size_t)b * nX * nX + i * nX + k]);
    if (i == k) v = csub2(v, HY[(size_t)b * nY * nY + l * nY + j]);
  } else {
    const double2* H = H1 + (size_t)b * n * n;
    const double w = rw ? rw[b] : 0.0;
    if (j == l) v = cadd(v, hblock_entry(H, n, X, w, i, k));
    if (i == k) v = csub2(v, hblock_entry(H, n, Y, w, l, j));
  }
  const double2 val = make_double2(unit * v.y, -unit * v.x);   // -i unit v
  double2* Lb = L + (size_t)b * M * M;
  if (heisenberg) Lb[(size_t)q * M + p] = val;
  else Lb[e] = val;
}

int launch_herm_eigh(const double2* H1, int n, int X, const double* rw, const double2* Hd, int nbd,
                     int nb, int B, double* E, double2* U, hipStream_t s) {
  const int m = nb + (nb & 1);
  const size_t lds = 2 * (size_t)m * m * sizeof(double2);
  if (lds > 64 * 1024)
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)herm_eigh_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(herm_eigh_kernel, dim3(B), dim3(kThreads), lds, s, H1, n, X, rw, Hd, nbd, E,
                     U);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

}  // namespace
}  // namespace qsx

using namespace qsx;

extern "C" int32_t qsx_block_eigh_hermitian(const void* H1, int32_t n, int32_t B, int32_t X,
                                            const double* rw, double* E, void* U, void* stream) {
  QSX_CHECK_ARG(H1 && E && U, "qsx_block_eigh_hermitian: null pointer");
  QSX_CHECK_ARG(n >= 1 && B >= 1 && B <= 65535 && X >= 0 && X <= 2,
                "qsx_block_eigh_hermitian: bad sizes n=%d B=%d X=%d", n, B, X);
  const int nb = hblock_size(X, n);
  QSX_CHECK_ARG(nb >= 1 && nb <= kMaxBlock, "qsx_block_eigh_hermitian: block of %d states (max %d)",
                nb, kMaxBlock);
  return launch_herm_eigh(static_cast<const double2*>(H1), n, X, rw, nullptr, 0, nb, B, E,
                          static_cast<double2*>(U), as_stream(stream));
}

extern "C" int32_t qsx_dense_eigh_hermitian(const void* H, int32_t nb, int32_t B, double* E,
                                            void* U, void* stream) {
  QSX_CHECK_ARG(H && E && U, "qsx_dense_eigh_hermitian: null pointer");
  QSX_CHECK_ARG(nb >= 1 && nb <= kMaxBlock && B >= 1 && B <= 65535,
                "qsx_dense_eigh_hermitian: bad sizes nb=%d B=%d", nb, B);
  return launch_herm_eigh(nullptr, 1, 0, nullptr, static_cast<const double2*>(H), nb, nb, B, E,
                          static_cast<double2*>(U), as_stream(stream));
}

extern "C" int32_t qsx_unitary_block_hermitian(const void* H1, int32_t n, int32_t B,
                                               const double* rw, double unit_convert, int32_t X,
                                               int32_t Y, int32_t heisenberg, void* L,
                                               void* stream) {
  QSX_CHECK_ARG(H1 && L, "qsx_unitary_block_hermitian: null pointer");
  QSX_CHECK_ARG(n >= 1 && B >= 1 && B <= 65535 && X >= 0 && X <= 2 && Y >= 0 && Y <= 2,
                "qsx_unitary_block_hermitian: bad sizes n=%d B=%d X=%d Y=%d", n, B, X, Y);
  const int M = hblock_size(X, n) * hblock_size(Y, n);
  QSX_CHECK_ARG(M >= 1, "qsx_unitary_block_hermitian: empty block");
  const int64_t blocks = ((int64_t)M * M + kThreads - 1) / kThreads;
  QSX_CHECK_ARG(blocks < (1ll << 31), "qsx_unitary_block_hermitian: block too large (M=%d)", M);
  hipLaunchKernelGGL(herm_unitary_kernel, dim3((unsigned)blocks, B), dim3(kThreads), 0,
                     as_stream(stream), static_cast<const double2*>(H1), n, rw, X, Y,
                     (const double2*)nullptr, (const double2*)nullptr, 0, 0, unit_convert,
                     heisenberg, static_cast<double2*>(L));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_unitary_block_dense_hermitian(const void* HX, const void* HY, int32_t nX,
                                                     int32_t nY, int32_t B, double unit_convert,
                                                     int32_t heisenberg, void* L, void* stream) {
  QSX_CHECK_ARG(HX && HY && L, "qsx_unitary_block_dense_hermitian: null pointer");
  QSX_CHECK_ARG(nX >= 1 && nY >= 1 && B >= 1 && B <= 65535,
                "qsx_unitary_block_dense_hermitian: bad sizes nX=%d nY=%d B=%d", nX, nY, B);
  const int M = nX * nY;
  const int64_t blocks = ((int64_t)M * M + kThreads - 1) / kThreads;
  QSX_CHECK_ARG(blocks < (1ll << 31), "qsx_unitary_block_dense_hermitian: block too large");
  hipLaunchKernelGGL(herm_unitary_kernel, dim3((unsigned)blocks, B), dim3(kThreads), 0,
                     as_stream(stream), (const double2*)nullptr, 1, (const double*)nullptr, 0, 0,
                     static_cast<const double2*>(HX), static_cast<const double2*>(HY), nX, nY,
                     unit_convert, heisenberg, static_cast<double2*>(L));
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}
