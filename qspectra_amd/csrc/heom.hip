// HEOM (hierarchical equations of motion) right-hand side and fused RK4
// stages for a Drude bath, batched over static-disorder samples (gfx950).
//
// State layout (qspectra/dynamics/heom.py:137-151, 278-286): per sample,
// ADO-major [n_ado][M] with ADO 0 the reduced density matrix.  In the site
// basis every bath term of heom.py:395-442 is diagonal in the Liouville
// index p = (r, c):  [P_j, rho]_p = (n_j(r) - n_j(c)) rho_p,
// (c_k P_j^L - c_k^* P_j^R) rho = (c_k n_j(r) - c_k^* n_j(c)) rho_p and the
// low-temperature double commutator is Delta sum_j (n_j(r) + n_j(c) -
// 2 n_j(r) n_j(c)).  So
//   d rho_n / dt = Lh rho_n + (shift_n + dvec) o rho_n
//                  + sum_{bins b = (j, k)} [ coef_{n,b,0} (n_j(r) - n_j(c)) rho_{src_{n,b,0}}
//                       + coef_{n,b,1} (c_k n_j(r) - c_k^* n_j(c)) rho_{src_{n,b,1}} ]_p
// with Lh the (sample-dependent) -i unit [H, .] block, shift_n the ADO
// damping and binned neighbour tables built on the host: kind 0 the n + e_b
// neighbour, kind 1 the n - e_b neighbour (or, in the Heisenberg picture,
// the ADO that has n as that neighbour; one edge per ADO, bin and kind
// either way).  The weights vanish unless site j is occupied in the row or
// the column state of element p, so each element walks only the bins of its
// occupied sites (a host list per element): <= 8 of the 28 edges of a
// 7-site, K = 1 hierarchy on the 'ee' block.
//
// Lh is applied either as a dense M x M block (any subspace, M <= 64) or,
// for a single Liouville block XY (M = nX nY), as the commutator
// -i unit (H_X rho - rho H_Y) with the Hilbert blocks H_X (nX x nX) and H_Y
// (nY x nY): nX + nY instead of M products per element.
//
// One RK4 stage is one launch over (sample, ADO tile).  Classical RK4 is
// stored without an update accumulator: with z1 = y, z2 = y + h/2 k1,
// z3 = y + h/2 k2, z4 = y + h k3, the step is
//   y' = y + ((z2 - y) + 2 (z3 - y) + (z4 - y)) / 3 + h/6 k4,
// so FP64 keeps y and the stage vectors z2, z3, z4 (208 bytes per element
// and step instead of 288 with a separate accumulator).  MIXED keeps the
// stage vectors and the stage derivatives k1..k3 in complex64 (a complex64
// copy z1 of the state serves the neighbour gathers of stage 1) and forms
// y' = y + h/6 (k1 + 2 k2 + 2 k3 + k4) in FP64 (160 bytes instead of 224).
// Work-groups are mapped so that every tile of one sample lands on the same
// XCD (work-groups are dispatched round-robin over the 8 XCDs): a sample's
// hierarchy and its neighbour re-reads stay in one L2.

#include "qsx_common.h"

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <vector>

namespace qsx {
namespace {

constexpr int kThreads = 256;
constexpr int kXcds = 8;
#ifndef QSX_HEOM_ELEMS_PER_THREAD
#define QSX_HEOM_ELEMS_PER_THREAD 3
#endif
// Liouville elements per thread and stage: a work-group covers
// kElemsPerThread * 256 / M ADOs (with their edge tables in LDS).  Measured
// on the 7-site, 3060-ADO hierarchy (MI355X, passes of 24 samples): 2
// elements per thread 9% faster than 1 (the work-group's set-up loads are
// amortised over more elements); with the compile-time commutator kernel 3
// are 3-5% faster than 2, 4 no faster than 3.
constexpr int kElemsPerThread = QSX_HEOM_ELEMS_PER_THREAD;

struct HeomArgs {
  int B, n_ado, M, nX, nY, dense, nbins, K1, nrel, tiles, ado_per_tile, xcd_map;
  int N = 0;               // HILBERT form: Hilbert-space dimension (0 = other forms)
  const double2* Hop;      // COMM: [B][nX*nX + nY*nY] = -i unit H_X, H_Y (or their
                           // transposes); DENSE: [B][M][M]; HILBERT: [B][N][N]
  const int* est = nullptr;   // HILBERT: [M][2] row and column state of element p
  const int* spos = nullptr;  // HILBERT: [N][N] element of (row, column) (-1 absent)
  const int* sblk = nullptr;  // HILBERT: [N][2] first and end state of a state's block
  const double* shift;     // [n_ado]   -unit * sum m nu
  const double2* dvec;     // [M]       -unit * Delta * D_p
  const int* bsrc;         // [n_ado][nbins][2] neighbour ADO (-1 = none)
  const double2* bcoef;    // [n_ado][nbins][2] coefficient (includes unit)
  const double2* ck;       // [K1]      correlation-function coefficients c_k
  const int* rsite;        // [M][nrel] sites occupied in the row or column state (-1 pad)
  const double* rocc;      // [M][nrel][2] their occupations (row, column)
};

template <typename T>
__device__ inline double2 ld(const cplx_t<T>* p) { return to_d2(*p); }

// operator entries per sample: the dense block, or H_X and H_Y (one block
// for a square subspace XX, where H_Y = H_X)
__host__ __device__ inline int heom_opsz(const HeomArgs& a) {
  if (a.N > 0) return a.N * a.N;
  return a.dense ? a.M * a.M : (a.nX == a.nY ? a.nX * a.nX : a.nX * a.nX + a.nY * a.nY);
}

// LDS layout of a stage work-group: operator block, the tile's stage
// vectors, and the tile's edge tables (so that the neighbour gathers do not
// wait on dependent global loads of the table).
inline size_t heom_lds_bytes(int opsz, int ado_per_tile, int M, int nbins, size_t zbytes,
                             int N = 0) {
  size_t off = (size_t)opsz * sizeof(double2);
  off += (size_t)ado_per_tile * M * zbytes;
  off = (off + 15) & ~(size_t)15;
  off += 2 * (size_t)ado_per_tile * nbins * sizeof(double2);
  off += 2 * (size_t)ado_per_tile * nbins * sizeof(int);
  off += (size_t)N * N * sizeof(int);                     // HILBERT: element positions
  return off;
}

// Neighbour terms of element p with NR occupied-site slots and KB bins per
// site known at compile time: every gather's offset and weight first
// (absent edges read the element's own ADO-0 entry with weight 0), then all
// the loads, then the products -- so the loads of one element issue together
// instead of one dependent round trip per edge.
template <typename T, int NR, int KB>
__device__ inline double2 neighbour_sum(const HeomArgs& a, const int* ts, const double2* tc,
                                        const cplx_t<T>* zs, int p, int M) {
  constexpr int NE = 2 * NR * KB;
  int64_t off[NE];
  double2 w[NE];
#pragma unroll
  for (int t = 0; t < NR; ++t) {
    const int j = a.rsite[p * NR + t];
    const bool okj = j >= 0 && (j + 1) * KB <= a.nbins;
    const double orow = okj ? a.rocc[(p * NR + t) * 2] : 0.0;
    const double ocol = okj ? a.rocc[(p * NR + t) * 2 + 1] : 0.0;
    const double wp = orow - ocol;
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
      const int bi = okj ? 2 * (j * KB + kk) : 0;
      const int s0 = okj ? ts[bi] : -1, s1 = okj ? ts[bi + 1] : -1;
      const bool v0 = s0 >= 0 && s0 < a.n_ado && wp != 0.0;
      const bool v1 = s1 >= 0 && s1 < a.n_ado;
      const double2 c0 = tc[bi], c1 = tc[bi + 1], c = a.ck[kk];
      // c_k n_j(r) - conj(c_k) n_j(c)
      const double2 wm = make_double2(c.x * orow - c.x * ocol, c.y * orow + c.y * ocol);
      const int e0 = 2 * (t * KB + kk);
      off[e0] = v0 ? (int64_t)s0 * M : 0;
      w[e0] = v0 ? make_double2(c0.x * wp, c0.y * wp) : make_double2(0.0, 0.0);
      off[e0 + 1] = v1 ? (int64_t)s1 * M : 0;
      w[e0 + 1] = v1 ? cmul(c1, wm) : make_double2(0.0, 0.0);
    }
  }
  cplx_t<T> v[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) v[e] = zs[off[e]];
  double2 k = make_double2(0.0, 0.0);
#pragma unroll
  for (int e = 0; e < NE; ++e) k = cadd(k, cmul(w[e], to_d2(v[e])));
  return k;
}

// Buffers of the RK4 stages (see the file comment): y the FP64 state, z[1..3]
// the stage vectors z2..z4, z[0] the complex64 copy of y (MIXED only), k the
// complex64 stage derivatives k1..k3 (MIXED only).
template <typename T>
struct HeomStageBufs {
  double2* y;
  cplx_t<T>* z[4];
  cplx_t<T>* k[3];
};

// The stage's output for element gi given its derivative k = F(zin)[gi] and
// zp = zin[gi] (see the file comment for the buffers): st = 4 writes
// zout = k; otherwise the RK4 stage st.
template <typename T>
__device__ __forceinline__ void stage_epilogue(double2 k, double2 zp, int64_t gi, int st, double h,
                                               const HeomStageBufs<T>& bf,
                                               cplx_t<T>* __restrict__ zout) {
  if (st == 4) {                           // plain application: zout = F(zin)
    zout[gi] = from_d2<T>(k);
    return;
  }
  if constexpr (std::is_same<T, double>::value) {
    if (st == 0) {                         // zin is y: z2 = y + h/2 k1
      bf.z[1][gi] = caxpy(0.5 * h, k, zp);
    } else if (st < 3) {                   // z3 = y + h/2 k2, z4 = y + h k3
      bf.z[st + 1][gi] = caxpy(st == 1 ? 0.5 * h : h, k, bf.y[gi]);
    } else {                               // y' from the stage vectors and k4
      const double2 yv = bf.y[gi], z2 = bf.z[1][gi], z3 = bf.z[2][gi];
      const double dr = ((z2.x - yv.x) + 2.0 * (z3.x - yv.x) + (zp.x - yv.x)) / 3.0;
      const double di = ((z2.y - yv.y) + 2.0 * (z3.y - yv.y) + (zp.y - yv.y)) / 3.0;
      bf.y[gi] = caxpy(h / 6.0, k, make_double2(yv.x + dr, yv.y + di));
    }
  } else {
    if (st < 3) {                          // k_{st+1} and the next stage vector
      bf.k[st][gi] = from_d2<T>(k);
      const double2 y1 = st == 0 ? zp : ld<T>(bf.z[0] + gi);
      bf.z[st + 1][gi] = from_d2<T>(caxpy(st == 2 ? h : 0.5 * h, k, y1));
    } else {
      const double2 k1 = ld<T>(bf.k[0] + gi), k2 = ld<T>(bf.k[1] + gi),
                    k3 = ld<T>(bf.k[2] + gi);
      const double2 sum = make_double2(k1.x + 2.0 * (k2.x + k3.x) + k.x,
                                       k1.y + 2.0 * (k2.y + k3.y) + k.y);
      const double2 yn = caxpy(h / 6.0, sum, bf.y[gi]);
      bf.y[gi] = yn;
      bf.z[0][gi] = from_d2<T>(yn);
    }
  }
}

// Stage st (0..3) of RK4 on input zin (y itself for stage 0 in FP64, z[st]
// otherwise), or st = 4: zout = F(zin).  NR, KB > 0: the neighbour terms
// with that many occupied-site slots and bins per site (neighbour_sum), each
// shape its own kernel so that its register count is its own; NR = 0: any
// shape, edge by edge.
template <typename T, int NR, int KB>
__global__ __launch_bounds__(kThreads) void heom_stage_kernel(
    HeomArgs a, const cplx_t<T>* __restrict__ zin, cplx_t<T>* __restrict__ zout,
    HeomStageBufs<T> bf, double h, int st) {
  using T2 = cplx_t<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int opsz = heom_opsz(a);
  const int nb2 = 2 * a.nbins;
  double2* Hs = reinterpret_cast<double2*>(smem);                 // H_X, H_Y (COMM) or M*M
  T2* zt = reinterpret_cast<T2*>(Hs + opsz);                      // [ado_per_tile][M]
  size_t off = (size_t)opsz * sizeof(double2) + (size_t)a.ado_per_tile * a.M * sizeof(T2);
  off = (off + 15) & ~(size_t)15;
  double2* tcf = reinterpret_cast<double2*>(smem + off);          // [ado_per_tile][nbins][2]
  int* tsrc = reinterpret_cast<int*>(tcf + a.ado_per_tile * nb2);  // [ado_per_tile][nbins][2]
  int* sp = tsrc + a.ado_per_tile * nb2;                          // HILBERT: [N][N]

  int b, tile;
  if (a.xcd_map) {
    const int x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    b = x + kXcds * (j / a.tiles);
    tile = j % a.tiles;
  } else {
    b = blockIdx.x / a.tiles;
    tile = blockIdx.x % a.tiles;
  }
  if (b >= a.B) return;                      // uniform per work-group
  const int tid = threadIdx.x;
  const int M = a.M;
  const int ado0 = tile * a.ado_per_tile;
  const int nloc = min(a.ado_per_tile, a.n_ado - ado0);
  const int64_t base = ((int64_t)b * a.n_ado + ado0) * M;
  const double2* Hb = a.Hop + (int64_t)b * opsz;
  for (int e = tid; e < opsz; e += kThreads) Hs[e] = Hb[e];
  for (int e = tid; e < nloc * M; e += kThreads) zt[e] = zin[base + e];
  for (int e = tid; e < nloc * nb2; e += kThreads) {
    const int64_t g = (int64_t)ado0 * nb2 + e;
    tsrc[e] = a.bsrc[g];
    tcf[e] = a.bcoef[g];
  }
  for (int e = tid; e < a.N * a.N; e += kThreads) sp[e] = a.spos[e];
  __syncthreads();

  for (int e = tid; e < nloc * M; e += kThreads) {
    const int la = e / M, p = e % M, ado = ado0 + la;
    const T2* z = zt + la * M;
    double2 k = make_double2(0.0, 0.0);
    if (a.dense) {
      const double2* row = Hs + p * M;
      for (int q = 0; q < M; ++q) k = cadd(k, cmul(row[q], ld<T>(z + q)));
    } else if (a.N > 0) {
      // (H rho - rho H)_{ra,cb} over the Hilbert blocks of the row state ra
      // and of the column state cb: every (q, cb) and (ra, q) lies in the
      // Liouville block of element p, so it is in the subspace
      const int N = a.N, ra = a.est[2 * p], cb = a.est[2 * p + 1];
      for (int q = a.sblk[2 * ra]; q < a.sblk[2 * ra + 1]; ++q)
        k = cadd(k, cmul(Hs[ra * N + q], ld<T>(z + sp[q * N + cb])));
      for (int q = a.sblk[2 * cb]; q < a.sblk[2 * cb + 1]; ++q) {
        const double2 t = cmul(ld<T>(z + sp[ra * N + q]), Hs[q * N + cb]);
        k.x -= t.x;
        k.y -= t.y;
      }
    } else {   // (H_X rho - rho H_Y)_{rc}, p = r + nX c, column-major rho
      const int nX = a.nX, nY = a.nY, r = p % nX, c = p / nX;
      const double2* HY = nX == nY ? Hs : Hs + nX * nX;     // square blocks: H_Y = H_X
      for (int q = 0; q < nX; ++q) k = cadd(k, cmul(Hs[r * nX + q], ld<T>(z + q + nX * c)));
      for (int q = 0; q < nY; ++q) {
        const double2 t = cmul(ld<T>(z + r + nX * q), HY[q * nY + c]);
        k.x -= t.x;
        k.y -= t.y;
      }
    }
    const double2 zp = ld<T>(z + p);
    const double2 dg = make_double2(a.shift[ado] + a.dvec[p].x, a.dvec[p].y);
    k = cadd(k, cmul(dg, zp));
    // neighbour terms: only the bins of the sites occupied in row or column
    const int* ts = tsrc + la * nb2;
    const double2* tc = tcf + la * nb2;
    const T2* zs = zin + (int64_t)b * a.n_ado * M + p;
    if constexpr (NR > 0) {
      k = cadd(k, neighbour_sum<T, NR, KB>(a, ts, tc, zs, p, M));
    } else {
      for (int t = 0; t < a.nrel; ++t) {
        const int j = a.rsite[p * a.nrel + t];
        if (j < 0 || (j + 1) * a.K1 > a.nbins) continue;
        const double orow = a.rocc[(p * a.nrel + t) * 2], ocol = a.rocc[(p * a.nrel + t) * 2 + 1];
        const double wp = orow - ocol;
        for (int kk = 0; kk < a.K1; ++kk) {
          const int bi = 2 * (j * a.K1 + kk);
          const int s0 = ts[bi], s1 = ts[bi + 1];
          if (s0 >= 0 && s0 < a.n_ado && wp != 0.0) {
            const double2 w = make_double2(tc[bi].x * wp, tc[bi].y * wp);
            k = cadd(k, cmul(w, ld<T>(zs + (int64_t)s0 * M)));
          }
          if (s1 >= 0 && s1 < a.n_ado) {
            const double2 c = a.ck[kk];
            // c_k n_j(r) - conj(c_k) n_j(c)
            const double2 wm = make_double2(c.x * orow - c.x * ocol, c.y * orow + c.y * ocol);
            k = cadd(k, cmul(cmul(tc[bi + 1], wm), ld<T>(zs + (int64_t)s1 * M)));
          }
        }
      }
    }
    stage_epilogue<T>(k, zp, base + e, st, h, bf, zout);
  }
}

// The commutator form with its block sizes known at compile time (NX x NY
// elements per ADO, NR occupied-site slots, KB bins per site): the same
// stage as heom_stage_kernel with every index division by a constant, the
// 2 (NX + NY) commutator terms unrolled with LDS offsets fixed at compile
// time, and the per-element neighbour weights (site bin, n_j(r) - n_j(c),
// c_k n_j(r) - conj(c_k) n_j(c)) tabulated once per work-group.  Absent
// edges point at the element's own ADO with a zero coefficient, so the
// gathers need no tests.  Counters on the 7-site, 3060-ADO hierarchy showed
// the generic kernel issuing ~490 vector instructions per element (the
// vector units busy 60% of the launch); this one issues about half as many,
// which made the MIXED stage 13% faster and the FP64 stage 2-5%: what is
// left is the latency of the neighbour gathers at four waves per SIMD (the
// 120 registers; forcing six or eight waves spills).
template <typename T, int NX, int NY, int NR, int KB>
__global__ __launch_bounds__(kThreads) void heom_comm_kernel(
    HeomArgs a, const cplx_t<T>* __restrict__ zin, cplx_t<T>* __restrict__ zout,
    HeomStageBufs<T> bf, double h, int st) {
  using T2 = cplx_t<T>;
  constexpr int M = NX * NY;
  constexpr int OPSZ = NX == NY ? NX * NX : NX * NX + NY * NY;
  __shared__ int pbin[M][NR];           // 2 * (first bin of the slot's site)
  __shared__ double pwp[M][NR];         // n_j(r) - n_j(c)
  __shared__ double2 pwm[M][NR][KB];    // c_k n_j(r) - conj(c_k) n_j(c)
  extern __shared__ __align__(16) unsigned char smem[];
  const int nb2 = 2 * a.nbins;
  double2* Hs = reinterpret_cast<double2*>(smem);
  T2* zt = reinterpret_cast<T2*>(Hs + OPSZ);
  size_t off = (size_t)OPSZ * sizeof(double2) + (size_t)a.ado_per_tile * M * sizeof(T2);
  off = (off + 15) & ~(size_t)15;
  double2* tcf = reinterpret_cast<double2*>(smem + off);
  int* tsrc = reinterpret_cast<int*>(tcf + a.ado_per_tile * nb2);

  int b, tile;
  if (a.xcd_map) {
    const int x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    b = x + kXcds * (j / a.tiles);
    tile = j % a.tiles;
  } else {
    b = blockIdx.x / a.tiles;
    tile = blockIdx.x % a.tiles;
  }
  if (b >= a.B) return;                      // uniform per work-group
  const int tid = threadIdx.x;
  const int ado0 = tile * a.ado_per_tile;
  const int nloc = min(a.ado_per_tile, a.n_ado - ado0);
  const int64_t base = ((int64_t)b * a.n_ado + ado0) * M;
  const double2* Hb = a.Hop + (int64_t)b * OPSZ;
  for (int e = tid; e < OPSZ; e += kThreads) Hs[e] = Hb[e];
  for (int e = tid; e < nloc * M; e += kThreads) zt[e] = zin[base + e];
  for (int e = tid; e < nloc * nb2; e += kThreads) {
    const int64_t g = (int64_t)ado0 * nb2 + e;
    const int src = a.bsrc[g];
    const bool ok = src >= 0 && src < a.n_ado;
    tsrc[e] = ok ? src : ado0 + e / nb2;
    tcf[e] = ok ? a.bcoef[g] : make_double2(0.0, 0.0);
  }
  for (int e = tid; e < M * NR; e += kThreads) {
    const int p = e / NR, t = e % NR;
    const int j = a.rsite[p * NR + t];
    const bool okj = j >= 0 && (j + 1) * KB <= a.nbins;
    const double orow = okj ? a.rocc[(p * NR + t) * 2] : 0.0;
    const double ocol = okj ? a.rocc[(p * NR + t) * 2 + 1] : 0.0;
    pbin[p][t] = okj ? 2 * j * KB : 0;
    pwp[p][t] = orow - ocol;
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
      const double2 c = a.ck[kk];
      pwm[p][t][kk] = make_double2(c.x * orow - c.x * ocol, c.y * orow + c.y * ocol);
    }
  }
  __syncthreads();

  const double2* HY = NX == NY ? Hs : Hs + NX * NX;   // square blocks: H_Y = H_X
  const T2* zb = zin + (int64_t)b * a.n_ado * M;
  for (int e = tid; e < nloc * M; e += kThreads) {
    const int la = e / M, p = e - la * M, ado = ado0 + la;
    const int r = p % NX, c = p / NX;
    const T2* z = zt + la * M;
    // (H_X rho - rho H_Y)_{rc}, p = r + NX c, column-major rho
    double2 k = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NX; ++q) k = cadd(k, cmul(Hs[r * NX + q], ld<T>(z + q + NX * c)));
#pragma unroll
    for (int q = 0; q < NY; ++q) {
      const double2 t = cmul(ld<T>(z + r + NX * q), HY[q * NY + c]);
      k.x -= t.x;
      k.y -= t.y;
    }
    const double2 zp = ld<T>(z + p);
    const double2 dg = make_double2(a.shift[ado] + a.dvec[p].x, a.dvec[p].y);
    k = cadd(k, cmul(dg, zp));
    // neighbour terms: all offsets and weights, then all gathers, then the sum
    const int* ts = tsrc + la * nb2;
    const double2* tc = tcf + la * nb2;
    constexpr int NE = 2 * NR * KB;
    int src[NE];
    double2 w[NE];
#pragma unroll
    for (int t = 0; t < NR; ++t) {
      const int b0 = pbin[p][t];
      const double wp = pwp[p][t];
#pragma unroll
      for (int kk = 0; kk < KB; ++kk) {
        const int bi = b0 + 2 * kk, e0 = 2 * (t * KB + kk);
        src[e0] = ts[bi];
        src[e0 + 1] = ts[bi + 1];
        const double2 c0 = tc[bi];
        w[e0] = make_double2(c0.x * wp, c0.y * wp);
        w[e0 + 1] = cmul(tc[bi + 1], pwm[p][t][kk]);
      }
    }
    T2 v[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) v[q] = zb[src[q] * M + p];
#pragma unroll
    for (int q = 0; q < NE; ++q) k = cadd(k, cmul(w[q], to_d2(v[q])));
    stage_epilogue<T>(k, zp, base + e, st, h, bf, zout);
  }
}

template <typename T>
using StageKernel = void (*)(HeomArgs, const cplx_t<T>*, cplx_t<T>*, HeomStageBufs<T>, double,
                             int);

// the stage kernel for the hierarchy's neighbour shape (occupied-site slots
// per element, bins per site)
template <typename T>
StageKernel<T> pick_stage_kernel(const HeomArgs& a) {
  if (!a.dense && a.N == 0 && a.nX == a.nY && a.nrel == 2 && (a.K1 == 2 || a.K1 == 3)) {
#define QSX_COMM(n_)                                                               \
    if (a.nX == n_) return a.K1 == 2 ? heom_comm_kernel<T, n_, n_, 2, 2>           \
                                     : heom_comm_kernel<T, n_, n_, 2, 3>;
    QSX_COMM(2) QSX_COMM(3) QSX_COMM(4) QSX_COMM(5) QSX_COMM(6) QSX_COMM(7) QSX_COMM(8)
#undef QSX_COMM
  }
  if (a.nrel == 2 && a.K1 == 2) return heom_stage_kernel<T, 2, 2>;
  if (a.nrel == 1 && a.K1 == 2) return heom_stage_kernel<T, 1, 2>;
  if (a.nrel == 3 && a.K1 == 2) return heom_stage_kernel<T, 3, 2>;
  if (a.nrel == 4 && a.K1 == 2) return heom_stage_kernel<T, 4, 2>;
  if (a.nrel == 2 && a.K1 == 3) return heom_stage_kernel<T, 2, 3>;
  if (a.nrel == 1 && a.K1 == 3) return heom_stage_kernel<T, 1, 3>;
  return heom_stage_kernel<T, 0, 0>;
}

template <typename T>
__global__ void heom_init_kernel(const double2* __restrict__ y0, double2* __restrict__ y,
                                 cplx_t<T>* __restrict__ z, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) {
    y[i] = y0[i];
    if (z != nullptr) z[i] = from_d2<T>(y0[i]);
  }
}

// Classical fields on a dense hierarchy block (simulate_with_fields of a HEOM
// model, qspectra/simulate/eom.py:79-103 with heom.py:22-58: the dipole
// commutators act on every ADO): the stage operator of sample b at stage
// sample sidx of global step g is Hop[b] + sum_i f_i C_i[b], f_i = F[b][g][sidx][i].
struct HeomFields {
  int P = 0;                    // field terms (0: no fields)
  const double2* C = nullptr;   // [B|1][P][M][M] dense commutators (times unit)
  int64_t strideC = 0;          // P*M*M or 0
  const double2* F = nullptr;   // [B|1][nF][3][P] -i E(t) (conjugated for '+')
  int64_t strideF = 0;          // nF*3*P or 0
  double2* Hwork = nullptr;     // [B][M][M] stage operators
};

__global__ void heom_field_op_kernel(const double2* __restrict__ Hop,
                                     const double2* __restrict__ C, int64_t strideC,
                                     const double2* __restrict__ F, int64_t strideF, int64_t g,
                                     int sidx, int P, int B, int opsz,
                                     double2* __restrict__ Hout) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * opsz) return;
  const int64_t b = i / opsz, e = i % opsz;
  const double2* f = F + b * strideF + (g * 3 + sidx) * P;
  const double2* Cb = C + b * strideC;
  double2 v = Hop[i];
  for (int k = 0; k < P; ++k) v = cadd(v, cmul(f[k], Cb[(int64_t)k * opsz + e]));
  Hout[i] = v;
}

// out[b][s][0:M] = ADO 0 of sample b; status on non-finite values
__global__ void heom_save_kernel(const double2* __restrict__ y, int n_ado, int M, int B, int s,
                                 int nsave, double2* __restrict__ out, int* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * M) return;
  const int b = i / M, p = i % M;
  const double2 v = y[(int64_t)b * n_ado * M + p];
  out[((int64_t)b * nsave + s) * M + p] = v;
  if (!cfinite(v)) status[0] = 1;
}

// Samples per pass: the whole hierarchy of a pass (y, ynew and the two stage
// vectors) is kept small enough to stay largely resident in the 256 MB
// memory-attached last-level cache while all its steps run, so the stage
// vectors and the neighbour gathers are served from there instead of HBM;
// a multiple of the XCD count so that every XCD holds whole samples.
// QSX_HEOM_SAMPLES_PER_PASS overrides the choice (0 = all samples at once).
inline int heom_samples_per_pass(int B, int64_t bytes_per_sample) {
  const char* env = getenv("QSX_HEOM_SAMPLES_PER_PASS");
  if (env != nullptr) {
    const int v = atoi(env);
    return v <= 0 ? B : std::min(B, v);
  }
  // measured on the 7-site, 3060-ADO hierarchy (MI355X): passes of 24
  // samples are the fastest in FP64 (9.6 MB each) and within 2% of the
  // fastest in MIXED (10.8 MB each); 48 are 20-25% slower in both
  const int64_t budget = 260ll << 20;
  int c = (int)std::max<int64_t>(1, budget / std::max<int64_t>(1, bytes_per_sample));
  if (c >= kXcds) c -= c % kXcds;
  return std::min(B, c);
}

// workspace bytes per hierarchy element: FP64 y, z2, z3, z4; MIXED y and the
// complex64 z1..z4, k1..k3
inline int64_t heom_bytes_per_element(bool fp64) { return fp64 ? 4 * 16 : 16 + 7 * 8; }

template <typename T>
int run_heom(const HeomArgs& a0, const double2* y0, const int32_t* nsub, const double* h,
             int nsave, double2* out, double2* yfin, void* workspace, int* status,
             hipStream_t stream, const HeomFields& fl = HeomFields()) {
  using T2 = cplx_t<T>;
  constexpr bool kFp64 = std::is_same<T, double>::value;
  const int64_t per = (int64_t)a0.n_ado * a0.M;               // elements per sample
  const int Bc = heom_samples_per_pass(a0.B, per * heom_bytes_per_element(kFp64));
  const int64_t ctotal = (int64_t)Bc * per;
  HeomStageBufs<T> bf;
  bf.y = static_cast<double2*>(workspace);
  T2* zbase = reinterpret_cast<T2*>(bf.y + ctotal);
  bf.z[0] = kFp64 ? nullptr : zbase;
  T2* zn = kFp64 ? zbase : zbase + ctotal;
  for (int i = 1; i < 4; ++i) bf.z[i] = zn + (i - 1) * ctotal;
  for (int i = 0; i < 3; ++i) bf.k[i] = kFp64 ? nullptr : zn + (3 + i) * ctotal;
  // the stage inputs: y itself (FP64) or its complex64 copy, then z2..z4
  const T2* zins[4] = {kFp64 ? reinterpret_cast<const T2*>(bf.y) : bf.z[0], bf.z[1], bf.z[2],
                       bf.z[3]};
  const int opsz = heom_opsz(a0);
  const size_t shmem = heom_lds_bytes(opsz, a0.ado_per_tile, a0.M, a0.nbins, sizeof(T2), a0.N);
  QSX_CHECK_ARG(shmem <= 160 * 1024, "qsx_heom_propagate: %zu bytes of LDS", shmem);
  const StageKernel<T> kern = pick_stage_kernel<T>(a0);
  if (shmem > 64 * 1024)
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  static const int kStageSample[4] = {0, 1, 1, 2};   // t, t + h/2, t + h/2, t + h
  for (int b0 = 0; b0 < a0.B; b0 += Bc) {
    HeomArgs a = a0;
    a.B = std::min(Bc, a0.B - b0);
    a.xcd_map = a.B >= kXcds;
    a.Hop = a0.Hop + (int64_t)b0 * opsz;
    const double2* Hbase = a.Hop;
    if (fl.P > 0) a.Hop = fl.Hwork;
    const unsigned fblocks = (unsigned)(((int64_t)a.B * opsz + 255) / 256);
    int64_t g = 0;                                      // global step (row of F)
    const int64_t total = (int64_t)a.B * per;
    const int ngroups = a.xcd_map ? kXcds * ((a.B + kXcds - 1) / kXcds) * a.tiles
                                  : a.B * a.tiles;
    hipLaunchKernelGGL(heom_init_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, y0 + (int64_t)b0 * per, bf.y, bf.z[0], total);
    const int sblocks = (a.B * a.M + 255) / 256;
    double2* outc = out + (int64_t)b0 * nsave * a.M;
    for (int s = 0; s < nsave; ++s) {
      for (int step = 0; step < nsub[s]; ++step) {
        for (int st = 0; st < 4; ++st) {
          if (fl.P > 0)
            hipLaunchKernelGGL(heom_field_op_kernel, dim3(fblocks), dim3(256), 0, stream, Hbase,
                               fl.C + (int64_t)b0 * fl.strideC, fl.strideC,
                               fl.F + (int64_t)b0 * fl.strideF, fl.strideF, g,
                               kStageSample[st], fl.P, a.B, opsz, fl.Hwork);
          hipLaunchKernelGGL(kern, dim3(ngroups), dim3(kThreads), shmem, stream, a, zins[st],
                             (T2*)nullptr, bf, h[s], st);
        }
        ++g;
      }
      hipLaunchKernelGGL(heom_save_kernel, dim3(sblocks), dim3(256), 0, stream, bf.y, a.n_ado,
                         a.M, a.B, s, nsave, outc, status);
    }
    QSX_CHECK_HIP(hipGetLastError());
    if (yfin != nullptr)
      QSX_CHECK_HIP(hipMemcpyAsync(yfin + (int64_t)b0 * per, bf.y, total * sizeof(double2),
                                   hipMemcpyDeviceToDevice, stream));
  }
  return QSX_OK;
}

}  // namespace
}  // namespace qsx

static int fill_args(qsx::HeomArgs& a, int B, int n_ado, int M, int n, const void* Hop,
                     const double* shift, const void* dvec, int nbins, int K1,
                     const int32_t* bin_src, const void* bin_coef, const void* ck, int nrel,
                     const int32_t* rel_site, const double* rel_occ, int N = 0,
                     const int32_t* est = nullptr, const int32_t* spos = nullptr,
                     const int32_t* sblk = nullptr) {
  using namespace qsx;
  QSX_CHECK_ARG(Hop && shift && dvec && bin_src && bin_coef && ck && rel_site && rel_occ,
                "qsx_heom: null operator pointer");
  QSX_CHECK_ARG(B >= 1 && n_ado >= 1 && M >= 1 && B <= (1 << 24),
                "qsx_heom: bad sizes B=%d n_ado=%d M=%d", B, n_ado, M);
  QSX_CHECK_ARG(K1 >= 1 && nbins >= K1 && nbins % K1 == 0 && nrel >= 1,
                "qsx_heom: bad bin tables nbins=%d K1=%d nrel=%d", nbins, K1, nrel);
  const bool hilbert = N > 0;
  const bool dense = n == 0 && !hilbert;
  if (hilbert)
    QSX_CHECK_ARG(est && spos && sblk && N <= 128, "qsx_heom: Hilbert form needs its tables "
                  "and N <= 128 (N=%d)", N);
  else
    QSX_CHECK_ARG(dense ? M <= 64 : (n >= 1 && M % n == 0),
                  "qsx_heom: commutator form needs M = nX * nY with nX = n, dense form M <= 64");
  a.B = B;
  a.n_ado = n_ado;
  a.M = M;
  a.nX = (dense || hilbert) ? 0 : n;
  a.nY = (dense || hilbert) ? 0 : M / n;
  a.dense = dense;
  a.N = hilbert ? N : 0;
  a.est = est;
  a.spos = spos;
  a.sblk = sblk;
  a.nbins = nbins;
  a.K1 = K1;
  a.nrel = nrel;
  a.ado_per_tile = std::max(1, kElemsPerThread * kThreads / M);
  a.tiles = (n_ado + a.ado_per_tile - 1) / a.ado_per_tile;
  a.xcd_map = B >= kXcds;
  a.Hop = (const double2*)Hop;
  a.shift = shift;
  a.dvec = (const double2*)dvec;
  a.bsrc = bin_src;
  a.bcoef = (const double2*)bin_coef;
  a.ck = (const double2*)ck;
  a.rsite = rel_site;
  a.rocc = rel_occ;
  const int64_t groups = a.xcd_map ? (int64_t)kXcds * ((B + kXcds - 1) / kXcds) * a.tiles
                                   : (int64_t)B * a.tiles;
  QSX_CHECK_ARG(groups < (1LL << 31), "qsx_heom: too many work-groups");
  return QSX_OK;
}

static int heom_apply_impl(const qsx::HeomArgs& a, const void* y, void* out, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(y && out, "qsx_heom_apply: null pointer");
  const int opsz = heom_opsz(a);
  const size_t shmem = heom_lds_bytes(opsz, a.ado_per_tile, a.M, a.nbins, sizeof(double2), a.N);
  QSX_CHECK_ARG(shmem <= 160 * 1024, "qsx_heom_apply: %zu bytes of LDS", shmem);
  const StageKernel<double> kern = pick_stage_kernel<double>(a);
  if (shmem > 64 * 1024)
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  const int ngroups = a.xcd_map ? kXcds * ((a.B + kXcds - 1) / kXcds) * a.tiles : a.B * a.tiles;
  HeomStageBufs<double> none{};
  hipLaunchKernelGGL(kern, dim3(ngroups), dim3(kThreads), shmem, as_stream(stream), a,
                     (const double2*)y, (double2*)out, none, 0.0, 4);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int32_t qsx_heom_apply(int32_t B, int32_t n_ado, int32_t M, int32_t n, const void* Hop,
                                  const double* shift, const void* dvec, int32_t nbins,
                                  int32_t K1, const int32_t* bin_src, const void* bin_coef,
                                  const void* ck, int32_t nrel, const int32_t* rel_site,
                                  const double* rel_occ, const void* y, void* out, void* stream) {
  using namespace qsx;
  HeomArgs a;
  int st = fill_args(a, B, n_ado, M, n, Hop, shift, dvec, nbins, K1, bin_src, bin_coef, ck, nrel,
                     rel_site, rel_occ);
  if (st) return st;
  return heom_apply_impl(a, y, out, stream);
}

extern "C" int32_t qsx_heom_apply_hilbert(int32_t B, int32_t n_ado, int32_t M, int32_t N,
                                          const int32_t* elem_states, const int32_t* state_pos,
                                          const int32_t* state_block, const void* Hop,
                                          const double* shift, const void* dvec, int32_t nbins,
                                          int32_t K1, const int32_t* bin_src,
                                          const void* bin_coef, const void* ck, int32_t nrel,
                                          const int32_t* rel_site, const double* rel_occ,
                                          const void* y, void* out, void* stream) {
  using namespace qsx;
  HeomArgs a;
  int st = fill_args(a, B, n_ado, M, 0, Hop, shift, dvec, nbins, K1, bin_src, bin_coef, ck, nrel,
                     rel_site, rel_occ, N, elem_states, state_pos, state_block);
  if (st) return st;
  return heom_apply_impl(a, y, out, stream);
}

extern "C" int64_t qsx_heom_workspace_bytes(int32_t precision, int32_t B, int32_t n_ado,
                                            int32_t M) {
  if (B < 1 || n_ado < 1 || M < 1 || (precision != QSX_FP64 && precision != QSX_MIXED)) return -1;
  const int64_t total = (int64_t)B * n_ado * M;
  return total * qsx::heom_bytes_per_element(precision == QSX_FP64);
}

static int heom_propagate_impl(int precision, const qsx::HeomArgs& a, const void* y0,
                               const int32_t* nsub_host, const double* h_host, int nsave,
                               void* out, void* y_final, void* workspace, int32_t* status,
                               void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(y0 && nsub_host && h_host && out && workspace && status,
                "qsx_heom_propagate: null pointer");
  QSX_CHECK_ARG(nsave >= 1, "qsx_heom_propagate: nsave=%d", nsave);
  for (int s = 0; s < nsave; ++s)
    QSX_CHECK_ARG(nsub_host[s] >= 0, "qsx_heom_propagate: negative step count");
  hipStream_t st = as_stream(stream);
  if (precision == QSX_FP64)
    return run_heom<double>(a, (const double2*)y0, nsub_host, h_host, nsave, (double2*)out,
                            (double2*)y_final, workspace, status, st);
  return run_heom<float>(a, (const double2*)y0, nsub_host, h_host, nsave, (double2*)out,
                         (double2*)y_final, workspace, status, st);
}

extern "C" int32_t qsx_heom_propagate(int32_t precision, int32_t B, int32_t n_ado, int32_t M,
                                      int32_t n, const void* Hop, const double* shift,
                                      const void* dvec, int32_t nbins, int32_t K1,
                                      const int32_t* bin_src, const void* bin_coef,
                                      const void* ck, int32_t nrel, const int32_t* rel_site,
                                      const double* rel_occ, const void* y0,
                                      const int32_t* nsub_host, const double* h_host,
                                      int32_t nsave, void* out, void* y_final, void* workspace,
                                      int32_t* status, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED, "qsx_heom_propagate: precision");
  HeomArgs a;
  int rc = fill_args(a, B, n_ado, M, n, Hop, shift, dvec, nbins, K1, bin_src, bin_coef, ck, nrel,
                     rel_site, rel_occ);
  if (rc) return rc;
  return heom_propagate_impl(precision, a, y0, nsub_host, h_host, nsave, out, y_final,
                             workspace, status, stream);
}

extern "C" int32_t qsx_heom_propagate_hilbert(
    int32_t precision, int32_t B, int32_t n_ado, int32_t M, int32_t N,
    const int32_t* elem_states, const int32_t* state_pos, const int32_t* state_block,
    const void* Hop, const double* shift, const void* dvec, int32_t nbins, int32_t K1,
    const int32_t* bin_src, const void* bin_coef, const void* ck, int32_t nrel,
    const int32_t* rel_site, const double* rel_occ, const void* y0, const int32_t* nsub_host,
    const double* h_host, int32_t nsave, void* out, void* y_final, void* workspace,
    int32_t* status, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED, "qsx_heom_propagate: precision");
  HeomArgs a;
  int rc = fill_args(a, B, n_ado, M, 0, Hop, shift, dvec, nbins, K1, bin_src, bin_coef, ck, nrel,
                     rel_site, rel_occ, N, elem_states, state_pos, state_block);
  if (rc) return rc;
  return heom_propagate_impl(precision, a, y0, nsub_host, h_host, nsave, out, y_final,
                             workspace, status, stream);
}

extern "C" int32_t qsx_heom_propagate_driven(
    int32_t precision, int32_t B, int32_t n_ado, int32_t M, const void* Hop, const double* shift,
    const void* dvec, int32_t nbins, int32_t K1, const int32_t* bin_src, const void* bin_coef,
    const void* ck, int32_t nrel, const int32_t* rel_site, const double* rel_occ, int32_t P,
    const void* C, int64_t strideC, const void* F, int64_t strideF, int64_t nF, const void* y0,
    const int32_t* nsub_host, const double* h_host, int32_t nsave, void* out, void* y_final,
    void* Hwork, void* workspace, int32_t* status, void* stream) {
  using namespace qsx;
  QSX_CHECK_ARG(precision == QSX_FP64 || precision == QSX_MIXED,
                "qsx_heom_propagate_driven: precision");
  QSX_CHECK_ARG(y0 && nsub_host && h_host && out && workspace && status && C && F && Hwork,
                "qsx_heom_propagate_driven: null pointer");
  QSX_CHECK_ARG(nsave >= 1 && P >= 1 && P <= 4 && nF >= 1,
                "qsx_heom_propagate_driven: nsave=%d P=%d nF=%lld", nsave, P, (long long)nF);
  HeomArgs a;
  int rc = fill_args(a, B, n_ado, M, 0, Hop, shift, dvec, nbins, K1, bin_src, bin_coef, ck, nrel,
                     rel_site, rel_occ);
  if (rc) return rc;
  int64_t steps = 0;
  for (int s = 0; s < nsave; ++s) {
    QSX_CHECK_ARG(nsub_host[s] >= 0, "qsx_heom_propagate_driven: negative step count");
    steps += nsub_host[s];
  }
  QSX_CHECK_ARG(steps <= nF, "qsx_heom_propagate_driven: %lld steps but field samples for %lld",
                (long long)steps, (long long)nF);
  QSX_CHECK_ARG(strideC == 0 || strideC == (int64_t)P * M * M,
                "qsx_heom_propagate_driven: strideC must be 0 or P*M*M");
  QSX_CHECK_ARG(strideF == 0 || strideF == nF * 3 * P,
                "qsx_heom_propagate_driven: strideF must be 0 or nF*3*P");
  HeomFields fl;
  fl.P = P;
  fl.C = static_cast<const double2*>(C);
  fl.strideC = strideC;
  fl.F = static_cast<const double2*>(F);
  fl.strideF = strideF;
  fl.Hwork = static_cast<double2*>(Hwork);
  hipStream_t st = as_stream(stream);
  if (precision == QSX_FP64)
    return run_heom<double>(a, (const double2*)y0, nsub_host, h_host, nsave, (double2*)out,
                            (double2*)y_final, workspace, status, st, fl);
  return run_heom<float>(a, (const double2*)y0, nsub_host, h_host, nsave, (double2*)out,
                         (double2*)y_final, workspace, status, st, fl);
}
