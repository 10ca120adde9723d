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
// One RK4 stage is one launch over (sample, ADO tile); the four stages of
// classical RK4 use buffers y (FP64 state), ynew (FP64 accumulated update)
// and a double-buffered stage vector zA/zB (FP64, or FP32 in MIXED mode).
// Work-groups are mapped so that every tile of one sample lands on the same
// XCD (work-groups are dispatched round-robin over the 8 XCDs): a sample's
// hierarchy and its neighbour re-reads stay in one L2.

#include "qsx_common.h"

#include <algorithm>
#include <cstdlib>
#include <vector>

namespace qsx {
namespace {

constexpr int kThreads = 256;
constexpr int kXcds = 8;
#ifndef QSX_HEOM_ELEMS_PER_THREAD
#define QSX_HEOM_ELEMS_PER_THREAD 1
#endif
// Liouville elements per thread and stage: a work-group covers
// kElemsPerThread * 256 / M ADOs (with their edge tables in LDS)
constexpr int kElemsPerThread = QSX_HEOM_ELEMS_PER_THREAD;

struct HeomArgs {
  int B, n_ado, M, nX, nY, dense, nbins, K1, nrel, tiles, ado_per_tile, xcd_map;
  const double2* Hop;      // COMM: [B][nX*nX + nY*nY] = -i unit H_X, H_Y (or their
                           // transposes); DENSE: [B][M][M]
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
  return a.dense ? a.M * a.M : (a.nX == a.nY ? a.nX * a.nX : a.nX * a.nX + a.nY * a.nY);
}

// Stage st of classical RK4: z_in = input stage vector, writes z_out and
// updates ynew (st 0..2) or y (st 3); st = 4 only writes z_out = F(z_in).
// LDS layout of a stage work-group: operator block, the tile's stage
// vectors, and the tile's edge tables (so that the neighbour gathers do not
// wait on dependent global loads of the table).
inline size_t heom_lds_bytes(int opsz, int ado_per_tile, int M, int nbins, size_t zbytes) {
  size_t off = (size_t)opsz * sizeof(double2);
  off += (size_t)ado_per_tile * M * zbytes;
  off = (off + 15) & ~(size_t)15;
  off += 2 * (size_t)ado_per_tile * nbins * sizeof(double2);
  off += 2 * (size_t)ado_per_tile * nbins * sizeof(int);
  return off;
}

// Stage st of classical RK4: z_in = input stage vector, writes z_out and
// updates ynew (st 0..2) or y (st 3); st = 4 only writes z_out = F(z_in).
template <typename T>
__global__ __launch_bounds__(kThreads) void heom_stage_kernel(
    HeomArgs a, const cplx_t<T>* __restrict__ zin, cplx_t<T>* __restrict__ zout,
    double2* __restrict__ y, double2* __restrict__ ynew, double h, int st) {
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
  __syncthreads();

  for (int e = tid; e < nloc * M; e += kThreads) {
    const int la = e / M, p = e % M, ado = ado0 + la;
    const T2* z = zt + la * M;
    double2 k = make_double2(0.0, 0.0);
    if (a.dense) {
      const double2* row = Hs + p * M;
      for (int q = 0; q < M; ++q) k = cadd(k, cmul(row[q], ld<T>(z + q)));
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
    const int64_t gi = base + e;
    if (st == 4) {                           // plain application: zout = F(zin)
      zout[gi] = from_d2<T>(k);
      continue;
    }
    if (st == 0) {
      const double2 yv = y[gi];
      ynew[gi] = caxpy(h / 6.0, k, yv);
      zout[gi] = from_d2<T>(caxpy(0.5 * h, k, yv));
    } else if (st < 3) {
      const double2 yv = y[gi];
      ynew[gi] = caxpy(h / 3.0, k, ynew[gi]);
      zout[gi] = from_d2<T>(caxpy(st == 1 ? 0.5 * h : h, k, yv));
    } else {
      const double2 yn = caxpy(h / 6.0, k, ynew[gi]);
      y[gi] = yn;
      zout[gi] = from_d2<T>(yn);
    }
  }
}

template <typename T>
__global__ void heom_init_kernel(const double2* __restrict__ y0, double2* __restrict__ y,
                                 cplx_t<T>* __restrict__ z, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) {
    y[i] = y0[i];
    z[i] = from_d2<T>(y0[i]);
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
// vectors) is kept small enough to stay resident in the 256 MB
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
  const int64_t budget = 160ll << 20;
  int c = (int)std::max<int64_t>(1, budget / std::max<int64_t>(1, bytes_per_sample));
  if (c >= kXcds) c -= c % kXcds;
  return std::min(B, c);
}

template <typename T>
int run_heom(const HeomArgs& a0, const double2* y0, const int32_t* nsub, const double* h,
             int nsave, double2* out, double2* yfin, void* workspace, int* status,
             hipStream_t stream, const HeomFields& fl = HeomFields()) {
  using T2 = cplx_t<T>;
  const int64_t per = (int64_t)a0.n_ado * a0.M;               // elements per sample
  const int Bc = heom_samples_per_pass(a0.B, per * (2 * sizeof(double2) + 2 * sizeof(T2)));
  const int64_t ctotal = (int64_t)Bc * per;
  double2* y = static_cast<double2*>(workspace);
  double2* ynew = y + ctotal;
  T2* zA = reinterpret_cast<T2*>(ynew + ctotal);
  T2* zB = zA + ctotal;
  const int opsz = heom_opsz(a0);
  const size_t shmem = heom_lds_bytes(opsz, a0.ado_per_tile, a0.M, a0.nbins, sizeof(T2));
  QSX_CHECK_ARG(shmem <= 160 * 1024, "qsx_heom_propagate: %zu bytes of LDS", shmem);
  if (shmem > 64 * 1024)
    QSX_CHECK_HIP(hipFuncSetAttribute((const void*)heom_stage_kernel<T>,
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
                       stream, y0 + (int64_t)b0 * per, y, zA, total);
    const int sblocks = (a.B * a.M + 255) / 256;
    double2* outc = out + (int64_t)b0 * nsave * a.M;
    for (int s = 0; s < nsave; ++s) {
      for (int step = 0; step < nsub[s]; ++step) {
        const T2* zi = zA;
        T2* zo = zB;
        for (int st = 0; st < 4; ++st) {
          if (fl.P > 0)
            hipLaunchKernelGGL(heom_field_op_kernel, dim3(fblocks), dim3(256), 0, stream, Hbase,
                               fl.C + (int64_t)b0 * fl.strideC, fl.strideC,
                               fl.F + (int64_t)b0 * fl.strideF, fl.strideF, g,
                               kStageSample[st], fl.P, a.B, opsz, fl.Hwork);
          hipLaunchKernelGGL(heom_stage_kernel<T>, dim3(ngroups), dim3(kThreads), shmem, stream,
                             a, zi, zo, y, ynew, h[s], st);
          const T2* t = zi;
          zi = zo;
          zo = const_cast<T2*>(t);
        }
        ++g;
      }
      hipLaunchKernelGGL(heom_save_kernel, dim3(sblocks), dim3(256), 0, stream, y, a.n_ado, a.M,
                         a.B, s, nsave, outc, status);
    }
    QSX_CHECK_HIP(hipGetLastError());
    if (yfin != nullptr)
      QSX_CHECK_HIP(hipMemcpyAsync(yfin + (int64_t)b0 * per, y, total * sizeof(double2),
                                   hipMemcpyDeviceToDevice, stream));
  }
  return QSX_OK;
}

}  // namespace
}  // namespace qsx

static int fill_args(qsx::HeomArgs& a, int B, int n_ado, int M, int n, const void* Hop,
                     const double* shift, const void* dvec, int nbins, int K1,
                     const int32_t* bin_src, const void* bin_coef, const void* ck, int nrel,
                     const int32_t* rel_site, const double* rel_occ) {
  using namespace qsx;
  QSX_CHECK_ARG(Hop && shift && dvec && bin_src && bin_coef && ck && rel_site && rel_occ,
                "qsx_heom: null operator pointer");
  QSX_CHECK_ARG(B >= 1 && n_ado >= 1 && M >= 1 && B <= (1 << 24),
                "qsx_heom: bad sizes B=%d n_ado=%d M=%d", B, n_ado, M);
  QSX_CHECK_ARG(K1 >= 1 && nbins >= K1 && nbins % K1 == 0 && nrel >= 1,
                "qsx_heom: bad bin tables nbins=%d K1=%d nrel=%d", nbins, K1, nrel);
  const bool dense = n == 0;
  QSX_CHECK_ARG(dense ? M <= 64 : (n >= 1 && M % n == 0),
                "qsx_heom: commutator form needs M = nX * nY with nX = n, dense form M <= 64");
  QSX_CHECK_ARG(M <= kThreads, "qsx_heom: M=%d exceeds one work-group", M);
  a.B = B;
  a.n_ado = n_ado;
  a.M = M;
  a.nX = dense ? 0 : n;
  a.nY = dense ? 0 : M / n;
  a.dense = dense;
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
  QSX_CHECK_ARG(y && out, "qsx_heom_apply: null pointer");
  const int opsz = heom_opsz(a);
  const size_t shmem = heom_lds_bytes(opsz, a.ado_per_tile, a.M, a.nbins, sizeof(double2));
  QSX_CHECK_ARG(shmem <= 64 * 1024, "qsx_heom_apply: %zu bytes of LDS", shmem);
  const int ngroups = a.xcd_map ? kXcds * ((a.B + kXcds - 1) / kXcds) * a.tiles : a.B * a.tiles;
  hipLaunchKernelGGL(heom_stage_kernel<double>, dim3(ngroups), dim3(kThreads), shmem,
                     as_stream(stream), a, (const double2*)y, (double2*)out, nullptr, nullptr,
                     0.0, 4);
  QSX_CHECK_HIP(hipGetLastError());
  return QSX_OK;
}

extern "C" int64_t qsx_heom_workspace_bytes(int32_t precision, int32_t B, int32_t n_ado,
                                            int32_t M) {
  if (B < 1 || n_ado < 1 || M < 1 || (precision != QSX_FP64 && precision != QSX_MIXED)) return -1;
  const int64_t total = (int64_t)B * n_ado * M;
  return total * (2 * 16 + 2 * (precision == QSX_FP64 ? 16 : 8));
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
  QSX_CHECK_ARG(y0 && nsub_host && h_host && out && workspace && status,
                "qsx_heom_propagate: null pointer");
  QSX_CHECK_ARG(nsave >= 1, "qsx_heom_propagate: nsave=%d", nsave);
  HeomArgs a;
  int rc = fill_args(a, B, n_ado, M, n, Hop, shift, dvec, nbins, K1, bin_src, bin_coef, ck, nrel,
                     rel_site, rel_occ);
  if (rc) return rc;
  for (int s = 0; s < nsave; ++s)
    QSX_CHECK_ARG(nsub_host[s] >= 0, "qsx_heom_propagate: negative step count");
  hipStream_t st = as_stream(stream);
  if (precision == QSX_FP64)
    return run_heom<double>(a, (const double2*)y0, nsub_host, h_host, nsave, (double2*)out,
                            (double2*)y_final, workspace, status, st);
  return run_heom<float>(a, (const double2*)y0, nsub_host, h_host, nsave, (double2*)out,
                         (double2*)y_final, workspace, status, st);
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
