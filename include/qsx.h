/*
 * qsx.h - C ABI of libqsx.so, the MI355X (gfx950) density-matrix propagation
 * library behind qspectra_amd.
 *
 * The reference (qspectra) has no native boundary: its hot path is the Python
 * operator API
 *   DynamicalModel.equation_of_motion(subspace, heisenberg_picture) -> eom(t, y)
 *       qspectra/dynamics/base.py:72-84, liouville_space.py:316-341,
 *       heom.py:228-244
 *   integrate(f, y0, t, t0, method_name, f_params, save_func, **kwargs)
 *       qspectra/simulate/utils.py:16-109
 *   RedfieldModel.evolution_super_operator (redfield.py:9-152),
 *   Hamiltonian.eig (hamiltonian.py:310-328).
 * The entry points below replace the arithmetic under that API; the Python
 * package qspectra_amd binds them with ctypes (qspectra_amd/_native.py).
 *
 * Conventions
 *   - every function returns int32 status: QSX_OK or a negative QSX_E* code;
 *     qsx_last_error() returns the message of the calling thread's last error;
 *   - pointers are device pointers unless stated; complex numbers are
 *     interleaved (re, im) pairs, double2 for complex128 and float2 for
 *     complex64, the layout of numpy/torch complex arrays;
 *   - all matrices are row-major; batch strides are in elements;
 *   - work is enqueued on `stream` (a hipStream_t, NULL = default stream);
 *     nothing synchronises the host and nothing allocates device memory
 *     unless the comment says so;
 *   - sizes are validated on the host before any launch.
 */
#ifndef QSX_H
#define QSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSX_OK 0
#define QSX_EINVAL (-1)      /* bad argument (size, null pointer, unsupported mode) */
#define QSX_EHIP (-2)        /* a HIP runtime call failed */

#define QSX_FP64 0           /* complex128 Liouvillian products, FP64 state */
#define QSX_MIXED 1          /* complex64 products (FP32 accumulate), FP64 state */

/* Library version as 10000*major + 100*minor + patch. */
int32_t qsx_version(void);

/* Copies the calling thread's last error message (NUL-terminated) into buf. */
int32_t qsx_last_error(char* buf, size_t len);

/*
 * out[b][r][:] = L_b . y[b][r][:]  (complex128).
 * Replaces the eom(t, y) closure of LiouvilleSpaceModel.equation_of_motion
 * (qspectra/dynamics/liouville_space.py:339-340: evolve_matrix.dot(rho)).
 *   L    [B][M][M] with batch stride strideL (0: one L shared by the batch)
 *   y    [B][R][M] contiguous, out [B][R][M] contiguous.
 */
int32_t qsx_apply_dense_c128(const void* L, int64_t strideL, int32_t M,
                             const void* y, int32_t R, int32_t B, void* out,
                             void* stream);

/*
 * Fused fixed-step classical RK4 propagation of R right-hand sides per
 * problem, for B problems, with a linear projection at every save point.
 * Replaces integrate(eom, y0, t, t0, save_func=V.commutator / bra, ...)
 * (qspectra/simulate/utils.py:16-109; the per-row loop of :103-109 becomes
 * the R axis, the ensemble loop of simulate/decorators.py:51-63 the B axis).
 *
 *   precision  QSX_FP64: L, proj complex128.
 *              QSX_MIXED: L, proj complex64; the diagonal of L is taken in
 *              FP64 from `diag` (complex128 [B][M], batch stride strideD) and
 *              the state and RK update stay complex128.
 *   L        [B][M][M], batch stride strideL (0 = shared)
 *   y0       [B][R][M] complex128, contiguous
 *   nsub,h   [nsave] int32 / double (device): save interval s is covered by
 *            nsub[s] RK4 steps of size h[s]; nsub[s] = 0 saves the current
 *            state again (the reference's y[0] = save(y0) when t[0] == t0).
 *   proj     [B][K][M] (stride strideP, 0 = shared) or NULL for the identity
 *            (then K must equal M)
 *   out      [B][R][nsave][K] complex128 (rows first, as utils.py:103-109)
 *   y_final  optional [B][R][M] complex128 final state (may be NULL)
 *   status   device int32[1]; set to 1 if a saved state is not finite
 *            (the host raises IntegratorError, simulate/utils.py:46-49)
 */
int32_t qsx_rk4_dense(int32_t precision, const void* L, int64_t strideL,
                      const void* diag, int64_t strideD, int32_t M,
                      const void* y0, int32_t R, int32_t B,
                      const int32_t* nsub, const double* h, int32_t nsave,
                      const void* proj, int64_t strideP, int32_t K,
                      void* out, void* y_final, int32_t* status, void* stream);

/* ---------------------------------------------------------------------------
 * Batched FP64 construction of Liouvillian blocks.
 *
 * Hilbert blocks are numbered 0 = g (1 state), 1 = e (n sites), 2 = f
 * (n(n-1)/2 site pairs i<j).  Inputs per problem b (B problems):
 *   H1  [B][n][n] double, the non-rotating one-exciton Hamiltonian (real
 *       symmetric; a disorder sample per problem),
 *   rw  [B] double rotating-frame frequency (NULL = 0).
 * The Liouville block XY acts on rho_ab, a in X, b in Y; M = nX*nY and the
 * local index is a + nX*b (column stacking).  Hilbert blocks of up to 64
 * states are supported.
 */

/* Eigensystem of Hilbert block X of H1 - X*rw (Hamiltonian.eig,
 * qspectra/hamiltonian.py:310-328, per excitation block): E [B][nb]
 * ascending, U [B][nb][nb] with eigenvectors in columns, each with its
 * largest-magnitude component positive.  Batched two-sided Jacobi in LDS.
 * workspace: qsx_redfield_workspace_bytes(n, X, X, B) bytes. */
int32_t qsx_block_eigh(const double* H1, int32_t n, int32_t B, int32_t X,
                       const double* rw, double* E, double* U, void* workspace,
                       void* stream);

/* The operators of linear absorption for a ragged batch of aggregates (up to
 * qsx_absorption_operators_max_sites() sites each), one launch: per problem b
 * the Heisenberg 'eg' Redfield block of the one-exciton Hamiltonian H1_b (Debye
 * bath T, lam, gam; the block qsx_redfield_block builds with X = e, Y = g and
 * QSX_HEISENBERG) in the rotating frame rw[b], times unit_convert, and its RK4
 * propagator P1_b = P4(h L_b) with PJ_b = P1_b^J (J = 1, 2 or 4).
 * table int64 [B][3] (device): n_b, offset of H1_b (doubles, n_b x n_b
 * row-major) in H1, offset of P1_b and PJ_b (complex128, n_b x n_b) in P1 and
 * PJ; rw [B]; bath [B][3]; flags: QSX_SECULAR, QSX_REAL_CORR, QSX_HEISENBERG.
 * max_n bounds every n_b (problems above it are skipped) and sizes the
 * launch's LDS: grouping problems by size keeps small ones at high occupancy.
 * Replaces, for absorption_spectra over an ensemble of unequal site counts,
 * RedfieldModel.evolution_super_operator['eg'] per member (qspectra/dynamics/
 * redfield.py:9-152, liouville_space.py:316-341) and the RK4 step of
 * simulate/utils.py:16-109. */
int32_t qsx_absorption_operators(const int64_t* table, int32_t B, int32_t max_n,
                                 const double* H1, const double* rw, const double* bath,
                                 int32_t flags, double unit_convert, double h, int32_t J,
                                 void* P1, void* PJ, void* stream);
int32_t qsx_absorption_operators_max_sites(void);

/* Bytes of device workspace needed by qsx_redfield_block / qsx_block_eigh
 * (or -1 for invalid arguments). */
int64_t qsx_redfield_workspace_bytes(int32_t n, int32_t X, int32_t Y, int32_t B);

#define QSX_SECULAR 1        /* secular (Bloch) approximation, redfield.py:77-83 */
#define QSX_REAL_CORR 2      /* discard_imag_corr: real C(w), bath.py:17-31 */
#define QSX_EIGEN_BASIS 4    /* return the block in the energy eigenbasis */
#define QSX_HEISENBERG 8     /* return the transpose (Heisenberg picture) */

/* Redfield Liouvillian block XY (complex128 [B][M][M]) for a Debye bath
 * (temperature T, reorganisation energy lam, cutoff gam, one identical
 * uncorrelated bath per site), times unit_convert.  Replaces
 * RedfieldModel.evolution_super_operator[idx][:, idx]
 * (qspectra/dynamics/redfield.py:9-152, liouville_space.py:316-341). */
int32_t qsx_redfield_block(const double* H1, int32_t n, int32_t B, const double* rw,
                           double T, double lam, double gam, int32_t flags,
                           double unit_convert, int32_t X, int32_t Y, void* L,
                           void* workspace, void* stream);

/* The same block for any bath, its correlation function tabulated by the
 * caller: ctabX [B][nX][nX], ctabY [B][nY][nY] complex128 with
 * ctab[b][d][c] = C(E_d - E_c) over the eigen-energies of Hilbert block X
 * (Y) of problem b (as qsx_block_eigh returns them).  Replaces the
 * bath.corr_func_real / corr_func_complex calls of redfield_tensor
 * (qspectra/dynamics/redfield.py:48-53) for ArbitraryBath and other
 * Python-defined baths (qspectra/bath.py:8-57). */
int32_t qsx_redfield_block_tabulated(const double* H1, int32_t n, int32_t B, const double* rw,
                                     const void* ctabX, const void* ctabY, int32_t flags,
                                     double unit_convert, int32_t X, int32_t Y, void* L,
                                     void* workspace, void* stream);

/* Reverse mode of qsx_redfield_block (Debye bath, same flags and blocks):
 * given a cotangent gL [B][M][M] of the block it returned, writes
 * gH1 [B][n][n], the cotangent of each one-exciton Hamiltonian (symmetric),
 * and glam [B] (may be null), the derivative with respect to the
 * reorganization energy, both in torch's convention dl = Re sum conj(g) dx.
 * The reference computes the tensor only (qspectra/dynamics/redfield.py:9-104,
 * qspectra/bath.py:84-102); this is the derivative of that same map, for
 * gradients of spectra with respect to site energies, couplings and lam.
 * Workspace: qsx_redfield_vjp_workspace_bytes. */
int64_t qsx_redfield_vjp_workspace_bytes(int32_t n, int32_t X, int32_t Y, int32_t B);
int32_t qsx_redfield_block_vjp(const double* H1, int32_t n, int32_t B, const double* rw,
                               double T, double lam, double gam, int32_t flags,
                               double unit_convert, int32_t X, int32_t Y, const void* gL,
                               double* gH1, double* glam, void* workspace, void* stream);

/* Unitary Liouvillian block -i*unit_convert*[H, .] in the rotating frame
 * (UnitaryModel.evolution_super_operator, qspectra/dynamics/unitary.py:5-9);
 * heisenberg != 0 returns the transpose. */
int32_t qsx_unitary_block(const double* H1, int32_t n, int32_t B, const double* rw,
                          double unit_convert, int32_t X, int32_t Y, int32_t heisenberg,
                          void* L, void* stream);

/* ---------------------------------------------------------------------------
 * Register-resident fixed-step propagation (one workgroup per right-hand
 * side; blocks up to M = 160).  Replaces the zvode loop of
 * qspectra/simulate/utils.py:16-109 like qsx_rk4_dense, but saves full
 * states (projections are applied afterwards with qsx_zgemm).
 *
 *   mode   QSX_MODE_RK4: A = L, four products per step of size h[s];
 *          QSX_MODE_PROPAGATOR: A = P = P4(h L) (FP64) or P - I (MIXED),
 *          one product per step; every h[s] must equal the h of P.
 *   precision / diag as in qsx_rk4_dense (MIXED: A off-diagonal complex64,
 *          diag the complex128 diagonal of L or of P - I).
 *   out    [B][R][nsave][M] complex128.
 */
#define QSX_MODE_RK4 0
#define QSX_MODE_PROPAGATOR 1
int32_t qsx_propagate(int32_t mode, int32_t precision, const void* A, int64_t strideA,
                      const void* diag, int64_t strideD, int32_t M, const void* y0,
                      int32_t R, int32_t B, const int32_t* nsub, const double* h,
                      int32_t nsave, void* out, void* y_final, int32_t* status, void* stream);

/* Time-chunked trajectory of one right-hand side per problem in one launch
 * (QSX_MODE_PROPAGATOR): work-group j of problem b computes the chunk start
 * (Pc)^j y0[b] with the FP64 power Pc = P^(q c) [B|1][M][M], then steps the
 * chunk with A (P in FP64, P - I complex64 with the FP64 diagonal diag in
 * MIXED) over the save intervals nsub/h (device, nsave entries): out
 * [B][nc][nsave][M] complex128, the same values as qsx_propagate with Pc on
 * the chain of chunk starts followed by qsx_propagate with A on the nc
 * starts as right-hand sides.  Like qsx_propagate it replaces the zvode
 * trajectories of qspectra/simulate/utils.py:16-109 (the t1 and t3
 * propagations of qspectra/simulate/response.py:269-337), for trajectories
 * saved at equally spaced points. */
int32_t qsx_propagate_chunks(int32_t precision, const void* Pc, int64_t strideP, const void* A,
                             int64_t strideA, const void* diag, int64_t strideD, int32_t M,
                             const void* y0, int32_t nc, int32_t B, const int32_t* nsub,
                             const double* h, int32_t nsave, void* out, int32_t* status,
                             void* stream);

/* Reverse mode of qsx_propagate in QSX_MODE_PROPAGATOR (FP64): given the
 * one-step propagator A = P [B|1][M][M], y0 [B][R][M], the step counts nsub
 * (device) and the cotangents gout [B][R][nsave][M] of the saved states
 * (dLoss/d conj(y)), writes gA [B][R][M][M] = sum over steps of
 * lambda y^H (one gradient per right-hand side; the caller sums over R,
 * and over B when A is shared) and gy0 [B][R][M] (optional).  The forward
 * states are recomputed from checkpoints every ckpt steps held in the
 * workspace (qsx_adjoint_workspace_bytes; total_steps = sum of nsub).
 * status: 1 non-finite gradient, 2 workspace smaller than the checkpoints.
 * Blocks up to M = 160 (above 64 the workspace also holds the operators'
 * conjugate transposes); ckpt * M * 16 bytes must fit the LDS.  The gradient
 * path of the reference's spectra (absorption_spectra,
 * qspectra/simulate/response.py:103-154), which has no reverse mode. */
int64_t qsx_adjoint_workspace_bytes(int32_t M, int32_t R, int32_t B, int64_t total_steps,
                                    int32_t ckpt);
int32_t qsx_propagate_adjoint(const void* A, int64_t strideA, int32_t M, const void* y0,
                              int32_t R, int32_t B, const int32_t* nsub, int32_t nsave,
                              const void* gout, int32_t ckpt, void* workspace,
                              int64_t workspace_bytes, void* gA, void* gy0, int32_t* status,
                              void* stream);

/* Linear-response signals of many problems of unequal size in one launch:
 * for problem b, y_{k+1} = P_b y_k from y0_b and s_k = v_b . y_k,
 * k = 0 .. nt-1, written to row r_b of out [B][nt] complex128 (FP64
 * products), J steps (1, 4 or 8) per pass: P holds P_b^J, W the projection
 * rows w_j = (P_b^j)^T v_b, j < J (qsx_response_bucket_weights writes them
 * from P_b and v_b; for J = 1, W is v).  waves int32 [nwaves][4] (device):
 * first problem, problem count, M, G per wavefront, with G a power of two,
 * G >= M + J, G * count <= 64 and M at most qsx_response_buckets_max_states();
 * for J > 1, G = 16 with M + J <= 16 keeps each problem's state in registers
 * and G twice the power of two at or above M + J gives two lanes per row;
 * offsets int64 [B][3] (device): offset of P_b^J (M^2 complex elements,
 * row-major) in P, of y0_b (M) in y0 (the rows of W at J times it, [J][M]),
 * and the output row r_b (a permutation of 0 .. B-1, so the signals can leave
 * in the caller's order).  status int32 [B]: 1 at r_b where a signal value
 * was not finite.  Replaces the per-member zvode runs of linear_response /
 * absorption_spectra (qspectra/simulate/response.py:15-43, :103-154;
 * simulate/utils.py:16-109) over an ensemble: P_b = P4(h L_b)^q of the
 * member's Heisenberg 'eg' block, y0_b = -bra, v_b = V rho0. */
int32_t qsx_response_buckets(const int32_t* waves, int32_t nwaves, const int64_t* offsets,
                             const void* P, const void* y0, const void* W, int32_t J,
                             int32_t nt, void* out, int32_t* status, void* stream);
/* The projection rows of qsx_response_buckets: W[J o_b + j M + c] =
 * ((P_b^j)^T v_b)[c] for j < J (1 <= J <= 4), from the one-step propagators P
 * (same offsets) and v [o_b + c] (o_b the problem's y0 offset). */
int32_t qsx_response_bucket_weights(const int32_t* waves, int32_t nwaves,
                                    const int64_t* offsets, const void* P, const void* v,
                                    int32_t J, void* W, void* stream);
int32_t qsx_response_buckets_max_states(void);

/* Batched complex GEMM, row-major storage:
 *   C[b] = alpha * op(A[b]) (m x k) * op(B[b]) (k x n) + beta * C[b] + gamma * I
 * with op(A) = A^T when trans & 1 (A stored k x m) and op(B) = B^T when
 * trans & 2 (B stored n x k); gamma only for m == n.  precision QSX_FP64:
 * complex128 operands; QSX_MIXED: complex64 A and B, FP32 products,
 * complex128 C.  Leading dimensions lda/ldb/ldc (elements, of the stored
 * operands), batch strides sA/sB/sC (0 = shared; sC must be nonzero when
 * batch > 1).  alpha, beta, gamma are complex (re, im) pairs passed by value.
 * Used for the RK4 propagator (Horner: P = I + hL(I + hL/2 (I + hL/3 (I +
 * hL/4)))), its powers, the dipole maps and response contractions of
 * response.py:309-336, and the Fourier transforms of two_dimensional_spectra
 * (simulate/utils.py:154-219) as products with the transform matrices. */
int32_t qsx_zgemm(int32_t precision, int32_t trans, int32_t m, int32_t n, int32_t k,
                  double alpha_re, double alpha_im, const void* A, int64_t lda, int64_t sA,
                  const void* B, int64_t ldb, int64_t sB, double beta_re, double beta_im,
                  void* C, int64_t ldc, int64_t sC, double gamma_re, double gamma_im,
                  int32_t batch, void* stream);

/* ---------------------------------------------------------------------------
 * Propagation operators of a time-invariant block, built on the device and
 * enqueued on `stream` (a captured graph rebuilds them on every replay).
 * The reference has no such operators: its zvode integration (qspectra/
 * simulate/utils.py:16-50) applies L = evolution_super_operator[idx][:, idx]
 * (qspectra/dynamics/liouville_space.py:316-341) at every internal step.  A
 * classical RK4 step of a constant L is exactly y <- P4(h L) y, so these
 * replace the step loop's operator; the trajectories they produce are the
 * RK4 ones up to rounding.
 *
 * Workspace of each: qsx_propagator_workspace_bytes(M, B) bytes
 * (4 B M^2 complex128).  All matrices complex128 [B][M][M], row-major. */
int64_t qsx_propagator_workspace_bytes(int32_t M, int32_t B);

/* P[b] = P4(h L[b])^q, P4(X) = I + X + X^2/2 + X^3/6 + X^4/24 (Horner: three
 * GEMMs), q >= 1 RK4 steps of size h per application (repeated squaring).
 * strideL = 0: one L shared by the batch. */
int32_t qsx_rk4_propagator(const void* L, int64_t strideL, int32_t M, int32_t B, double h,
                           int32_t q, void* P, void* workspace, int64_t workspace_bytes,
                           void* stream);

/* No counterpart in the reference: qsx_rk4_propagator for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * Same result as qsx_rk4_propagator (the same kernels in the same order); its
 * workspace is a device array inside the library, QSX_PROPAGATOR_SCRATCH_BYTES
 * (32 MiB) per GPU, held until the process ends.  Rejects problems whose
 * qsx_propagator_workspace_bytes(M, B) exceeds it (e.g. M = 147 with B > 24).
 * Calls no allocating routine.  All calls on one GPU share that array:
 * calls that overlap in time on one GPU (concurrent streams, threads) race
 * and must be serialised by the caller.  The stream must belong to the
 * current device. */
#define QSX_PROPAGATOR_SCRATCH_BYTES (32ll << 20)
int32_t qsx_rk4_propagator_noalloc(const void* L, int64_t strideL, int32_t M, int32_t B,
                                   double h, int32_t q, void* P, void* stream);

/* out[b][s] = A[b]^exps[s] for s < ns (exps: HOST array, >= 0; exponent 0
 * gives the identity), one shared chain of squarings. */
int32_t qsx_matrix_powers(const void* A, int64_t strideA, int32_t M, int32_t B,
                          const int64_t* exps, int32_t ns, void* out, void* workspace,
                          int64_t workspace_bytes, void* stream);

/* The MIXED operands of a propagator P [B][M][M]: off [B][M][M] complex64 =
 * P with its diagonal zeroed, diag [B][M] complex128 = diag(P) - 1 (the
 * stepper applies y + off y + diag * y). */
int32_t qsx_split_increment(const void* P, int32_t M, int32_t B, void* off, void* diag,
                            void* stream);

/* MIXED operands (off complex64 [B][M][M], diag complex128 [B][M], as
 * qsx_split_increment) of P4(h L)^q built on the FP32 matrix cores without an
 * FP64 P: with L = D + O, Horner's rule runs on the differences from the
 * exact FP64 diagonal recursion (three complex FP32-MFMA GEMMs whose
 * epilogues add the column-scaled O), then q (a power of two) RK4 steps per
 * application by squaring the split.  Workspace:
 * qsx_split_propagator_workspace_bytes(M, B). */
int64_t qsx_split_propagator_workspace_bytes(int32_t M, int32_t B);
int32_t qsx_rk4_propagator_split(const void* L, int64_t strideL, int32_t M, int32_t B,
                                 double h, int32_t q, void* off, void* diag, void* workspace,
                                 int64_t workspace_bytes, void* stream);

/* Small blocks (M <= 64), one work-group per batch entry with the matrices in
 * LDS, so the whole operator program is one launch:
 *   qsx_small_propagator: P = P4(h L)^q into P [B][M][M] (may be NULL), its
 *     power P^c into Pc [B][M][M] (c = 0: none), and the MIXED split of P
 *     into off / diag as qsx_split_increment (both NULL: none);
 *   qsx_propagate_powers: out[b][s][r] = P4(h L[b])^exps[s] y0[b][r] for R
 *     states y0 [B][R][M] and ns <= 8 cumulative step counts exps (HOST
 *     array): the few-save-point propagation of integrate
 *     (qspectra/simulate/utils.py:16-109, e.g. the t2 stage of
 *     qspectra/simulate/response.py:321-322), the squarings shared by all
 *     save points.
 * qsx_small_propagator_fits(M, R, ns) is 1 when a block fits these kernels
 * (R = 0: qsx_small_propagator). */
int32_t qsx_small_propagator_fits(int32_t M, int32_t R, int32_t ns);
int32_t qsx_small_propagator(const void* L, int64_t strideL, int32_t M, int32_t B, double h,
                             int32_t q, int64_t c, void* P, void* Pc, void* off, void* diag,
                             void* stream);
int32_t qsx_propagate_powers(const void* L, int64_t strideL, int32_t M, int32_t B, double h,
                             const int64_t* exps, int32_t ns, const void* y0, int32_t R, void* out,
                             void* stream);

/* ---------------------------------------------------------------------------
 * HEOM for a Drude bath (qspectra/dynamics/heom.py:177-443), batched over B
 * samples that share the hierarchy structure.  State per sample: ADO-major
 * [n_ado][M] complex128.  Operator tables (device pointers):
 *   Hop    n > 0: the subspace is the single block XY, nX = n, nY = M / n
 *          (M <= 256): [B][nX*nX + nY*nY] = -i*unit*H_X followed by
 *          -i*unit*H_Y, or for a square block (nX = nY) [B][n][n] = H_X
 *          alone (their transposes for the Heisenberg picture);
 *          n = 0: [B][M][M] dense -i*unit*[H, .] block (M <= 64)
 *   shift  [n_ado] double, -unit * sum_jk m_jk nu_k
 *   dvec   [M] complex, -unit * Delta * sum_j (n_j(r) + n_j(c) - 2 n_j(r) n_j(c))
 *   bin_src/bin_coef [n_ado][nbins][2] (nbins = N * K1, bin b = j * K1 + k):
 *          the neighbour coupled through bin b (kind 0: n + e_b with weight
 *          n_j(r) - n_j(c); kind 1: n - e_b with weight c_k n_j(r) - c_k^*
 *          n_j(c); in the Heisenberg picture the transposed edges), -1 for
 *          none, and its coefficient (includes unit)
 *   ck     [K1] complex correlation-function coefficients c_k (heom.py:69-89)
 *   rel_site [M][nrel] sites occupied in the row or the column state of
 *          element p (-1 padded); rel_occ [M][nrel][2] their occupations
 *          (row, column): the weights vanish for every other site
 */

/* out[b] = F(y[b]) (the hierarchy's eom(t, y), heom.py:241-242). */
int32_t qsx_heom_apply(int32_t B, int32_t n_ado, int32_t M, int32_t n, const void* Hop,
                       const double* shift, const void* dvec, int32_t nbins, int32_t K1,
                       const int32_t* bin_src, const void* bin_coef, const void* ck,
                       int32_t nrel, const int32_t* rel_site, const double* rel_occ,
                       const void* y, void* out, void* stream);

/* Device workspace bytes of qsx_heom_propagate (-1 for invalid arguments). */
int64_t qsx_heom_workspace_bytes(int32_t precision, int32_t B, int32_t n_ado, int32_t M);

/* Classical RK4 over the hierarchy (integrate on the HEOM eom,
 * simulate/utils.py:16-109): save interval s has nsub_host[s] steps of size
 * h_host[s] (HOST arrays); out [B][nsave][M] receives ADO 0 (the reduced
 * density matrix) at each save point; y_final (optional) the final states.
 * Each step is four stage launches; MIXED keeps the stage vector in
 * complex64 (state and update complex128).  Allocates nothing: workspace
 * holds qsx_heom_workspace_bytes bytes. */
int32_t qsx_heom_propagate(int32_t precision, int32_t B, int32_t n_ado, int32_t M, int32_t n,
                           const void* Hop, const double* shift, const void* dvec, int32_t nbins,
                           int32_t K1, const int32_t* bin_src, const void* bin_coef,
                           const void* ck, int32_t nrel, const int32_t* rel_site,
                           const double* rel_occ, const void* y0, const int32_t* nsub_host,
                           const double* h_host, int32_t nsave, void* out, void* y_final,
                           void* workspace, int32_t* status, void* stream);

/* The same two operations with the system part in Hilbert-commutator form,
 * for Liouville subspaces of several blocks beyond the dense form's 64
 * elements (the reference's HEOM_tensor on any subspace, heom.py:298-443):
 * Hop [B][N][N] = -i unit H over the Hilbert space (H^T in the Heisenberg
 * picture, block diagonal by excitation number), elem_states [M][2] the row
 * and column state of element p, state_pos [N][N] the element of (row,
 * column) or -1, state_block [N][2] the first and end state of each state's
 * Hilbert block; every (q, column) and (row, q) with q in the block of the
 * row (column) state must be in the subspace.  N <= 128. */
int32_t qsx_heom_apply_hilbert(int32_t B, int32_t n_ado, int32_t M, int32_t N,
                               const int32_t* elem_states, const int32_t* state_pos,
                               const int32_t* state_block, const void* Hop, const double* shift,
                               const void* dvec, int32_t nbins, int32_t K1,
                               const int32_t* bin_src, const void* bin_coef, const void* ck,
                               int32_t nrel, const int32_t* rel_site, const double* rel_occ,
                               const void* y, void* out, void* stream);
int32_t qsx_heom_propagate_hilbert(int32_t precision, int32_t B, int32_t n_ado, int32_t M,
                                   int32_t N, const int32_t* elem_states,
                                   const int32_t* state_pos, const int32_t* state_block,
                                   const void* Hop, const double* shift, const void* dvec,
                                   int32_t nbins, int32_t K1, const int32_t* bin_src,
                                   const void* bin_coef, const void* ck, int32_t nrel,
                                   const int32_t* rel_site, const double* rel_occ,
                                   const void* y0, const int32_t* nsub_host,
                                   const double* h_host, int32_t nsave, void* out,
                                   void* y_final, void* workspace, int32_t* status,
                                   void* stream);

/* qsx_heom_propagate under classical fields for a dense hierarchy block
 * (M <= 64, the multi-block subspaces such as 'gg,ge,eg,ee'): every ADO
 * evolves with the stage operator Hop[b] + sum_i f_i(t) C_i, C [B|1][P][M][M]
 * the dense dipole commutators (strideC 0 or P*M*M), f_i = -i E_i(t)
 * (conjugated for a '+' interaction) sampled by the host at the stage times
 * t, t + h/2, t + h of every step: F [B|1][nF][3][P] (strideF 0 or nF*3*P,
 * nF >= the total step count).  Hwork [B][M][M] complex128 holds the stage
 * operators.  simulate_with_fields / simulate_pump of a HEOM model
 * (qspectra/simulate/eom.py:79-103 with the HEOMSpaceOperator commutator of
 * qspectra/dynamics/heom.py:22-58, which maps over ADOs). */
int32_t qsx_heom_propagate_driven(int32_t precision, int32_t B, int32_t n_ado, int32_t M,
                                  const void* Hop, const double* shift, const void* dvec,
                                  int32_t nbins, int32_t K1, const int32_t* bin_src,
                                  const void* bin_coef, const void* ck, int32_t nrel,
                                  const int32_t* rel_site, const double* rel_occ, int32_t P,
                                  const void* C, int64_t strideC, const void* F, int64_t strideF,
                                  int64_t nF, const void* y0, const int32_t* nsub_host,
                                  const double* h_host, int32_t nsave, void* out, void* y_final,
                                  void* Hwork, void* workspace, int32_t* status, void* stream);

/* ---------------------------------------------------------------------------
 * Fixed-step RK4 under classical fields (rotating wave approximation):
 *     y' = L y + sum_i f_i(t) C_i y,   f_i(t) = -i E_i(t) (conjugated for a
 * '+' interaction), C_i the commutator superoperator of field term i's
 * dipole operator.  Replaces the zvode integration of the field-driven
 * equation of motion in simulate_with_fields / simulate_pump
 * (qspectra/simulate/eom.py:79-103, the callback f at :86-93).
 *
 *   A      L [B|1][M][M] (strideA 0 = shared); MIXED: off-diagonal complex64
 *          plus diag, the complex128 diagonal [B|1][M] (strideD).
 *   K, col, coef   the C_i in ELL form over the union of their sparsity
 *          patterns (over all terms and batch entries): col int32 [M][K]
 *          column of slot k of each row, coef [B|1][M][K][P] (complex64
 *          for MIXED) the P coefficients of that slot (zero in padding
 *          slots); strideE = slots per batch entry of coef (M*K, or 0 =
 *          shared); 1 <= P <= 4.
 *   F      complex128 [B|1][nF][3][P] (strideF): f_i at the stage times
 *          t_g, t_g + h/2, t_g + h of global step g (steps numbered across
 *          the save intervals in order).
 *   nsub, h, nsave, y0 [B][R][M], out [B][R][nsave][M] as qsx_propagate.
 *   status 1: non-finite state; 2: the schedule has more steps than nF.
 * Blocks up to M = 64 (one work-group per right-hand side); K up to 16, or
 * up to the row width. */
int32_t qsx_propagate_driven(int32_t precision, const void* A, int64_t strideA,
                             const void* diag, int64_t strideD, int32_t K, const int32_t* col,
                             const void* coef, int64_t strideE, int32_t P, const void* F,
                             int64_t strideF, int64_t nF, int32_t M, const void* y0, int32_t R,
                             int32_t B, const int32_t* nsub, const double* h, int32_t nsave,
                             void* out, int32_t* status, void* stream);

/* ---------------------------------------------------------------------------
 * ZOFE master equation (qspectra/dynamics/zofe.py:121-225): fixed-step RK4
 * of the density matrix rho (N x N) and the auxiliary operators O[p, s]
 * (P pseudomodes, S sites), optionally driven by classical fields as
 * qsx_propagate_driven (the commutator acting on rho and on every O[p, s],
 * zofe.py:24-34).  One work-group per right-hand side; the state must fit
 * the LDS (N^2 (1 + P S) <= 8192 elements).
 *
 *   H       [B|1][N][N] rotating-frame Hamiltonian of the Hilbert subspace
 *           (complex128, not multiplied by unit_convert; strideH)
 *   L       [B|1][S][N][N] coupling operators (the negated system-bath
 *           couplings, zofe.py:214-215; strideL)
 *   Gamma, w  [B|1][P][S] Omega^2 huang and i Omega + gamma (strideG)
 *   flags   bit 0 ham_hermit, bit 1 rho_hermit (zofe.py:170-186)
 *   nfield, V [B|1][nfield][N][N], F [B|1][nF][3][nfield]: field terms
 *   y0      [B][R][N^2 (1 + P S)] states in the reference's layout (vec of
 *           rho, then vec of O with shape (P, S, N, N), both column-major)
 *   nsub, h, nsave  schedule as qsx_propagate
 *   K, proj [B|1][K][N^2] rows applied to vec(rho) at every save (for
 *           expectation values, zofe.py:36-39), or K = 0: full states
 *   out     [B][R][nsave][K or size]
 *   workspace  qsx_zofe_workspace_bytes(B, R, N, S, P) bytes
 *   status  1 non-finite state, 2 fewer field samples than steps.
 * Replaces the zvode integration of ZOFEModel.equation_of_motion
 * (simulate/utils.py:16-109 with zofe.py:204-225). */
int64_t qsx_zofe_workspace_bytes(int32_t B, int32_t R, int32_t N, int32_t S, int32_t P);
int32_t qsx_zofe_propagate(int32_t B, int32_t R, int32_t N, int32_t S, int32_t P, const void* H,
                           int64_t strideH, const void* L, int64_t strideL, const void* Gamma,
                           const void* w, int64_t strideG, int32_t flags, double unit_convert,
                           int32_t nfield, const void* V, int64_t strideV, const void* F,
                           int64_t strideF, int64_t nF, const void* y0, const int32_t* nsub,
                           const double* h, int32_t nsave, int32_t K, const void* proj,
                           int64_t strideP, void* out, void* workspace, int32_t* status,
                           void* stream);
/* eom(t, y) of the ZOFE model (zofe.py:211-225) for R states per sample:
 * out [B][R][size] in the reference's layout. */
int32_t qsx_zofe_apply(int32_t B, int32_t R, int32_t N, int32_t S, int32_t P, const void* H,
                       int64_t strideH, const void* L, int64_t strideL, const void* Gamma,
                       const void* w, int64_t strideG, int32_t flags, double unit_convert,
                       const void* y, void* out, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Dense-block Liouvillian builders, for Hamiltonians whose Hilbert blocks
 * are not built from a one-exciton matrix (VibronicHamiltonian,
 * qspectra/hamiltonian.py:633-807).  HX [B][nX][nX], HY [B][nY][nY]: the
 * rotating-frame Hamiltonian restricted to Hilbert blocks X and Y (real
 * symmetric, <= 64 states; pass the same pointers for X == Y); occX
 * [n][nX], occY [n][nY]: the diagonals of the n system-bath coupling
 * operators over the block bases.  Same output and flags as
 * qsx_redfield_block / qsx_unitary_block (RedfieldModel / UnitaryModel
 * evolution_super_operator, redfield.py:9-152, unitary.py:5-9).  The
 * workspace holds qsx_redfield_dense_workspace_bytes(n, nX, nY, B) bytes. */
int64_t qsx_redfield_dense_workspace_bytes(int32_t n, int32_t nX, int32_t nY, int32_t B);
int32_t qsx_redfield_block_dense(const double* HX, const double* HY, int32_t nX, int32_t nY,
                                 const double* occX, const double* occY, int32_t n, int32_t B,
                                 double T, double lam, double gam, int32_t flags,
                                 double unit_convert, void* L, void* workspace, void* stream);
int32_t qsx_unitary_block_dense(const double* HX, const double* HY, int32_t nX, int32_t nY,
                                int32_t B, double unit_convert, int32_t heisenberg, void* L,
                                void* stream);
/* Eigensystems of B dense real symmetric matrices H [B][nb][nb] (nb <= 64):
 * E [B][nb] ascending, U [B][nb][nb] (columns, canonical signs) as
 * qsx_block_eigh; workspace qsx_redfield_dense_workspace_bytes(1, nb, nb, B).
 * Replaces Hamiltonian.eig (hamiltonian.py:310-328) for dense Hamiltonians. */
int32_t qsx_dense_eigh(const double* H, int32_t nb, int32_t B, double* E, double* U,
                       void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Complex-Hermitian Hamiltonians (check_hermitian, qspectra/hamiltonian.py:
 * 18-22).  H1 [B][n][n] complex128: the non-rotating one-exciton Hamiltonian
 * of each problem; Hilbert block X (0 g, 1 e, 2 f) as for the real entry
 * points, diagonal shifted by -X rw[b] (rw may be null).
 *
 * qsx_block_eigh_hermitian: E [B][nb] ascending, U [B][nb][nb] complex128
 * (eigenvectors as columns, each with its largest-magnitude component real
 * and positive), by complex Jacobi rotations in LDS, nb <= 64.  Replaces
 * Hamiltonian.eig (hamiltonian.py:310-328, scipy eigh) for complex H.
 * qsx_dense_eigh_hermitian: the same for dense Hermitian H [B][nb][nb].
 * qsx_unitary_block_hermitian: the XY block of -i unit [H, .] (complex128
 * [B][M][M], M = nX nY, column-stacked; its transpose for heisenberg != 0),
 * qspectra/dynamics/unitary.py:5-9 with liouville_space.py:93-131.
 * qsx_unitary_block_dense_hermitian: the same from dense blocks HX [B][nX][nX],
 * HY [B][nY][nY] complex128. */
int32_t qsx_block_eigh_hermitian(const void* H1, int32_t n, int32_t B, int32_t X,
                                 const double* rw, double* E, void* U, void* stream);
int32_t qsx_dense_eigh_hermitian(const void* H, int32_t nb, int32_t B, double* E, void* U,
                                 void* stream);
int32_t qsx_unitary_block_hermitian(const void* H1, int32_t n, int32_t B, const double* rw,
                                    double unit_convert, int32_t X, int32_t Y,
                                    int32_t heisenberg, void* L, void* stream);
int32_t qsx_unitary_block_dense_hermitian(const void* HX, const void* HY, int32_t nX,
                                          int32_t nY, int32_t B, double unit_convert,
                                          int32_t heisenberg, void* L, void* stream);

/* ---------------------------------------------------------------------------
 * fourier_transform (qspectra/simulate/utils.py:128-219): x [batch][n]
 * complex128 sampled at dt, zero-padded to the grid of N points symmetric
 * about t = 0 with the samples at lo .. lo+n-1, scaled by dt, ifftshift, DFT,
 * fftshift, reversed for sign = +1; out [batch][N].  With adjoint != 0 the
 * conjugate transpose: x [batch][N] -> out [batch][n] (the backward of the
 * transform).  A direct sum (N*n multiply-adds per row, exact integer
 * reduction of the exponent, a twiddle table of N entries in the workspace of
 * qsx_fourier_workspace_bytes(N) bytes); nothing is compiled at run time. */
int64_t qsx_fourier_workspace_bytes(int32_t N);
int32_t qsx_fourier_transform(const void* x, int64_t batch, int32_t n, int32_t N, int32_t lo,
                              int32_t sign, double dt, int32_t adjoint, void* workspace,
                              void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QSX_H */
