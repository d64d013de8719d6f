/*
 * gradwave_hip -- C ABI of the MI355X (gfx950) kernels behind gradwave's
 * plane-wave H|psi> apply and batched block Davidson.
 *
 * Conventions (all entry points):
 *   - plain pointers and sizes; device pointers are HIP device memory owned by
 *     the caller (no ownership transfer); `stream` is a hipStream_t or NULL;
 *   - complex data is interleaved (re, im), row-major, in the reference's
 *     padded k-batch layout (nk, nb, npw_max) with zeros in padded slots
 *     (core/batch.py:1-14);
 *   - precision GW_FP64 = complex128/float64 data, GW_FP32 = complex64/float32;
 *   - return 0 on success, a GW_ERR_* code otherwise; gw_last_error() gives
 *     the message of the last failure on the calling thread.
 *
 * The reference (gradwave, pure Python/PyTorch) has no C ABI; each entry point
 * names the reference function whose tensor contract it implements. A
 * reference-side binding is a ctypes stub that declares these signatures.
 */
#ifndef GRADWAVE_HIP_H
#define GRADWAVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GW_OK = 0,
  GW_ERR_ARG = 1,         /* bad argument / shape */
  GW_ERR_HIP = 2,         /* HIP runtime error */
  GW_ERR_UNSUPPORTED = 3, /* box length or size outside what the kernels handle */
  GW_ERR_WORKSPACE = 4    /* workspace smaller than gw_*_workspace_bytes */
};
enum { GW_FP64 = 0, GW_FP32 = 1 };
/* terms of H for gw_hpsi_apply (bit mask) */
enum { GW_TERM_KINETIC = 1, GW_TERM_LOCAL = 2, GW_TERM_NONLOCAL = 4, GW_TERM_ALL = 7 };

typedef struct gw_plan gw_plan;

const char* gw_last_error(void);
int gw_version(void);

/* ---- operator set-up: BatchedHamiltonian.__init__ / _tables (core/batch.py:249-331) ---- */

/* Geometry from the padded BatchedK maps (core/batch.py:126-205): flat_idx
 * (nk, npw_max) int64 host array, 0 in padding; npw (nk) int64 host array.
 * Any box is accepted. The box FFT kernels need every length to factor as a*b
 * with a, b in {1..10,12,14,15,16} (the three-factor splits a*b*c are built but
 * not enabled); on a box with another length (good_fft_size, grids.py:32-41,
 * also returns 125, 162, ..., every 7-smooth length) the plan serves the
 * kinetic and nonlocal terms and gw_sphere_to_box / gw_box_to_sphere, and the
 * local term, psi(r), the V(r) adjoint, the density and ragged batches return
 * GW_ERR_UNSUPPORTED naming the length (info[14] of gw_plan_info holds it). */
int gw_plan_create(gw_plan** out, int n1, int n2, int n3, int nk, int npw_max,
                   const int64_t* flat_idx, const int64_t* npw);
/* The split a plan takes for box length n: f[0..2] with f[0]*f[1]*f[2] = n and
 * f[2] = 1 (two codelets); GW_ERR_UNSUPPORTED when there is none (e.g. 125,
 * 162, 200, 216, 243, 250 and every length above 256). Host only, no GPU. */
int gw_box_split(int n, int* f);
int gw_plan_destroy(gw_plan* plan);
/* plan options (no reference counterpart): GW_PLAN_LEGACY selects the generic
 * LDS kernels (any 7-smooth box) instead of the wave-line kernels of
 * cubic 24/36/90/120 boxes; GW_PLAN_FORCE_UNFUSED then splits each x-plane pass
 * into stick and column-chunk kernels (both for testing). GW_PLAN_PENCIL runs
 * the local term as z-stick transforms around one (y, x) slab kernel per
 * (z, band) wherever a slab of the active planes fits LDS (moving sticks, not
 * planes, between kernels). */
enum { GW_PLAN_FORCE_UNFUSED = 1, GW_PLAN_TIMING = 2, GW_PLAN_LEGACY = 4, GW_PLAN_PENCIL = 8 };
int gw_plan_set_flags(gw_plan* plan, int flags);
/* With GW_PLAN_TIMING set, gw_hpsi_apply records HIP events around each of its
 * stages on the launch stream: 0 x-plane inverse (pack, z, y), 1 x pass
 * (x inverse, V, x forward), 2 x-plane forward (y, z, gather, kinetic),
 * 3 projection <p|c>, 4 D multiply, 5 back-projection. This returns the summed
 * milliseconds and call counts per stage (synchronising on pending events). */
int gw_plan_stage_times(gw_plan* plan, double* ms, long* calls, int n, int reset);
/* info[0..14] = active x-planes, max sticks per k, fused plane path (0/1) for
 * `precision` (2 wave-line, 1 LDS plane tile, 0 split), LDS bytes per
 * workgroup, nproj, the first two factors of n1, n2, n3 (pairs), the third
 * factor of n1, n2, n3 (1 for a two-factor split), then 0 or the first box
 * length with no split (no box FFTs on this plan). */
int gw_plan_info(const gw_plan* plan, int precision, int* info, int n);
/* kinetic diagonal t = HBAR2_2M |k+G|^2 (nk, npw_max) float64, device (BatchedK.t) */
int gw_plan_set_kinetic(gw_plan* plan, const double* t, void* stream);
/* local potential v_eff_r (n1, n2, n3) float64, device (BatchedHamiltonian.v_eff_r) */
int gw_plan_set_potential(gw_plan* plan, const double* v_r, void* stream);
/* projectors p (nk, nproj, npw_max) complex128 device (projectors_b, core/batch.py:229-235)
 * and D (nproj, nproj) float64 host (BatchedK.dij_full); nproj may be 0. */
int gw_plan_set_projectors(gw_plan* plan, int nproj, const void* p, const double* dij,
                           void* stream);

/* ---- H|psi>: BatchedHamiltonian.apply (core/batch.py:480-505) ----
 * out = (t*c + FFT^-1[V FFT c] + p D <p|c>) * mask for the terms in `terms`.
 * c, out: (nk, nb, npw_max) complex device, may not alias. Workspace: device
 * scratch of at least gw_hpsi_workspace_bytes(plan, nb, precision) bytes. */
size_t gw_hpsi_workspace_bytes(const gw_plan* plan, int nb, int precision);
int gw_hpsi_apply(gw_plan* plan, int precision, int terms, const void* c, void* out, int nb,
                  void* workspace, size_t workspace_bytes, void* stream);
/* Same on strided blocks: row (band) stride ldc/ldo and k stride sc/so, in
 * complex elements (e.g. a band slice v[:, m0:m1] of a Davidson basis). */
int gw_hpsi_apply_strided(gw_plan* plan, int precision, int terms, const void* c, long ldc,
                          long sc, void* out, long ldo, long so, int nb, void* workspace,
                          size_t workspace_bytes, void* stream);
/* No-allocation variant of gw_hpsi_apply_strided. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its scratch is a fixed array of gw_hpsi_noalloc_max_bytes() bytes (64 MiB)
 * in the library's device memory, one per GPU, held until the process ends;
 * a call whose gw_hpsi_workspace_bytes exceed it fails with GW_ERR_WORKSPACE
 * before any launch. It calls no hipMalloc/hipFree and no library routine that
 * allocates: the projector contractions always take the MFMA kernels here
 * (never rocBLAS), and plan timing is skipped for the call. Every call on a
 * GPU shares that GPU's scratch, so calls that overlap on one GPU (other
 * streams, other threads) race; the scratch keeps the last call's contents and
 * needs no clearing (every byte a call reads it has written first). */
size_t gw_hpsi_noalloc_max_bytes(void);
int gw_hpsi_apply_noalloc(gw_plan* plan, int precision, int terms, const void* c, long ldc, long sc,
                          void* out, long ldo, long so, int nb, void* stream);

/* psi(r) = N ifftn(scatter(c)) on the natural box (nk, nb, n1, n2, n3) complex:
 * core/fftbox.py:63-70 g_to_r / core/batch.py:208-216 g_to_r_b. */
size_t gw_wfc_to_real_workspace_bytes(const gw_plan* plan, int nb, int precision);
int gw_wfc_to_real(gw_plan* plan, int precision, const void* c, void* psi_r, int nb,
                   void* workspace, size_t workspace_bytes, void* stream);

/* Pack/unpack through the plan's plane/stick maps (bit-exact index parity):
 * box (nk, nb, n1*n2*n3) complex = scatter of c by flat_idx (core/fftbox.py:29-37);
 * c_out (nk, nb, npw_max) = gather of box by flat_idx, 0 in padding (fftbox.py:40-43). */
int gw_sphere_to_box(gw_plan* plan, int precision, const void* c, void* box, int nb, void* stream);
int gw_box_to_sphere(gw_plan* plan, int precision, const void* box, void* c_out, int nb,
                     void* stream);

/* ---- dense complex contractions on MFMA (solvers/davidson.py:428-432, 648-654) ----
 * inner:   C[k][i][j] = sum_g conj(X[k][i][g]) * Y[k][j][g]
 *          X (nk, M, K) row stride ldx, batch stride sx; Y (nk, N, K); C (nk, M, N).
 *          e.g. s = v^H hv (davidson.py:648), becp (core/batch.py:590-595).
 * combine: Z[k][a][g] = beta*Z[k][a][g] + sum_j U'[k][j][a] * Y[k][j][g],
 *          U' = U or conj(U); U (nk, J, A); Y (nk, J, G); Z (nk, A, G).
 *          e.g. Ritz vectors x = u^T v (davidson.py:653), back-projection (batch.py:500).
 * Strides are in complex elements. The inner product needs scratch for its
 * split-K partial sums: gw_zgemm_inner_workspace_bytes. An operand of zero size
 * may be null; K = 0 writes C = 0 and J = 0 leaves Z = beta*Z (zero for beta = 0). */
size_t gw_zgemm_inner_workspace_bytes(int precision, int nk, int M, int N, int K);
int gw_zgemm_inner(int precision, int nk, int M, int N, int K, const void* X, long ldx, long sx,
                   const void* Y, long ldy, long sy, void* C, long ldc, long sc,
                   void* workspace, size_t workspace_bytes, void* stream);
int gw_zgemm_combine(int precision, int nk, int A, int G, int J, const void* U, long ldu,
                     long su, int conj_u, const void* Y, long ldy, long sy, void* Z, long ldz,
                     long sz, double beta, void* stream);

/* ---- real contractions on MFMA (the Gamma layout's real-embedded Davidson,
 * core/gamma.py:291-320 over solvers/davidson.py:329-395) ----
 * inner:   C[k][i][j] = sum_g X[k][i][g] Y[k][j][g]      (split-K, fp64 partial sums)
 * combine: Z[k][a][g] = beta Z[k][a][g] + alpha sum_j U[k][j][a] Y[k][j][g]
 * Real float64 / float32 data, strides in elements. */
size_t gw_rgemm_inner_workspace_bytes(int precision, int nk, int M, int N, int K);
int gw_rgemm_inner(int precision, int nk, int M, int N, int K, const void* X, long ldx, long sx,
                   const void* Y, long ldy, long sy, void* C, long ldc, long sc,
                   void* workspace, size_t workspace_bytes, void* stream);
int gw_rgemm_combine(int precision, int nk, int A, int G, int J, const void* U, long ldu, long su,
                     const void* Y, long ldy, long sy, void* Z, long ldz, long sz, double alpha,
                     double beta, void* stream);

/* ---- Davidson vector kernels ----
 * residual (davidson.py:656-657): r = hx - eig*x, rn[k][b] = ||r[k][b]|| (float64 out).
 * x, hx, r: (nk, nb, npw) complex with row stride ld and batch stride s; eig (nk, nb) float64. */
int gw_residual(int precision, int nk, int nb, int npw, const void* x, const void* hx,
                long ld, long s, const double* eig, void* r, double* rn, void* stream);
/* Teter preconditioned, unit-normalised expansion vectors (davidson.py:729-743,
 * precond.py:20-39): for each selected band sel[k][j] (int64, (nk, nsel)):
 *   tb = <x|T|x>, d = K(t/tb) * r[sel], d /= ||d|| (if > 1e-300).
 * d: (nk, nsel, npw) complex out with row stride ldd, batch stride sd. */
int gw_teter_expand(int precision, int nk, int nb, int nsel, int npw, const void* x,
                    const void* r, long ld, long s, const double* t, long ldt,
                    const int64_t* sel, void* d, long ldd, long sd, void* stream);
/* row norms ||v[k][i]|| for (nk, m, npw) complex, float64 out (nk, m) */
int gw_row_norms(int precision, int nk, int m, int npw, const void* v, long ld, long s,
                 double* out, void* stream);

/* ---- batched subspace eigensolve ----
 * Replaces the subspace eigensolve of davidson_batched (davidson.py:640-652,
 * torch.linalg.eigh of the Hermitised (nk, m, m) matrix, of which only the
 * first nb Ritz pairs are used, davidson.py:651-652) for many independent
 * matrices of different sizes: the lowest nev eigenpairs of each, one
 * workgroup per matrix, all matrices in one launch (32 per launch).
 * a: n x n complex128, row-major with leading dimension lda, Hermitised as
 *    (A + A^H) / 2 on load (read only);
 * w: nev float64, ascending; z: n x nev complex128 row-major (ldz), column j
 *    the unit eigenvector of w[j];
 * ws: gw_heev_workspace_bytes(n, nev) bytes of device memory per job
 *    (16-byte aligned, not shared between jobs of one call).
 * 1 <= n <= gw_heev_max_n() (GW_ERR_UNSUPPORTED above), 1 <= nev <= n.
 * Householder tridiagonalisation, multisection on Sturm counts, inverse
 * iteration with reorthogonalisation inside clusters closer than 1e-5 |T|,
 * back-transformation; eigenvalues to a few ulp of |T|, eigenvectors
 * orthonormal to ~eps |T| / (gap to the nearest eigenvalue outside the
 * cluster), at most ~1e-11. The call copies the job array into the launch
 * arguments, so the host array may be reused as soon as it returns. */
typedef struct gw_heev_job {
  const void* a;
  void* w;
  void* z;
  void* ws;
  int n, lda, nev, ldz;
} gw_heev_job;
size_t gw_heev_workspace_bytes(int n, int nev);
int gw_heev_max_n(void);
int gw_heev_batched(int count, const gw_heev_job* jobs, void* stream);
/* Batched Cholesky factor inverse, the CholQR step of the orthonormalisation
 * (davidson.py:195-222: G = V V^H = L L^H, Q = L^-1 V): for each Hermitian
 * positive definite G (n x n complex128 row-major, leading dimension lda, lower
 * triangle read), linv = L^-1 (n x n row-major, ldl, zeros above the
 * diagonal) and info = 0, or the 1-based index of the first non-positive
 * pivot (linv is then not usable), as torch.linalg.cholesky_ex reports it.
 * 1 <= n <= gw_chol_max_n(); one workgroup per matrix, 32 per launch; the
 * job array is copied into the launch arguments. */
typedef struct gw_chol_job {
  const void* a;
  void* linv;
  void* info; /* int32 */
  int n, lda, ldl;
} gw_chol_job;
int gw_chol_max_n(void);
int gw_chol_inv_batched(int count, const gw_chol_job* jobs, void* stream);

/* ---- Sternheimer conjugate-gradient column kernels ----
 * Replace the per-column updates of cg_sternheimer / _cg_percolumn
 * (postscf/_response.py:502-690). The CG vectors x, p, r, z, ap live in the
 * padded layout: column c = k * nocc + n is the row c of a (nk * nocc, ld)
 * complex array. Each call works on the nrows columns listed in rows (int64
 * column indices), one workgroup per listed column; sums are fp64.
 *
 * gw_cg_update: for listed column j (c = rows[j], k = kcol[j], a = alpha[j]):
 *   x_c += a p_c, r_c -= a ap_c (both skipped when a == 0),
 *   z_c = K(t[k] / t_band[c]) r_c with the Teter filter K of solvers/precond.py:20-39,
 *   rn2[j] = ||r_c||^2, rz[j] = Re <r_c, z_c>.
 *   t: (nk, ldt) float64 kinetic diagonal; t_band: (nk * nocc) float64.
 * gw_cg_direction: p_c = z_c + beta[j] p_c; when pc is not null the new p_c is
 *   also written to the compact row j of pc (row stride ldc) for a column apply.
 * gw_row_dots: out[j] = Re sum_g conj(a_c[g]) w[wrow[j]][g] b_c[g] with c = rows[j]
 *   (c = j when rows is null, w = 1 when null, w row 0 when wrow is null). */
int gw_cg_update(int precision, int nrows, int npw, const int64_t* rows, const int64_t* kcol,
                 const double* alpha, void* x, const void* p, void* r, const void* ap, long ld,
                 const double* t, long ldt, const double* t_band, void* z, double* rn2, double* rz,
                 void* stream);
int gw_cg_direction(int precision, int nrows, int npw, const int64_t* rows, const double* beta,
                    const void* z, void* p, long ld, void* pc, long ldc, void* stream);
int gw_row_dots(int precision, int nrows, int npw, const int64_t* rows, const void* a, const void* b,
                long ld, const double* w, long ldw, const int64_t* wrow, double* out, void* stream);

/* ---- adjoints (vector-Jacobian products) of H|psi> ----
 * The reference differentiates HamiltonianK.apply (core/hamiltonian.py:179-191),
 * projectors (108-112) and local_potential_g (core/energies/local_pp.py:22-38)
 * with torch autograd (postscf/forces.py:168-215). Gradients follow torch's
 * complex convention: for a real loss L, grad z = dL/dRe z + i dL/dIm z.
 *
 * V(r): grad_v (n1, n2, n3) float64, natural order, is ACCUMULATED:
 *   grad_v(r) += (1/N) sum_{k,b} Re[conj(phi_kb(r)) psi_kb(r)],
 *   psi = N ifftn(scatter c), phi = N ifftn(scatter g); c, g as in gw_hpsi_apply_strided.
 *   Needs a cubic 24/36/90/120 box (gw_plan_info fused_planes == 2). */
size_t gw_hpsi_adjoint_v_workspace_bytes(const gw_plan* plan, int nb, int precision);
int gw_hpsi_adjoint_v(gw_plan* plan, int precision, const void* c, long ldc, long sc, const void* g,
                      long ldg, long sg, int nb, double* grad_v, void* workspace,
                      size_t workspace_bytes, void* stream);
/* band density (core/batch.py:562-587 before its division by the volume):
 * rho (n1, n2, n3) float64 natural order, ACCUMULATED: rho(r) += sum_{k,b} w[k][b] |psi_kb(r)|^2,
 * psi = N ifftn(scatter c); weights (nk, nb) float64 on the device (kweight * occupation).
 * Needs a cubic 24/36/90/120 box (gw_plan_info fused_planes == 2). */
size_t gw_density_workspace_bytes(const gw_plan* plan, int nb, int precision);
int gw_density(gw_plan* plan, int precision, const void* c, long ldc, long sc, int nb,
               const double* weights, double* rho, void* workspace, size_t workspace_bytes, void* stream);
/* projector phases (core/batch.py:229-235): p[k][q][g] = pf[k][q][g] exp(-i kpg[k][g] . tau[a(q)]),
 * 0 for g >= npw[k]. pf, p (nk, nproj, npw_max) complex128; kpg (nk, npw_max, 3) float64 A^-1;
 * positions (na, 3) A; atom_index (nproj,) int64; npw (nk,) int32 on the device. */
int gw_proj_phase(int nk, int nproj, int npw_max, const int* npw, const void* pf, const double* kpg,
                  const double* positions, const int64_t* atom_index, void* p, void* stream);
/* its adjoint: grad_tau[a][x] += sum_{k, q: a(q)=a, g} Im[conj(grad_p) p] kpg_x (grad_tau (na, 3)) */
size_t gw_proj_phase_grad_workspace_bytes(int nk, int nproj);
int gw_proj_phase_grad(int nk, int nproj, int npw_max, int na, const void* grad_p, const void* p,
                       const double* kpg, const int64_t* atom_index, double* grad_tau, void* workspace,
                       size_t workspace_bytes, void* stream);
/* local potential on the full box (local_pp.py:22-38): vloc_g (npts) complex128 =
 * (1/volume) sum_a tables[species_index[a]][G] exp(-i G . tau_a);
 * tables (nspecies, npts) float64 eV A^3; g_cart (npts, 3) A^-1; npts = n1 n2 n3. */
int gw_vloc_g(int na, long npts, const double* positions, const int64_t* species_index,
              const double* tables, const double* g_cart, double volume, void* vloc_g, void* stream);
/* its adjoint: grad_tau[a][x] += (1/volume) sum_G tab(G) G_x Im[S_a(G) conj(grad_vloc_g(G))] */
size_t gw_vloc_g_grad_workspace_bytes(int na);
int gw_vloc_g_grad(int na, long npts, const double* positions, const int64_t* species_index,
                   const double* tables, const double* g_cart, double volume, const void* grad_vloc_g,
                   double* grad_tau, void* workspace, size_t workspace_bytes, void* stream);

/* ---- tuning knobs (no reference counterpart; written by gradwave_amd/autotune.py) ----
 * GW_TUNE_RADIX: for box length `arg`, the first factor `value` of the two-codelet
 *   split used by plans created afterwards (0 = the default most balanced split).
 * GW_TUNE_GEMM_WGS: the number of workgroups the split-K inner products aim at
 *   (`arg` ignored; 0 = default 2048).
 * GW_TUNE_GEMM_TILE: MFMA blocks per wave of the complex GEMMs; `arg` = precision
 *   (GW_FP64 / GW_FP32) + 256 * (0: both contractions, 1: inner only, 2: combine
 *   only; gw_tune_get with 0 reports the inner one); value 0 = 2x2, 1 = 4x2,
 *   2 = 2x4, 3 = 4x4; 4 and 5 cover 160
 *   bands (80 complex columns of the inner product, 160 rows of the
 *   back-projection) per workgroup: 4 = inner 2x5 / combine 5x2, 5 = inner 4x5 /
 *   combine 5x4. 6-13 are the complex-accumulator kernels (workgroup tile
 *   32 BM x 32 BN complex, double-buffered LDS): 6-9 with four real MFMA
 *   products per complex one (inner BM x BN = 2x2, 3x3, 2x5, 4x3; combine 2x2,
 *   3x2, 5x2, 3x4), 10-13 with three (Gauss: inner 2x2, 3x2, 2x3, 3x3; combine
 *   2x2, 3x2, 2x3, 5x2).
 * GW_TUNE_XPASS: launch shape of the wave-line x pass for precision `arg` (GW_FP64 /
 *   GW_FP32; any other `arg` sets both), value = mode * 10000 + waves per workgroup
 *   (2, 4, 8) * 100 + bands per workgroup (1..64); mode 0 reads the x-lines straight
 *   from the box, mode 1 stages each band's planes through LDS with row-contiguous
 *   loads and stores (used when the tiles fit in LDS); default 408 for FP64,
 *   10416 for FP32.
 * GW_TUNE_COMBINE_LIB: for precision `arg`, 1 runs gw_zgemm_combine (and the
 *   back-projection) on rocBLAS strided-batched GEMM, 0 on the kernels here (default).
 * GW_TUNE_PLANE_BANDS: bands one workgroup of the wave-line plane passes transforms in
 *   turn for one x-plane, 1..64; `arg` = precision + 256 * (0: both passes, 1: P1 only,
 *   2: P3 only; gw_tune_get with 0 reports P1). 0, the default: 4 where the plane's
 *   LDS allows one workgroup per CU (P1 of both precisions, P3 in FP64), else 1. The
 *   geometry is set up once per workgroup and the next band's points load during a
 *   band's transforms.
 * GW_TUNE_BAND_CHUNK: for precision `arg`, bands per chunk of the local term of
 *   gw_hpsi_apply (0, the default: the whole block); each chunk runs the three
 *   passes back to back on a box slice of that many bands, which also sets the
 *   box part of gw_hpsi_workspace_bytes.
 * GW_TUNE_INNER_LIB: for precision `arg`, 1 runs the K-chunks of gw_zgemm_inner (and
 *   of the projection <beta|psi>) as one rocBLAS strided-batched GEMM whose partials the
 *   FP64 split-K reduction here sums, 0 the MFMA kernels here (default). */
enum { GW_TUNE_RADIX = 1, GW_TUNE_GEMM_WGS = 2, GW_TUNE_GEMM_TILE = 3, GW_TUNE_XPASS = 4,
       GW_TUNE_COMBINE_LIB = 5, GW_TUNE_INNER_LIB = 6, GW_TUNE_BAND_CHUNK = 7, GW_TUNE_PLANE_BANDS = 8 };
int gw_tune_set(int key, int arg, int value);
int gw_tune_get(int key, int arg);

/* ---- ragged batches of independent operators ----
 * No reference counterpart: the reference applies one operator per call
 * (core/batch.py:480-505); a batch of independent cells (different boxes,
 * spheres, potentials, projector and band counts) runs here as ONE launch per
 * local-term pass over all instances' work items, then the projector GEMMs of
 * every instance as one launch per contraction (inner product, split-K sum, D
 * multiply, back-projection; job tables uploaded by bind). With a rocBLAS
 * contraction selected (GW_TUNE_*_LIB) the instances' GEMM chains run instead
 * on the launch stream and three side streams of the batch (default priority;
 * forked from and joined back into the launch stream by events, so a graph
 * capture of apply_bound records them too). Calls through one batch must not
 * overlap. Plans must take the generic-size plane kernels (x-plane fits LDS).
 * Each call's descriptor tables (per-instance block pointers, strides and band
 * counts; nb[i] = 0 leaves instance i out of the call) reach the device by an
 * asynchronous copy on the launch stream from pinned staging, so they are
 * ordered after every earlier launch on that stream.
 * gw_batch_apply: eager call; its tables go to one of a few slots owned by the
 *   batch (a slot is reused only after the launch that read it has finished).
 * gw_batch_bind: creates a binding with device tables of its own, uploaded on
 *   `stream`; gw_batch_apply_bound(binding, stream) only launches, so it can be
 *   recorded in a HIP graph, and the graph stays valid while the binding, its
 *   blocks and its workspace live, whatever later calls through the batch do.
 *   The batch must outlive its bindings.
 * Terms must include GW_TERM_LOCAL. Workspace: gw_batch_workspace_bytes. */
typedef struct gw_batch gw_batch;
typedef struct gw_batch_binding gw_batch_binding;
int gw_batch_create(gw_batch** out, gw_plan* const* plans, int n);
int gw_batch_destroy(gw_batch* batch);
size_t gw_batch_workspace_bytes(const gw_batch* batch, const int* nb, int precision);
int gw_batch_bind(gw_batch* batch, int precision, int terms, const void* const* c, const long* ldc,
                  const long* sc, void* const* out, const long* ldo, const long* so, const int* nb,
                  void* workspace, size_t workspace_bytes, void* stream, gw_batch_binding** binding);
int gw_batch_apply_bound(gw_batch_binding* binding, void* stream);
int gw_batch_binding_destroy(gw_batch_binding* binding);
int gw_batch_apply(gw_batch* batch, int precision, int terms, const void* const* c, const long* ldc,
                   const long* sc, void* const* out, const long* ldo, const long* so, const int* nb,
                   void* workspace, size_t workspace_bytes, void* stream);

/* ---- small-cell local term as a matrix (core/batch.py:333-464) ----
 * M[k][i][j] = vhat[(G_i - G_j) wrapped into the box] for i, j < npw[k], else 0;
 * vhat (n1 n2 n3) complex = fftn(v_eff) / N; npw (nk) and flat_idx (nk, npw_max)
 * int32 on the device; M (nk, npw_max, npw_max) complex. The local term is then
 * out[k][b][i] = sum_j M[k][i][j] c[k][b][j] (gw_zgemm_inner of conj(c) with M). */
int gw_toeplitz_build(int precision, int nk, int npw_max, int n1, int n2, int n3, const int* npw,
                      const int* flat_idx, const void* vhat, void* M, void* stream);

/* ---- Gamma-point real layout (core/gamma.py) ----
 * Half-sphere vectors travel in the real embedding of core/gamma.py:242-262:
 * phi[0] = Re c(0), phi[2h-1] = sqrt2 Re c(h), phi[2h] = sqrt2 Im c(h), length npw = 2 nhalf - 1.
 * pack: z[j] = full(phi[2j]) + i full(phi[2j+1]) on the full sphere (npw points,
 *   zero up to npw_max), full(c)[g] = c[src_half[g]], conjugated where conj_full[g];
 *   z ((nb+1)/2, ldz) complex; a missing odd partner band counts as zero.
 * unpack: H phi[2j] / H phi[2j+1] from Hz = H z: (Hz(G) +- conj Hz(-G)) / (2 or 2i)
 *   on the half sphere (rep_full[h] = full index of half slot h, partner[g] = index of -G),
 *   written in the embedding, out (nb, ldo) real. Together with gw_hpsi_apply_strided
 *   on a one-k plan of the full sphere this is GammaHamiltonian.apply (core/gamma.py:206-228)
 *   at two bands per complex apply. int32 maps, uint8 flags. */
int gw_gamma_pack(int precision, int nb, int npw, int npw_max, const void* phi, long ldp,
                  const int* src_half, const unsigned char* conj_full, void* z, long ldz, void* stream);
int gw_gamma_unpack(int precision, int nb, int nhalf, const void* hz, long ldz, const int* rep_full,
                    const int* partner, void* out, long ldo, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GRADWAVE_HIP_H */
