/*
 * smh: MI355X (gfx950) kernels for smuthi's lookup-table particle-coupling
 * solve and plane-wave-expansion near field.
 *
 * Plain C ABI: opaque handles, plain pointers and sizes, no framework types.
 * Complex numbers cross the boundary as interleaved (re, im) doubles, i.e.
 * numpy complex128 memory. System vectors use smuthi's layout: particle after
 * particle, each particle's blocksize(l_max_i, m_max_i) coefficients in the
 * (tau, l, m) order of smuthi.fields.multi_to_single_index.
 *
 * Every call returns 0 on success or one of the SMH_ERR_* codes; the message
 * of the last failure on a context is available from smh_last_error. Calls on
 * one context (and the operators created from it) share one HIP stream, are
 * synchronous and are not reentrant.
 */
#ifndef SMH_H
#define SMH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMH_OK 0
#define SMH_ERR_ARG 1         /* invalid argument or size */
#define SMH_ERR_HIP 2         /* HIP runtime error */
#define SMH_ERR_UNSUPPORTED 3 /* configuration the library has no kernel for */
#define SMH_ERR_STATE 4       /* call out of order (e.g. matvec before a table) */
#define SMH_ERR_NOCONV 5      /* solver stopped before reaching the tolerance */

#define SMH_INTERP_LINEAR 0      /* linear interpolation (smuthi 'linear', CPU and GPU) */
#define SMH_INTERP_CATMULL_ROM 1 /* 4-tap cubic Hermite, as smuthi's CUDA 'cubic' kernel */
#define SMH_INTERP_SPLINE 2      /* not-a-knot cubic spline, as smuthi's CPU 'cubic' (scipy interp1d) */

#define SMH_FP32 0 /* FP32 complex arithmetic, FP32 accumulation */
#define SMH_FP64 1 /* FP64 complex arithmetic */
#define SMH_FP16 2 /* FP16 operands, FP32 accumulation (8-RHS matvecs of lookup operators; a few 1e-4 relative error) */

#define SMH_SOLVE_FP64 0  /* restarted GMRES, FP64 matvec, FP64 Arnoldi */
#define SMH_SOLVE_MIXED 1 /* FP64 iterative refinement around FP32-matvec GMRES */
#define SMH_SOLVE_MIXED16 2 /* FP64 iterative refinement around FP16-matvec GMRES (batched lookup solves) */
#define SMH_SOLVE_LU 3      /* LU factorization of the explicit master matrix (solver_type 'LU') */

typedef struct smh_ctx smh_ctx;
typedef struct smh_op smh_op;

/* ---- runtime ----------------------------------------------------------
 * Replaces smuthi.utility.cuda.enable_gpu's device probe (utility/cuda.py:10-68). */
int smh_init(int device, smh_ctx **out);
int smh_destroy(smh_ctx *ctx);
const char *smh_last_error(const smh_ctx *ctx);
int smh_device_info(smh_ctx *ctx, char *name, int name_len, int *num_cu);

/* Device memory on the context's GPU for callers that keep right-hand sides
 * and solutions in HBM (smh_solve accepts device pointers): allocate, free
 * (after the context's stream drains) and copy in any direction (host or
 * device pointers on either side, synchronous on the context's stream). No
 * counterpart in smuthi, whose vectors live in host memory. */
int smh_device_alloc(smh_ctx *ctx, int64_t bytes, void **out);
int smh_device_free(smh_ctx *ctx, void *p);
int smh_device_copy(smh_ctx *ctx, void *dst, const void *src, int64_t bytes);

/* ---- radial lookup coupling operator -----------------------------------
 * Replaces CouplingMatrixRadialLookupCUDA.__init__ (linearsystem/linear_system.py:826-886):
 * particle coordinates and multipole cut-offs, table grid
 * rho_j = rho_min + j * resolution (j < len_rho), interpolation kind.
 * xy: [n_part][2] doubles (all particles share one z). The call bins every
 * particle pair by distance for the matvec. */
int smh_radial_create(smh_ctx *ctx, int64_t n_part, const double *xy, const int *lmax, const int *mmax,
                      int l_max, int m_max, double rho_min, double resolution, int64_t len_rho, int interp,
                      smh_op **out);

/* Volume lookup operator (CouplingMatrixVolumeLookupCUDA, linear_system.py:688-783):
 * xyz: [n_part][3] (all particles in one layer); complex64 tables over
 * (rho, z1+z2) ("plus") and (rho, z1-z2) ("minus") on the grids
 * rho_min + i h (i < len_rho), sz_min + j h and dz_min + j h (j < len_z),
 * interpolated with interp = SMH_INTERP_LINEAR (bilinear: the reference's GPU
 * 'linear' and CPU RectBivariateSpline kx=ky=1), SMH_INTERP_SPLINE (tensor
 * not-a-knot spline: CPU RectBivariateSpline kx=ky=3, linear_system.py:652-663;
 * needs 7 nodes per axis) or SMH_INTERP_CATMULL_ROM (tensor Catmull-Rom: the
 * reference's CUDA 'cubic' kernel, linear_system_cuda.py:149-256). The call
 * bins every particle pair by table cell for the matvec. */
int smh_volume_create(smh_ctx *ctx, int64_t n_part, const double *xyz, const int *lmax, const int *mmax,
                      int l_max, int m_max, double rho_min, double sz_min, double dz_min, double resolution,
                      int64_t len_rho, int64_t len_z, int interp, smh_op **out);

/* Several particle ensembles in one operator (no counterpart in smuthi,
 * which runs one Simulation per ensemble, simulation.py:363-414, and builds
 * its table each time, linear_system.py:183-225): ensemble e holds particles
 * [ens_off[e], ens_off[e+1]) of xy / xyz (ens_off[0] = 0, n_ens + 1 offsets,
 * every ensemble non-empty). Pairs couple particles of one ensemble only; all
 * ensembles share the table (sized by the caller for the largest), every
 * matvec launch covers all of them, and smh_solve treats each (ensemble,
 * right-hand side) as a system of its own with its own convergence. System
 * vectors concatenate the ensembles' vectors. Other arguments as the
 * single-ensemble constructors. */
int smh_radial_create_ensembles(smh_ctx *ctx, int n_ens, const int64_t *ens_off, const double *xy, const int *lmax,
                                const int *mmax, int l_max, int m_max, double rho_min, double resolution,
                                int64_t len_rho, int interp, smh_op **out);
int smh_volume_create_ensembles(smh_ctx *ctx, int n_ens, const int64_t *ens_off, const double *xyz, const int *lmax,
                                const int *mmax, int l_max, int m_max, double rho_min, double sz_min, double dz_min,
                                double resolution, int64_t len_rho, int64_t len_z, int interp, smh_op **out);

/* Right block-sparse preconditioner for smh_solve (no counterpart in smuthi,
 * whose GMRES is unpreconditioned, linear_system.py:271-286):
 *   (P v)_i = sum_{k < k1} blocks[i][k] v_{nbr[i][k]}
 * blocks: complex128 [n_part][k1][B][B], B = blocksize(l_max, m_max) of the
 * operator, on the padded multipole index (entries of multipoles a particle
 * does not have act on zeros); nbr: [n_part][k1] emitter particles (-1: empty
 * slot); k1 = 1 with nbr[i][0] = i is a block-diagonal preconditioner.
 * smh_solve then runs GMRES on M P y = b and returns x = P y, with the same
 * FP64 stopping rule on ||b - M x||. blocks = NULL removes it. Needs the
 * batched kernels (radial or volume operators). */
int smh_op_set_preconditioner(smh_op *op, int k1, const int *nbr, const double *blocks);

/* ---- learned near-field preconditioner (no counterpart in smuthi) ---------
 * Near-field blocks A_ij = diag(t_i) W_ij of every particle i and the
 * emitters nbr[i][k] (k < k1, nbr[i][0] = i, -1: empty; the K nearest
 * neighbours in the caller's use), W_ij the block the operator's lookup
 * applies to that pair (the taps and weights of its FP64 matvec, phases
 * e^{i (m2 - m1) phi} included), t_diag: complex128 [n_unk] T-matrix
 * diagonal in smuthi's layout. Kept on the operator; the download returns
 * complex128 [n_part][k1][B][B] on the padded index. */
int smh_op_near_blocks(smh_op *op, int k1, const int *nbr, const double *t_diag);
int smh_op_download_near_blocks(smh_op *op, double *out);
/* Preconditioner from the near blocks and per-channel scales:
 *   P[i][k] = [k == 0] 1 + diag(s[i][k][ch(n1)]) A[i][k],
 * ch(n) = tau l_max + l - 1 the (tau, l) channel of the receiving multipole;
 * s: complex64 as float [n_part][k1][2 l_max][2]. s = 1 everywhere is the
 * first-order near-field inverse of M = 1 - T W. NULL removes it. */
int smh_op_set_near_scaling(smh_op *op, const float *s);

/* The network that predicts the scales: x -> relu(x W1 + b1) -> relu(. W2 +
 * b2) -> . W3 + b3 (s = 1 + output), parameters FP32 (W1 [n_in][n_hidden],
 * b1, W2 [n_hidden][n_hidden], b2, W3 [n_hidden][n_out], b3, contiguous),
 * rounded to BF16 where they enter the products (as are the layer inputs and
 * back-propagated gradients), FP32 accumulation on the BF16 MFMA. The output
 * layer starts at zero. Rows of x and y are particles. */
typedef struct smh_mlp smh_mlp;
int smh_mlp_create(smh_ctx *ctx, int n_in, int n_hidden, int n_out, uint64_t seed, smh_mlp **out);
int smh_mlp_destroy(smh_mlp *mlp);
int64_t smh_mlp_num_params(const smh_mlp *mlp);
int smh_mlp_get_params(smh_mlp *mlp, float *out);
int smh_mlp_set_params(smh_mlp *mlp, const float *params);
/* y [rows][n_out] = network(x [rows][n_in]) */
int smh_mlp_forward(smh_mlp *mlp, int64_t rows, const float *x, float *y);
/* grad [n_params] = d(sum dy . y) / d params at x */
int smh_mlp_backward(smh_mlp *mlp, int64_t rows, const float *x, const float *dy, float *grad);
/* Add one ensemble's particles to the training set: features feat
 * [n_part][n_in] (float), right-hand sides rhs and the solver's solutions sol
 * = M^{-1} rhs, complex128 [nrhs][n_unk]; the operator's near blocks must be
 * computed. The loss is L = sum_{i,r} |(P rhs_r)_i - sol_{i,r}|^2 /
 * sum |sol|^2 + (w / rows) sum |y|^2 with P from the network's scales
 * s = 1 + y; the second term (weight w, smh_mlp_set_prior_weight, default 0)
 * pulls the scales towards the first-order structure s = 1. */
int smh_mlp_add_data(smh_mlp *mlp, smh_op *op, const float *feat, int64_t nrhs, const double *rhs, const double *sol);
/* L and dL/dy over the training set at outputs y [rows][n_out] (NULL: the
 * network's outputs); dy may be NULL. */
int smh_mlp_data_loss(smh_mlp *mlp, const float *y, float *dy, double *loss);
int smh_mlp_set_prior_weight(smh_mlp *mlp, double w);
/* steps of full-batch Adam (beta 0.9 / 0.999) at learning rate lr on the
 * training set; losses (may be NULL) receives L before each step. */
int smh_mlp_train(smh_mlp *mlp, int steps, double lr, double *losses);

/* Build both volume tables on the GPU (volumetric_coupling_lookup_table,
 * prepare_lookup.py:16-250, and its helper kernel prepare_lookup_cuda.py:6-52),
 * FP64 arithmetic, stored complex64:
 *   plus[r][j]  = c64( sum_k J_|dm|(kappa_k rho_r) (C[k][0] e^{i kz_k (sz_j - 2 z_s)} + C[k][1] e^{-i kz_k (sz_j - 2 z_s)}) )
 *   minus[r][j] = c64(direct) + c64( sum_k J_|dm|(kappa_k rho_r) (C[k][2] e^{i kz_k dz_j} + C[k][3] e^{-i kz_k dz_j}) )
 *   direct = sum_ld direct_coef[n1][n2][ld] h_ld(k d) P_ld^|dm|(dz/d), d = sqrt(rho^2 + dz^2),
 *            zero where rho < mask_radius.
 * direct_coef: complex128 [B][B][2 l_max + 1]; kappa, kz: complex128 [n_kappa];
 * layer_coef (C): complex128 [n_kappa][4][B][B] with contour weights and
 * layer response folded in, or n_kappa = 0 for a degenerate layer system. */
int smh_volume_build_table(smh_op *op, double k_re, double k_im, double mask_radius, double z_s,
                           const double *direct_coef, int64_t n_kappa, const double *kappa, const double *kz,
                           const double *layer_coef);

int smh_op_destroy(smh_op *op);

/* Upload a host table (prepare_lookup.py:253-464 / :16-250 output):
 * radial: complex128 [len_rho][B][B]; volume: two complex128 [len_rho][len_z][B][B]
 * (plus, minus), stored complex64. B = blocksize(l_max, m_max). */
int smh_op_upload_table(smh_op *op, const double *table, const double *table_minus);

/* Explicit coupling operator (CouplingMatrixExplicit.__init__, linearsystem/linear_system.py:453-486):
 * the dense coupling matrix of every particle pair, built on the GPU in FP64
 * at the pairs' exact positions (no lookup). Block of receiver i, emitter j:
 *   W_ij[n1][n2] = e^{i (m2 - m1) phi_ij} (D_ij + R_ij)
 *   D_ij = sum_ld direct_coef[n1][n2][ld] h_ld(k_{L_i} d_ij) P_ld^|m1-m2|(dz_ij / d_ij)
 *          (direct_coupling_block, direct_coupling.py:18-105), for i != j in one
 *          layer, or every i != j when direct_all is set (a layer system of one medium);
 *   R_ij = sum_k J_|m1-m2|(kappa_k rho_ij) sum_{a,b in 0,1} layer_coef[L_i][L_j][k][2a+b][n1][n2]
 *          e^{i s_a kz[L_i][k] z_rel_i} e^{-i s_b kz[L_j][k] z_rel_j},   s_0 = +1, s_1 = -1
 *          (layer_mediated_coupling_block, layer_mediated_coupling.py:215-340; the
 *          Jacobian, layer response and quadrature weights folded into layer_coef).
 * layer: [n_part] index of each particle's layer among n_layers occupied layers;
 * k_layer: complex128 [n_layers]; z_rel: [n_part] height above the layer's
 * reference; direct_coef: complex128 [B][B][2 l_max + 1]; kappa: complex128
 * [n_kappa]; kz: complex128 [n_layers][n_kappa]; layer_coef: complex128
 * [n_layers][n_layers][n_kappa][4][B][B] (n_kappa = 0 and NULL pointers: no
 * layer-mediated part). l_max <= 8. The matrix needs 24 bytes per
 * entry of device memory ((n_part B)^2 entries); build_ms (may be NULL)
 * receives the build kernels' time. smh_op_matvec and smh_solve apply it as a
 * dense matrix; smh_op_download_table returns it on the padded index
 * i * B + n (B = blocksize(l_max, m_max)), complex128 [n_part B][n_part B]. */
int smh_explicit_create(smh_ctx *ctx, int64_t n_part, const double *xyz, const int *lmax, const int *mmax, int l_max,
                        int m_max, const int *layer, int n_layers, const double *k_layer, const double *z_rel,
                        int direct_all, const double *direct_coef, int64_t n_kappa, const double *kappa,
                        const double *kz, const double *layer_coef, double *build_ms, smh_op **out);

/* Explicit blocks between two particle sets, returned to the host: the
 * same W_ij as smh_explicit_create for receivers i = 0 .. n_recv - 1 and
 * emitters j = n_recv .. n_recv + n_emit - 1 of one particle array (xyz,
 * layer, z_rel of n_recv + n_emit entries), all padded to (l_max, m_max).
 * With x == NULL, out receives W: complex128 [n_recv B][n_emit B]; else
 * out = W x (x: complex128 [n_emit B], out: [n_recv B]) formed on the
 * device. Replaces the per-source block calls of
 * DipoleSource.spherical_wave_expansion (initial_field.py:390-411):
 * (direct_coupling_block + layer_mediated_coupling_block)(particle, dipole)
 * applied to the dipole's outgoing SWE. */
int smh_explicit_blocks(smh_ctx *ctx, int64_t n_recv, int64_t n_emit, const double *xyz, int l_max, int m_max,
                        const int *layer, int n_layers, const double *k_layer, const double *z_rel, int direct_all,
                        const double *direct_coef, int64_t n_kappa, const double *kappa, const double *kz,
                        const double *layer_coef, const double *x, double *out, double *build_ms);

/* Build the radial table on the GPU in FP64 (replaces radial_coupling_lookup_table,
 * prepare_lookup.py:253-464, and its helper kernel, prepare_lookup_cuda.py:55-95):
 *   L[j,n1,n2] = c64(direct[j,n1,n2]) + sum_k J_|m1-m2|(kappa_k rho_j) * lm_coef[k,n1,n2]
 *   direct[j,n1,n2] = sum_ld direct_coef[n1,n2,ld] * h_ld(k_medium rho_j)   if rho_j >= mask_radius, else 0
 * direct_coef: complex128 [B][B][2 l_max + 1]; kappa: complex128 [n_kappa];
 * lm_coef: complex128 [n_kappa][B][B] (trapezoid weights, Jacobian and the
 * layer response already folded in) or NULL for a degenerate layer system. */
int smh_radial_build_table(smh_op *op, double k_re, double k_im, double mask_radius, const double *direct_coef,
                           int64_t n_kappa, const double *kappa, const double *lm_coef);

/* Copy the FP64 table back to the host (the .lookup_table attribute). */
int smh_op_download_table(smh_op *op, double *table, double *table_minus);

/* y = W x for nrhs vectors. x, y: host complex128 [nrhs][n_unk] in smuthi's
 * system-vector layout (replaces the CUDA matvec closure, linear_system.py:888-898).
 * Radial operators apply nrhs > 1 in blocks of 8 with one batched launch each. */
int smh_op_matvec(smh_op *op, const double *x, double *y, int64_t nrhs, int precision);

/* y = W x like smh_op_matvec for one vector, allocating nothing: no
 * counterpart in smuthi; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). x, y: DEVICE pointers to complex128 [n_unk] in smuthi's
 * system-vector layout, owned by the caller. Radial lookup operators, FP32 or
 * FP64. Its staging (padded vectors, input bound, per-work-item partials)
 * lives in a fixed 64 MiB device array per GPU, held until the process ends;
 * operators whose padded vectors and one work item's partial do not fit it
 * are rejected with SMH_ERR_ARG (about 1 million padded unknowns in FP64).
 * Every call on a GPU shares that array, so calls that overlap on one GPU
 * (other threads, other contexts of the device) race; the call is
 * synchronous on the context's stream. Same result as smh_op_matvec (the
 * same kernels; the work items run in ranges that fit the array and are
 * summed in the same order). */
int smh_op_matvec_noalloc(smh_op *op, const double *x, double *y, int precision);

/* Time `iters` matvecs on device-resident synthetic vectors; returns the
 * average time of the coupling kernel (HIP events on the operator's stream)
 * and of the whole matvec in milliseconds. */
int smh_op_bench_matvec(smh_op *op, int precision, int iters, double *kernel_ms, double *total_ms);

/* The same for the batched kernel: one launch applies W to 8 right-hand
 * sides (smh_op_matvec uses it for nrhs > 1 on lookup operators). No
 * counterpart in smuthi, which has no multi-RHS path. */
int smh_op_bench_matvec_batch(smh_op *op, int precision, int iters, double *kernel_ms, double *total_ms);

/* Launch-shape autotuner of a lookup operator's matvec (radial or volume,
 * table set). For each of n_cand candidates c it regroups the pair bins into
 * work items (about items_per_cu[c] workgroups per compute unit) launched in
 * the order launch_order[c]: 0 chunk order, 1 sorted by table row and dealt
 * to the XCDs in contiguous shares, n in [2, 4096] items also cut at the
 * boundaries of n equal bands of table rows and launched band after band;
 * times `iters` matvecs of the given shape (batch != 0: the 8-RHS kernel in
 * SMH_FP16 / FP32 / FP64; else the single-vector kernel in FP32 / FP64; the
 * kernels and the reduction of their partial sums) into matvec_ms[c] (ms per
 * matvec), and keeps the fastest grouping (*best = its index). The FP16
 * and FP32 8-RHS kernels share one grouping. Results stay within each
 * kernel's accuracy (the grouping changes summation order and fixed-point
 * scales only). Drops the solver's captured graphs. No counterpart in
 * smuthi, whose CUDA kernels launch one thread per matrix row
 * (linear_system_cuda.py:26-405, cuda_blocksize). */
int smh_op_autotune(smh_op *op, int batch, int precision, int iters, int n_cand, const int *items_per_cu,
                    const int *launch_order, double *matvec_ms, int *best);

/* Interpolation of the FP16 operator used by the mixed16 inner solves (and
 * smh_op_matvec with SMH_FP16): *terms = 2 when it interpolates linearly
 * between the table nodes (chosen when that error, measured on the table,
 * stays within 2^-10 of each used bin's scale), else the operator's own
 * number of Horner terms (0 before a table is set); *lin_ratio = the
 * measured largest error / scale (-1 when not measured). No counterpart in
 * smuthi, which has no reduced-precision operator. */
int smh_op_half_interp(smh_op *op, int *terms, double *lin_ratio);

/* Solve (1 - T W) b = T a (LinearSystem.solve, linear_system.py:254-344)
 * for nrhs right-hand sides.
 * t_diag: complex128 [n_unk] diagonal sphere T-matrix; rhs, x: complex128
 * [nrhs][n_unk]. t_diag, rhs and x may each be host or device memory (HIP's
 * unified addressing decides the copy direction), so callers that keep the
 * right-hand sides in HBM solve without PCIe traffic. SMH_SOLVE_FP64
 * (solver_type 'gmres') starts from x0 = rhs as smuthi does;
 * SMH_SOLVE_MIXED / MIXED16 refine from x0 = 0. They stop at
 * ||rhs - M x|| <= tol ||rhs|| measured in FP64 and return SMH_ERR_NOCONV
 * when maxiter Arnoldi steps do not get there. SMH_SOLVE_LU (solver_type
 * 'LU', linear_system.py:259-268, explicit operators only) factors M on the
 * device once per T (kept on the operator) and solves directly; tol, restart
 * and maxiter are not used. stats (out, 8 doubles): Arnoldi steps, FP64
 * matvecs, FP32 matvecs, worst relative FP64 residual, solve seconds,
 * refinement sweeps, right-hand sides per batched matvec (0 when the solve
 * was not batched), FP16 matvecs. */
int smh_solve(smh_op *op, const double *t_diag, const double *rhs, double *x, int64_t nrhs, double tol,
              int restart, int maxiter, int mode, double *stats);

/* ---- plane wave expansion fields ---------------------------------------
 * Replaces PlaneWaveExpansion.electric_field / magnetic_field on the GPU
 * (fields/expansions.py:635-782, kernels fields/expansions_cuda.py:5-421).
 * kind 0: E, 1: H*omega (smuthi multiplies by 1/omega on the host).
 * kp, kz: complex128 [n_k]; alpha: double [n_a]; g: complex128 [2][n_k][n_a];
 * xyz: [n_pts][3] positions relative to the expansion's reference point;
 * out: complex128 [3][n_pts]. Nested trapezoid rule, alpha inner, kappa outer. */
int smh_pwe_field(smh_ctx *ctx, int kind, int64_t n_k, const double *kp, const double *kz, int64_t n_a,
                  const double *alpha, const double *g, double k_re, double k_im, int64_t n_pts,
                  const double *xyz, double *out, int precision, double *kernel_ms);

/* Plane wave expansion of the scattered field of particles in one layer
 * (the particle sum of swe_to_pwe_conversion, fields/transformations.py:133-226,
 * as used by scattered_field_pwe, postprocessing/scattered_field.py:80-150):
 *   g_dir(pol, kappa_k, alpha_a) = sum_p e^{-i kappa (x_p cos a + y_p sin a)} e^{+-i kz (z_ref - z_p)}
 *                                  sum_m e^{i m a} sum_{tau,l} coef[p][n(tau,l,m)] btab[k][dir][pol][n]
 * xyz: [n_part][3]; coef: complex128 [n_part][B] in the (l_max, m_max) layout
 * (l_max <= 64; more than 9 orders m run as several launches); kp, kz: complex128 [n_k] (kz upgoing); btab: complex128
 * [n_k][2 dir (up, down)][2 pol][B], the transformation coefficients
 * B_{n,pol}(+-kz/k, kappa/k) / (2 pi kz k); alpha: [n_a];
 * g: complex128 [2 dir][2 pol][n_k][n_a]. */
int smh_pwe_assemble(smh_ctx *ctx, int64_t n_part, const double *xyz, int l_max, int m_max, const double *coef,
                     double z_ref, int64_t n_k, const double *kp, const double *kz, const double *btab, int64_t n_a,
                     const double *alpha, double *g, int precision, double *kernel_ms);

/* ---- spherical wave expansion fields ----------------------------------
 * Direct field of the particles' outgoing spherical wave expansions, summed
 * at points: replaces SphericalWaveExpansion.electric_field / magnetic_field
 * (fields/expansions.py:348-424, SVWFs of fields/vector_wave_functions.py:64-155)
 * over the scattered-field SWEs that PiecewiseFieldExpansion adds
 * (postprocessing/scattered_field.py:72-75). pos: [n_part][3]; radius:
 * [n_part]; zlim: [n_part][2] (z range of the particle's layer, +-inf
 * allowed); coef: [n_part][blocksize(l_max, m_max)] complex in (tau, l, m)
 * order (zeros for multipoles a particle does not have). A particle's SWE
 * counts at points with zlim[0] <= z < zlim[1] strictly outside its sphere.
 * field 0: E = sum b N; 1: H = -i k / omega sum b_(1-tau) N_tau. out:
 * [3][n_pts] complex. 1 <= l_max <= 20, 0 <= m_max <= l_max (unrolled
 * kernels up to l_max 4, a runtime-order kernel above). kernel_ms (may be
 * NULL): device time. */
int smh_swe_field(smh_ctx *ctx, int64_t n_part, const double *pos, const double *radius, const double *zlim,
                  int l_max, int m_max, const double *coef, double k_re, double k_im, int64_t n_pts,
                  const double *pts, int field, double omega, int precision, double *out, double *kernel_ms);

/* Plane wave expansion -> regular SWE at n_pts reference points
 * (pwe_to_swe_conversion, fields/transformations.py:77-130, continuous
 * branch), FP64:
 *   out[p][n] = sum_k w_k e^{i kz_k dz} sum_pol bdag[k][pol][n]
 *               sum_j w_a[j] e^{-i m_n alpha_j} coef[pol][k][j] e^{i kappa_k (cos alpha_j dx + sin alpha_j dy)}
 * with (dx, dy, dz) = pts[p] (the points relative to the expansion's
 * reference point), kz signed by the expansion's kind, w_k = 4 kappa_k times
 * the kappa trapezoid weight, w_a the alpha trapezoid weights, bdag the
 * B-dagger coefficients of (l_max, m_max) at (kz / k, kappa / k).
 * kp, kz, w_k, bdag, coef, out: complex128; out [n_pts][blocksize].
 * 1 <= l_max <= 20, 0 <= m_max <= l_max (unrolled kernels up to l_max 4, a
 * runtime-order kernel above, five orders m per launch). */
int smh_pwe_to_swe(smh_ctx *ctx, int64_t n_k, const double *kp, const double *kz, const double *w_k, int64_t n_a,
                   const double *alpha, const double *w_a, const double *coef, const double *bdag, int l_max,
                   int m_max, int64_t n_pts, const double *pts, double *out, double *kernel_ms);

/* ---- measurement helper (no counterpart in smuthi) ----------------------
 * Copy `bytes` (a multiple of 16) between two fresh device buffers with one
 * plain streaming kernel and return its time; used to calibrate the HBM
 * traffic counters against a kernel whose traffic is known (2 * bytes). */
int smh_bench_copy(smh_ctx *ctx, int64_t bytes, double *ms);

/* ---- host evaluation of the device special functions -------------------
 * Runs the same code the table kernels use, on the CPU, so it can be checked
 * without a device: out[j][n] = J_n(z_j) (complex128 [count][nmax+1]) and
 * out[j][n] = h_n^(1)(z_j) spherical Hankel. z: complex128 [count]. */
int smh_host_bessel_j(int nmax, int64_t count, const double *z, double *out);
int smh_host_spherical_hankel(int nmax, int64_t count, const double *z, double *out);

#ifdef __cplusplus
}
#endif

#endif /* SMH_H */
