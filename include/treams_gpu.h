/*
 * treams_gpu.h -- C ABI of the MI355X (gfx950) cluster-coupling library.
 *
 * The library replaces the CPU hot path of treams-rs for clusters of homogeneous
 * spheres in vacuum: the pairwise spherical-wave translation blocks, the cluster
 * operator (I - T C), its solves (GMRES, LU) and adjoint, and the cross sections.
 * Reference paths below are relative to crates/treams-core/src/ of treams-rs unless
 * they start with python/.
 *
 * Conventions (identical to the reference):
 *  - modes per sphere in sw::modes order (sw/mod.rs:114-125): l = 1..lmax, m = -l..l,
 *    pol 1 (positive helicity) before pol 0; M = 2 lmax (lmax + 2) modes per sphere,
 *    n = N M unknowns, particle-major (global index p * M + mode);
 *  - complex numbers are interleaved (re, im) doubles (tg_c128), matrices column-major;
 *  - translations use displacement = destination - source (sw/mod.rs:19-22);
 *  - C has exactly zero self blocks (cluster/spheres.rs:186-220).
 *
 * Memory: vectors and matrices passed to compute calls are DEVICE pointers owned by
 * the caller (e.g. torch tensors); geometry passed to tg_cluster_create is host
 * memory, copied into the handle.  Every call runs on the handle's HIP stream and
 * returns after its work is complete on that stream.
 *
 * Errors: every function returns a tg_status; tg_last_error() gives the message of
 * the calling thread's last failure (lib.rs:379-411 variants).  Inputs are validated
 * before any device work (cluster/spheres.rs:99-133, linalg/gmres.rs:64-78).
 */
#ifndef TREAMS_GPU_H
#define TREAMS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } tg_c128;

typedef enum {
    TG_OK = 0,
    TG_INVALID_INPUT = 1,     /* Error::InvalidInput */
    TG_SPECIAL_FUNCTION = 2,  /* Error::SpecialFunction */
    TG_NON_FINITE = 3,        /* Error::NonFinite */
    TG_NOT_CONVERGED = 4,     /* Error::NotConverged */
    TG_SINGULAR = 5,          /* Error::Singular */
    TG_OUT_OF_MEMORY = 6,     /* Error::OutOfMemory */
    TG_DEVICE_ERROR = 7       /* HIP runtime failure (no reference counterpart) */
} tg_status;

/* Storage precision of the assembled coupling matrix. */
typedef enum { TG_STORE_NONE = 0, TG_STORE_C64 = 1, TG_STORE_C128 = 2 } tg_storage;

/* Operator used by tg_cluster_apply / the solvers. */
typedef enum {
    TG_OP_MATRIX_FREE_F64 = 0, /* FP64 matrix-free (cluster/iterative.rs:187-221): pair-table GEMMs */
    TG_OP_STORED = 1,          /* stream the stored coupling (c64 or c128) */
    TG_OP_MATRIX_FREE_C64 = 2, /* matrix-free from c64 pair tables on FP32 MFMA (no reference counterpart) */
    TG_OP_MATRIX_FREE_C64_HI = 3 /* the same with split-FP16 tables multiplying the hi halves only: the relaxed
                                    applies of tg_cluster_set_inner_relaxation (no reference counterpart) */
} tg_operator;

/* Preconditioner of the GMRES solves (right preconditioning; no reference counterpart). */
typedef enum {
    TG_PC_NONE = 0,            /* the reference's unpreconditioned GMRES */
    TG_PC_BLOCK_JACOBI = 1,    /* exact (LU) inverses of spatial sphere groups */
    TG_PC_LEARNED = 2          /* near-field Neumann correction with MLP-weighted pair blocks */
} tg_preconditioner;

/* linalg/gmres.rs:22-43 plus the precision and preconditioner choices of this library. */
typedef struct {
    double rtol;            /* default 1e-10 */
    double atol;            /* default 0 */
    int restart;            /* default 30 */
    int max_iterations;     /* default 300 */
    int mixed_precision;    /* 1: inner solves on the c64 stored coupling, 2: inner solves on the
                               matrix-free c64 operator (FP32 MFMA); both refined by FP64 residuals */
    int preconditioner;     /* tg_preconditioner */
} tg_gmres_options;

/* linalg/gmres.rs:81-93. */
typedef struct {
    int iterations;
    int refinements;        /* outer FP64 refinement steps (mixed precision), else 0 */
    double residual_norm;   /* true residual |b - A x| recomputed in FP64 */
    double rhs_norm;
} tg_convergence;

typedef struct tg_cluster tg_cluster;

/* Message of the calling thread's last failure ("" if none). */
const char *tg_last_error(void);

/* Number of modes per sphere, 2 lmax (lmax + 2); replaces sw::modes (sw/mod.rs:114). */
int tg_modes_per_sphere(int lmax);

/*
 * Host-side translation plan between all modes up to lmax in the helicity basis;
 * replaces TranslationPlan::between (sw/plan.rs:71-170).  Fills, for each of the
 * H*H = (lmax (lmax+2))^2 (source (l,m), destination (lambda,mu)) pairs in
 * column-major order, the same-helicity coupling as
 *   C_pol = sum_same w T[idx] + (2 pol - 1) i sum_cross w T[idx],
 * T[p^2 + p + q] = z_p(kd) P_p^q(cos theta) e^{i q phi}.
 * starts has H*H+1 entries, nsame H*H; with w == NULL returns the term count only.
 */
int tg_plan_terms(int lmax, int32_t *starts, uint8_t *nsame, double *w, uint16_t *hidx, int64_t *count);

/* Wigner 3j symbol used by the plan (host; special/wigner.rs:263). */
double tg_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);

/*
 * Rotation matrix of z-y-z Euler angles (phi, theta, psi) in one spherical basis of
 * degrees 1 .. lmax (2 lmax (lmax + 2) modes, reference order): row (l, m', pol),
 * column (l, m, pol) holds D^l_(m' m) = e^(-i m' phi) d^l_(m' m)(theta) e^(-i m psi),
 * zero between different degrees or polarisations.  out (host) is M x M, row-major.
 * A T-matrix rotates as R T R^H.  Replaces sw_rotate / diff.rotation
 * (rotation.rs:458-488, special/wigner.rs:411-449, python/treams_rs/diff.py:1743) for
 * a single-origin basis (TMatrix.rotate, python/treams_rs/_framework_tmatrix.py:356).
 */
int tg_rotation_matrix(int lmax, double phi, double theta, double psi, tg_c128 *out_host);

/*
 * Create a sphere cluster: homogeneous nonmagnetic spheres in vacuum.  Replaces
 * IterativeSphereCluster::new (cluster/iterative.rs:53-86) and the shared parts of
 * sphere_cluster / sphere_cluster_factor (cluster/spheres.rs:56-184): validation,
 * Mie T-matrix blocks (coeffs/mie.rs:209; tmatrix/sphere.rs:43) and the plan.
 * stream: a hipStream_t (NULL = the null stream).
 */
int tg_cluster_create(int lmax, double k0, int n_spheres, const double *radii, const tg_c128 *epsilon,
                      const double *positions, void *stream, tg_cluster **out);
/*
 * Create a particle cluster: arbitrary local T-matrices at distinct positions coupled
 * through the singular expansion, replacing particle_cluster with spherical bases in
 * the helicity basis (cluster/particles.rs:65-123; python/treams_rs/diff.py:1473).
 * tmatrices: device buffer of n_particles column-major M x M blocks, M = 2 lmax
 * (lmax + 2), in the reference mode order (tg_modes_per_sphere).  Particles with
 * smaller bases or mode subsets are embedded with zero rows and columns (their modes
 * solve to zero; the other modes solve the reduced system).  One real wavenumber k0
 * for both helicities.  Solves, factors, the interacting and global T-matrices and
 * cross sections work as for spheres; the preconditioners and the pullbacks to radii
 * and permittivities need spheres and return TG_INVALID_INPUT.
 */
int tg_cluster_create_particles(int lmax, double k0, int n_particles, const double *positions,
                                const tg_c128 *tmatrices, void *stream, tg_cluster **out);
int tg_cluster_destroy(tg_cluster *cluster);
/* Unknowns n = N M. */
int64_t tg_cluster_dimension(const tg_cluster *cluster);

/*
 * Per-sphere, per-degree 2x2 Mie blocks in mode order (cluster/iterative.rs:99-103,
 * coeffs/mie.rs:123-130) copied to host: out has N * lmax * 4 entries, row-major 2x2.
 */
int tg_cluster_mie_blocks(tg_cluster *cluster, tg_c128 *out_host);

/*
 * Dense coupling C (n x n, column-major, c128) in the reference layout, into a
 * caller device buffer: replaces coupling_matrix (cluster/spheres.rs:188-220).
 * With regular != 0 writes the regular overlap R instead (sw/expansion.rs:116-130
 * with Radial::Regular; self blocks are the identity).
 */
int tg_pair_matrix(tg_cluster *cluster, int regular, tg_c128 *out_dev);

/*
 * Translation blocks of a list of displacements (device double[3 * count]), each
 * an M x M column-major c128 block: replaces TranslationPlan::evaluate (sw/plan.rs:282).
 */
int tg_translation_blocks(int lmax, double k, int64_t count, const double *displacements_dev, int regular,
                          tg_c128 *out_dev, void *stream);

/*
 * Directional derivatives of tg_translation_blocks: for each displacement d
 * (device double[3 count]) with tangent dd (device double[3 count]) and a wavenumber
 * tangent dk, the M x M block sum_a dB/dd_a dd_a + dB/dk dk (column-major c128).
 * Replaces TranslationPlan::pushforward / tangent_table (sw/plan.rs:552-595); the
 * wavenumber and its tangent are real here.
 */
int tg_translation_blocks_tangent(int lmax, double k, int64_t count, const double *displacements_dev,
                                  const double *tangent_displacements_dev, double tangent_k, int regular,
                                  tg_c128 *out_dev, void *stream);

/*
 * tg_translation_blocks without any device allocation: the same kernel and output,
 * with the plan staged in a fixed 4 MiB device array of the library (every lmax up
 * to 12 fits; a plan that would not is rejected with TG_INVALID_INPUT) instead of a
 * hipMalloc'ed buffer.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Every call on one GPU shares that GPU's array,
 * so calls that overlap on one GPU race; the call returns after its work is done.
 */
int tg_translation_blocks_noalloc(int lmax, double k, int64_t count, const double *displacements_dev, int regular,
                                  tg_c128 *out_dev, void *stream);

/*
 * tg_translation_blocks between bases of different degree: destination degrees up to
 * lmax_dst, source degrees up to lmax_src (lmax_dst + lmax_src <= 24); each block is
 * (2 Hd) x (2 Hs) column-major, destination inner, Hd = lmax_dst (lmax_dst + 2).
 * Replaces TranslationPlan::between on two spherical bases with TranslationPlan::evaluate
 * (sw/plan.rs:71-170,282-289), as sw::expansion assembles them (sw/expansion.rs:116).
 */
int tg_translation_blocks_rect(int lmax_dst, int lmax_src, double k, int64_t count, const double *displacements_dev,
                               int regular, tg_c128 *out_dev, void *stream);

/*
 * Assemble and keep the coupling on the device in the helicity-split layout
 * (two N H x N H column-major matrices, H = M / 2), in c64 or c128.  No reference
 * counterpart for the storage precision; the values are those of coupling_matrix.
 */
int tg_cluster_assemble(tg_cluster *cluster, int storage);

/*
 * Free the stored coupling, the pair tables of the matrix-free operators and the
 * preconditioners (including the learned one's near pairs; an attached network
 * stays attached); the next call that needs them rebuilds them.  No reference
 * counterpart (device memory management).
 */
int tg_cluster_release(tg_cluster *cluster);

/*
 * Mark the stored coupling, the pair tables and the preconditioners stale, keeping
 * their device memory: the next call that needs them recomputes them in place
 * (tg_cluster_release frees the memory instead).  For repeated whole jobs on one
 * cluster without re-allocating a stored coupling of tens of GB each time.  No
 * reference counterpart (device memory management).
 */
int tg_cluster_reset(tg_cluster *cluster);

/*
 * Largest pair-table set the matrix-free operators keep, in bytes (default 64 GiB).
 * One set holds N^2 (2 lmax + 1)^2 c128 values (2.8 GB at N = 1024, lmax = 6); the
 * singular coupling, its c64 copy and the regular overlap each keep one.  Clusters
 * whose set is larger recompute the translation coefficients of every pair in every
 * apply (slower, no table memory).  Frees the current tables.  No reference
 * counterpart (device memory policy).
 */
int tg_cluster_set_table_limit(tg_cluster *cluster, int64_t bytes);

/*
 * Arithmetic of the c64-precision coupling GEMMs of the inner mixed-precision solves
 * (TG_OP_MATRIX_FREE_C64 and mixed_precision 2): TG_INNER_SPLIT16 (default) holds
 * tables and sources as scaled FP16 hi + lo pairs and multiplies them on FP16 MFMA
 * (22 significant bits per value, FP32 accumulation); TG_INNER_FP32 rounds the tables
 * to c64 and multiplies on FP32 MFMA.  Frees the current c64-precision tables.  No
 * reference counterpart (device arithmetic choice; the FP64 residual certifies every
 * solve either way).
 */
typedef enum { TG_INNER_FP32 = 0, TG_INNER_SPLIT16 = 1 } tg_inner_gemm;
int tg_cluster_set_inner_gemm(tg_cluster *cluster, int kind);

/*
 * Relaxed inner applies of mixed-precision solves on split-FP16 tables (inexact
 * Krylov: Simoncini and Szyld 2003, Bouras and Fraysse 2005).  Before each Arnoldi
 * step of an inner solve the operator may err by tol / |r_j| relative (tol the inner
 * tolerance, |r_j| the step's residual estimate, the smallest bound over the columns
 * still extending); when that bound is at least `threshold` the step multiplies the
 * FP16 hi halves only (11 significant bits per value, a third of the MFMAs), otherwise
 * hi and lo.  threshold 0 switches relaxation off; the default is 1e-3 (the hi-only
 * coupling errs by about 3e-4 relative).  The FP64
 * residual of the refinement certifies every solve either way, so the threshold moves
 * iteration counts, never the accepted answer.  No reference counterpart (device
 * arithmetic choice).
 */
int tg_cluster_set_inner_relaxation(tg_cluster *cluster, double threshold);

/*
 * y = (I - T C) x, or (I - C^H T^H) x with adjoint != 0, for `cols` column-major
 * vectors of length n (device c128): replaces IterativeSphereCluster::apply
 * (cluster/iterative.rs:225-232).  op selects the matrix-free FP64 or stored path.
 */
int tg_cluster_apply(tg_cluster *cluster, const tg_c128 *x_dev, tg_c128 *y_dev, int cols, int adjoint, int op);

/* y = C x or C^H x (coupling only), cluster/iterative.rs:145-159. */
int tg_cluster_coupling(tg_cluster *cluster, const tg_c128 *x_dev, tg_c128 *y_dev, int cols, int adjoint, int op);

/* y = T x or T^H x with the Mie blocks, cluster/iterative.rs:99-136. */
int tg_cluster_local(tg_cluster *cluster, const tg_c128 *x_dev, tg_c128 *y_dev, int cols, int adjoint);

/*
 * Local regular coefficients of a unit plane wave of helicity pol (1 positive,
 * 0 negative) along the real direction dir (normalized): pw/expand.rs:23-66 times
 * the position phase of pw/field.rs:20-26.  out_dev has n entries.
 */
int tg_cluster_plane_wave(tg_cluster *cluster, const double *dir, int pol, tg_c128 *out_dev);

/*
 * Solve (I - T C) X = T B for `cols` incident columns (device c128, n x cols):
 * replaces IterativeSphereCluster::solve (cluster/iterative.rs:264-284).  With
 * options->mixed_precision the inner GMRES runs on the c64 stored coupling and an
 * FP64 matrix-free residual refines the answer; convergence is always certified
 * by the FP64 true residual.  conv_host has cols entries.  With adjoint != 0
 * solves (I - T C)^H X = B (cluster/iterative.rs:225-230, no T applied to B).
 */
int tg_cluster_solve(tg_cluster *cluster, const tg_c128 *incident_dev, int cols, const tg_gmres_options *options,
                     int adjoint, tg_c128 *x_dev, tg_convergence *conv_host);

/*
 * y = R x or R^H x with the regular overlap R applied matrix-free (self blocks I):
 * replaces `_radiation_overlap @ v` (python/treams_rs/_tmatrix.py:404-411).
 */
int tg_cluster_overlap(tg_cluster *cluster, const tg_c128 *x_dev, tg_c128 *y_dev, int cols, int adjoint);

/*
 * Cross sections of one illumination (python/treams_rs/_framework_tmatrix.py:231-289):
 * scattering = Re(x^H R (x / k^2)) 0.5 / flux, extinction = -Re(a^H x / k^2) 0.5 / flux,
 * with the regular overlap R applied matrix-free.
 */
int tg_cluster_cross_sections(tg_cluster *cluster, const tg_c128 *incident_dev, const tg_c128 *x_dev, double flux,
                              double *scattering, double *extinction);

/*
 * Gradients of a real loss L with dL = Re<G, dX> through X = (I - T C)^{-1} T B:
 * replaces IterativeResidual::pullback (cluster/iterative.rs:460-488) and
 * SphereClusterResidual::pullback (cluster/spheres.rs:321-378).  Outputs (host):
 * k0 gradient, radii[N], epsilon[N] (c128), positions[3N]; incident_grad_dev (n x cols)
 * receives T^H Y.  One adjoint solve per column with the given options.
 */
int tg_cluster_pullback(tg_cluster *cluster, const tg_c128 *incident_dev, const tg_c128 *x_dev,
                        const tg_c128 *cotangent_dev, int cols, const tg_gmres_options *options, double *grad_k0,
                        double *grad_radii, tg_c128 *grad_eps, double *grad_positions, tg_c128 *incident_grad_dev,
                        tg_convergence *conv_host);

/*
 * Directional derivative of a solved X = (I - T C)^{-1} T B (device, n x cols,
 * reference layout) along tangents of k0, the radii[N], permittivities[N] and
 * positions[3N] (host arrays; NULL means zero) and of the incident columns B (device
 * n x cols; NULL means zero): solves (I - T C) dX = dT (B + C X) + T (dB + dC X) with
 * GMRES under `options` (relative tolerance only, 1e-10 when rtol is 0).  dC X runs as
 * the pair-table GEMMs on the tables' tangents, so one FP64 table set must fit the
 * table limit.  Replaces IterativeResidual::pushforward (cluster/iterative.rs:362-447).
 * dx_dev: n x cols; conv_host: cols entries (the tangent solves' certificates).
 */
int tg_cluster_pushforward(tg_cluster *cluster, const tg_c128 *incident_dev, const tg_c128 *x_dev, int cols,
                           const tg_gmres_options *options, double tangent_k0, const double *tangent_radii,
                           const tg_c128 *tangent_epsilon, const double *tangent_positions,
                           const tg_c128 *tangent_incident_dev, tg_c128 *dx_dev, tg_convergence *conv_host);

/*
 * Position gradients of Re<Gbar, R> for the regular overlap with cotangent
 * Gbar = sum_c u_c v_c^H (device c128 n x cols each): the R-dependence of the
 * scattering cross section (sw/expansion.rs ExpansionResidual::pullback).
 */
int tg_cluster_overlap_pullback(tg_cluster *cluster, const tg_c128 *u_dev, const tg_c128 *v_dev, int cols,
                                double *grad_positions);

/* ---------------------------------------------------------------------------
 * Dense LU path.
 * ------------------------------------------------------------------------- */

typedef struct tg_factor tg_factor;

/*
 * Assemble A = I - T C and factor it on the device: replaces sphere_cluster_factor
 * (cluster/spheres.rs:85-94) with InteractionFactor::from_blocks / operator / Lu::new
 * (cluster/interaction.rs:155-205, linalg/mod.rs:398-428).  precision is TG_STORE_C64
 * (complex FP32 factors on MFMA, every solve refined in FP64) or TG_STORE_C128;
 * the factors take n^2 x 8 or 16 bytes.  Pivoting is partial within 64-row diagonal
 * blocks (every trailing update an MFMA GEMM); a zero pivot there refactors with
 * partial pivoting over whole columns, the reference's (faer) order, and a solve whose
 * FP64 refinement stalls or turns non-finite refactors once in c128 with whole-column
 * pivoting and repeats (tg_factor_info reports both).  TG_SINGULAR when whole-column
 * pivoting meets a zero pivot.  The cluster must outlive the factor.
 * An operator whose entries span more than a factor 10 while a diagonal entry is below
 * a tenth of the largest is equilibrated first (linalg/mod.rs:137-310: rows and
 * columns scaled by their largest magnitudes, rounded here to powers of two, then
 * balanced by power-of-two sweeps that even out the magnitude sums); its solves are
 * then certified by the residual of the scaled system.
 */
int tg_factor_create(tg_cluster *cluster, int precision, tg_factor **out);
int tg_factor_destroy(tg_factor *factor);

/*
 * The factor's state: precision (TG_STORE_C64 / TG_STORE_C128, after any fallback),
 * pivoting (0 within 64-row blocks, 1 over whole columns) and fallback (1 when a
 * stalled or singular solve refactored the operator in c128 with whole-column
 * pivoting).  No reference counterpart (reports this library's fallbacks).
 */
int tg_factor_info(const tg_factor *factor, int *precision, int *pivoting, int *fallback);

/*
 * LU factors of a caller's dense n x n matrix A (device, column-major c128), with the
 * same kernels, equilibration and fallbacks as tg_factor_create; A is copied (FP64,
 * the refinement residual) and factored in c64 or c128.  Replaces Lu::new
 * (linalg/mod.rs:398-428) for the dense interaction of python/treams_rs/diff.py:1638
 * (InteractionFactor::from_blocks with dense T and C, cluster/interaction.rs:180-205).
 */
int tg_dense_factor_create(int64_t n, const tg_c128 *a_dev, int precision, void *stream, tg_factor **out);
/*
 * x = A^{-1} b (adjoint != 0: A^{-H} b) for `cols` columns (device, n x cols), refined
 * in FP64 until |b - A x| <= rtol |b|: Lu::solve_in_place / solve_adjoint_in_place
 * (linalg/mod.rs:431-438).
 */
int tg_dense_factor_solve(tg_factor *factor, const tg_c128 *b_dev, int cols, double rtol, int adjoint,
                          tg_c128 *x_dev, tg_convergence *conv_host);

/* 1 when the factors are those of the equilibrated operator D_r A D_c, else 0. */
int tg_factor_equilibrated(const tg_factor *factor);

/*
 * adjoint == 0: X = (I - T C)^{-1} T A for `cols` incident columns (device c128,
 * n x cols, reference layout): InteractionFactor::solve (cluster/interaction.rs:213-218).
 * adjoint != 0: X = (I - T C)^{-H} A (Lu::solve_adjoint_in_place, linalg/mod.rs:436-438;
 * InteractionFactor::pullback_incident is T^H of it, cluster/interaction.rs:220-226).
 * Refined with the FP64 matrix-free residual until |b - A x| <= rtol |b|;
 * TG_NOT_CONVERGED otherwise.  conv_host (cols entries) gets the residuals and the
 * refinement count.
 */
int tg_factor_solve(tg_factor *factor, const tg_c128 *incident_dev, int cols, double rtol, int adjoint,
                    tg_c128 *x_dev, tg_convergence *conv_host);

/*
 * The cluster's T-matrix in one spherical basis at the origin with degrees up to
 * lmax_global: T_g = sum_ij R(0 <- p_i) X_ij R(p_j <- 0), X = (I - T C)^{-1} T, with
 * regular translations between the spheres' lmax and lmax_global; out_dev is Mg x Mg
 * c128 column-major (Mg = 2 lmax_global (lmax_global + 2), the sphere mode order).
 * Replaces TMatrix::expand(SphericalBasis.default(lmax_global)) of a solved sphere
 * cluster (python/treams_rs/_tmatrix.py:243-257 with sw/expansion.rs:116-132, on
 * sphere_cluster's X, cluster/spheres.rs:56-80).  Solves the Mg columns T R(p_j <- 0)
 * with the factors (conv_host: Mg entries); rtol as in tg_factor_solve.
 */
int tg_factor_global_tmatrix(tg_factor *factor, int lmax_global, double rtol, tg_c128 *out_dev,
                             tg_convergence *conv_host);

/*
 * tg_cluster_pullback with the adjoint solve by the LU factors: replaces
 * SphereClusterResidual::pullback (cluster/spheres.rs:321-378) through
 * InteractionResidual::pullback_blocks (cluster/interaction.rs:379-440).
 */
int tg_factor_pullback(tg_factor *factor, const tg_c128 *incident_dev, const tg_c128 *x_dev,
                       const tg_c128 *cotangent_dev, int cols, double rtol, double *grad_k0, double *grad_radii,
                       tg_c128 *grad_eps, double *grad_positions, tg_c128 *incident_grad_dev,
                       tg_convergence *conv_host);

/*
 * tg_cluster_pushforward with the tangent solve by the LU factors (one FP64-refined
 * solve to rtol): replaces SphereClusterResidual::pushforward (cluster/spheres.rs:263-300)
 * and IlluminateResidual::pushforward (cluster/interaction.rs:464-478) for physical
 * tangents; with B = I (n columns) it is the tangent of the interacting T-matrix,
 * InteractionResidual::pushforward_blocks (cluster/interaction.rs:359-368).
 */
int tg_factor_pushforward(tg_factor *factor, const tg_c128 *incident_dev, const tg_c128 *x_dev, int cols,
                          double rtol, double tangent_k0, const double *tangent_radii, const tg_c128 *tangent_epsilon,
                          const double *tangent_positions, const tg_c128 *tangent_incident_dev, tg_c128 *dx_dev,
                          tg_convergence *conv_host);

/*
 * Derivatives with the local T-matrices as independent inputs (dense M x M blocks per
 * particle, device, reference layout, column-major, N blocks): the k0 tangent or
 * gradient covers the coupling only.  This is the differentiation of a particle
 * cluster, ParticleClusterResidual::{pushforward, pullback} (cluster/particles.rs:
 * 206-250) for spherical bases, helicity and one real wavenumber, and of the dense
 * interaction, IlluminateResidual::{pushforward, pullback} (cluster/interaction.rs:
 * 464-558) with C given by the positions; it works for sphere clusters too.
 * pushforward: dX from tangents of k0, positions[3N] (host; NULL = 0), the local blocks
 * (NULL = 0) and the incident columns (NULL = 0).  pullback: the gradients of k0 and
 * positions (host), the local blocks T-bar = Y (B + C X)^H (device, N M x M blocks;
 * NULL skips them) and B-bar = T^H Y (device; C-bar = B-bar X^H).
 */
int tg_cluster_pushforward_local(tg_cluster *cluster, const tg_c128 *incident_dev, const tg_c128 *x_dev, int cols,
                                 const tg_gmres_options *options, double tangent_k0, const double *tangent_positions,
                                 const tg_c128 *tangent_local_dev, const tg_c128 *tangent_incident_dev, tg_c128 *dx_dev,
                                 tg_convergence *conv_host);
int tg_factor_pushforward_local(tg_factor *factor, const tg_c128 *incident_dev, const tg_c128 *x_dev, int cols,
                                double rtol, double tangent_k0, const double *tangent_positions,
                                const tg_c128 *tangent_local_dev, const tg_c128 *tangent_incident_dev, tg_c128 *dx_dev,
                                tg_convergence *conv_host);
int tg_cluster_pullback_local(tg_cluster *cluster, const tg_c128 *incident_dev, const tg_c128 *x_dev,
                              const tg_c128 *cotangent_dev, int cols, const tg_gmres_options *options, double *grad_k0,
                              double *grad_positions, tg_c128 *grad_local_dev, tg_c128 *incident_grad_dev,
                              tg_convergence *conv_host);
int tg_factor_pullback_local(tg_factor *factor, const tg_c128 *incident_dev, const tg_c128 *x_dev,
                             const tg_c128 *cotangent_dev, int cols, double rtol, double *grad_k0,
                             double *grad_positions, tg_c128 *grad_local_dev, tg_c128 *incident_grad_dev,
                             tg_convergence *conv_host);

/* ---------------------------------------------------------------------------
 * Fields
 * ------------------------------------------------------------------------- */

/*
 * Electric field of helicity vector spherical waves about every sphere of the
 * cluster, sum_i sum_modes x[i M + mode] wave_mode(point - p_i), at `npoints` points
 * (device FP64, xyz per point) for `cols` coefficient columns (device c128, n x cols,
 * reference layout): out_dev is (3 npoints) x cols column-major, rows 3 p + axis.
 * regular = 0: outgoing (singular) waves, h_l; points at a sphere centre give NaN
 * (the reference raises there).  regular != 0: regular (incoming) waves, j_l, finite
 * everywhere (the series of j_l(kr) / r^l below kr = 1/2, fields.rs:140-180).
 * Replaces fields::field (fields.rs:947) with the cluster's multi-centre helicity
 * basis (Wave.efield, python/treams_rs/_waves.py).
 */
int tg_cluster_field(tg_cluster *cluster, const tg_c128 *x_dev, int cols, int64_t npoints, const double *points_dev,
                     int regular, tg_c128 *out_dev);

/*
 * The same for one expansion of degree up to lmax (1 .. 12) about `center` (host,
 * xyz): x_dev holds 2 lmax (lmax + 2) modes per column in the reference order (a
 * global T-matrix's incident or scattered coefficients, sw/expansion.rs:116, or a
 * plane wave's regular expansion).  Waves of wavenumber k0 in vacuum, outgoing
 * (regular = 0) or regular.  Replaces fields::field (fields.rs:947) for a
 * single-origin SphericalBasis.
 */
int tg_expansion_field(int lmax, double k0, const double *center, const tg_c128 *x_dev, int cols, int64_t npoints,
                       const double *points_dev, int regular, tg_c128 *out_dev, void *stream);

/*
 * Field of a unit helicity plane wave, (N + (2 pol - 1) M) / sqrt(2) exp(i k.r), at
 * `npoints` points: out_dev (3 npoints) c128.  Replaces pw::field / field_value
 * (pw/field.rs:267, pw/polarization.rs:138-203).
 */
int tg_plane_wave_field(double k0, const double *direction, int pol, int64_t npoints, const double *points_dev,
                        tg_c128 *out_dev, void *stream);

/* ---------------------------------------------------------------------------
 * Preconditioners (no reference counterpart: the reference GMRES is
 * unpreconditioned).  Right preconditioning: iteration counts change, the answer
 * and its FP64 true-residual certificate do not.
 * ------------------------------------------------------------------------- */

/* Spheres per block-Jacobi group (spatial recursive bisection), default 4. */
int tg_cluster_set_group_size(tg_cluster *cluster, int spheres);

/*
 * Learned near-field preconditioner M^{-1} x = x + T sum_{(i,j)} D(alpha_ij) C_ij x_j
 * over the k nearest neighbours j of each sphere i; D(alpha) scales destination
 * degree l by alpha_ij[l].  tg_cluster_near_pairs builds the pair list, the pair
 * blocks and the 64 FP32 features per pair (k d, 1/(k d), direction cosines, Mie
 * blocks of both spheres for l <= 6).  Pair weights come from 16 FP32 outputs per
 * pair, alpha[l] = 1 + 0.5 tanh(out[2l]) + 0.5 i tanh(out[2l+1]) (out = 0: the
 * first-order Neumann preconditioner of the near field).
 */
int tg_cluster_near_pairs(tg_cluster *cluster, int k_nearest, int64_t *npairs);
/* (i, j) of every pair, host int32[2 * npairs]. */
int tg_cluster_pair_list(tg_cluster *cluster, int32_t *pairs_host);
/* Copy the pair features (device float[npairs * 64]). */
int tg_cluster_pair_features(tg_cluster *cluster, float *features_dev);
/* Set the pair weights from network outputs (device float[npairs * 16]). */
int tg_cluster_set_pair_output(tg_cluster *cluster, const float *out_dev);
/* y = M^{-1} x (adjoint 1: M^{-H} x), reference layout, device vectors; adjoint 2: M^{-1} x
 * as the mixed-precision inner solves apply it (same-helicity near blocks stored as
 * bfloat16 pairs, FP32 products). */
int tg_cluster_learned_apply(tg_cluster *cluster, const tg_c128 *x_dev, tg_c128 *y_dev, int cols, int adjoint);
/*
 * Training objective of the pair weights for probe columns V (device, n x cols):
 * loss = |A M^{-1} V - V|_F^2 / |V|_F^2 with A = I - T C in FP64, and, when
 * dout_dev != NULL, d loss / d out (device float[npairs * 16]).
 */
int tg_cluster_learned_loss(tg_cluster *cluster, const tg_c128 *probes_dev, int cols, double *loss, float *dout_dev);

/* Network evaluated for TG_PC_LEARNED solves (not owned; NULL detaches). */
typedef struct tg_mlp tg_mlp;
int tg_cluster_set_mlp(tg_cluster *cluster, tg_mlp *mlp);

/*
 * Multilayer perceptron on MFMA (BF16 operands, FP32 accumulation), FP32 master
 * weights: n_in -> width (x hidden_layers, ReLU) -> n_out.  The last layer starts
 * at zero.  Parameters are ordered layer by layer, W (row-major out x in) then b.
 */
int tg_mlp_create(int n_in, int width, int n_out, int hidden_layers, uint64_t seed, void *stream, tg_mlp **out);
int tg_mlp_destroy(tg_mlp *mlp);
int64_t tg_mlp_param_count(const tg_mlp *mlp);
int tg_mlp_get_params(tg_mlp *mlp, float *host);
int tg_mlp_set_params(tg_mlp *mlp, const float *host);
int tg_mlp_get_grads(tg_mlp *mlp, float *host);
int tg_mlp_zero_grad(tg_mlp *mlp);
/* y = network(x) for a batch (device float[batch * n_in] -> float[batch * n_out]). */
int tg_mlp_forward(tg_mlp *mlp, const float *x_dev, int64_t batch, float *y_dev);
/* Accumulate parameter gradients of sum(dy * y) for the last forward batch. */
int tg_mlp_backward(tg_mlp *mlp, const float *dy_dev);
int tg_mlp_adam_step(tg_mlp *mlp, float lr, float beta1, float beta2, float eps, int step);

/* Kernel timing (HIP events on the handle's stream) for measurement. */
int tg_cluster_timing_enable(tg_cluster *cluster, int enable);
/* Total milliseconds and launch count of a timed kernel by name; 0 launches if unknown. */
int tg_cluster_timing_get(tg_cluster *cluster, const char *kernel, double *total_ms, int64_t *launches);
int tg_cluster_timing_reset(tg_cluster *cluster);

#ifdef __cplusplus
}
#endif
#endif
