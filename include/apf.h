/*
 * apf.h -- C ABI of the MI355X (gfx950) hot path: saddle-point assembly,
 * equilibrated-flux estimator and preconditioned MINRES for the space-time
 * heat equation of Rsmeets99/DoubleAdapParabolicFEM.
 *
 * Plain pointers and sizes only.  The host adapter (C++ or Python/ctypes)
 * owns every host array; an apf_ctx owns device memory.  Calls on one ctx are
 * not thread-safe.  Every function returns APF_OK (0) or an error code and, on
 * error, leaves a message readable with apf_last_error().
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference repository):
 *   apf_space_create      -- FESpace::initialize + DOF distribution + the
 *                            two-pass symbolic pattern
 *                            (src/finite_element/fespace/fespace.hpp:111-758,
 *                             src/finite_element/assembly/main_system/
 *                             two_pass_full_saddle_2d.hpp:326-421) and the
 *                            vertex-patch sets
 *                            (src/finite_element/time_slabs/
 *                             time_slab_vertex_patch_builder.hpp:24-140)
 *   apf_assemble_main     -- assemble_main_system_blocks
 *                            (src/finite_element/system/
 *                             assemble_saddle_problem.hpp:131-392) incl.
 *                            compose_saddle_point_matrix
 *                            (src/linear_algebra/system/saddle_point_system.hpp:105-118)
 *   apf_minres_solve      -- SaddlePreconditionedSolveDispatcher<Backend>::solve ->
 *                            solve_minres_with_parabolic_graph_norm
 *                            (src/linear_algebra/eigen_backend/preconditioners/
 *                             solve_saddle_system.hpp:32-121,
 *                             parabolic_graph_norm_minres.hpp:232-369)
 *   apf_csr_minres_solve  -- the same dispatcher for a K, rhs and Hhat assembled by
 *                            the caller (Eigen compressed arrays)
 *   apf_estimate          -- compute_adaptive_test_space_iteration, estimator part
 *                            (src/finite_element/time_slabs/
 *                             time_slab_adaptive_loop.hpp:833-2026):
 *                            solve_all_slabs, solve_all_patches
 *                            (time_slab_equilibrated_flux_reconstruction.hpp:2515),
 *                            compute_equilibrated_flux_error_squared_by_source_cell_*
 *                            (time_slab_equilibrated_flux_indicators.hpp:34-525),
 *                            compute_Y_error_squared_by_source_cell
 *                            (time_slab_error_norms.hpp:371-537) and the X-marking
 *                            components (algorithm/detail/indicator_utilities.hpp:444-472)
 */
#ifndef APF_H
#define APF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum apf_status {
    APF_OK = 0,
    APF_BAD_ARG = 1,            /* std::invalid_argument / std::runtime_error on bad input */
    APF_NOT_CONVERGED = 2,      /* PreconditionedSolveFailure (MINRES) */
    APF_RESIDUAL_TOO_LARGE = 3, /* patch residual acceptance failed (dense_local_error_blocks_2d.hpp:923-949) */
    APF_SINGULAR = 4,           /* non-SPD patch factor */
    APF_HIP_ERROR = 5,
    APF_NO_DEVICE = 6
};

enum apf_precision { APF_FP64 = 0, APF_FP32_REFINE = 1 };
enum apf_assembly_mode { APF_ASSEMBLE_KRON = 0, APF_ASSEMBLE_SCATTER = 1 };
enum apf_precond_kind { APF_PRECOND_DIAGONAL = 0, APF_PRECOND_MULTIGRID = 1 };
/* slab reconstruction: lambda_tilde = lambda of the current solution (exact when the
 * main solve is converged), or solved slab by slab to 1e-12 relative */
enum apf_slab_mode { APF_SLAB_FROM_SOLUTION = 0, APF_SLAB_SOLVE = 1 };
/* Why an iterative slab solve (APF_SLAB_SOLVE, apf_g_estimate) stopped: the tolerance is
 * met, a restart from the true residual gained less than 10% (the FP64 round-off floor
 * of the system, reached above the tolerance), or the iteration limit.  The reference
 * solves these systems directly (time_slabs/time_slab_reconstruction_operator.hpp:810-814),
 * so a caller that needs the tolerance checks for APF_SLAB_STOP_TOL. */
enum apf_slab_stop { APF_SLAB_STOP_NONE = -1, APF_SLAB_STOP_TOL = 0, APF_SLAB_STOP_FLOOR = 1, APF_SLAB_STOP_MAX_IT = 2 };

typedef struct apf_ctx apf_ctx;
typedef struct apf_space apf_space;
typedef struct apf_system apf_system;

/* Tensor space-time mesh: a conforming spatial mesh (intervals in 1D,
 * counter-clockwise triangles in 2D) times a time partition. */
typedef struct {
    int dim;                 /* 1 or 2 */
    int p;                   /* polynomial degree in space and time, 1..4 (estimator: 1..3) */
    int n_vertices;
    const double* vertices;  /* [n_vertices * dim] */
    int n_cells;
    const int* cells;        /* [n_cells * (dim + 1)] */
    int n_times;
    const double* times;     /* [n_times], strictly increasing */
} apf_mesh_desc;

typedef struct {
    int64_t nY, nX, n, nnz;  /* sizes of Y, X, K = [A_y B^T; B C] and nnz(K) */
    int dim, p, Ns, Nt;      /* local spatial / temporal Lagrange basis sizes */
    int nf;                  /* free spatial nodes */
    int n_nodes, n_space_cells, n_intervals, n_vertices;
    int nnz_spatial;         /* spatial adjacency entries */
    int n_patches;           /* spatial vertex patches (one per vertex, shared by all slabs) */
    int max_patch_cells, max_patch_flux, max_patch_scalar;
    int n_bottom_points;     /* bottom (t = 0) quadrature points per spatial cell */
} apf_space_info;

/* Problem data.  M is one SPD tensor per spatial cell (piecewise constant,
 * time independent), ell a constant load, u0 sampled at the bottom quadrature
 * points (apf_space_bottom_points) of every spatial cell. */
typedef struct {
    const double* M;          /* [n_space_cells * dim * dim] row-major, NULL -> identity */
    double ell;
    const double* u0_bottom;  /* [n_space_cells * n_bottom_points] */
} apf_coeff;

typedef struct {
    int iterations;
    int converged;
    double true_relative_residual;  /* ||K x - b||_2 / ||b||_2 */
    double estimate;                /* MINRES preconditioned residual estimate */
    double seconds;
} apf_minres_stats;

typedef struct {
    double lambda_y2, trace2, eta2;          /* ||lambda||_Y^2, ||u0 - u(0)||^2, eta^2 */
    double flux2, recon2, y_est2;            /* estimator components */
    double div_residual_l2;
    double max_patch_residual;               /* max_i |K s - b|_inf / max(1, |b|_inf) over patches */
    double seconds_patches, seconds_diagnostics;
    int extra_refinements;                   /* refinement rounds beyond refine_steps, summed over
                                                patch blocks (each block refines until its problems pass
                                                the 1e-10 acceptance, at most 4 more rounds) */
    int slab_iterations;                     /* APF_SLAB_SOLVE: preconditioned iterations */
    double slab_relative_residual;           /* |f - B^T u - A_y lt| / |f - B^T u| (-1: not solved) */
    double seconds_slab;
    int slab_stop;                           /* enum apf_slab_stop; APF_SLAB_STOP_NONE when not solved */
    int fullpiv_patches;                     /* patches solved again with complete pivoting after missing the
                                                acceptance (dense_local_error_blocks_2d.hpp:1014-1048) */
    double seconds_total;                    /* device span of the whole estimate: slab reconstruction,
                                                patches and indicators, the phases possibly overlapping */
} apf_estimate_out;

/* ---- host-side integer work (no GPU needed) ---- */
int apf_space_create(apf_space** out, const apf_mesh_desc* mesh, int build_pattern);
void apf_space_destroy(apf_space* s);
int apf_space_get_info(const apf_space* s, apf_space_info* info);
/* physical bottom quadrature points [n_space_cells][nb][dim] */
int apf_space_bottom_points(const apf_space* s, double* xy);
/* spatial free-node id per local basis function [n_space_cells][Ns] (-1 = Dirichlet) */
int apf_space_cell_free(const apf_space* s, int* out);
/* spatial node coordinates of free nodes [nf][dim] */
int apf_space_free_coords(const apf_space* s, double* out);
/* CSR pattern of K: rowptr [n+1], col [nnz] */
int apf_space_pattern(const apf_space* s, int64_t* rowptr, int* col);
/* Append the next coarser nested spatial mesh (same dim; every cell of the
 * current coarsest mesh must lie in one cell of the new mesh) for the
 * multigrid preconditioner of apf_minres_solve.  Call fine-to-coarse before
 * apf_system_create. */
int apf_space_add_coarse_mesh(apf_space* s, const apf_mesh_desc* coarse);
/* patch topology: per patch n_flux, n_scalar, n_cells, boundary flag */
int apf_space_patch_sizes(const apf_space* s, int* n_flux, int* n_scalar, int* n_cells, int* boundary);
const char* apf_last_error(const apf_ctx* ctx); /* ctx may be NULL: last error of this thread */

/* ---- device ---- */
int apf_ctx_create(apf_ctx** out, int device);
void apf_ctx_destroy(apf_ctx* ctx);
int apf_system_create(apf_ctx* ctx, const apf_space* s, apf_system** out);
void apf_system_destroy(apf_system* sys);

/* Assemble K (CSR values), f, g on the device. */
int apf_assemble_main(apf_system* sys, const apf_coeff* coeff, int mode);
/* Copy K values [nnz] and b = [f; g] [n] to the host (either may be NULL). */
int apf_system_download(apf_system* sys, double* K_values, double* b);
/* Copy the device CSR pattern of K (row pointer [n+1], columns [nnz]) to the
 * host (either may be NULL).  The column indices are generated on the device
 * when the system is created.  Replaces reading the compressed index arrays of
 * the assembled SparseMatrix (src/linear_algebra/eigen_backend/sparse_matrix.hpp:29-37). */
int apf_system_download_pattern(apf_system* sys, int64_t* rowptr, int* col);

/* Preconditioned MINRES on K x = b (x0 may be NULL).  The preconditioner is
 * block diagonal: diag-in-space / exact-in-time blocks of A_y and of the
 * parabolic graph-norm Schur approximation, applied in FP64 (precond_prec =
 * APF_FP64) or FP32 (APF_FP32_REFINE).  Accepts when the true relative
 * residual <= tol, checked every check_interval iterations. */
int apf_minres_solve(apf_system* sys, double tol, int max_it, int check_interval, int precond_prec,
                     const double* x0, double* x_out, apf_minres_stats* stats);
/* Same with an explicit preconditioner: APF_PRECOND_DIAGONAL (time-exact,
 * space-diagonal blocks) or APF_PRECOND_MULTIGRID (Kronecker/time-eigenmode
 * blocks with spatial multigrid V-cycles; needs apf_space_add_coarse_mesh).
 * apf_minres_solve picks the multigrid one when coarse meshes are present. */
int apf_minres_solve_ex(apf_system* sys, double tol, int max_it, int check_interval, int precond_kind,
                        int precond_prec, const double* x0, double* x_out, apf_minres_stats* stats);
/* Average device time of `reps` back-to-back SpMVs y = K x (HIP events on the
 * system stream), for the HBM roofline of the MINRES inner loop. */
int apf_bench_spmv(apf_system* sys, int reps, double* ms_avg);
/* y = K x for host arrays x, y [n]: the assembled CSR K (matrix_free = 0), or the same
 * entries formed from K's Kronecker factors (spatial K, M and the time tables) on the
 * device (matrix_free = 1: the product of the MINRES iterations; the true residuals that
 * decide acceptance use the CSR K).  Replaces the Eigen sparse product K * x
 * (eigen_backend/sparse_matrix.hpp) on the hot path's operator. */
int apf_system_matvec(apf_system* sys, const double* x, double* y, int matrix_free);
/* Average device time of `reps` matrix-free products (as apf_bench_spmv). */
int apf_bench_kron_apply(apf_system* sys, int reps, double* ms_avg);
/* Average device time of `reps` launches of the multigrid's fine-level Chebyshev step
 * over all columns (the preconditioner's dominant kernel) in precision prec (APF_FP64 or
 * APF_FP32_REFINE: FP32 V-cycles), and its algorithmic bytes per launch.  Timing only: it
 * overwrites the multigrid's work vectors. */
int apf_bench_mg_step(apf_system* sys, int reps, int prec, double* ms_avg, double* bytes);
/* Device copy bandwidth (read + write bytes / s, in GB/s) of a plain copy kernel
 * over `bytes` bytes, `reps` launches: the measured HBM ceiling and the
 * calibration kernel for counter-based traffic. */
int apf_bench_stream(apf_ctx* ctx, int64_t bytes, int reps, double* gbs);
/* Device store bandwidth (bytes written / s, in GB/s) of a plain store-only
 * kernel over `bytes` bytes, `reps` launches: the measured ceiling of kernels
 * whose only large stream is written, like the matrix fill of the assembly. */
int apf_bench_store(apf_ctx* ctx, int64_t bytes, int reps, double* gbs);
/* Capture the MINRES iteration in a hipGraph (default on) or launch eagerly. */
int apf_system_use_graph(apf_system* sys, int on);
/* Host seconds of the last multigrid setup (Galerkin operators, time eigenmodes). */
int apf_system_mg_setup_seconds(apf_system* sys, double* sec);
/* Upload a solution x = [lambda; u] [n] (e.g. from another solver). */
int apf_system_set_solution(apf_system* sys, const double* x);

/* Estimator on the current solution: slab reconstruction, one mixed patch
 * problem per (slab, vertex), per-cell flux / residual / reconstruction terms.
 * prec: APF_FP64 (FP64 inverse) or APF_FP32_REFINE (FP32 inverse, FP64
 * residuals, at least `refine_steps` refinement steps, more where a patch
 * misses the 1e-10 acceptance).  Per-cell outputs (may be NULL) are indexed by
 * space-time cell k * n_space_cells + c. */
int apf_estimate(apf_system* sys, int prec, int refine_steps, apf_estimate_out* out,
                 double* flux_by_cell, double* resid_by_cell, double* recon_by_cell);
/* Same with the slab reconstruction mode (apf_slab_mode).  APF_SLAB_SOLVE
 * solves A_slab lt = f_slab - (B^T u)_slab for every slab
 * (time_slabs/time_slab_reconstruction.hpp:584-735) with the multigrid's
 * slab preconditioner; it needs apf_space_add_coarse_mesh. */
int apf_estimate_ex(apf_system* sys, int prec, int refine_steps, int slab_mode, apf_estimate_out* out,
                    double* flux_by_cell, double* resid_by_cell, double* recon_by_cell);

/* Device timing of the last call of each phase in milliseconds (HIP events):
 * [0] assembly, [1] patch kernel, [2] diagnostics, [3] one SpMV (last). */
int apf_system_timings(apf_system* sys, double* ms4);
/* Device time of the three patch phases of the last apf_estimate in
 * milliseconds: [0] patch matrices and inverses, [1] right-hand sides,
 * [2] batched solves with refinement. */
int apf_estimate_phase_timings(apf_system* sys, double* ms3);

/* ---- a K assembled by the caller ---- */
/* How the compressed arrays of apf_csr_matrix_create are read: every stored entry
 * (a symmetric matrix in full storage; row- and column-major then coincide), or only
 * the entries with inner index >= outer index, completed symmetrically (for Eigen's
 * default ColMajor storage that is the lower triangle, the part Eigen::MINRES<...,
 * Eigen::Lower, ...> reads, parabolic_graph_norm_minres.hpp:240-243). */
enum apf_csr_storage { APF_CSR_FULL = 0, APF_CSR_TRIANGLE = 1 };
typedef struct apf_csr_matrix apf_csr_matrix;
/* Upload a symmetric sparse matrix given by the compressed arrays of an
 * Eigen::SparseMatrix<double, ColMajor, int> after makeCompressed() (outerIndexPtr [n+1],
 * innerIndexPtr, valuePtr: src/linear_algebra/eigen_backend/sparse_matrix.hpp:29-37).
 * Replaces handing K (and the graph-norm Hhat) to the Eigen MINRES. */
/* Copy a matrix's CSR arrays (row pointer [n+1] as int64, sorted columns, values) to the
 * host (any may be NULL). */
int apf_csr_matrix_download(const apf_csr_matrix* M, int64_t* rowptr, int* col, double* values);
int apf_csr_matrix_create(apf_ctx* ctx, int64_t n, const int* outer, const int* inner, const double* values,
                          int storage, apf_csr_matrix** out);
void apf_csr_matrix_destroy(apf_csr_matrix* M);
/* rows (= columns) and stored entries after the symmetric completion */
int apf_csr_matrix_info(const apf_csr_matrix* M, int64_t* n, int64_t* nnz);
/* MINRES on K x = b for a caller's K = [A_y B^T; B C] whose first n_lambda unknowns
 * are lambda.  Replaces SaddlePreconditionedSolveDispatcher<Backend>::solve with the
 * ParabolicGraphNorm MINRES (src/linear_algebra/eigen_backend/preconditioners/
 * solve_saddle_system.hpp:32-121, parabolic_graph_norm_minres.hpp:46-369): K and rhs
 * are its arguments, n_lambda = blocks.A.rows(), Hhat = context.graph_norm->hhat (may
 * be NULL).  Preconditioner blockdiag(q(D^-1 A_y) D^-1, q(D^-1 Hhat) D^-1), q the
 * Chebyshev polynomial of `cheb_degree` steps (1..64) on [lambda_max / degree^2,
 * lambda_max] of the Jacobi-scaled block, applied in FP64 or FP32 (precond_prec);
 * without Hhat the X block is the same polynomial in D_S^-1 S for the Schur complement
 * S = -C + B diag(A_y)^-1 B^T, applied from K's blocks (D_S = diag(S)).  Accepts when the true
 * relative residual <= tol, checked every check_interval iterations
 * (iterative_residual_acceptance.hpp:117-216); APF_NOT_CONVERGED otherwise (x then
 * holds the last iterate).  b, x, x0 (may be NULL) are host arrays of length n. */
int apf_csr_minres_solve(apf_ctx* ctx, apf_csr_matrix* K, int64_t n_lambda, const apf_csr_matrix* Hhat,
                         const double* b, double* x, const double* x0, double tol, int max_it, int check_interval,
                         int cheb_degree, int precond_prec, apf_minres_stats* stats);
/* y = A x (host vectors of length n).  Replaces la::ops::matvec / transpose_matvec on a
 * symmetric SparseMatrix (src/linear_algebra/eigen_backend/operations.hpp:103-128); the
 * G estimator forms its right-hand side f_G - B(X,G)^T u_delta with it
 * (algorithm/adaptive_driver.hpp:2790-2798). */
int apf_csr_spmv(apf_ctx* ctx, const apf_csr_matrix* A, const double* x, double* y);
/* d^T B d for the leading block B = A[0:n_block, 0:n_block] (host d of length n_block),
 * reduced in a fixed order.  Replaces la::ops::quadratic_form(A_g, d)
 * (eigen_backend/operations.hpp:129-150; adaptive_driver.hpp:2860-2866). */
int apf_csr_block_energy(apf_ctx* ctx, const apf_csr_matrix* A, int64_t n_block, const double* d, double* out);
/* Conjugate gradients on B x = b for the SPD leading block B = A[0:n_block, 0:n_block],
 * preconditioned by q(D^-1 B) D^-1 (the Chebyshev polynomial of cheb_degree steps, as in
 * apf_csr_minres_solve).  Replaces the direct solve of the G system
 * (solver_.compute(A_g); solver_.solve(rhs_g, lambda_g), adaptive_driver.hpp:2806-2814).
 * The recursive residual is read every check_interval iterations; at a tenth of tol the
 * true relative residual ||b - B x|| / ||b|| decides (restarting from it, at most 8 times);
 * APF_NOT_CONVERGED when it stays above tol.  b, x, x0 (may be NULL) are host arrays. */
int apf_csr_block_pcg_solve(apf_ctx* ctx, const apf_csr_matrix* A, int64_t n_block, const double* b, double* x,
                            const double* x0, double tol, int max_it, int check_interval, int cheb_degree,
                            apf_minres_stats* stats);

/* ---- space-time meshes that are not tensor products ---- */
/* One space (X or Y) as the reference's FESpace hands it to the assembler: the active
 * cells (prisms = triangle x [t0, t1], or rectangles in 1+1D) and each cell's restriction
 * rows u_cell = C_c u_true (src/finite_element/assembly/constrained_dofs.hpp:36-56,
 * build_cell_restriction :133-226): per (cell, local dof) a row of (true dof, weight),
 * true dofs strictly increasing; [(d, 1.0)] for a free dof, empty for a Dirichlet dof,
 * the masters and weights of a hanging dof.  Local dof = i_space * (p+1) + i_time
 * (basis/functions/triangular_prism_lagrange.hpp:32-35), spatial nodes in the
 * equidistant order (nodes/triangle_equidistant.hpp:39-53). */
typedef struct {
    int dim, p;                  /* 1 or 2; 1..3 */
    int64_t n_vertices;
    const double* vertices;      /* [n_vertices * dim] spatial vertex coordinates */
    int64_t n_cells;
    const int* cells;            /* [n_cells * (dim+1)] spatial vertices (counter-clockwise in 2D) */
    const double* times;         /* [n_cells * 2] t0 < t1 */
    const int64_t* restr_ptr;    /* [n_cells * n_local + 1], n_local = Ns * (p+1) */
    const int* restr_dof;        /* true dof of each row entry */
    const double* restr_weight;
    int64_t n_dofs;              /* true dofs of the space */
} apf_cell_space_desc;
/* Coefficients of the non-tensor path.  M is evaluated at every quadrature point (the
 * reference calls its M callback per point, dense_local_main_blocks_2d.hpp:234-240): the
 * identity (grid_n = 0) or piecewise constant on a grid_n^dim grid of (0,1)^dim, tensor of
 * grid cell (iy, ix) at (iy * grid_n + ix) * dim * dim.  u0 at the bottom rule points of
 * the X cells with t0 = 0 (apf_general_bottom_points). */
typedef struct {
    int grid_n;
    const double* M_grid;
    double ell;
    const double* u0_bottom;     /* [n_bottom_cells * n_bottom_points] */
} apf_general_coeff;
typedef struct apf_general_system apf_general_system;
/* Host only (no GPU): the CSR pattern of K = [A_y B^T; B C] (row pointer [n+1], sorted
 * columns) for the given spaces, y_ancestor[c] = the X cell containing Y cell c
 * (assembly/detail/active_cell_locator.hpp).  Replaces the symbolic pass of the two-pass
 * saddle assembly (assembly/main_system/two_pass_full_saddle_2d.hpp:326-459).  With rowptr
 * and col NULL it returns nnz only. */
int apf_general_pattern(const apf_cell_space_desc* X, const apf_cell_space_desc* Y, const int* y_ancestor,
                        int64_t* nnz, int64_t* rowptr, int* col);
/* Device system of a non-tensor mesh: uploads cells, restriction rows and the pattern. */
int apf_general_system_create(apf_ctx* ctx, const apf_cell_space_desc* X, const apf_cell_space_desc* Y,
                              const int* y_ancestor, apf_general_system** out);
void apf_general_system_destroy(apf_general_system* sys);
int apf_general_system_info(const apf_general_system* sys, int64_t* n, int64_t* n_lambda, int64_t* nnz,
                            int64_t* n_bottom_cells, int* n_bottom_points);
/* physical bottom rule points [n_bottom_cells * n_bottom_points * dim] */
int apf_general_bottom_points(const apf_general_system* sys, double* pts);
/* Element-local quadrature assembly of K and b = [f; g] on the device, with the X basis of
 * each Y cell's ancestor evaluated at the mapped quadrature points.  Replaces
 * assemble_main_system_blocks (system/assemble_saddle_problem.hpp:131-392) with
 * assemble_main_y_dense_local_blocks_2d (assembly/main_system/dense_local_main_blocks_2d.hpp:48-640),
 * the trace blocks and scatter_matrix (assembly/scatter.hpp:66-115).  ms (may be NULL):
 * device time of the assembly kernels. */
int apf_general_assemble(apf_general_system* sys, const apf_general_coeff* coeff, double* ms);
/* Copy the pattern, values and b to the host (any may be NULL). */
int apf_general_download(apf_general_system* sys, int64_t* rowptr, int* col, double* values, double* b);
/* A view of the system's K for apf_csr_minres_solve (destroy it before the system). */
int apf_general_matrix(apf_general_system* sys, apf_csr_matrix** out);
/* The two parts of the X-marking indicator of the solution x = [lambda; u] (host, length n)
 * (compute_x_marking_indicator_components, algorithm/detail/indicator_utilities.hpp:435-473):
 * lambda_y2[c] = int_{Y cell c} M grad_x lambda . grad_x lambda on the cell rule
 * (compute_spatial_energy_squared_by_active_cell, :109-171), and trace2[a] =
 * int_{t=0} (u0 - u)^2 on X cell a if t0 = 0, else 0 (compute_initial_trace_mismatch_squared_
 * by_x_cell, :333-396).  M and u0 (at apf_general_bottom_points) as for apf_general_assemble.
 * The caller sums lambda_y2 into the Y cells' X ancestors (aggregate_to_x_active_cells,
 * :398-433) and adds trace2 to get eta^2 per X cell. */
int apf_general_x_indicators(apf_general_system* sys, const double* x, const apf_general_coeff* coeff,
                             double* lambda_y2, double* trace2);
/* The parabolic graph-norm approximation Hhat = -C_signed + B_dt diag(A_y)^-1 B_dt^T + A_x
 * (X x X, symmetric positive definite, full storage) of an assembled system, built on the
 * device: B_dt = int dpsi/dt phi and A_x = int M grad_x psi . grad_x psi by element
 * quadrature on the Y cells (X basis of each ancestor at the mapped points), the outer
 * products of B_dt's columns scaled by 1 / A_y(j, j), and C from K.  Replaces
 * assemble_parabolic_graph_norm_approximation (assembly/main_system/parabolic_graph_norm.hpp:
 * 22-65) with diagonal_schur_approximation_with_diagnostics and
 * parabolic_graph_norm_approximation (linear_algebra/preconditioners/schur_approximations.hpp:
 * 80-287); APF_BAD_ARG with the reference's message if an A_y diagonal is <= 1e-14.  The
 * result is apf_csr_minres_solve's Hhat (destroy it with apf_csr_matrix_destroy). */
int apf_general_graph_norm(apf_general_system* sys, const apf_general_coeff* coeff, apf_csr_matrix** out);

/* Estimator on a non-tensor mesh.  The estimator works per time slab (the slabs between
 * the global time endpoints, time_slabs/time_slab_builder.hpp:53-120); slabs that share one
 * spatial mesh form a group, given as a tensor space (apf_space_create on that mesh and the
 * group's time nodes, with apf_space_add_coarse_mesh for the slab solve) whose space-time
 * cells are the Y cells first_y_cell[g] + k * n_space_cells + c (slab k, spatial cell c, the
 * general system's Y cells in that order, no constrained Y dofs inside a slab).  On each
 * group lambda is restricted and u_delta interpolated through the X ancestors, and the
 * estimator with the slab solve runs on it (the per-slab reconstruction of
 * time_slab_reconstruction.hpp:584-735, the patches and indicators of
 * time_slab_adaptive_loop.hpp:833-2026). */
typedef struct apf_general_estimator apf_general_estimator;
int apf_general_estimator_create(apf_general_system* gsys, int n_groups, apf_space* const* spaces,
                                 const int64_t* first_y_cell, apf_general_estimator** out);
/* The same for Y cells that span several time slabs (adaptive meshes): the slabs lie between
 * consecutive time endpoints of the Y cells and each Y cell is sliced into the slabs it spans
 * (time_slab_builder.hpp:53-120).  Cell (slab k, spatial cell c) of group g is a slice of the
 * general system's Y cell cell_source[cell_ptr[g] + k * n_space_cells + c] with the same
 * spatial cell; lambda at the group's Y nodes is that cell's polynomial at the slab's time
 * nodes, and the per-cell outputs of apf_general_estimate are summed into the source cells
 * (by_source_cell).  A group space without coarse meshes solves its slab problems on one
 * multigrid level (a dense inverse per time shift, at most 4096 coupled nodes). */
int apf_general_estimator_create_sliced(apf_general_system* gsys, int n_groups, apf_space* const* spaces,
                                        const int64_t* cell_ptr, const int* cell_source, apf_general_estimator** out);
void apf_general_estimator_destroy(apf_general_estimator* ge);
/* coefficients of group g's tensor system (M per spatial cell of the group mesh, u0 at
 * its bottom points; only the group that starts at t = 0 uses u0) */
int apf_general_estimator_assemble(apf_general_estimator* ge, int group, const apf_coeff* coeff);
/* estimate on the solution x [n] (host) of the general system; totals as apf_estimate_out
 * (slab iterations and residual: the largest over the groups); per-cell outputs (may be
 * NULL) are indexed by the general system's Y cells. */
int apf_general_estimate(apf_general_estimator* ge, const double* x, int prec, int refine_steps, apf_estimate_out* out,
                         double* flux_by_cell, double* resid_by_cell, double* recon_by_cell);

/* Largest n accepted by apf_csr_minres_solve_noalloc: its work buffers live in one
 * file-scope device array of 13 * 8 * APF_NOALLOC_MAX_N bytes plus 36 KiB (27.3 MB) per
 * GPU, held until the process ends. */
#define APF_NOALLOC_MAX_N 262144
/* No counterpart in the reference.  The same solve as apf_csr_minres_solve (same kernels,
 * bit-identical result) for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  It calls no
 * hipMalloc / hipFree or allocating routine; its work buffers are the file-scope scratch
 * array above (at most APF_NOALLOC_MAX_N unknowns, else APF_BAD_ARG).  d_b and d_x are
 * device arrays of length n; with use_x0 != 0, d_x holds the initial guess on entry.
 * Without Hhat (the X block is the Schur-complement polynomial).  Every call on one GPU
 * shares that GPU's scratch array and leaves its last contents there: calls that overlap
 * on one GPU (two threads, or two contexts on one device) race. */
int apf_csr_minres_solve_noalloc(apf_ctx* ctx, const apf_csr_matrix* K, int64_t n_lambda, const double* d_b,
                                 double* d_x, int use_x0, double tol, int max_it, int check_interval,
                                 int cheb_degree, int precond_prec, apf_minres_stats* stats);
/* Device memory and copies for the device vectors of apf_csr_minres_solve_noalloc (no
 * counterpart in the reference: the reference's vectors live in host memory). */
int apf_device_alloc(apf_ctx* ctx, int64_t bytes, void** out);
void apf_device_free(apf_ctx* ctx, void* p);
int apf_device_upload(apf_ctx* ctx, void* dst, const void* src, int64_t bytes);
int apf_device_download(apf_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* ---- G-estimator ---- */
typedef struct {
    double lambda_difference_squared;       /* G^2 = max(0, raw) (g_lambda_difference_squared) */
    double lambda_difference;               /* G */
    double raw;                             /* (lambda_G - I lambda)^T A_G (lambda_G - I lambda) */
    double coefficient_difference_squared;  /* |lambda_G - I lambda|^2 over the coefficients */
    double solver_relative_residual;        /* |rhs - A_G lambda_G| / |rhs| */
    int solver_iterations;
    int64_t g_true_dofs;                    /* dimension of G */
    double seconds;
    int solver_stop;                        /* apf_slab_stop: why the slab solve ended */
} apf_g_estimate_out;
/* G-estimator of the degree-p solution held by `sys` (after apf_minres_solve or
 * apf_system_set_solution).  Replaces compute_g_estimator_for_accepted_inner_iteration_
 * (algorithm/adaptive_driver.hpp:2680-2905).  `gsys` is a system of degree p+1
 * on the same mesh, assembled with the same coefficients and given the coarse
 * meshes of its multigrid (apf_space_add_coarse_mesh); G is its Y space.  Its
 * solution vector is overwritten: [I lambda; I u] (nodal interpolation, as
 * prolong_true_coefficients_nodal).  lambda_G solves A_G lambda_G = f_G - B^T I u
 * to the relative residual `tol` (slab solve, at most `max_it` iterations).
 * A raw value below -1000 eps max(1, |lambda_G - I lambda|^2) returns APF_BAD_ARG,
 * as the reference's runtime_error at adaptive_driver.hpp:2869-2883. */
int apf_g_estimate(apf_system* sys, apf_system* gsys, double tol, int max_it, apf_g_estimate_out* out);

/* ---- 1+1D adaptive meshes and their X / Y spaces (host; no GPU needed) ---- */
/* The reference's cell tree, vertex ids, refinement closure and DoF numbering for 1+1D
 * rectangles, so that cell ids, true-DoF numbers and restriction rows are those of the
 * reference.  A space references its mesh (the X and Y spaces share one mesh, as the
 * reference's FESpaces do); destroy the spaces before the mesh. */
typedef struct apf_mesh1d apf_mesh1d;
typedef struct apf_fespace1d apf_fespace1d;
/* Root cell [x0, x1] x [t0, t1] (Mesh(x0, x1, t0, t1), mesh/mesh.hpp:52-59 with
 * mesh/initialization/root_cell.hpp:60-107); cell 0, spatial vertices 0, 1, temporal 0, 1. */
int apf_mesh1d_create(double x0, double x1, double t0, double t1, apf_mesh1d** out);
/* The mesh of one time slab: the source cells (in the given order) sliced to
 * [t_begin, t_end], each a new leaf cell (ids 0..n-1) with the source's generation and
 * levels and spatial boundary flags, both temporal faces on the slab's boundary, vertices
 * found or created by coordinate (TimeSlab::append_slice_from_source,
 * time_slabs/time_slab.hpp:180-249; mesh/refinement/time_slicing/slice_cell_in_time_1d.hpp:14-113). */
int apf_mesh1d_create_slab(const apf_mesh1d* source, const int* source_cells, int64_t n, double t_begin,
                           double t_end, apf_mesh1d** out);
void apf_mesh1d_destroy(apf_mesh1d* m);
int apf_mesh1d_info(const apf_mesh1d* m, int64_t* n_cells, int64_t* n_spatial_vertices, int64_t* n_temporal_vertices);
/* vertex coordinates by id (either may be NULL) */
int apf_mesh1d_vertices(const apf_mesh1d* m, double* x, double* t);
/* per mesh cell (any may be NULL): parent (-1 for the root), generation, leaf flag,
 * [x0 id, x1 id, t0 id, t1 id], box [x0, x1, t0, t1] (mesh/cell.hpp). */
int apf_mesh1d_cells(const apf_mesh1d* m, int* parent, int* generation, int* is_leaf, int* vertex_ids, double* box);
/* FE space of degree p (1..3 in space and time) on the given active cells: the X space
 * (continuous_in_time != 0: SpaceTimePolicy) or the Y space (0: SpaceOnlyPolicy,
 * fespace/policies.hpp:9-28).  FESpace::initialize (fespace/fespace.hpp:115-121):
 * adjacency and distribute_dofs_1d (fespace/dofs/dof_distribution_1d.hpp:655-671). */
int apf_fespace1d_create(apf_mesh1d* m, int p, int continuous_in_time, const int* active_cells, int64_t n_active,
                         apf_fespace1d** out);
void apf_fespace1d_destroy(apf_fespace1d* s);
/* FESpace::refine (fespace/fespace.hpp:347-371 -> fespace/refinement/refine_1d.hpp:282-321):
 * the marked active cells split in the given order (the Doerfler order), then the
 * one-level grading closure; the DoFs are redistributed. */
int apf_fespace1d_refine(apf_fespace1d* s, const int* marked, int64_t n_marked);
/* active cells, true (free) DoFs, constrained DoFs, entries of the restriction rows */
int apf_fespace1d_info(const apf_fespace1d* s, int64_t* n_active, int64_t* n_true_dofs, int64_t* n_constrained_dofs,
                       int64_t* n_restriction_entries);
/* active cell ids in the partition's order (fespace/partition_view.hpp) */
int apf_fespace1d_active_cells(const apf_fespace1d* s, int* cells);
/* [n_active][(p+1)^2] global DoF ids per local DoF (local = i_space * (p+1) + j_time):
 * true DoFs 0..n_true-1, constrained DoFs after them, -1 for a Dirichlet DoF
 * (DoFHandler::cell_dofs, fespace/dofs/dof_handler.hpp) */
int apf_fespace1d_cell_dofs(const apf_fespace1d* s, int* gids);
/* constrained DoF n_true + k: masters[ptr[k] .. ptr[k+1]) (true DoFs) and weights */
int apf_fespace1d_constraints(const apf_fespace1d* s, int64_t* ptr, int* masters, double* weights);
/* restriction rows of the active cells, in active order ([n_active * (p+1)^2 + 1] ptr):
 * build_cell_restriction (assembly/constrained_dofs.hpp:133-226), the rows that
 * apf_cell_space_desc takes */
int apf_fespace1d_restriction(const apf_fespace1d* s, int64_t* ptr, int* dof, double* weight);
/* Neighbours of each active cell across its spatial (temporal = 0) or temporal faces in
 * the layout of the reference's mesh file (build_*_neighbour_lists + flatten_neighbours,
 * fespace/io/detail/write_mesh_binary_detail.hpp:60-186): *max_per_face = the longest
 * list; flat (may be NULL) [n_active][2][max_per_face] cell ids, -1 padded. */
int apf_fespace1d_neighbours(const apf_fespace1d* s, int temporal, int* max_per_face, int* flat);
/* for each active cell of `fine`, the position in coarse's active list of its first
 * active ancestor (find_active_ancestor_cell, assembly/detail/active_cell_locator.hpp) */
int apf_fespace1d_ancestors(const apf_fespace1d* coarse, const apf_fespace1d* fine, int* index);

/* ---- marking (host; no GPU needed) ---- */
typedef struct {
    double total, target;          /* Neumaier total over ids ascending, theta * total */
    double cutoff_value, cutoff_margin, max_indicator;
    int64_t cutoff_index;          /* position of the last marked entry (n if none) */
    int64_t near_tie_count;        /* entries within near_tie_tolerance of the cutoff value */
} apf_doerfler_diag;

/* estimator^2 = flux^2 + reconstruction^2 per cell; every input (resid2 may be
 * NULL) must be finite and >= -1e-14.  Replaces
 * combine_time_slab_estimator_squared_errors_deterministic
 * (src/finite_element/time_slabs/time_slab_estimator_combination.hpp:88-250). */
int apf_combine_estimator(int64_t n, const double* flux2, const double* recon2, const double* resid2, double* est2);
/* Deterministic Doerfler marking: cell ids (NULL -> 0..n-1, unique) sorted by
 * value descending, ties to the lower id, marked until the compensated
 * cumulative sum reaches theta * total; marked[] receives the ids in marking
 * order.  Replaces CellwiseSquaredError::doerfler_marking_deterministic
 * (src/finite_element/time_slabs/time_slab_cellwise_errors.hpp:156-280). */
int apf_doerfler_mark(int64_t n, const int64_t* ids, const double* values, double theta, double near_tie_tolerance,
                      int64_t* marked, int64_t* n_marked, apf_doerfler_diag* diag);
/* rho^2 (lambda_Y^2 + (1 + 1/(rho + 1/2)^2) trace^2).  Replaces
 * compute_y_estimator_threshold_squared (algorithm/detail/indicator_utilities.hpp:475-492). */
int apf_y_estimator_threshold(double rho, double lambda_y2, double trace2, double* out);

#ifdef __cplusplus
}
#endif
#endif /* APF_H */
