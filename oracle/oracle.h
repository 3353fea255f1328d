/*
 * oracle.h - CPU restatement of utopia's sparse-solver hot path (FP64 / FP32).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it to compare against the
 * HIP backend. No product code path calls it.
 *
 * Each function restates one reference algorithm (zulianp/utopia, paths
 * relative to the repository root) in plain C with the reference's operation
 * order. Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#ifndef UTOPIA_ORACLE_H
#define UTOPIA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_csr {
    int32_t nrows, ncols;
    int64_t nnz;
    int32_t *rowptr;
    int32_t *col;
    double *val;
} or_csr;

/* ---- kernels ---- */
void or_spmv(const or_csr *A, const double *x, double *y);
void or_spmv_f32(int32_t n, const int32_t *rowptr, const int32_t *col, const float *val, const float *x, float *y);
double or_dot(int64_t n, const double *x, const double *y);
double or_norm2(int64_t n, const double *x);
void or_transpose(const or_csr *A, or_csr *B);
void or_spgemm(const or_csr *A, const or_csr *B, or_csr *C);
void or_ptap(const or_csr *A, const or_csr *P, or_csr *Ac);
void or_csr_free(or_csr *A);
/* dense Cholesky (row-major, lower) and solve */
int or_cholesky(int32_t n, double *A);
void or_cholesky_solve(int32_t n, const double *L, const double *b, double *x);

/* ---- solvers ---- */
typedef struct or_status {
    int64_t iterates;
    double gradient_norm;
    int reason; /* utopia ConvergenceReason: 2 FNORM_ABS, 3 FNORM_RELATIVE, 4 SNORM_RELATIVE, -5 DIVERGED_MAX_IT */
    int converged;
} or_status;

/* ConjugateGradient<M,V,HOMEMADE>, preconditioned with InvDiagPreconditioner
 * (precond=1) or none (precond=0). */
int or_cg_solve(const or_csr *A, const double *b, double *x, int precond, double atol, double rtol, double stol,
                int64_t max_it, or_status *st);

/* Multigrid with Chebyshev3level (smoother=0), Jacobi (smoother=1), multicolour
 * GaussSeidel (smoother=2) or PointJacobi (smoother=3), dense Cholesky coarse solve,
 * IPRTransfer, masks, Galerkin assembly. */
typedef struct or_mg or_mg;
or_mg *or_mg_create(int nlevels, const or_csr *const *interp /* coarse->fine, nlevels-1 */, int smoother,
                    int pre_steps, int post_steps, int mg_type, int masks, int fix_semidefinite);
/* cycle_type 1 (MULTIPLICATIVE_CYCLE) or 3 (FULL_CYCLE), v_cycle_repetition, use_line_search
 * (defaults 1, 10, 0; utopia_MultiLevelBase.hpp:38-40, utopia_Multigrid.hpp:74) */
void or_mg_set_cycle(or_mg *mg, int cycle_type, int64_t v_cycle_repetition, int line_search);
/* Galerkin assembly + smoother/coarse setup (PowerMethod uses rand()). */
int or_mg_update(or_mg *mg, const or_csr *A);
/* Multigrid::apply with max_it iterations, atol/rtol */
int or_mg_apply(or_mg *mg, const double *rhs, double *x, int64_t max_it, double atol, double rtol);
/* Preconditioned CG with the MG as preconditioner */
int or_cg_mg_solve(or_mg *mg, const double *b, double *x, double atol, int64_t max_it, or_status *st);
int or_mg_nlevels(const or_mg *mg);
const or_csr *or_mg_level(const or_mg *mg, int l);
const double *or_mg_mask(const or_mg *mg, int l);
void or_mg_eigs(const or_mg *mg, int l, double *eig_max, double *eig_min);
/* Override the Chebyshev bounds of level l (tests comparing cycles bit-for-bit). */
void or_mg_set_eigs(or_mg *mg, int l, double eig_max, double eig_min);
/* Jones-Plassmann colouring (smallest available colour, splitmix64 hash of i ^ seed) */
void or_jp_color(const or_csr *A, uint64_t seed, int32_t *color, int32_t *ncolors);
/* smoother 2 (multicolour Gauss-Seidel): the colouring of level l, after or_mg_update */
void or_mg_set_colors(or_mg *mg, int l, const int32_t *color, int32_t ncolors);
void or_gs_smooth(const or_csr *A, const double *dinv, const int32_t *color, int32_t ncolors, const double *rhs,
                  double *x, int sweeps, int symmetric);
void or_mg_destroy(or_mg *mg);

/* Jacobi::update + smooth (utopia_Jacobi.hpp:99-140) and PointJacobi::update + smooth
 * (utopia_PointJacobi.hpp:61-142): sweeps from the given x */
void or_jacobi_smooth(const or_csr *A, const double *rhs, double *x, int sweeps);
void or_point_jacobi_smooth(const or_csr *A, const double *rhs, double *x, int sweeps);
/* PowerMethod::get_max_eig (rand() initial vector) */
double or_power_method(const or_csr *A, double tol, int64_t max_it);
/* Chebyshev3level::solve with explicit bounds, nu sweeps (smooth mode) */
void or_chebyshev_smooth(const or_csr *A, const double *b, double *x, int nu, double eig_max, double eig_min);

/* MPRGP (solvers/nonlinear/constrained/quadratic_programming/utopia_MPRGP.hpp:57-387) */
int or_mprgp_solve(const or_csr *A, const double *b, const double *lb, const double *ub, double *x, double atol,
                   double rtol, int64_t max_it, double eps_eig_est, int64_t power_method_max_it, or_status *st);
/* Agglomerate::create_transfer_matrix: piecewise-constant aggregation P (n x returned count) */
int32_t or_agglomerate(const or_csr *A, int64_t bmax, double weight, or_csr *P);
/* MonotoneMultigrid with multicolour projected Gauss-Seidel (JP colouring, seed 0x5eed),
 * GaussSeidel-smoothed coarse Multigrid; interp: nlevels-1 interpolations, coarse to fine */
int or_mmg_solve(const or_csr *A, const double *b, const double *lb, const double *ub, double *x, int nlevels,
                 const or_csr *const *interp, int pre, int post, int64_t coarse_space_it, double active_tol,
                 double atol, double rtol, double stol, int64_t max_it, or_status *st);
/* ProjectedConjugateGradient / ProjectedGradient (quadratic_programming/); NULL bound = none */
int or_pcg_solve(const or_csr *A, const double *b, const double *lb, const double *ub, double *x, double atol,
                 double rtol, double stol, int64_t max_it, or_status *st);
int or_pg_solve(const or_csr *A, const double *b, const double *lb, const double *ub, double *x, double atol,
                double rtol, double stol, int64_t max_it, or_status *st);
/* ProjectedGaussSeidel<PetscMatrix,PetscVector,PETSC>::apply with the scalar sweeper */
int or_pgs_solve(const or_csr *A, const double *b, const double *lb, const double *ub, double *x, double stol,
                 int64_t max_it, int symmetric, int64_t check_each, or_status *st);

/* OpenMP thread count used by the oracle (0 = runtime default). */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
