/*
 * oracle.c - CPU restatement of utopia's sparse-solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * by the smoke check and by bench.py's CPU-baseline leg.
 *
 * Parity status of each function (how it is pinned) is listed in
 * tests/test_oracle_pins.py; every restated function has at least one
 * independent check there (closed form, brute force or reference fixture).
 *
 * Compiled with -ffp-contract=off: products are rounded before they are added,
 * exactly as written. Dots use a fixed 256-chunk split (the result does not
 * depend on the OpenMP thread count).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_CHUNKS 256

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ kernels */

/* y = A x, row by row, columns in stored order: PetscMatrix::multiply -> MatMult
 * (backend/petsc/utopia_petsc_Matrix.cpp:1326-1358; AIJ MatMult sums each row
 * in stored column order). */
void or_spmv(const or_csr *A, const double *x, double *y) {
    const int32_t n = A->nrows;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j) s += A->val[j] * x[A->col[j]];
        y[i] = s;
    }
}

void or_spmv_f32(int32_t n, const int32_t *rowptr, const int32_t *col, const float *val, const float *x, float *y) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (int32_t j = rowptr[i]; j < rowptr[i + 1]; ++j) s += val[j] * x[col[j]];
        y[i] = s;
    }
}

/* dot(x, y): VecDot (backend/petsc/utopia_petsc_Vector.cpp:59), fixed chunking */
double or_dot(int64_t n, const double *x, const double *y) {
    double part[OR_CHUNKS];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < OR_CHUNKS; ++c) {
        const int64_t a = n * c / OR_CHUNKS, b = n * (c + 1) / OR_CHUNKS;
        double s = 0.0;
        for (int64_t i = a; i < b; ++i) s += x[i] * y[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < OR_CHUNKS; ++c) s += part[c];
    return s;
}

/* norm2: VecNorm NORM_2 (backend/petsc/utopia_petsc_Vector.hpp:301-305) */
double or_norm2(int64_t n, const double *x) { return sqrt(or_dot(n, x, x)); }

static double *vnew(int64_t n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static void vcopy(int64_t n, const double *x, double *y) { memcpy(y, x, sizeof(double) * (size_t)n); }

void or_csr_free(or_csr *A) {
    if (!A) return;
    free(A->rowptr);
    free(A->col);
    free(A->val);
    A->rowptr = NULL;
    A->col = NULL;
    A->val = NULL;
}

static void csr_copy(const or_csr *A, or_csr *B) {
    *B = *A;
    B->rowptr = (int32_t *)malloc(sizeof(int32_t) * ((size_t)A->nrows + 1));
    B->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(A->nnz > 0 ? A->nnz : 1));
    B->val = (double *)malloc(sizeof(double) * (size_t)(A->nnz > 0 ? A->nnz : 1));
    memcpy(B->rowptr, A->rowptr, sizeof(int32_t) * ((size_t)A->nrows + 1));
    memcpy(B->col, A->col, sizeof(int32_t) * (size_t)A->nnz);
    memcpy(B->val, A->val, sizeof(double) * (size_t)A->nnz);
}

/* B = A^T: MatTranspose (utopia_petsc_Matrix.cpp:462), stable counting sort so
 * each row of B lists A's rows in increasing order. */
void or_transpose(const or_csr *A, or_csr *B) {
    const int32_t n = A->nrows, m = A->ncols;
    const int64_t nnz = A->nnz;
    B->nrows = m;
    B->ncols = n;
    B->nnz = nnz;
    B->rowptr = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    B->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    B->val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    for (int64_t e = 0; e < nnz; ++e) B->rowptr[A->col[e] + 1]++;
    for (int32_t j = 0; j < m; ++j) B->rowptr[j + 1] += B->rowptr[j];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * ((size_t)m + 1));
    memcpy(fill, B->rowptr, sizeof(int32_t) * ((size_t)m + 1));
    for (int32_t i = 0; i < n; ++i)
        for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e) {
            const int32_t j = A->col[e];
            B->col[fill[j]] = i;
            B->val[fill[j]] = A->val[e];
            fill[j]++;
        }
    free(fill);
}

static int cmp_i32(const void *a, const void *b) {
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* C = A*B, Gustavson row by row with a dense accumulator: C_ij accumulates
 * a_ik*b_kj over k in A's row order, starting from the first product. The
 * pattern is the structural product (no numerical dropping, sorted columns),
 * as MatMatMult/MatPtAP produce it (backend/petsc/eval/utopia_petsc_Eval_Multiply.cpp:12-61). */
void or_spgemm(const or_csr *A, const or_csr *B, or_csr *C) {
    const int32_t n = A->nrows, m = B->ncols;
    int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    double *acc = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    for (int32_t j = 0; j < m; ++j) mark[j] = -1;
    C->nrows = n;
    C->ncols = m;
    C->rowptr = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    /* symbolic */
    int64_t nnz = 0;
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e) {
            const int32_t k = A->col[e];
            for (int32_t q = B->rowptr[k]; q < B->rowptr[k + 1]; ++q) {
                const int32_t j = B->col[q];
                if (mark[j] != i) {
                    mark[j] = i;
                    ++nnz;
                }
            }
        }
        C->rowptr[i + 1] = (int32_t)nnz;
    }
    C->nnz = nnz;
    C->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    C->val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    for (int32_t j = 0; j < m; ++j) mark[j] = -1;
    /* numeric */
    for (int32_t i = 0; i < n; ++i) {
        int32_t cnt = C->rowptr[i];
        for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e) {
            const int32_t k = A->col[e];
            const double a = A->val[e];
            for (int32_t q = B->rowptr[k]; q < B->rowptr[k + 1]; ++q) {
                const int32_t j = B->col[q];
                const double p = a * B->val[q];
                if (mark[j] != i) {
                    mark[j] = i;
                    acc[j] = p;
                    C->col[cnt++] = j;
                } else {
                    acc[j] += p;
                }
            }
        }
        qsort(C->col + C->rowptr[i], (size_t)(C->rowptr[i + 1] - C->rowptr[i]), sizeof(int32_t), cmp_i32);
        for (int32_t e = C->rowptr[i]; e < C->rowptr[i + 1]; ++e) C->val[e] = acc[C->col[e]];
    }
    free(mark);
    free(acc);
}

/* Ac = P^T A P as R*(A*P) with R = P^T: IPRTransfer::restrict(M) -> ptap
 * (solvers/multilevel/transfer/utopia_IPRTransfer.hpp:183-187). */
void or_ptap(const or_csr *A, const or_csr *P, or_csr *Ac) {
    or_csr R, AP;
    or_transpose(P, &R);
    or_spgemm(A, P, &AP);
    or_spgemm(&R, &AP, Ac);
    or_csr_free(&R);
    or_csr_free(&AP);
}

/* Dense Cholesky A = L L^T, row-major, lower triangle; returns 0 or the
 * 1-based index of the first non-positive pivot. (Replaces the coarse LU of
 * Factorization<PETSC>, backend/petsc/solvers/linear/utopia_petsc_Factorization.hpp:13-84.) */
int or_cholesky(int32_t n, double *A) {
    for (int32_t j = 0; j < n; ++j) {
        double d = A[(int64_t)j * n + j];
        for (int32_t k = 0; k < j; ++k) d -= A[(int64_t)j * n + k] * A[(int64_t)j * n + k];
        if (!(d > 0.0)) return j + 1;
        d = sqrt(d);
        A[(int64_t)j * n + j] = d;
        /* rows below j are independent: each keeps its own sequential sum */
#pragma omp parallel for schedule(static) if (n - j > 256)
        for (int32_t i = j + 1; i < n; ++i) {
            double s = A[(int64_t)i * n + j];
            for (int32_t k = 0; k < j; ++k) s -= A[(int64_t)i * n + k] * A[(int64_t)j * n + k];
            A[(int64_t)i * n + j] = s / d;
        }
        for (int32_t k = j + 1; k < n; ++k) A[(int64_t)j * n + k] = 0.0;
    }
    return 0;
}

void or_cholesky_solve(int32_t n, const double *L, const double *b, double *x) {
    for (int32_t i = 0; i < n; ++i) {
        double s = b[i];
        for (int32_t k = 0; k < i; ++k) s -= L[(int64_t)i * n + k] * x[k];
        x[i] = s / L[(int64_t)i * n + i];
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int32_t k = i + 1; k < n; ++k) s -= L[(int64_t)k * n + i] * x[k];
        x[i] = s / L[(int64_t)i * n + i];
    }
}

/* ---------------------------------------------------- convergence semantics */

/* IterativeSolver::check_convergence (solvers/linear/utopia_IterativeSolver.hpp:148-191)
 * with norm_frequency 1. */
static int check_convergence(int64_t it, double g_norm, double r_norm, double s_norm, double atol, double rtol,
                             double stol, int64_t max_it, or_status *st) {
    int converged = 0, reason = 0;
    if (g_norm <= atol) converged = 1, reason = 2;  /* CONVERGED_FNORM_ABS */
    if (s_norm <= stol) converged = 1, reason = 4;  /* CONVERGED_SNORM_RELATIVE */
    if (r_norm <= rtol) converged = 1, reason = 3;  /* CONVERGED_FNORM_RELATIVE */
    if (it >= max_it) converged = 1, reason = -5;   /* DIVERGED_MAX_IT */
    if (converged && st) {
        st->iterates = it;
        st->gradient_norm = g_norm;
        st->reason = reason;
    }
    return converged;
}

/* ------------------------------------------------------------ CG (HOMEMADE) */

/* ConjugateGradient<M,V,HOMEMADE>::unpreconditioned_solve
 * (solvers/smoothers/utopia_ConjugateGradient_impl.hpp:90-192) */
static int cg_unprec(const or_csr *A, const double *b, double *x, double atol, double rtol, double stol,
                     int64_t max_it, or_status *st) {
    const int64_t n = A->nrows;
    double *r = vnew(n), *p = vnew(n), *q = vnew(n);
    int64_t it = 0;
    double rho = 1., rho_1 = 1., beta = 0., alpha = 1., r_norm;
    or_spmv(A, x, r);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    r_norm = or_norm2(n, r);
    it++;
    int converged = 0;
    while (!converged) {
        rho = or_dot(n, r, r);
        if (rho == 0.) {
            converged = 1;
            if (st) st->iterates = it, st->gradient_norm = r_norm, st->reason = 2;
            break;
        }
        if (it > 1) {
            beta = rho / rho_1;
            for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        } else {
            vcopy(n, r, p);
        }
        or_spmv(A, p, q);
        const double dot_pq = or_dot(n, p, q);
        if (dot_pq == 0.) {
            converged = 1;
            break;
        }
        alpha = rho / dot_pq;
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] = r[i] - alpha * q[i];
        rho_1 = rho;
        r_norm = or_norm2(n, r);
        converged = check_convergence(it, r_norm, 1, 1, atol, rtol, stol, max_it, st);
        it++;
    }
    free(r);
    free(p);
    free(q);
    if (st) st->converged = converged;
    return converged;
}

typedef void (*prec_fn)(void *ctx, const double *r, double *z);

/* ConjugateGradient<M,V,HOMEMADE>::preconditioned_solve
 * (solvers/smoothers/utopia_ConjugateGradient_impl.hpp:195-294) */
static int cg_prec(const or_csr *A, const double *b, double *x, prec_fn prec, void *pctx, double atol, double rtol,
                   double stol, int64_t max_it, or_status *st) {
    const int64_t n = A->nrows;
    double *r = vnew(n), *p = vnew(n), *Ap = vnew(n), *r_new = vnew(n), *z = vnew(n), *z_new = vnew(n);
    int64_t it = 0;
    double beta = 0., alpha = 1., r_norm = 9e9;
    int stop = 0;
    or_spmv(A, x, r);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    prec(pctx, r, z);
    vcopy(n, z, p);
    if (st) memset(st, 0, sizeof(*st));
    while (!stop) {
        or_spmv(A, p, Ap);
        alpha = or_dot(n, r, z) / or_dot(n, p, Ap);
        if (isinf(alpha) || isnan(alpha)) {
            stop = check_convergence(it, r_norm, 1, 1, atol, rtol, stol, max_it, st);
            break;
        }
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r_new[i] = r[i] - alpha * Ap[i];
        r_norm = or_norm2(n, r_new);
        if (r_norm <= atol) {
            stop = check_convergence(it, r_norm, 1, 1, atol, rtol, stol, max_it, st);
            break;
        }
        memset(z_new, 0, sizeof(double) * (size_t)n);
        prec(pctx, r_new, z_new);
        beta = or_dot(n, z_new, r_new) / or_dot(n, z, r);
        for (int64_t i = 0; i < n; ++i) p[i] = z_new[i] + beta * p[i];
        vcopy(n, r_new, r);
        vcopy(n, z_new, z);
        stop = check_convergence(it, r_norm, 1, 1, atol, rtol, stol, max_it, st);
        it++;
    }
    free(r);
    free(p);
    free(Ap);
    free(r_new);
    free(z);
    free(z_new);
    const int ok = r_norm <= atol;
    if (st) st->converged = ok, st->gradient_norm = r_norm;
    return ok;
}

/* InvDiagPreconditioner (solvers/preconditioners/utopia_TrivialPreconditioners.hpp:20-53):
 * d = pseudo_inv(diag(A)) with tol = machine epsilon, apply z = d .* r. */
typedef struct {
    int64_t n;
    double *d;
} invdiag;

static void invdiag_apply(void *c, const double *r, double *z) {
    invdiag *p = (invdiag *)c;
    for (int64_t i = 0; i < p->n; ++i) z[i] = p->d[i] * r[i];
}

static void diag_of(const or_csr *A, double *d) {
    for (int32_t i = 0; i < A->nrows; ++i) {
        d[i] = 0.0;
        for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j)
            if (A->col[j] == i) d[i] = A->val[j];
    }
}

/* ElementWisePseudoInverse (utopia_core/utopia_ElementWisePseudoInverse.hpp:20-56) */
static void pseudo_inv(int64_t n, double *d, double tol) {
    for (int64_t i = 0; i < n; ++i) d[i] = (fabs(d[i]) > tol) ? 1. / d[i] : 0.0;
}

int or_cg_solve(const or_csr *A, const double *b, double *x, int precond, double atol, double rtol, double stol,
                int64_t max_it, or_status *st) {
    if (!precond) return cg_unprec(A, b, x, atol, rtol, stol, max_it, st);
    invdiag pc;
    pc.n = A->nrows;
    pc.d = vnew(pc.n);
    diag_of(A, pc.d);
    pseudo_inv(pc.n, pc.d, DBL_EPSILON);
    const int ok = cg_prec(A, b, x, invdiag_apply, &pc, atol, rtol, stol, max_it, st);
    free(pc.d);
    return ok;
}

/* --------------------------------------------------- PowerMethod, Chebyshev */

/* PowerMethod::get_max_eig + compute_max_eig (solvers/eigensolvers/utopia_PowerMethod.hpp:95-146):
 * v_i = rand()/RAND_MAX + 1, v *= 1/|v|, then v = A v, lambda = |v|, v *= 1/lambda
 * until |lambda_old - lambda| < tol or it > max_it. */
double or_power_method(const or_csr *A, double tol, int64_t max_it) {
    const int64_t n = A->nrows;
    double *v = vnew(n), *w = vnew(n);
    for (int64_t i = 0; i < n; ++i) v[i] = ((double)rand() / (RAND_MAX)) + 1;
    const double s0 = 1. / or_norm2(n, v);
    for (int64_t i = 0; i < n; ++i) v[i] = s0 * v[i];
    int64_t it = 0;
    int converged = 0;
    double lambda = 0.0, lambda_old = 0.0;
    while (!converged) {
        or_spmv(A, v, w);
        vcopy(n, w, v);
        lambda = or_norm2(n, v);
        const double s = 1. / lambda;
        for (int64_t i = 0; i < n; ++i) v[i] = s * v[i];
        it = it + 1;
        converged = ((fabs(lambda_old - lambda) < tol) || it > max_it) ? 1 : 0;
        lambda_old = lambda;
    }
    free(v);
    free(w);
    return lambda;
}

/* Chebyshev3level::smooth -> solve (solvers/smoothers/utopia_Chebyshev3level.hpp:109-178)
 * with norm_frequency 0 and max_it = nu sweeps. */
void or_chebyshev_smooth(const or_csr *A, const double *b, double *x, int nu, double eig_max, double eig_min) {
    const int64_t n = A->nrows;
    double *r = vnew(n), *p = vnew(n), *help = vnew(n);
    const double avg_eig = (eig_max + eig_min) / 2.0;
    const double diff_eig = (eig_max - eig_min) / 2.0;
    or_spmv(A, x, help);
    for (int64_t i = 0; i < n; ++i) r[i] = help[i] - b[i];
    int64_t it = 0;
    double alpha = 0.0, beta = 0.0;
    for (;;) {
        if (it == 0) {
            for (int64_t i = 0; i < n; ++i) p[i] = -1.0 * r[i];
            alpha = 1.0 / avg_eig;
        } else if (it == 1) {
            const double dea = (diff_eig * alpha);
            beta = 0.5 * dea * dea;
            alpha = 1.0 / (avg_eig - (beta / alpha));
            for (int64_t i = 0; i < n; ++i) p[i] = -1.0 * r[i] + (beta * p[i]);
        } else {
            const double dea = (diff_eig * alpha);
            beta = 0.25 * dea * dea;
            alpha = 1.0 / (avg_eig - (beta / alpha));
            for (int64_t i = 0; i < n; ++i) p[i] = -1.0 * r[i] + (beta * p[i]);
        }
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + (alpha * p[i]);
        it++;
        if (it >= nu) break;
        or_spmv(A, p, help);
        for (int64_t i = 0; i < n; ++i) r[i] = r[i] + (alpha * help[i]);
    }
    free(r);
    free(p);
    free(help);
}

/* ---------------------------------------------------------------- Multigrid */

#define OR_MAXL 16

struct or_mg {
    int L;
    or_csr I[OR_MAXL]; /* I[l]: level l -> l+1 (coarse to fine), l < L-1 */
    or_csr R[OR_MAXL]; /* R[l] = I[l]^T */
    or_csr A[OR_MAXL]; /* level 0 coarsest */
    double *mask[OR_MAXL];
    double eig_max[OR_MAXL], eig_min[OR_MAXL];
    double *dinv[OR_MAXL];
    double *chol;
    int smoother, pre, post, mg_type, masks, fix_semi;
    int cycle_type, line_search; /* MultiLevelBase::cycle_type (1 or 3), Multigrid::use_line_search */
    int64_t v_cycle_repetition;
    or_csr LU[OR_MAXL];          /* PointJacobi: A - D (smoother 3) */
    double *r[OR_MAXL], *c[OR_MAXL], *cI[OR_MAXL], *rR[OR_MAXL], *f[OR_MAXL];
    int32_t *color[OR_MAXL]; /* multicolour Gauss-Seidel: colour of each row (smoother 2) */
    int32_t ncolors[OR_MAXL];
};

or_mg *or_mg_create(int nlevels, const or_csr *const *interp, int smoother, int pre_steps, int post_steps,
                    int mg_type, int masks, int fix_semidefinite) {
    if (nlevels < 2 || nlevels > OR_MAXL) return NULL;
    or_mg *mg = (or_mg *)calloc(1, sizeof(or_mg));
    mg->L = nlevels;
    for (int l = 0; l < nlevels - 1; ++l) {
        csr_copy(interp[l], &mg->I[l]);
        or_transpose(&mg->I[l], &mg->R[l]); /* IPRTransfer(I): R = transpose(I) (utopia_IPRTransfer.hpp:28-39) */
    }
    mg->smoother = smoother;
    mg->pre = pre_steps;
    mg->post = post_steps;
    mg->mg_type = mg_type;
    mg->masks = masks;
    mg->fix_semi = fix_semidefinite;
    mg->cycle_type = 1;          /* MULTIPLICATIVE_CYCLE (utopia_MultiLevelBase.hpp:39) */
    mg->v_cycle_repetition = 10; /* (utopia_MultiLevelBase.hpp:40) */
    mg->line_search = 0;         /* (utopia_Multigrid.hpp:74) */
    return mg;
}

void or_mg_set_cycle(or_mg *mg, int cycle_type, int64_t v_cycle_repetition, int line_search) {
    mg->cycle_type = cycle_type;
    mg->v_cycle_repetition = v_cycle_repetition;
    mg->line_search = line_search;
}

int or_mg_nlevels(const or_mg *mg) { return mg->L; }
const or_csr *or_mg_level(const or_mg *mg, int l) { return &mg->A[l]; }
const double *or_mg_mask(const or_mg *mg, int l) { return mg->mask[l]; }
void or_mg_eigs(const or_mg *mg, int l, double *emax, double *emin) {
    *emax = mg->eig_max[l];
    *emin = mg->eig_min[l];
}
void or_mg_set_eigs(or_mg *mg, int l, double emax, double emin) {
    mg->eig_max[l] = emax;
    mg->eig_min[l] = emin;
}

void or_mg_set_colors(or_mg *mg, int l, const int32_t *color, int32_t ncolors) {
    const int64_t n = mg->A[l].nrows;
    free(mg->color[l]);
    mg->color[l] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    memcpy(mg->color[l], color, sizeof(int32_t) * (size_t)n);
    mg->ncolors[l] = ncolors;
}

/* Jones-Plassmann colouring with the smallest available colour: in each round
 * every uncoloured row whose (hash, index) exceeds that of all its uncoloured
 * neighbours takes the smallest colour none of its coloured neighbours has.
 * hash(i) = splitmix64 finaliser of (i xor seed). Needs a structurally
 * symmetric pattern. (The colouring that replaces the reference's
 * lexicographic Gauss-Seidel order; no reference counterpart.) */
static uint64_t jp_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_jp_color(const or_csr *A, uint64_t seed, int32_t *color, int32_t *ncolors) {
    const int32_t n = A->nrows;
    int32_t *next = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; ++i) color[i] = -1;
    int32_t left = n, maxc = -1;
    while (left > 0) {
        for (int32_t i = 0; i < n; ++i) {
            next[i] = color[i];
            if (color[i] >= 0) continue;
            const uint64_t hi = jp_mix((uint64_t)i ^ seed);
            int best = 1;
            for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1] && best; ++e) {
                const int32_t j = A->col[e];
                if (j == i || color[j] >= 0) continue;
                const uint64_t hj = jp_mix((uint64_t)j ^ seed);
                if (hj > hi || (hj == hi && j > i)) best = 0;
            }
            if (!best) continue;
            int32_t c = 0;
            for (;;) {
                int used = 0;
                for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1] && !used; ++e)
                    used = (A->col[e] != i) && (color[A->col[e]] == c);
                if (!used) break;
                ++c;
            }
            next[i] = c;
        }
        for (int32_t i = 0; i < n; ++i) {
            if (color[i] < 0 && next[i] >= 0) {
                --left;
                if (next[i] > maxc) maxc = next[i];
            }
            color[i] = next[i];
        }
    }
    *ncolors = maxc + 1;
    free(next);
}

/* GaussSeidel::smooth with local_sweeps (solvers/smoothers/utopia_GaussSeidel.hpp:45-57,133-183):
 * per sweep r = rhs - A x; c = 0; one forward then (symmetric) one backward pass,
 * c_i = d_inv_i (r_i - sum_{j != i} a_ij c_j); x += c. The row order of a pass is
 * the colour order (colour 0 first going forward, last colour first going
 * backward); with color[i] = i it is the reference's lexicographic order. */
void or_gs_smooth(const or_csr *A, const double *dinv, const int32_t *color, int32_t ncolors, const double *rhs,
                  double *x, int sweeps, int symmetric) {
    const int64_t n = A->nrows;
    double *r = vnew(n), *c = vnew(n);
    for (int s = 0; s < sweeps; ++s) {
        or_spmv(A, x, r);
        for (int64_t i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
        for (int64_t i = 0; i < n; ++i) c[i] = 0.0;
        for (int pass = 0; pass < (symmetric ? 2 : 1); ++pass) {
            for (int32_t t = 0; t < ncolors; ++t) {
                const int32_t k = pass == 0 ? t : ncolors - 1 - t;
                for (int64_t i = 0; i < n; ++i) {
                    if (color[i] != k) continue;
                    double v = r[i];
                    for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e)
                        if (A->col[e] != i) v -= A->val[e] * c[A->col[e]];
                    c[i] = dinv[i] * v;
                }
            }
        }
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + c[i];
    }
    free(r);
    free(c);
}

/* LinearMultiLevel::fix_semidefinite_operator (solvers/multilevel/utopia_LinearMultiLevel.hpp:89-113) */
static void fix_semidefinite(or_csr *A) {
    for (int32_t i = 0; i < A->nrows; ++i)
        for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j)
            if (A->col[j] == i && fabs(A->val[j]) < 1e-12) A->val[j] += 1.0;
}

/* MultiLevelMask::generate_masks (solvers/multilevel/utopia_MultiLevelMask.hpp:39-62,69-115)
 * with IPRTransfer::boolean_restrict_or (utopia_IPRTransfer.hpp:144-173). */
static void generate_masks(or_mg *mg, const or_csr *A) {
    const double tol = DBL_EPSILON * 1e6;
    const int L = mg->L;
    for (int l = 0; l < L; ++l) {
        free(mg->mask[l]);
        mg->mask[l] = NULL;
    }
    /* finest: off_value 1 everywhere, on_value 0 where the row has an off-diagonal > tol */
    double *m = vnew(A->nrows);
    for (int32_t i = 0; i < A->nrows; ++i) {
        m[i] = 1.0;
        for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j)
            if (A->col[j] != i && fabs(A->val[j]) > tol) m[i] = 0.0;
    }
    mg->mask[L - 1] = m;
    for (int l = L - 1; l > 0; --l) {
        const or_csr *R = &mg->R[l - 1];
        double *mc = vnew(R->nrows);
        for (int32_t i = 0; i < R->nrows; ++i) {
            double s = 0.0;
            for (int32_t j = R->rowptr[i]; j < R->rowptr[i + 1]; ++j)
                s += (double)(fabs(R->val[j]) > tol) * mg->mask[l][R->col[j]];
            mc[i] = s > 1. ? 1. : s;
        }
        mg->mask[l - 1] = mc;
    }
    for (int l = 0; l < L; ++l) {
        const int64_t n = (l == L - 1) ? A->nrows : mg->R[l].nrows;
        for (int64_t i = 0; i < n; ++i) mg->mask[l][i] = fabs(mg->mask[l][i] + (-1.0));
    }
}

/* Multigrid::update (solvers/multilevel/utopia_Multigrid.hpp:129-153) with
 * LinearMultiLevel::galerkin_assembly (utopia_LinearMultiLevel.hpp:201-226). */
int or_mg_update(or_mg *mg, const or_csr *A) {
    const int L = mg->L;
    for (int l = 0; l < L; ++l) or_csr_free(&mg->A[l]);
    if (mg->masks) generate_masks(mg, A);
    /* levels from fine (index L-1) to coarse */
    csr_copy(A, &mg->A[L - 1]);
    for (int l = L - 1; l > 0; --l) {
        or_ptap(&mg->A[l], &mg->I[l - 1], &mg->A[l - 1]);
        if (mg->fix_semi) fix_semidefinite(&mg->A[l - 1]);
    }
    for (int l = 1; l < L; ++l) {
        const int64_t n = mg->A[l].nrows;
        if (mg->smoother == 0) {
            /* Chebyshev3level::update -> init_eigs (utopia_Chebyshev3level.hpp:90-121) */
            const double lam = or_power_method(&mg->A[l], 1e-2, 30);
            mg->eig_max[l] = 1.2 * lam;
            mg->eig_min[l] = 0.06 * mg->eig_max[l];
        } else if (mg->smoother == 2) {
            /* GaussSeidel::init (utopia_GaussSeidel.hpp:203-214): d_inv = 1/d (0 kept for d = 0) */
            free(mg->dinv[l]);
            mg->dinv[l] = vnew(n);
            diag_of(&mg->A[l], mg->dinv[l]);
            for (int64_t i = 0; i < n; ++i) mg->dinv[l][i] = mg->dinv[l][i] != 0.0 ? 1.0 / mg->dinv[l][i] : 0.0;
        } else if (mg->smoother == 3) {
            /* PointJacobi::update (solvers/smoothers/utopia_PointJacobi.hpp:84-101): d = diag(A);
             * LU = A - diag(d) (MatAXPY: the diagonal entries stay, as A_ii - d_i = 0);
             * d_inv = 1. / d evaluated as VecReciprocal (zeros kept,
             * backend/petsc/utopia_petsc_Vector.hpp:711-720); check_indef: zeros become ones */
            free(mg->dinv[l]);
            mg->dinv[l] = vnew(n);
            diag_of(&mg->A[l], mg->dinv[l]);
            or_csr_free(&mg->LU[l]);
            csr_copy(&mg->A[l], &mg->LU[l]);
            for (int32_t i = 0; i < mg->LU[l].nrows; ++i)
                for (int32_t j = mg->LU[l].rowptr[i]; j < mg->LU[l].rowptr[i + 1]; ++j)
                    if (mg->LU[l].col[j] == i) mg->LU[l].val[j] = mg->LU[l].val[j] + (-1.0 * mg->dinv[l][i]);
            for (int64_t i = 0; i < n; ++i) {
                const double d = mg->dinv[l][i];
                mg->dinv[l][i] = d != 0.0 ? 1.0 / d : d;
                if (mg->dinv[l][i] == 0) mg->dinv[l][i] = 1;
            }
        } else {
            /* Jacobi::update (solvers/smoothers/utopia_Jacobi.hpp:99-110) */
            free(mg->dinv[l]);
            mg->dinv[l] = vnew(n);
            diag_of(&mg->A[l], mg->dinv[l]);
            pseudo_inv(n, mg->dinv[l], 0.0);
        }
    }
    const int32_t n0 = mg->A[0].nrows;
    free(mg->chol);
    mg->chol = (double *)calloc((size_t)n0 * n0, sizeof(double));
    for (int32_t i = 0; i < n0; ++i)
        for (int32_t j = mg->A[0].rowptr[i]; j < mg->A[0].rowptr[i + 1]; ++j)
            mg->chol[(int64_t)i * n0 + mg->A[0].col[j]] += mg->A[0].val[j];
    const int info = or_cholesky(n0, mg->chol);
    for (int l = 0; l < L; ++l) {
        const int64_t n = mg->A[l].nrows;
        free(mg->r[l]);
        free(mg->c[l]);
        free(mg->cI[l]);
        free(mg->rR[l]);
        free(mg->f[l]);
        mg->r[l] = vnew(n);
        mg->c[l] = vnew(n);
        mg->cI[l] = vnew(n);
        mg->rR[l] = vnew(n);
        mg->f[l] = vnew(n);
    }
    return info;
}

/* PointJacobi::smooth -> sweep (solvers/smoothers/utopia_PointJacobi.hpp:61-66,132-142):
 * r = rhs - LU x; x = d_inv .* r */
static void point_jacobi_smooth(const or_csr *LU, const double *dinv, const double *rhs, double *x, int sweeps) {
    const int64_t n = LU->nrows;
    double *r = vnew(n);
    for (int s = 0; s < sweeps; ++s) {
        or_spmv(LU, x, r);
        for (int64_t i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
        for (int64_t i = 0; i < n; ++i) x[i] = dinv[i] * r[i];
    }
    free(r);
}

/* Jacobi::smooth (solvers/smoothers/utopia_Jacobi.hpp:89-140): r = rhs - A x; x += d_inv .* r */
static void jacobi_smooth(const or_csr *A, const double *dinv, const double *rhs, double *x, int sweeps) {
    const int64_t n = A->nrows;
    double *r = vnew(n);
    for (int s = 0; s < sweeps; ++s) {
        or_spmv(A, x, r);
        for (int64_t i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + dinv[i] * r[i];
    }
    free(r);
}

/* Jacobi / PointJacobi smoothing on a matrix (update + smooth), for pins */
void or_jacobi_smooth(const or_csr *A, const double *rhs, double *x, int sweeps) {
    const int64_t n = A->nrows;
    double *d = vnew(n);
    diag_of(A, d);
    pseudo_inv(n, d, 0.0);
    jacobi_smooth(A, d, rhs, x, sweeps);
    free(d);
}

void or_point_jacobi_smooth(const or_csr *A, const double *rhs, double *x, int sweeps) {
    const int64_t n = A->nrows;
    double *d = vnew(n);
    or_csr LU;
    diag_of(A, d);
    csr_copy(A, &LU);
    for (int32_t i = 0; i < LU.nrows; ++i)
        for (int32_t j = LU.rowptr[i]; j < LU.rowptr[i + 1]; ++j)
            if (LU.col[j] == i) LU.val[j] = LU.val[j] + (-1.0 * d[i]);
    for (int64_t i = 0; i < n; ++i) {
        const double v = d[i];
        d[i] = v != 0.0 ? 1.0 / v : v;
        if (d[i] == 0) d[i] = 1;
    }
    point_jacobi_smooth(&LU, d, rhs, x, sweeps);
    or_csr_free(&LU);
    free(d);
}

static void smoothing(or_mg *mg, int l, const double *rhs, double *x, int nu) {
    if (mg->smoother == 0)
        or_chebyshev_smooth(&mg->A[l], rhs, x, nu, mg->eig_max[l], mg->eig_min[l]);
    else if (mg->smoother == 2)
        or_gs_smooth(&mg->A[l], mg->dinv[l], mg->color[l], mg->ncolors[l], rhs, x, nu, 1);
    else if (mg->smoother == 3)
        point_jacobi_smooth(&mg->LU[l], mg->dinv[l], rhs, x, nu);
    else
        jacobi_smooth(&mg->A[l], mg->dinv[l], rhs, x, nu);
}

/* Multigrid::standard_cycle (solvers/multilevel/utopia_Multigrid.hpp:252-367) */
static void standard_cycle(or_mg *mg, int l) {
    const int64_t n = mg->A[l].nrows;
    double *r = mg->r[l], *c = mg->c[l], *cI = mg->cI[l], *rR = mg->rR[l];
    memset(c, 0, sizeof(double) * (size_t)n);
    if (l == 0) {
        or_cholesky_solve((int32_t)n, mg->chol, r, c);
        return;
    }
    for (int k = 0; k < mg->mg_type; ++k) {
        smoothing(mg, l, r, c, mg->pre);
        or_spmv(&mg->A[l], c, rR);
        for (int64_t i = 0; i < n; ++i) rR[i] = r[i] - rR[i];
        or_spmv(&mg->R[l - 1], rR, mg->r[l - 1]);
        if (mg->masks) {
            const int64_t nc = mg->A[l - 1].nrows;
            for (int64_t i = 0; i < nc; ++i) mg->r[l - 1][i] = mg->mask[l - 1][i] * mg->r[l - 1][i];
        }
        standard_cycle(mg, l - 1);
        or_spmv(&mg->I[l - 1], mg->c[l - 1], cI);
        if (mg->line_search) {
            /* use_line_search, symmetric form (utopia_Multigrid.hpp:324-341):
             * alpha = dot(c_I, r_R) / dot(A c_I, c_I); c += alpha c_I, or c += c_I if alpha <= 0 */
            double *Ac = vnew(n);
            or_spmv(&mg->A[l], cI, Ac);
            const double alpha = or_dot(n, cI, rR) / or_dot(n, Ac, cI);
            free(Ac);
            if (alpha <= 0.)
                for (int64_t i = 0; i < n; ++i) c[i] = c[i] + cI[i];
            else
                for (int64_t i = 0; i < n; ++i) c[i] = c[i] + alpha * cI[i];
        } else {
            for (int64_t i = 0; i < n; ++i) c[i] = c[i] + cI[i];
        }
        smoothing(mg, l, r, c, mg->post);
    }
}

/* Multigrid::full_cycle (solvers/multilevel/utopia_Multigrid.hpp:381-415): rhss[l] = r; restrict
 * and mask down to level 0; coarse solve; c[1] = I c[0]; per level i = 1..l-1: v_cycle_repetition
 * times (r[i] = rhss[i]; standard_cycle(i)), then c[i+1] = I c[i]; finally v_cycle_repetition
 * standard cycles on level l. */
static void full_cycle(or_mg *mg, int l) {
    double **rhss = mg->f;
    memcpy(rhss[l], mg->r[l], sizeof(double) * (size_t)mg->A[l].nrows);
    for (int i = l - 1; i >= 0; --i) {
        or_spmv(&mg->R[i], rhss[i + 1], rhss[i]);
        if (mg->masks)
            for (int64_t k = 0; k < mg->A[i].nrows; ++k) rhss[i][k] = mg->mask[i][k] * rhss[i][k];
    }
    memset(mg->c[0], 0, sizeof(double) * (size_t)mg->A[0].nrows);
    or_cholesky_solve(mg->A[0].nrows, mg->chol, rhss[0], mg->c[0]);
    or_spmv(&mg->I[0], mg->c[0], mg->c[1]);
    for (int i = 1; i < l; ++i) {
        for (int64_t j = 0; j < mg->v_cycle_repetition; ++j) {
            memcpy(mg->r[i], rhss[i], sizeof(double) * (size_t)mg->A[i].nrows);
            standard_cycle(mg, i);
        }
        or_spmv(&mg->I[i], mg->c[i], mg->c[i + 1]);
    }
    for (int64_t j = 0; j < mg->v_cycle_repetition; ++j) standard_cycle(mg, l);
}

/* Multigrid::apply (solvers/multilevel/utopia_Multigrid.hpp:162-238), MULTIPLICATIVE_CYCLE */
int or_mg_apply(or_mg *mg, const double *rhs, double *x, int64_t max_it, double atol, double rtol) {
    const int l = mg->L - 1;
    const int64_t n = mg->A[l].nrows;
    double *r = mg->r[l];
    or_spmv(&mg->A[l], x, r);
    for (int64_t i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
    double r_norm = or_norm2(n, r);
    const double r0_norm = r_norm;
    int64_t it = 1;
    int converged = 0;
    if (mg->cycle_type == 3) max_it = 1; /* FULL_CYCLE sets max_it(1) (utopia_Multigrid.hpp:191-193) */
    while (!converged) {
        if (mg->cycle_type == 3)
            full_cycle(mg, l);
        else
            standard_cycle(mg, l);
        for (int64_t i = 0; i < n; ++i) x[i] = x[i] + mg->c[l][i];
        or_spmv(&mg->A[l], x, r);
        for (int64_t i = 0; i < n; ++i) r[i] = rhs[i] - r[i];
        r_norm = or_norm2(n, r);
        const double rel_norm = r_norm / r0_norm;
        converged = check_convergence(it, r_norm, rel_norm, 1, atol, rtol, 1e-11, max_it, NULL);
        it++;
    }
    return 1;
}

typedef struct {
    or_mg *mg;
} mg_prec;

static void mg_prec_apply(void *c, const double *r, double *z) {
    mg_prec *p = (mg_prec *)c;
    /* the preconditioner keeps IterativeSolver defaults (atol = rtol = 1e-9) and max_it = 1 */
    or_mg_apply(p->mg, r, z, 1, 1e-9, 1e-9);
}

int or_cg_mg_solve(or_mg *mg, const double *b, double *x, double atol, int64_t max_it, or_status *st) {
    mg_prec pc = {mg};
    return cg_prec(&mg->A[mg->L - 1], b, x, mg_prec_apply, &pc, atol, 1e-9, 1e-11, max_it, st);
}

void or_mg_destroy(or_mg *mg) {
    if (!mg) return;
    for (int l = 0; l < OR_MAXL; ++l) {
        or_csr_free(&mg->I[l]);
        or_csr_free(&mg->R[l]);
        or_csr_free(&mg->A[l]);
        free(mg->mask[l]);
        free(mg->dinv[l]);
        free(mg->r[l]);
        free(mg->c[l]);
        free(mg->cI[l]);
        free(mg->rR[l]);
        free(mg->f[l]);
        free(mg->color[l]);
        or_csr_free(&mg->LU[l]);
    }
    free(mg->chol);
    free(mg);
}

/* -------------------------------------------------------------------- MPRGP */

/* get_fi (utopia_MPRGP.hpp:266-287) */
static void get_fi(int64_t n, const double *x, const double *g, const double *lb, const double *ub, double *fi) {
    for (int64_t i = 0; i < n; ++i) fi[i] = (lb[i] < x[i] && x[i] < ub[i]) ? g[i] : 0.0;
}

/* get_beta (utopia_MPRGP.hpp:324-348) */
static void get_beta(int64_t n, const double *x, const double *g, const double *lb, const double *ub, double *beta) {
    for (int64_t i = 0; i < n; ++i) {
        const double li = lb[i], ui = ub[i], xi = x[i], gi = g[i];
        beta[i] = (fabs(li - xi) < 1e-14) ? fmin(0.0, gi) : ((fabs(ui - xi) < 1e-14) ? fmax(0.0, gi) : 0.0);
    }
}

/* get_alpha_f (utopia_MPRGP.hpp:289-322): multi_min of the two step-length vectors */
static double get_alpha_f(int64_t n, const double *x, const double *p, const double *lb, const double *ub) {
    double m = DBL_MAX;
    for (int64_t i = 0; i < n; ++i) {
        const double h1 = (p[i] > 0) ? ((x[i] - lb[i]) / p[i]) : 1e15;
        const double h2 = (p[i] < 0) ? ((x[i] - ub[i]) / p[i]) : 1e15;
        if (h1 < m) m = h1;
        if (h2 < m) m = h2;
    }
    return m;
}

/* VariableBoundSolverInterface::project (solvers/nonlinear/constrained/utopia_VariableBoundSolverInterface.hpp:107-155) */
static void project(int64_t n, const double *x, const double *lb, const double *ub, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        const double li = lb[i], ui = ub[i], xi = x[i];
        out[i] = (li >= xi) ? li : ((ui <= xi) ? ui : xi);
    }
}

/* MPRGP::power_method (utopia_MPRGP.hpp:351-387) */
static double mprgp_power(const or_csr *A, double eps, int64_t max_it, double *h1, double *h2, double *fi) {
    const int64_t n = A->nrows;
    for (int64_t i = 0; i < n; ++i) h2[i] = 1.0;
    int64_t it = 0;
    int converged = 0;
    double gnorm, lambda = 0.0, lambda_old;
    while (!converged) {
        vcopy(n, h2, h1);
        or_spmv(A, h1, h2);
        const double s = 1.0 / or_norm2(n, h2);
        for (int64_t i = 0; i < n; ++i) h2[i] = s * h2[i];
        lambda_old = lambda;
        or_spmv(A, h2, h1);
        lambda = or_dot(n, h2, h1);
        for (int64_t i = 0; i < n; ++i) fi[i] = h2[i] - h1[i];
        gnorm = or_norm2(n, fi);
        converged = ((gnorm < eps) || (fabs(lambda_old - lambda) < eps) || it > max_it) ? 1 : 0;
        it = it + 1;
    }
    return lambda;
}

/* MPRGP::solve, non-"hardik" variant without preconditioner (utopia_MPRGP.hpp:57-262) */
int or_mprgp_solve(const or_csr *A, const double *b, const double *lb_in, const double *ub_in, double *x, double atol,
                   double rtol, int64_t max_it, double eps_eig_est, int64_t pm_max_it, or_status *st) {
    const int64_t n = A->nrows;
    double *lb = vnew(n), *ub = vnew(n);
    for (int64_t i = 0; i < n; ++i) {
        lb[i] = lb_in ? lb_in[i] : -DBL_MAX; /* BoxConstraints::fill_empty_bounds (utopia_BoxConstraints.hpp:281-300) */
        ub[i] = ub_in ? ub_in[i] : DBL_MAX;
    }
    double *fi = vnew(n), *beta = vnew(n), *gp = vnew(n), *p = vnew(n), *y = vnew(n), *Ap = vnew(n),
           *Abeta = vnew(n), *Ax = vnew(n), *g = vnew(n), *h1 = vnew(n), *h2 = vnew(n);
    const double gamma = 1.0;
    const double alpha_bar = 1.95 / mprgp_power(A, eps_eig_est, pm_max_it, h1, h2, fi);
    double pAp, beta_beta, fi_fi, gp_dot, g_betta, beta_Abeta;
    int64_t it = 0;
    int converged = 0, ret = 0;
    double gnorm = -1, gnorm0 = -1, alpha_cg, alpha_f, beta_sc;
    if (st) memset(st, 0, sizeof(*st));
    project(n, x, lb, ub, x);
    or_spmv(A, x, Ax);
    for (int64_t i = 0; i < n; ++i) g[i] = Ax[i] - b[i];
    get_fi(n, x, g, lb, ub, fi);
    get_beta(n, x, g, lb, ub, beta);
    for (int64_t i = 0; i < n; ++i) gp[i] = fi[i] + beta[i];
    vcopy(n, fi, p);
    beta_beta = or_dot(n, beta, beta);
    fi_fi = or_dot(n, fi, fi);
    while (!converged) {
        if (beta_beta <= (gamma * gamma * fi_fi)) {
            or_spmv(A, p, Ap);
            pAp = or_dot(n, p, Ap);
            gp_dot = or_dot(n, g, p);
            if (pAp <= 0.0) {
                ret = 1;
                goto done;
            }
            alpha_cg = gp_dot / pAp;
            alpha_f = get_alpha_f(n, x, p, lb, ub);
            for (int64_t i = 0; i < n; ++i) y[i] = x[i] - alpha_cg * p[i];
            if (alpha_cg <= alpha_f) {
                vcopy(n, y, x);
                for (int64_t i = 0; i < n; ++i) g[i] = g[i] - alpha_cg * Ap[i];
                get_fi(n, x, g, lb, ub, fi);
                beta_sc = or_dot(n, fi, Ap) / pAp;
                for (int64_t i = 0; i < n; ++i) p[i] = fi[i] - beta_sc * p[i];
            } else {
                for (int64_t i = 0; i < n; ++i) x[i] = x[i] - alpha_f * p[i];
                for (int64_t i = 0; i < n; ++i) g[i] = g[i] - alpha_f * Ap[i];
                get_fi(n, x, g, lb, ub, fi);
                for (int64_t i = 0; i < n; ++i) h1[i] = x[i] - (alpha_bar * fi[i]);
                project(n, h1, lb, ub, x);
                or_spmv(A, x, Ax);
                for (int64_t i = 0; i < n; ++i) g[i] = Ax[i] - b[i];
                get_fi(n, x, g, lb, ub, p);
            }
        } else {
            or_spmv(A, beta, Abeta);
            g_betta = or_dot(n, g, beta);
            beta_Abeta = or_dot(n, beta, Abeta);
            if (beta_Abeta <= 0.0) {
                ret = 1;
                goto done;
            }
            alpha_cg = g_betta / beta_Abeta;
            for (int64_t i = 0; i < n; ++i) x[i] = x[i] - alpha_cg * beta[i];
            for (int64_t i = 0; i < n; ++i) g[i] = g[i] - alpha_cg * Abeta[i];
            get_fi(n, x, g, lb, ub, p);
        }
        get_fi(n, x, g, lb, ub, fi);
        get_beta(n, x, g, lb, ub, beta);
        for (int64_t i = 0; i < n; ++i) gp[i] = fi[i] + beta[i];
        beta_beta = or_dot(n, beta, beta);
        fi_fi = or_dot(n, fi, fi);
        gnorm = sqrt(or_dot(n, gp, gp));
        if (it == 0) gnorm0 = gnorm;
        it++;
        converged = check_convergence(it, gnorm, gnorm / gnorm0, 1, atol, rtol, 1e-11, max_it, st);
    }
    ret = converged;
done:
    if (st) st->converged = converged, st->gradient_norm = gnorm;
    free(lb), free(ub), free(fi), free(beta), free(gp), free(p), free(y), free(Ap), free(Abeta), free(Ax), free(g),
        free(h1), free(h2);
    return ret;
}

/* ------------------------------------------------------ ProjectedGaussSeidel */

/* ------------------------------------------------------- Agglomerate (AMG) */

/* Agglomerate::create_transfer_matrix (solvers/multilevel/agglomerate/utopia_Agglomerate.hpp:50-178),
 * use_additive_correction off: rows with one stored value get no aggregate (-2); a row not yet
 * aggregated opens aggregate k and claims up to bmax rows j with weight*max_{j!=i}|a_ij| < |a_ij|
 * that are not yet aggregated, in storage order; claiming none, it takes the aggregate of the
 * largest |a_ij| of its row (diagonal included) instead; rows still unassigned become singletons.
 * P(i, parent_i) = 1 where max_{j!=i}|a_ij| != 0, else 0. Returns the number of aggregates. */
int32_t or_agglomerate(const or_csr *A, int64_t bmax, double weight, or_csr *P) {
    const int32_t n = A->nrows;
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    double *amax = vnew(n);
    for (int32_t i = 0; i < n; ++i) parent[i] = (A->rowptr[i + 1] - A->rowptr[i] == 1) ? -2 : -1;
    for (int32_t i = 0; i < n; ++i)
        for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e)
            if (A->col[e] != i && fabs(A->val[e]) > amax[i]) amax[i] = fabs(A->val[e]);
    int32_t k = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (parent[i] != -1) continue;
        parent[i] = k;
        int64_t claimed = 0;
        for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1] && claimed < bmax; ++e) {
            const int32_t j = A->col[e];
            if (parent[j] == -1 && fabs(A->val[e]) > weight * amax[i]) {
                parent[j] = k;
                ++claimed;
            }
        }
        if (claimed) {
            ++k;
        } else {
            int32_t jmax = -1;
            double vmax = 0;
            for (int32_t e = A->rowptr[i]; e < A->rowptr[i + 1]; ++e)
                if (fabs(A->val[e]) > vmax) vmax = fabs(A->val[e]), jmax = A->col[e];
            parent[i] = (jmax >= 0 && jmax != i) ? parent[jmax] : -1;
        }
    }
    for (int32_t i = 0; i < n; ++i)
        if (parent[i] == -1) parent[i] = k++;
    P->nrows = n;
    P->ncols = k;
    P->rowptr = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    P->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    P->val = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int64_t z = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (parent[i] >= 0) {
            P->col[z] = parent[i];
            P->val[z] = amax[i] != 0 ? 1.0 : 0.0;
            ++z;
        }
        P->rowptr[i + 1] = (int32_t)z;
    }
    P->nnz = z;
    free(parent);
    free(amax);
    return k;
}

/* ---------------------------------------------------- MonotoneMultigrid */

/* One projected Gauss-Seidel step in colour order (ProjectedGaussSeidel::step,
 * backend/petsc/.../utopia_petsc_ProjectedGaussSeidel.cpp:331-412, rows visited
 * colour by colour as the GPU's multicolour sweep): r = b - A x; c = 0; symmetric
 * projected sweeps with bounds lb - x, ub - x; x += c. */
static void pgs_colored_step(const or_csr *A, const double *dinv, const int32_t *color, int32_t ncolors,
                             const double *b, const double *lb, const double *ub, int constrained, double *x,
                             double *r, double *c) {
    const int32_t n = A->nrows;
    or_spmv(A, x, r);
    for (int32_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    memset(c, 0, sizeof(double) * (size_t)n);
    for (int pass = 0; pass < 2; ++pass)
        for (int32_t t = 0; t < ncolors; ++t) {
            const int32_t k = pass == 0 ? t : ncolors - 1 - t;
            for (int32_t i = 0; i < n; ++i) {
                if (color[i] != k) continue;
                const double lbl = lb[i] - x[i], ubl = ub[i] - x[i];
                double val = r[i];
                if (constrained && dinv[i] <= 0.0) {
                    c[i] = (val == 0.0) ? 0.0 : ((val > 0.0) ? lbl : ubl);
                } else {
                    for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j)
                        if (A->col[j] != i && A->val[j] != 0.0) val -= A->val[j] * c[A->col[j]];
                    /* unconstrained_step (:414-455): no projection */
                    c[i] = constrained ? fmax(lbl, fmin(dinv[i] * val, ubl)) : dinv[i] * val;
                }
            }
        }
    for (int32_t i = 0; i < n; ++i) x[i] = x[i] + c[i];
}

typedef struct {
    int nlevels;
    const or_csr *const *interp;
    int pre, post;
    or_csr It, Rt, AH;
    or_mg *mg;   /* coarse-space Multigrid (nlevels > 2) */
    double *chol; /* dense Cholesky of AH (nlevels == 2) */
} mmg_state;

/* galerkin_assembly (utopia_MonotoneMultigrid.hpp:217-231) */
static void mmg_galerkin(mmg_state *m, const or_csr *A) {
    or_csr_free(&m->Rt);
    or_csr_free(&m->AH);
    or_transpose(&m->It, &m->Rt);
    or_ptap(A, &m->It, &m->AH);
    fix_semidefinite(&m->AH);
    if (m->nlevels > 2) {
        if (m->mg) or_mg_destroy(m->mg);
        /* Multigrid(GaussSeidel, coarse solver): masks off, semidefinite fix on (:59-62) */
        m->mg = or_mg_create(m->nlevels - 1, m->interp, 2, m->pre, m->post, 1, 0, 1);
        or_mg_update(m->mg, &m->AH);
        for (int l = 1; l < m->nlevels - 1; ++l) {
            const or_csr *G = or_mg_level(m->mg, l);
            int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(G->nrows > 0 ? G->nrows : 1));
            int32_t nc = 0;
            or_jp_color(G, 0x5eedull, col, &nc);
            or_mg_set_colors(m->mg, l, col, nc);
            free(col);
        }
    } else {
        const int32_t nh = m->AH.nrows;
        free(m->chol);
        m->chol = (double *)calloc((size_t)nh * nh + 1, sizeof(double));
        for (int32_t i = 0; i < nh; ++i)
            for (int32_t j = m->AH.rowptr[i]; j < m->AH.rowptr[i + 1]; ++j)
                m->chol[(int64_t)i * nh + m->AH.col[j]] += m->AH.val[j];
        or_cholesky(nh, m->chol);
    }
}

/* MonotoneMultigrid::apply / step (solvers/multilevel/utopia_MonotoneMultigrid.hpp:158-318)
 * with ProjectedGaussSeidel as the QP smoother, ActiveSet tol 0, the finest
 * interpolation truncated (IPTruncatedTransfer), coarse_space_it Multigrid cycles
 * (GaussSeidel smoothing, dense Cholesky coarse solve, IterativeSolver tolerances
 * 1e-9) or, with two levels, the coarse solve. interp: nlevels-1 interpolations,
 * coarse to fine. */
int or_mmg_solve(const or_csr *A, const double *b, const double *lb_in, const double *ub_in, double *x,
                 int nlevels, const or_csr *const *interp, int pre, int post, int64_t coarse_space_it,
                 double active_tol, double atol, double rtol, double stol, int64_t max_it, or_status *st) {
    const int32_t n = A->nrows;
    const int has_bound = lb_in || ub_in;
    double *lb = vnew(n), *ub = vnew(n), *dinv = vnew(n), *r = vnew(n), *c = vnew(n), *ind = vnew(n),
           *x_old = vnew(n), *help = vnew(n), *grad = vnew(n);
    int32_t *color = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t ncolors = 0;
    for (int32_t i = 0; i < n; ++i) {
        lb[i] = lb_in ? lb_in[i] : -DBL_MAX;
        ub[i] = ub_in ? ub_in[i] : DBL_MAX;
    }
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j)
            if (A->col[j] == i) dinv[i] = (fabs(A->val[j]) > 0.0) ? (1 / A->val[j]) : 0.0;
    or_jp_color(A, 0x5eedull, color, &ncolors);
    mmg_state m;
    memset(&m, 0, sizeof(m));
    m.nlevels = nlevels;
    m.interp = interp;
    m.pre = pre;
    m.post = post;
    csr_copy(interp[nlevels - 2], &m.It);
    mmg_galerkin(&m, A);
    memset(st, 0, sizeof(*st));
    vcopy(n, x, x_old);
    int converged = 0;
    for (int64_t it = 1; it < max_it && !converged; ++it) {
        for (int s = 0; s < (has_bound ? pre : pre + 1); ++s)
            pgs_colored_step(A, dinv, color, ncolors, b, lb, ub, has_bound, x, r, c);
        if (has_bound) { /* ActiveSet::determine (utopia_ActiveSet.hpp:41-79) */
            int changed = 0;
            for (int32_t i = 0; i < n; ++i) {
                const double v = (x[i] <= lb[i] + active_tol || x[i] >= ub[i] - active_tol) ? 1.0 : 0.0;
                changed += v != ind[i];
                ind[i] = v;
            }
            if (changed) {
                or_csr_free(&m.It);
                csr_copy(interp[nlevels - 2], &m.It);
                for (int32_t i = 0; i < m.It.nrows; ++i)
                    if (ind[i] == 1.0)
                        for (int32_t j = m.It.rowptr[i]; j < m.It.rowptr[i + 1]; ++j) m.It.val[j] = 0.0;
                mmg_galerkin(&m, A);
            }
        }
        or_spmv(A, x, r);
        for (int32_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        const int32_t nh = m.Rt.nrows;
        double *rc = vnew(nh), *cc = vnew(nh), *cf = vnew(n);
        or_spmv(&m.Rt, r, rc);
        if (nlevels > 2) or_mg_apply(m.mg, rc, cc, coarse_space_it, 1e-9, 1e-9);
        else or_cholesky_solve(nh, m.chol, rc, cc);
        or_spmv(&m.It, cc, cf);
        for (int32_t i = 0; i < n; ++i) x[i] = x[i] + cf[i];
        free(rc), free(cc), free(cf);
        for (int s = 0; s < (has_bound ? post : post + 1); ++s)
            pgs_colored_step(A, dinv, color, ncolors, b, lb, ub, has_bound, x, r, c);
        for (int32_t i = 0; i < n; ++i) help[i] = x[i] - x_old[i];
        const double x_diff = or_norm2(n, help);
        or_spmv(A, x, grad);
        for (int32_t i = 0; i < n; ++i) grad[i] = grad[i] - b[i];
        for (int32_t i = 0; i < n; ++i) {
            const double v = x[i] - grad[i];
            help[i] = ((lb[i] >= v) ? lb[i] : ((ub[i] <= v) ? ub[i] : v)) - x[i];
        }
        const double g_norm = or_norm2(n, help);
        converged = check_convergence(it, g_norm, 1, x_diff, atol, rtol, stol, max_it, st);
        if (!converged) vcopy(n, x, x_old);
    }
    st->converged = converged;
    if (m.mg) or_mg_destroy(m.mg);
    or_csr_free(&m.It), or_csr_free(&m.Rt), or_csr_free(&m.AH);
    free(m.chol);
    free(lb), free(ub), free(dinv), free(r), free(c), free(ind), free(x_old), free(help), free(grad), free(color);
    return converged;
}

/* ProjectedConjugateGradient::apply
 * (solvers/nonlinear/constrained/quadratic_programming/utopia_ProjectedConjugateGradient.hpp:33-123).
 * The element loop reads z_k where the reference's x_view points (a view of zk). */
int or_pcg_solve(const or_csr *A, const double *b, const double *lb_in, const double *ub_in, double *x, double atol,
                 double rtol, double stol, int64_t max_it, or_status *st) {
    const int64_t n = A->nrows;
    double *lb = vnew(n), *ub = vnew(n), *x_old = vnew(n), *x_half = vnew(n), *uk = vnew(n), *wk = vnew(n),
           *zk = vnew(n), *pk = vnew(n), *Apk = vnew(n), *d = vnew(n);
    for (int64_t i = 0; i < n; ++i) {
        lb[i] = lb_in ? lb_in[i] : -DBL_MAX;
        ub[i] = ub_in ? ub_in[i] : DBL_MAX;
        zk[i] = 0.0;
    }
    memset(st, 0, sizeof(*st));
    memcpy(x_old, x, sizeof(double) * (size_t)n);
    or_spmv(A, x, uk);
    for (int64_t i = 0; i < n; ++i) uk[i] = b[i] - uk[i];
    for (int64_t i = 0; i < n; ++i) pk[i] = -uk[i];
    int converged = 0;
    int64_t it = 0;
    const double eps = DBL_EPSILON;
    while (!converged) {
        or_spmv(A, pk, Apk);
        const double pAp = or_dot(n, pk, Apk);
        const double alpha = or_dot(n, uk, pk) / pAp;
        if (alpha == 0. || isinf(alpha) || isnan(alpha)) break;
        for (int64_t i = 0; i < n; ++i) x_half[i] = x_old[i] + alpha * pk[i];
        for (int64_t i = 0; i < n; ++i) {
            const double v = x_half[i] < ub[i] ? x_half[i] : ub[i];
            x[i] = v > lb[i] ? v : lb[i];
        }
        or_spmv(A, x, uk);
        for (int64_t i = 0; i < n; ++i) uk[i] = b[i] - uk[i];
        for (int64_t i = 0; i < n; ++i) {
            const double elem = zk[i];
            double val = (fabs(elem - ub[i]) <= eps || fabs(elem - lb[i]) <= eps) ? fmax(uk[i], 0.0) : uk[i];
            zk[i] = (val == 0) ? fmax(pk[i], 0.0) : pk[i];
            wk[i] = val;
        }
        /* Apk = A * pk again in the reference: pk is unchanged, so Apk is */
        const double beta = or_dot(n, wk, Apk) / pAp;
        for (int64_t i = 0; i < n; ++i) pk[i] = wk[i] + beta * zk[i];
        if (it % 10 == 0 || isinf(beta) || isnan(beta)) {
            for (int64_t i = 0; i < n; ++i) d[i] = x_old[i] - x[i];
            converged = check_convergence(it, 1, 1, or_norm2(n, d), atol, rtol, stol, max_it, st);
        }
        ++it;
        if (converged) break;
        memcpy(x_old, x, sizeof(double) * (size_t)n);
    }
    st->converged = converged;
    free(lb), free(ub), free(x_old), free(x_half), free(uk), free(wk), free(zk), free(pk), free(Apk), free(d);
    return converged;
}

/* ProjectedGradient::solve / solve_unconstrained
 * (solvers/nonlinear/constrained/quadratic_programming/utopia_ProjectedGradient.hpp:54-213).
 * Unconstrained when both bounds are NULL. */
int or_pg_solve(const or_csr *A, const double *b, const double *lb_in, const double *ub_in, double *x, double atol,
                double rtol, double stol, int64_t max_it, or_status *st) {
    const int64_t n = A->nrows;
    const int bounded = lb_in || ub_in;
    double *lb = vnew(n), *ub = vnew(n), *x_old = vnew(n), *x_half = vnew(n), *u = vnew(n), *p = vnew(n),
           *Ap = vnew(n), *d = vnew(n);
    for (int64_t i = 0; i < n; ++i) {
        lb[i] = lb_in ? lb_in[i] : -DBL_MAX;
        ub[i] = ub_in ? ub_in[i] : DBL_MAX;
    }
    memset(st, 0, sizeof(*st));
    memcpy(x_old, x, sizeof(double) * (size_t)n);
    or_spmv(A, x, u);
    for (int64_t i = 0; i < n; ++i) u[i] = b[i] - u[i];
    memcpy(p, u, sizeof(double) * (size_t)n);
    double alpha = 1.;
    int converged = 0;
    int64_t it = 0;
    const double tol = 10. * DBL_EPSILON;
    while (!converged) {
        if (bounded) {
            for (int64_t i = 0; i < n; ++i) x_half[i] = x[i] + alpha * p[i];
            for (int64_t i = 0; i < n; ++i) {
                const double v = lb[i] > x_half[i] ? lb[i] : x_half[i];
                x[i] = ub[i] < v ? ub[i] : v;
            }
            or_spmv(A, x, u);
            for (int64_t i = 0; i < n; ++i) u[i] = b[i] - u[i];
            for (int64_t i = 0; i < n; ++i)
                p[i] = (fabs(x[i] - ub[i]) <= tol || fabs(x[i] - lb[i]) <= tol) ? 0.0 : u[i];
        } else {
            for (int64_t i = 0; i < n; ++i) x[i] = x[i] + alpha * p[i];
            or_spmv(A, x, u);
            for (int64_t i = 0; i < n; ++i) p[i] = b[i] - u[i];
        }
        if (it % 20 == 0) {
            for (int64_t i = 0; i < n; ++i) d[i] = x_old[i] - x[i];
            converged = check_convergence(it, 1, 1, or_norm2(n, d), atol, rtol, stol, max_it, st);
        }
        ++it;
        if (converged) break;
        memcpy(x_old, x, sizeof(double) * (size_t)n);
        or_spmv(A, p, Ap);
        alpha = (bounded ? or_dot(n, u, p) : or_dot(n, p, p)) / or_dot(n, p, Ap);
        if (isinf(alpha) || alpha == 0. || isnan(alpha)) {
            for (int64_t i = 0; i < n; ++i) d[i] = x_old[i] - x[i];
            converged = check_convergence(it, 1, 1, or_norm2(n, d), atol, rtol, stol, max_it, st);
            break;
        }
    }
    st->converged = converged;
    free(lb), free(ub), free(x_old), free(x_half), free(u), free(p), free(Ap), free(d);
    return converged;
}

/* ProjectedGaussSeidel<PetscMatrix,PetscVector,PETSC>::apply/step
 * (backend/petsc/solvers/nonlinear/constrained/quadratic_programming/utopia_petsc_ProjectedGaussSeidel.cpp:134-178,369-455)
 * with the scalar sweeper ProjectedScalarGaussSeidelSweep
 * (solvers/nonlinear/constrained/quadratic_programming/utopia_ProjectedGaussSeidelSweep.hpp:81-132,191-235),
 * lexicographic order, n_local_sweeps = 1, no line search. */
int or_pgs_solve(const or_csr *A, const double *b, const double *lb_in, const double *ub_in, double *x, double stol,
                 int64_t max_it, int symmetric, int64_t check_each, or_status *st) {
    const int32_t n = A->nrows;
    double *lb = vnew(n), *ub = vnew(n), *dinv = vnew(n), *r = vnew(n), *c = vnew(n), *lbl = vnew(n),
           *ubl = vnew(n), *x_old = vnew(n), *diff = vnew(n);
    for (int32_t i = 0; i < n; ++i) {
        lb[i] = lb_in ? lb_in[i] : -DBL_MAX;
        ub[i] = ub_in ? ub_in[i] : DBL_MAX;
    }
    /* init_from_local_matrix: d_inv and the off-diagonal nonzeros */
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j)
            if (A->col[j] == i) dinv[i] = (fabs(A->val[j]) > 0.0) ? (1 / A->val[j]) : 0.0;
    if (st) memset(st, 0, sizeof(*st));
    vcopy(n, x, x_old);
    int converged = 0;
    int64_t iteration = 0;
    while (!converged) {
        or_spmv(A, x, r);
        for (int32_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        for (int32_t i = 0; i < n; ++i) {
            ubl[i] = ub[i] - x[i];
            lbl[i] = lb[i] - x[i];
        }
        memset(c, 0, sizeof(double) * (size_t)n);
        for (int pass = 0; pass < (symmetric ? 2 : 1); ++pass) {
            for (int32_t t = 0; t < n; ++t) {
                const int32_t i = pass == 0 ? t : n - 1 - t;
                double val = r[i];
                if (dinv[i] <= 0.0) {
                    c[i] = (val == 0.0) ? 0.0 : ((val > 0.0) ? lbl[i] : ubl[i]);
                } else {
                    for (int32_t j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j)
                        if (A->col[j] != i && A->val[j] != 0.0) val -= A->val[j] * c[A->col[j]];
                    c[i] = fmax(lbl[i], fmin(dinv[i] * val, ubl[i]));
                }
            }
        }
        for (int32_t i = 0; i < n; ++i) x[i] = x[i] + c[i];
        if (iteration % check_each == 0) {
            for (int32_t i = 0; i < n; ++i) diff[i] = x[i] - x_old[i];
            const double d = or_norm2(n, diff);
            converged = check_convergence(iteration, 1, 1, d, 1e-9, 1e-9, stol, max_it, st);
        }
        ++iteration;
        if (converged) break;
        vcopy(n, x, x_old);
    }
    if (st) st->converged = converged;
    free(lb), free(ub), free(dinv), free(r), free(c), free(lbl), free(ubl), free(x_old), free(diff);
    return converged;
}
