// Box-constrained QP kernels: projection onto [lb, ub], the element-wise MPRGP
// operators (free gradient fi, chopped gradient beta, feasible step length), and
// the multicolour projected Gauss-Seidel sweep with its graph colouring.
#include <hipcub/hipcub.hpp>

#include <cfloat>

#include "uhip_common.hpp"

namespace uhip {

// VariableBoundSolverInterface::project (utopia_VariableBoundSolverInterface.hpp:107-155)
__device__ __forceinline__ double clamp_box(double x, double l, double u) { return (l >= x) ? l : ((u <= x) ? u : x); }

// ProjectedConjugateGradient::apply, element loop (utopia_ProjectedConjugateGradient.hpp:71-97).
// The active test reads z (the reference's x_view is a view of zk): elem = z_i,
// val = (|elem - ub_i| <= eps or |elem - lb_i| <= eps) ? max(u_i, 0) : u_i,
// z_i = (val == 0) ? max(p_i, 0) : p_i, w_i = val.
__global__ __launch_bounds__(kBlock) void k_pcg_wz(int64_t n, const double *__restrict__ u,
                                                   const double *__restrict__ lb, const double *__restrict__ ub,
                                                   const double *__restrict__ p, double *__restrict__ z,
                                                   double *__restrict__ w) {
    const double eps = DBL_EPSILON;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double elem = z[i];
        const double ui = u[i];
        const double val = (fabs(elem - ub[i]) <= eps || fabs(elem - lb[i]) <= eps) ? fmax(ui, 0.0) : ui;
        z[i] = (val == 0.0) ? fmax(p[i], 0.0) : p[i];
        w[i] = val;
    }
}

// ProjectedGradient::solve direction (utopia_ProjectedGradient.hpp:157-170):
// p_i = 0 where x_i is within 10 eps of a bound, else u_i.
__global__ __launch_bounds__(kBlock) void k_pg_dir(int64_t n, const double *__restrict__ x,
                                                   const double *__restrict__ u, const double *__restrict__ lb,
                                                   const double *__restrict__ ub, double *__restrict__ p) {
    const double tol = 10.0 * DBL_EPSILON;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double xi = x[i];
        p[i] = (fabs(xi - ub[i]) <= tol || fabs(xi - lb[i]) <= tol) ? 0.0 : u[i];
    }
}

// MPRGP::get_fi (utopia_MPRGP.hpp:266-287)
__device__ __forceinline__ double free_grad(double x, double g, double l, double u) {
    return (l < x && x < u) ? g : 0.0;
}

// MPRGP::get_beta (utopia_MPRGP.hpp:324-348)
__device__ __forceinline__ double chopped_grad(double x, double g, double l, double u) {
    return (fabs(l - x) < 1e-14) ? fmin(0.0, g) : ((fabs(u - x) < 1e-14) ? fmax(0.0, g) : 0.0);
}

__global__ __launch_bounds__(kBlock) void k_project(int64_t n, const double *__restrict__ x,
                                                    const double *__restrict__ lb, const double *__restrict__ ub,
                                                    double *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        out[i] = clamp_box(x[i], lb[i], ub[i]);
}

// fi, beta, gp = fi + beta, and the three dots beta.beta, fi.fi, gp.gp (utopia_MPRGP.hpp:238-245)
__global__ __launch_bounds__(kBlock) void k_mprgp_state(int64_t n, const double *__restrict__ x,
                                                        const double *__restrict__ g, const double *__restrict__ lb,
                                                        const double *__restrict__ ub, double *__restrict__ fi,
                                                        double *__restrict__ beta, double *__restrict__ gp,
                                                        double *__restrict__ partials, int64_t stride) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double xi = x[i], gi = g[i], li = lb[i], ui = ub[i];
        const double f = free_grad(xi, gi, li, ui), b = chopped_grad(xi, gi, li, ui);
        const double q = f + b;
        if (fi) fi[i] = f;
        if (beta) beta[i] = b;
        if (gp) gp[i] = q;
        s0 += b * b;
        s1 += f * f;
        s2 += q * q;
    }
    s0 = block_sum<kBlock>(s0);
    s1 = block_sum<kBlock>(s1);
    s2 = block_sum<kBlock>(s2);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s0;
        partials[stride + blockIdx.x] = s1;
        partials[2 * stride + blockIdx.x] = s2;
    }
}

__global__ __launch_bounds__(kBlock) void k_free_grad(int64_t n, const double *__restrict__ x,
                                                      const double *__restrict__ g, const double *__restrict__ lb,
                                                      const double *__restrict__ ub, double *__restrict__ fi) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        fi[i] = free_grad(x[i], g[i], lb[i], ub[i]);
}

// MPRGP::get_alpha_f (utopia_MPRGP.hpp:289-322): min over (x-l)/p for p > 0 and (x-u)/p for p < 0 (1e15 otherwise)
__global__ __launch_bounds__(kBlock) void k_alpha_f(int64_t n, const double *__restrict__ x,
                                                    const double *__restrict__ p, const double *__restrict__ lb,
                                                    const double *__restrict__ ub, double *__restrict__ partials) {
    double m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double pi = p[i], xi = x[i];
        const double h1 = (pi > 0) ? ((xi - lb[i]) / pi) : 1e15;
        const double h2 = (pi < 0) ? ((xi - ub[i]) / pi) : 1e15;
        m = fmin(m, fmin(h1, h2));
    }
    __shared__ double red[kBlock];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_min_finish(const double *__restrict__ partials, int64_t np,
                                                       double *__restrict__ out) {
    double m = INFINITY;
    for (int64_t b = threadIdx.x; b < np; b += kBlock) m = fmin(m, partials[b]);
    __shared__ double red[kBlock];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// CG step of MPRGP (utopia_MPRGP.hpp:167-182): x -= a p; g -= a Ap; fi = get_fi(x, g);
// partial fi.Ap. a = s[ia].
__global__ __launch_bounds__(kBlock) void k_mprgp_cg_step(int64_t n, double *__restrict__ x,
                                                          double *__restrict__ g, const double *__restrict__ p,
                                                          const double *__restrict__ Ap, const double *__restrict__ lb,
                                                          const double *__restrict__ ub, double *__restrict__ fi,
                                                          double a, double *__restrict__ partials) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double xi = x[i] - a * p[i];
        const double gi = g[i] - a * Ap[i];
        x[i] = xi;
        g[i] = gi;
        const double f = free_grad(xi, gi, lb[i], ub[i]);
        fi[i] = f;
        acc += f * Ap[i];
    }
    acc = block_sum<kBlock>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// CG step fused with the state that follows it (utopia_MPRGP.hpp:168-183 and the
// get_fi / chopped-gradient pass at the end of the iteration): x -= a p, g -= a Ap,
// then fi, beta, gp = fi + beta from the new x, g, and the dots beta.beta, fi.fi,
// gp.gp, fi.Ap. The grid and each thread's entries are those of k_mprgp_cg_step and
// k_mprgp_state, so the four sums are bit-identical to the two kernels'.
__global__ __launch_bounds__(kBlock) void k_mprgp_cg_step_state(int64_t n, double *__restrict__ x,
                                                                double *__restrict__ g, const double *__restrict__ p,
                                                                const double *__restrict__ Ap,
                                                                const double *__restrict__ lb,
                                                                const double *__restrict__ ub, double *__restrict__ fi,
                                                                double *__restrict__ beta, double *__restrict__ gp,
                                                                double a, double *__restrict__ partials,
                                                                int64_t stride) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double api = Ap[i];
        const double xi = x[i] - a * p[i];
        const double gi = g[i] - a * api;
        x[i] = xi;
        g[i] = gi;
        const double li = lb[i], ui = ub[i];
        const double f = free_grad(xi, gi, li, ui), b = chopped_grad(xi, gi, li, ui);
        const double q = f + b;
        fi[i] = f;
        beta[i] = b;
        gp[i] = q;
        s0 += b * b;
        s1 += f * f;
        s2 += q * q;
        s3 += f * api;
    }
    s0 = block_sum<kBlock>(s0);
    s1 = block_sum<kBlock>(s1);
    s2 = block_sum<kBlock>(s2);
    s3 = block_sum<kBlock>(s3);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s0;
        partials[stride + blockIdx.x] = s1;
        partials[2 * stride + blockIdx.x] = s2;
        partials[3 * stride + blockIdx.x] = s3;
    }
}

// p = fi - (s[inum]/s[iden]) p
__global__ __launch_bounds__(kBlock) void k_mprgp_pupdate(int64_t n, double *__restrict__ p,
                                                          const double *__restrict__ fi, const double *__restrict__ s,
                                                          int inum, int iden) {
    const double beta_sc = s[inum] / s[iden];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        p[i] = fi[i] - beta_sc * p[i];
}

// Expansion step (utopia_MPRGP.hpp:184-198): x -= af p; g -= af Ap; fi = get_fi(x, g);
// x = P(x - abar fi).
__global__ __launch_bounds__(kBlock) void k_mprgp_expand(int64_t n, double *__restrict__ x,
                                                         const double *__restrict__ g, const double *__restrict__ p,
                                                         const double *__restrict__ Ap, const double *__restrict__ lb,
                                                         const double *__restrict__ ub, double af, double abar) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double li = lb[i], ui = ub[i];
        const double xi = x[i] - af * p[i];
        const double gi = g[i] - af * Ap[i];
        const double f = free_grad(xi, gi, li, ui);
        x[i] = clamp_box(xi - (abar * f), li, ui);
    }
}

// Proportioning step (utopia_MPRGP.hpp:223-235): x -= a beta; g -= a Abeta; p = get_fi(x, g)
__global__ __launch_bounds__(kBlock) void k_mprgp_prop(int64_t n, double *__restrict__ x, double *__restrict__ g,
                                                       const double *__restrict__ beta,
                                                       const double *__restrict__ Abeta,
                                                       const double *__restrict__ lb, const double *__restrict__ ub,
                                                       double *__restrict__ p, double a) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double xi = x[i] - a * beta[i];
        const double gi = g[i] - a * Abeta[i];
        x[i] = xi;
        g[i] = gi;
        p[i] = free_grad(xi, gi, lb[i], ub[i]);
    }
}

// ActiveSet::determine (solvers/nonlinear/constrained/utopia_ActiveSet.hpp:41-79):
// ind_i = (x_i <= l_i + tol or x_i >= u_i - tol); partial count of changed entries.
__global__ __launch_bounds__(kBlock) void k_active_set(int64_t n, const double *__restrict__ x,
                                                       const double *__restrict__ lb, const double *__restrict__ ub,
                                                       double tol, double *__restrict__ ind,
                                                       double *__restrict__ partials) {
    double changed = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double xi = x[i];
        const double v = (xi <= lb[i] + tol || xi >= ub[i] - tol) ? 1.0 : 0.0;
        changed += (v != ind[i]) ? 1.0 : 0.0;
        ind[i] = v;
    }
    changed = block_sum<kBlock>(changed);
    if (threadIdx.x == 0) partials[blockIdx.x] = changed;
}

// IPTruncatedTransfer::truncate_interpolation (transfer/utopia_IPTruncatedTransfer.hpp:203-223):
// the values of rows i with ind_i == 1 become zero (pattern kept).
__global__ __launch_bounds__(kBlock) void k_zero_rows(int32_t n, const int32_t *__restrict__ rowptr,
                                                      const double *__restrict__ ind, double *__restrict__ val) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n || ind[r] != 1.0) return;
    for (int32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) val[e] = 0.0;
}

// ---- multicolour projected Gauss-Seidel ----

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Jones-Plassmann round: an uncoloured row whose (hash, index) beats every
// uncoloured neighbour takes the smallest colour none of its coloured
// neighbours has. Rows chosen in one round are never adjacent, so they cannot
// conflict. Needs a structurally symmetric pattern.
__global__ __launch_bounds__(kBlock) void k_jp_round(int32_t n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col, int32_t *__restrict__ color,
                                                     uint64_t seed, unsigned int *__restrict__ remaining) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || color[i] >= 0) return;
    const uint64_t hi = mix64((uint64_t)i ^ seed);
    bool best = true;
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
        const int32_t j = col[e];
        if (j == i) continue;
        if (color[j] >= 0) continue;  // coloured in an earlier round
        const uint64_t hj = mix64((uint64_t)j ^ seed);
        if (hj > hi || (hj == hi && j > i)) {
            best = false;
            break;
        }
    }
    if (!best) {
        atomicAdd(remaining, 1u);
        return;
    }
    // smallest colour not used by a coloured neighbour (neighbours chosen in
    // this round are marked pending, <= -2, and are never adjacent to i)
    int32_t c = 0;
    for (;;) {
        bool used = false;
        for (int32_t e = rowptr[i]; e < rowptr[i + 1] && !used; ++e) used = (col[e] != i) && (color[col[e]] == c);
        if (!used) break;
        ++c;
    }
    color[i] = -2 - c;  // pending until k_jp_commit, after every row of the round has decided
}

__global__ __launch_bounds__(kBlock) void k_jp_commit(int32_t n, int32_t *__restrict__ color) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && color[i] <= -2) color[i] = -2 - color[i];
}

// One colour of the projected sweep (ProjectedScalarGaussSeidelSweep::apply,
// utopia_ProjectedGaussSeidelSweep.hpp:191-235), bounds relative to x:
// lbl = lb - x, ubl = ub - x (ProjectedGaussSeidel::step, utopia_petsc_ProjectedGaussSeidel.cpp:384-385).
__global__ __launch_bounds__(kBlock) void k_pgs_color(int32_t cnt, const int32_t *__restrict__ rows,
                                                      const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col, const double *__restrict__ val,
                                                      const double *__restrict__ dinv, const double *__restrict__ r,
                                                      const double *__restrict__ x, const double *__restrict__ lb,
                                                      const double *__restrict__ ub, double *__restrict__ c,
                                                      int constrained) {
    const int32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= cnt) return;
    const int32_t i = rows[t];
    double v = r[i];
    const double d = dinv[i];
    if (constrained) {
        const double lbl = lb[i] - x[i], ubl = ub[i] - x[i];
        if (d <= 0.0) {
            c[i] = (v == 0.0) ? 0.0 : ((v > 0.0) ? lbl : ubl);
            return;
        }
        for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            const int32_t j = col[e];
            const double a = val[e];
            if (j != i && a != 0.0) v -= a * c[j];
        }
        c[i] = fmax(lbl, fmin(d * v, ubl));
    } else {
        for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            const int32_t j = col[e];
            const double a = val[e];
            if (j != i && a != 0.0) v -= a * c[j];
        }
        c[i] = d * v;
    }
}

// d_inv_i = 1/a_ii where |a_ii| > 0, else 0 (init_from_local_matrix, :95-104)
__global__ __launch_bounds__(kBlock) void k_pgs_dinv(int32_t n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col, const double *__restrict__ val,
                                                     double *__restrict__ dinv) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
        if (col[e] == i) d = (fabs(val[e]) > 0.0) ? (1 / val[e]) : 0.0;
    dinv[i] = d;
}

}  // namespace uhip

using namespace uhip;

extern "C" {

int uhip_active_set_f64(uhip_ctx ctx, int64_t n, const double *x, const double *lb, const double *ub, double tol,
                        double *indicator, double *changed_d) {
    UHIP_REQUIRE(ctx && x && lb && ub && indicator && changed_d, "uhip_active_set_f64: null argument");
    const int g = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, g));
    hipLaunchKernelGGL(k_active_set, dim3(g), dim3(kBlock), 0, ctx->stream, n, x, lb, ub, tol, indicator,
                       ctx->partials);
    UHIP_CHECK_LAUNCH();
    return finish_sums(ctx, 1, g, ctx->max_partials, changed_d, 0u);
}

int uhip_csr_zero_rows(uhip_ctx ctx, uhip_csr *A, const double *indicator) {
    UHIP_REQUIRE(ctx && A && indicator && A->dtype == UHIP_F64, "uhip_csr_zero_rows: bad argument");
    if (A->nrows == 0) return UHIP_OK;
    hipLaunchKernelGGL(k_zero_rows, dim3((A->nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, A->nrows,
                       A->rowptr, indicator, (double *)A->val);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_pcg_wz_f64(uhip_ctx ctx, int64_t n, const double *u, const double *lb, const double *ub, const double *p,
                    double *z, double *w) {
    UHIP_REQUIRE(ctx && u && lb && ub && p && z && w, "uhip_pcg_wz_f64: null argument");
    if (n <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_pcg_wz, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, u, lb, ub, p, z, w);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_pg_direction_f64(uhip_ctx ctx, int64_t n, const double *x, const double *u, const double *lb,
                          const double *ub, double *p) {
    UHIP_REQUIRE(ctx && x && u && lb && ub && p, "uhip_pg_direction_f64: null argument");
    if (n <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_pg_dir, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, x, u, lb, ub, p);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_box_project_f64(uhip_ctx ctx, int64_t n, const double *x, const double *lb, const double *ub, double *out) {
    UHIP_REQUIRE(ctx && x && lb && ub && out, "uhip_box_project_f64: null argument");
    if (n <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_project, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, x, lb, ub, out);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_mprgp_state_f64(uhip_ctx ctx, int64_t n, const double *x, const double *g, const double *lb, const double *ub,
                         double *fi, double *beta, double *gp, double *dots_d) {
    UHIP_REQUIRE(ctx && x && g && lb && ub && dots_d, "uhip_mprgp_state_f64: null argument");
    const int gr = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, gr));
    hipLaunchKernelGGL(k_mprgp_state, dim3(gr), dim3(kBlock), 0, ctx->stream, n, x, g, lb, ub, fi, beta, gp,
                       ctx->partials, ctx->max_partials);
    UHIP_CHECK_LAUNCH();
    return finish_sums(ctx, 3, gr, ctx->max_partials, dots_d, 0u);
}

int uhip_mprgp_free_grad_f64(uhip_ctx ctx, int64_t n, const double *x, const double *g, const double *lb,
                             const double *ub, double *fi) {
    UHIP_REQUIRE(ctx && x && g && lb && ub && fi, "uhip_mprgp_free_grad_f64: null argument");
    if (n <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_free_grad, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, x, g, lb, ub, fi);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_mprgp_alpha_f_f64(uhip_ctx ctx, int64_t n, const double *x, const double *p, const double *lb,
                           const double *ub, double *result_d) {
    UHIP_REQUIRE(ctx && x && p && lb && ub && result_d, "uhip_mprgp_alpha_f_f64: null argument");
    const int gr = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, gr));
    hipLaunchKernelGGL(k_alpha_f, dim3(gr), dim3(kBlock), 0, ctx->stream, n, x, p, lb, ub, ctx->partials);
    UHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_min_finish, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, (int64_t)gr, result_d);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_mprgp_cg_step_f64(uhip_ctx ctx, int64_t n, double *x, double *g, const double *p, const double *Ap,
                           const double *lb, const double *ub, double *fi, double alpha, double *scalars,
                           int i_fiAp) {
    UHIP_REQUIRE(ctx && x && g && p && Ap && lb && ub && fi && scalars, "uhip_mprgp_cg_step_f64: null argument");
    const int gr = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, gr));
    hipLaunchKernelGGL(k_mprgp_cg_step, dim3(gr), dim3(kBlock), 0, ctx->stream, n, x, g, p, Ap, lb, ub, fi, alpha,
                       ctx->partials);
    UHIP_CHECK_LAUNCH();
    return finish_sums(ctx, 1, gr, ctx->max_partials, scalars + i_fiAp, 0u);
}

int uhip_mprgp_cg_step_state_f64(uhip_ctx ctx, int64_t n, double *x, double *g, const double *p, const double *Ap,
                                 const double *lb, const double *ub, double *fi, double *beta, double *gp,
                                 double alpha, double *dots_d) {
    UHIP_REQUIRE(ctx && x && g && p && Ap && lb && ub && fi && beta && gp && dots_d,
                 "uhip_mprgp_cg_step_state_f64: null argument");
    const int gr = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, gr));
    hipLaunchKernelGGL(k_mprgp_cg_step_state, dim3(gr), dim3(kBlock), 0, ctx->stream, n, x, g, p, Ap, lb, ub, fi, beta,
                       gp, alpha, ctx->partials, ctx->max_partials);
    UHIP_CHECK_LAUNCH();
    return finish_sums(ctx, 4, gr, ctx->max_partials, dots_d, 0u);
}

int uhip_mprgp_pupdate_f64(uhip_ctx ctx, int64_t n, double *p, const double *fi, const double *scalars, int i_num,
                           int i_den) {
    UHIP_REQUIRE(ctx && p && fi && scalars, "uhip_mprgp_pupdate_f64: null argument");
    if (n <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_mprgp_pupdate, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, p, fi, scalars,
                       i_num, i_den);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_mprgp_expand_f64(uhip_ctx ctx, int64_t n, double *x, const double *g, const double *p, const double *Ap,
                          const double *lb, const double *ub, double alpha_f, double alpha_bar) {
    UHIP_REQUIRE(ctx && x && g && p && Ap && lb && ub, "uhip_mprgp_expand_f64: null argument");
    if (n <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_mprgp_expand, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, x, g, p, Ap, lb,
                       ub, alpha_f, alpha_bar);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_mprgp_proportion_f64(uhip_ctx ctx, int64_t n, double *x, double *g, const double *beta, const double *Abeta,
                              const double *lb, const double *ub, double *p, double alpha) {
    UHIP_REQUIRE(ctx && x && g && beta && Abeta && lb && ub && p, "uhip_mprgp_proportion_f64: null argument");
    if (n <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_mprgp_prop, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, x, g, beta, Abeta,
                       lb, ub, p, alpha);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_graph_color(uhip_ctx ctx, const uhip_csr *A, uint64_t seed, int32_t *color, int32_t *ncolors) {
    UHIP_REQUIRE(ctx && A && color && ncolors, "uhip_graph_color: null argument");
    UHIP_REQUIRE(A->nrows == A->ncols, "uhip_graph_color: square pattern required");
    const int32_t n = A->nrows;
    *ncolors = 0;
    if (n == 0) return UHIP_OK;
    UHIP_CHECK_HIP(hipMemsetAsync(color, 0xff, sizeof(int32_t) * (size_t)n, ctx->stream));  // -1
    unsigned int *rem = ctx->counters + 1;
    dim3 g((n + kBlock - 1) / kBlock);
    // every round colours at least the uncoloured row of largest (hash, index)
    for (int32_t round = 0; round < n; ++round) {
        UHIP_CHECK_HIP(hipMemsetAsync(rem, 0, sizeof(unsigned int), ctx->stream));
        hipLaunchKernelGGL(k_jp_round, g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col, color, seed, rem);
        UHIP_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_jp_commit, g, dim3(kBlock), 0, ctx->stream, n, color);
        UHIP_CHECK_LAUNCH();
        unsigned int h = 0;
        UHIP_CHECK_HIP(hipMemcpyAsync(&h, rem, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        if (h == 0) {
            int32_t *mx = nullptr;
            void *tmp = nullptr;
            size_t tb = 0;
            UHIP_CHECK_HIP(hipMalloc(&mx, sizeof(int32_t)));
            UHIP_CHECK_HIP(hipcub::DeviceReduce::Max(nullptr, tb, color, mx, n, ctx->stream));
            UHIP_CHECK_HIP(hipMalloc(&tmp, tb));
            UHIP_CHECK_HIP(hipcub::DeviceReduce::Max(tmp, tb, color, mx, n, ctx->stream));
            int32_t m = 0;
            UHIP_CHECK_HIP(hipMemcpyAsync(&m, mx, sizeof(m), hipMemcpyDeviceToHost, ctx->stream));
            UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
            hipFree(mx);
            hipFree(tmp);
            *ncolors = m + 1;
            return UHIP_OK;
        }
    }
    return set_error(UHIP_ERR_NUMERIC, "uhip_graph_color: did not terminate");
}

int uhip_pgs_dinv(uhip_ctx ctx, const uhip_csr *A, double *dinv) {
    UHIP_REQUIRE(ctx && A && dinv && A->dtype == UHIP_F64, "uhip_pgs_dinv: bad argument");
    const int32_t n = A->nrows;
    if (n == 0) return UHIP_OK;
    hipLaunchKernelGGL(k_pgs_dinv, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                       (const double *)A->val, dinv);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_pgs_color_sweep(uhip_ctx ctx, const uhip_csr *A, int32_t cnt, const int32_t *rows, const double *dinv,
                         const double *r, const double *x, const double *lb, const double *ub, double *c,
                         int constrained) {
    UHIP_REQUIRE(ctx && A && rows && dinv && r && c && A->dtype == UHIP_F64, "uhip_pgs_color_sweep: bad argument");
    UHIP_REQUIRE(!constrained || (x && lb && ub), "uhip_pgs_color_sweep: bounds needed");
    if (cnt <= 0) return UHIP_OK;
    hipLaunchKernelGGL(k_pgs_color, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, cnt, rows,
                       A->rowptr, A->col, (const double *)A->val, dinv, r, x, lb, ub, c, constrained);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

// rows grouped by colour (stable), offsets[k] .. offsets[k+1]
int uhip_color_groups(uhip_ctx ctx, int32_t n, const int32_t *color, int32_t ncolors, int32_t *rows,
                      int32_t *offsets_h) {
    UHIP_REQUIRE(ctx && color && rows && offsets_h && ncolors >= 0, "uhip_color_groups: bad argument");
    if (n == 0) {
        offsets_h[0] = 0;
        return UHIP_OK;
    }
    int32_t *keys_out = nullptr, *idx_in = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&keys_out, sizeof(int32_t) * n));
    UHIP_CHECK_HIP(hipMalloc(&idx_in, sizeof(int32_t) * n));
    std::vector<int32_t> seq(n);
    for (int32_t i = 0; i < n; ++i) seq[(size_t)i] = i;
    UHIP_CHECK_HIP(hipMemcpyAsync(idx_in, seq.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    void *tmp = nullptr;
    size_t tb = 0;
    int bits = 1;
    while ((1 << bits) <= ncolors) ++bits;
    UHIP_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, color, keys_out, idx_in, rows, n, 0, bits,
                                                      ctx->stream));
    UHIP_CHECK_HIP(hipMalloc(&tmp, tb));
    UHIP_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, color, keys_out, idx_in, rows, n, 0, bits, ctx->stream));
    std::vector<int32_t> kh(n);
    UHIP_CHECK_HIP(hipMemcpyAsync(kh.data(), keys_out, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(keys_out);
    hipFree(idx_in);
    hipFree(tmp);
    for (int32_t k = 0; k <= ncolors; ++k) offsets_h[k] = 0;
    for (int32_t i = 0; i < n; ++i) offsets_h[kh[(size_t)i] + 1]++;
    for (int32_t k = 0; k < ncolors; ++k) offsets_h[k + 1] += offsets_h[k];
    return UHIP_OK;
}

}  // extern "C"
