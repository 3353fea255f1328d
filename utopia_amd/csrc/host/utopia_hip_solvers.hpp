// utopia solver classes over the HIP backend.
//
// Same names, parameters, defaults and update()/apply() semantics as the
// reference's templates; the arithmetic goes through fused HIP kernels:
//   IterativeSolver        solvers/linear/utopia_IterativeSolver.hpp:20-242
//   InvDiagPreconditioner  solvers/preconditioners/utopia_TrivialPreconditioners.hpp:20-53
//   ConjugateGradient      solvers/smoothers/utopia_ConjugateGradient_impl.hpp:51-331 (HOMEMADE)
//   PowerMethod            solvers/eigensolvers/utopia_PowerMethod.hpp:18-157
//   Chebyshev3level        solvers/smoothers/utopia_Chebyshev3level.hpp:16-189
//   Jacobi                 solvers/smoothers/utopia_Jacobi.hpp:16-141
//   Factorization          backend/petsc/solvers/linear/utopia_petsc_Factorization.hpp:13-84
//   Multigrid              solvers/multilevel/utopia_Multigrid.hpp:27-519,
//                          utopia_LinearMultiLevel.hpp:24-229, utopia_MultiLevelMask.hpp:34-115,
//                          transfer/utopia_IPRTransfer.hpp:22-274
#pragma once

#include <cstdio>
#include <cstdlib>
#include <limits>

#include "utopia_hip_backend.hpp"

namespace utopia {
// the solvers of the HIP backend (the tensor types, HipVector and HipCrsMatrix, are in utopia)
namespace hip {

// ConvergenceReason codes (solvers/utopia_ConvergenceReason.hpp:84-100)
struct ConvergenceReason {
    static const int CONVERGED_ITERATING = 0;
    static const int CONVERGED_FNORM_ABS = 2;
    static const int CONVERGED_FNORM_RELATIVE = 3;
    static const int CONVERGED_SNORM_RELATIVE = 4;
    static const int DIVERGED_MAX_IT = -5;
};

struct SolutionStatus {
    int64_t iterates = 0;
    double gradient_norm = 0, relative_gradient_norm = 0, step_norm = 0;
    int reason = 0;
    void clear() { *this = SolutionStatus(); }
};

template <class Matrix, class Vector>
class LinearSolver {
public:
    virtual ~LinearSolver() = default;
    virtual void update(const std::shared_ptr<const Matrix> &op) { op_ = op; }
    virtual bool apply(const Vector &rhs, Vector &sol) = 0;
    // apply() with sol known to be zero on entry (what CG does: z_new.set(0); precond->apply(r, z_new)).
    virtual bool apply_zero_initial_guess(const Vector &rhs, Vector &sol) {
        if (sol.size() != rhs.size() || sol.empty()) sol.zeros_like(rhs);
        else sol.set(0);
        return apply(rhs, sol);
    }
    // The same, followed by dot_d[0] = sol . rhs (device scalar); solvers that
    // produce sol in a fused kernel compute the dot while writing it.
    virtual bool apply_zero_initial_guess_dot(const Vector &rhs, Vector &sol, double *dot_d) {
        const bool ok = apply_zero_initial_guess(rhs, sol);
        const typename Vector::Scalar *a[1] = {sol.data()}, *b[1] = {rhs.data()};
        if (Vector::dtype == UHIP_F64)
            hip_check(uhip_vec_dots_f64(sol.ctx(), sol.size(), 1, (const double *const *)a, (const double *const *)b,
                                        dot_d),
                      "dot(z, r)");
        else
            hip_check(uhip_vec_dots_f32(sol.ctx(), sol.size(), 1, (const float *const *)a, (const float *const *)b,
                                        dot_d),
                      "dot(z, r)");
        return ok;
    }
    // FP32 copy of the right-hand side that a reduced-precision solver reads
    // instead of rhs (nullptr: not wanted). A caller that fills it with
    // (float)rhs passes it to apply_zero_initial_guess_dot_f32.
    virtual float *rhs_f32_buffer(int64_t) { return nullptr; }
    // z = M rhs (zero guess) left in FP32 storage owned by the preconditioner, with
    // dot_d[0] = z.rhs; nullptr when the preconditioner has no FP32 output
    virtual const float *apply_zero_initial_guess_dot_z32(const Vector &, const float *, double *) { return nullptr; }
    virtual bool apply_zero_initial_guess_dot_f32(const Vector &rhs, const float *, Vector &sol, double *dot_d) {
        return apply_zero_initial_guess_dot(rhs, sol, dot_d);
    }
    bool solve(const Matrix &A, const Vector &b, Vector &x) {
        update(std::shared_ptr<const Matrix>(&A, [](const Matrix *) {}));
        return apply(b, x);
    }
    const std::shared_ptr<const Matrix> &get_operator() const { return op_; }
    void verbose(bool v) { verbose_ = v; }
    bool verbose() const { return verbose_; }

protected:
    std::shared_ptr<const Matrix> op_;
    bool verbose_ = false;
};

template <class Matrix, class Vector>
class IterativeSolver : public LinearSolver<Matrix, Vector> {
public:
    double atol() const { return atol_; }
    double rtol() const { return rtol_; }
    double stol() const { return stol_; }
    int64_t max_it() const { return max_it_; }
    void atol(double v) { atol_ = v; }
    void rtol(double v) { rtol_ = v; }
    void stol(double v) { stol_ = v; }
    void max_it(int64_t v) { max_it_ = v; }
    int64_t sweeps() const { return max_it_; }
    void sweeps(int64_t v) { max_it_ = v; }
    const SolutionStatus &solution_status() const { return status_; }

protected:
    // IterativeSolver::check_convergence (utopia_IterativeSolver.hpp:148-191), norm_frequency 1
    bool check_convergence(int64_t it, double g_norm, double r_norm, double s_norm) {
        bool converged = false;
        if (g_norm <= atol_) converged = true, status_.reason = ConvergenceReason::CONVERGED_FNORM_ABS;
        if (s_norm <= stol_) converged = true, status_.reason = ConvergenceReason::CONVERGED_SNORM_RELATIVE;
        if (r_norm <= rtol_) converged = true, status_.reason = ConvergenceReason::CONVERGED_FNORM_RELATIVE;
        if (it >= max_it_) converged = true, status_.reason = ConvergenceReason::DIVERGED_MAX_IT;
        if (converged) {
            status_.iterates = it;
            status_.gradient_norm = g_norm;
            status_.relative_gradient_norm = r_norm;
            status_.step_norm = s_norm;
        }
        return converged;
    }
    void print_iter(int64_t it, double v) const {
        if (this->verbose_) std::printf("%8lld  %.6e\n", (long long)it, v);
    }

    double atol_ = 1e-9, rtol_ = 1e-9, stol_ = 1e-11;
    int64_t max_it_ = 1000;
    SolutionStatus status_;
};

// z = d .* r with d = pseudo_inv(diag(A)), tol = machine epsilon
template <class Matrix, class Vector>
class InvDiagPreconditioner final : public LinearSolver<Matrix, Vector> {
public:
    using T = typename Vector::Scalar;
    bool apply(const Vector &rhs, Vector &sol) override {
        if (sol.size() != rhs.size() || sol.empty()) sol.zeros_like(rhs);
        if (Vector::dtype == UHIP_F64)
            hip_check(uhip_vec_emul_f64(d_.ctx(), d_.size(), (const double *)d_.data(), (const double *)rhs.data(),
                                        (double *)sol.data()),
                      "invdiag");
        else
            hip_check(uhip_vec_emul_f32(d_.ctx(), d_.size(), (const float *)d_.data(), (const float *)rhs.data(),
                                        (float *)sol.data()),
                      "invdiag");
        return true;
    }
    bool apply_zero_initial_guess(const Vector &rhs, Vector &sol) override { return apply(rhs, sol); }
    void update(const std::shared_ptr<const Matrix> &op) override {
        LinearSolver<Matrix, Vector>::update(op);
        op->build_diag(d_);
        pseudo_inv(d_, std::numeric_limits<T>::epsilon());
    }
    const Vector &get_d() const { return d_; }

    static void pseudo_inv(Vector &d, T tol) {
        if (Vector::dtype == UHIP_F64)
            hip_check(uhip_vec_pseudo_inv_f64(d.ctx(), d.size(), (const double *)d.data(), (double *)d.data(), tol),
                      "pseudo_inv");
        else
            hip_check(uhip_vec_pseudo_inv_f32(d.ctx(), d.size(), (const float *)d.data(), (float *)d.data(), tol),
                      "pseudo_inv");
    }

private:
    Vector d_;
};

// ---------------------------------------------------------------------------
// ConjugateGradient<Matrix, Vector, HOMEMADE> (FP64 vectors)
// ---------------------------------------------------------------------------
template <class Matrix, class Vector>
class ConjugateGradient : public IterativeSolver<Matrix, Vector> {
public:
    void set_preconditioner(const std::shared_ptr<LinearSolver<Matrix, Vector>> &p) { precond_ = p; }
    const std::shared_ptr<LinearSolver<Matrix, Vector>> &get_preconditioner() const { return precond_; }
    bool has_preconditioner() const { return (bool)precond_; }
    void reset_initial_guess(bool v) { reset_initial_guess_ = v; }
    // After reaching atol on the recurrence residual, recompute b - A x in FP64
    // and restart until the true residual also satisfies atol (iterative
    // refinement for inexact - e.g. FP32 - preconditioners). Off by default.
    void refine(bool v) { refine_ = v; }
    int64_t refinements() const { return refinements_; }
    int64_t total_iterations() const { return total_its_; }

    // PreconditionedSolver::update (solvers/linear/utopia_PreconditionedSolver.hpp:37-46)
    void update(const std::shared_ptr<const Matrix> &op) override {
        IterativeSolver<Matrix, Vector>::update(op);
        if (precond_) precond_->update(op);
        init_memory(op->rows(), op->context());
    }
    // PreconditionedSolver::update(op, prec) (utopia_PreconditionedSolver.hpp:48-61): the
    // iteration runs on op, the preconditioner is built from prec
    void update(const std::shared_ptr<const Matrix> &op, const std::shared_ptr<const Matrix> &prec) {
        IterativeSolver<Matrix, Vector>::update(op);
        if (precond_) precond_->update(prec);
        init_memory(op->rows(), op->context());
    }

    bool apply(const Vector &b, Vector &x) override {
        const Matrix &A = *this->op_;
        if (x.empty() || x.size() != b.size()) x.zeros_like(b);
        else if (reset_initial_guess_) x.set(0);
        refinements_ = 0;
        total_its_ = 0;
        bool ok = precond_ ? preconditioned_solve(A, b, x) : unpreconditioned_solve(A, b, x);
        total_its_ = this->status_.iterates + 1;
        while (ok && refine_ && precond_) {
            A.residual(x, b, r_);
            const double true_norm = r_.norm2();
            if (true_norm <= this->atol_ || refinements_ >= 10) break;
            ++refinements_;
            ok = preconditioned_solve(A, b, x);
            total_its_ += this->status_.iterates + 1;
        }
        return ok;
    }

private:
    enum { S_RZ = 0, S_PAP = 1, S_ALPHA = 2, S_RNORM = 3, S_RZNEW = 4 };

    void init_memory(int64_t n, const HipContextPtr &ctx) {
        if (r_.size() == n && !r_.empty()) return;
        for (Vector *v : {&r_, &p_, &Ap_, &r_new_, &z_, &z_new_}) *v = Vector(ctx, n, 0.0);
    }

    // preconditioned_solve (utopia_ConjugateGradient_impl.hpp:195-294)
    bool preconditioned_solve(const Matrix &A, const Vector &b, Vector &x) {
        int64_t it = 0;
        double r_norm = 9e9;
        this->status_.clear();
        uhip_ctx c = A.c();
        double *s = A.context()->scalars();
        A.residual(x, b, r_);                        // r = b - A x
        precond_->apply_zero_initial_guess(r_, z_);  // z.set(0); precond->apply(r, z)
        p_.copy_from(z_);
        const double *rz[1] = {r_.data()}, *zz[1] = {z_.data()};
        hip_check(uhip_vec_dots_f64(c, b.size(), 1, rz, zz, s + S_RZ), "dot(r,z)");
        bool stop = false;
        float *r32 = precond_->rhs_f32_buffer(b.size());
        int i_rz = S_RZ, i_rznew = S_RZNEW;  // the two r.z slots swap roles every iteration
        // alpha and |r_new| come back asynchronously: while the host waits for them the stream
        // already holds the preconditioner and the x/p update of the same iteration, so the GPU
        // does not idle on the read-back. Those are enqueued before the convergence test; when it
        // stops the loop, x += alpha p was what the iteration had left to do (the update kernel
        // leaves x alone when alpha is not finite) and the preconditioner's output is unused. When
        // the residual is expected to reach atol in this iteration (its last reduction factor
        // applied once more) the test comes first instead, which saves that last preconditioner.
        double r_prev = -1.0, r_prev2 = -1.0;
        while (!stop) {
            hip_check(uhip_spmv_dot(c, &A.raw(), p_.data(), Ap_.data(), s + S_PAP), "Ap, dot(p,Ap)");
            // alpha, r_new = r - alpha Ap, |r_new|; x += alpha p is applied with the p update
            hip_check(uhip_cg_residual_f64(c, b.size(), r_.data(), Ap_.data(), r_new_.data(), r32, s, i_rz, S_PAP,
                                           S_ALPHA, S_RNORM),
                      "cg residual");
            A.context()->fetch_scalars(0, s + S_ALPHA, 2);
            const bool near = r_prev > 0 && r_prev2 > 0 && r_prev * (r_prev / r_prev2) <= 2.0 * this->atol_;
            double h[2];
            if (near) {
                A.context()->fetch_wait(0, 2, h);
                if (std::isinf(h[0]) || std::isnan(h[0])) {  // stops before touching x
                    stop = this->check_convergence(it, r_norm, 1, 1);
                    break;
                }
                if (h[1] <= this->atol_) {
                    r_norm = h[1];
                    hip_check(uhip_cg_xpupdate_f64(c, b.size(), x.data(), p_.data(), nullptr, s, S_ALPHA, 0, 0),
                              "x update");
                    this->print_iter(it, r_norm);
                    stop = this->check_convergence(it, r_norm, 1, 1);
                    break;
                }
            }
            // z_new.set(0); precond->apply(r_new, z_new); dot(z_new, r_new)
            const float *z32 = r32 ? precond_->apply_zero_initial_guess_dot_z32(r_new_, r32, s + i_rznew) : nullptr;
            if (!z32) {
                if (r32) precond_->apply_zero_initial_guess_dot_f32(r_new_, r32, z_new_, s + i_rznew);
                else precond_->apply_zero_initial_guess_dot(r_new_, z_new_, s + i_rznew);
            }
            // x += alpha p; p = z_new + beta p
            hip_check(uhip_cg_xpupdate2(c, b.size(), x.data(), p_.data(), z32 ? (const void *)z32 : z_new_.data(),
                                        z32 ? UHIP_F32 : UHIP_F64, s, S_ALPHA, i_rznew, i_rz),
                      "x, p update");
            if (!near) {
                A.context()->fetch_wait(0, 2, h);
                if (std::isinf(h[0]) || std::isnan(h[0])) {  // the update above left x alone
                    stop = this->check_convergence(it, r_norm, 1, 1);
                    break;
                }
                if (h[1] <= this->atol_) {  // x += alpha p is in the stream
                    r_norm = h[1];
                    this->print_iter(it, r_norm);
                    stop = this->check_convergence(it, r_norm, 1, 1);
                    break;
                }
            }
            r_norm = h[1];
            r_prev2 = r_prev;
            r_prev = r_norm;
            std::swap(r_, r_new_);
            std::swap(z_, z_new_);
            std::swap(i_rz, i_rznew);
            this->print_iter(it, r_norm);
            stop = this->check_convergence(it, r_norm, 1, 1);
            it++;
        }
        this->status_.gradient_norm = r_norm;
        return r_norm <= this->atol_;
    }

    // unpreconditioned_solve (utopia_ConjugateGradient_impl.hpp:90-192)
    bool unpreconditioned_solve(const Matrix &A, const Vector &b, Vector &x) {
        int64_t it = 0;
        double rho = 1., rho_1 = 1., beta = 0., alpha = 1., r_norm = 9e9;
        this->status_.clear();
        A.residual(x, b, r_);
        r_norm = r_.norm2();
        it++;
        bool converged = false;
        while (!converged) {
            rho = r_.dot(r_);
            if (rho == 0.) {
                converged = true;
                this->status_.iterates = it;
                this->status_.reason = ConvergenceReason::CONVERGED_FNORM_ABS;
                break;
            }
            if (it > 1) {
                beta = rho / rho_1;
                p_.axpby(1.0, r_, beta);  // p = r + beta p
            } else {
                p_.copy_from(r_);
            }
            A.multiply(p_, Ap_);
            const double dot_pq = p_.dot(Ap_);
            if (dot_pq == 0.) {
                converged = true;
                break;
            }
            alpha = rho / dot_pq;
            x.axpy(alpha, p_);
            r_.axpy(-alpha, Ap_);
            rho_1 = rho;
            r_norm = r_.norm2();
            this->print_iter(it, r_norm);
            converged = this->check_convergence(it, r_norm, 1, 1);
            it++;
        }
        this->status_.gradient_norm = r_norm;
        return converged;
    }

    std::shared_ptr<LinearSolver<Matrix, Vector>> precond_;
    bool reset_initial_guess_ = false, refine_ = false;
    int64_t refinements_ = 0, total_its_ = 0;
    Vector r_, p_, Ap_, r_new_, z_, z_new_;
};

// ---------------------------------------------------------------------------
// PowerMethod (utopia_PowerMethod.hpp:95-146): rand() initial vector drawn on
// the host in index order, then v = A v / |A v| until |lambda - lambda_old| < tol.
// ---------------------------------------------------------------------------
template <class Matrix, class Vector>
class PowerMethod {
public:
    double tol() const { return tol_; }
    void tol(double t) { tol_ = t; }
    int64_t max_it() const { return max_it_; }
    void max_it(int64_t m) { max_it_ = m; }
    int64_t iterations() const { return its_; }

    double get_max_eig(const Matrix &A) {
        const int64_t n = A.rows();
        std::vector<typename Vector::Scalar> h((size_t)n);
        for (int64_t i = 0; i < n; ++i) h[(size_t)i] = ((double)std::rand() / (RAND_MAX)) + 1;
        Vector v(A.context(), n), w(A.context(), n);
        v.upload(h.data());
        v.scale(1. / v.norm2());
        int64_t it = 0;
        bool converged = false;
        double lambda = 0.0, lambda_old = 0.0;
        while (!converged) {
            A.multiply(v, w);
            std::swap(v, w);
            lambda = v.norm2();
            v.scale(1. / lambda);
            it = it + 1;
            converged = ((std::abs(lambda_old - lambda) < tol_) || it > max_it_);
            lambda_old = lambda;
        }
        its_ = it;
        return lambda;
    }

private:
    double tol_ = 1e-2;
    int64_t max_it_ = 30, its_ = 0;
};

// Chebyshev three-term coefficients of Chebyshev3level::solve
// (utopia_Chebyshev3level.hpp:138-152) for nu steps.
// Chebyshev3level::solve updates x before its first convergence check
// (utopia_Chebyshev3level.hpp:137-160), so sweeps(0) still takes one step: nu < 1 counts as 1.
inline void chebyshev_coefficients(double eig_max, double eig_min, int nu, std::vector<double> &alpha,
                                   std::vector<double> &beta) {
    nu = std::max(nu, 1);
    const double avg_eig = (eig_max + eig_min) / 2.0;
    const double diff_eig = (eig_max - eig_min) / 2.0;
    alpha.assign((size_t)std::max(nu, 1), 0.0);
    beta.assign((size_t)std::max(nu, 1), 0.0);
    double a = 0.0;
    for (int k = 0; k < nu; ++k) {
        if (k == 0) {
            a = 1.0 / avg_eig;
        } else if (k == 1) {
            const double dea = (diff_eig * a);
            beta[1] = 0.5 * dea * dea;
            a = 1.0 / (avg_eig - (beta[1] / a));
        } else {
            const double dea = (diff_eig * a);
            beta[(size_t)k] = 0.25 * dea * dea;
            a = 1.0 / (avg_eig - (beta[(size_t)k] / a));
        }
        alpha[(size_t)k] = a;
    }
}

// nu Chebyshev steps on level operator A (values T) smoothing A x = b.
// b has dtype b_dt; x (level dtype) is the initial guess unless zero_guess; the
// result goes to out (dtype out_dt; may be x itself). Work: rk, pa, pb (level dtype).
// With dot_with (FP64, only when out is FP64) the last step also writes
// dot_d[0] = out . dot_with; returns false when it could not (nu == 1).
template <typename T>
bool chebyshev_smooth_fused(const HipCrsMatrixT<T> &A, const void *b, int b_dt, T *x, bool zero_guess, void *out,
                            int out_dt, int nu, const std::vector<double> &alpha, const std::vector<double> &beta,
                            T *rk, T *pa, T *pb, const double *dot_with = nullptr, double *dot_d = nullptr) {
    constexpr int dt = HipDtype<T>::value;
    nu = std::max(nu, 1);  // sweeps(0) takes one step, as chebyshev_coefficients
    uhip_ctx c = A.c();
    const int64_t n = A.rows();
    const void *pprev;
    int p_dt;
    if (zero_guess) {  // r0 = A*0 - b = -b, p0 = b
        pprev = b;
        p_dt = b_dt;
    } else {  // p0 = b - A x  (= -(A x - b))
        hip_check(uhip_spmv_residual(c, &A.raw(), x, b, b_dt, pa), "cheb r0");
        pprev = pa;
        p_dt = dt;
    }
    if (nu == 1) {  // x = x + alpha0 p0
        hip_check(uhip_vec_axpy_cast(c, n, alpha[0], pprev, p_dt, zero_guess ? nullptr : x, dt, out, out_dt),
                  "cheb nu=1");
        return false;
    }
    T *pbuf[2] = {pb, pa};
    for (int k = 1; k < nu; ++k) {
        const bool last = (k == nu - 1);
        uhip_cheb_step st;
        std::memset(&st, 0, sizeof(st));
        st.pprev = pprev;
        st.p_dtype = p_dt;
        st.rprev = (k == 1) ? nullptr : rk;
        st.xin = (k == 1 && zero_guess) ? nullptr : x;
        st.xout = last ? out : (void *)x;
        st.x_dtype = last ? out_dt : dt;
        T *pnext = pbuf[(k - 1) & 1];
        if (pnext == pprev) pnext = pbuf[k & 1];
        st.rout = rk;
        st.pout = pnext;
        st.a_prev = alpha[(size_t)k - 1];
        st.beta = beta[(size_t)k];
        st.a_cur = alpha[(size_t)k];
        st.last = last ? 1 : 0;
        if (last && dot_with) {  // z.r in the same pass (z of either precision, r FP64)
            st.dot_with = dot_with;
            st.dot_d = dot_d;
        }
        hip_check(uhip_cheb_step_apply(c, &A.raw(), &st), "cheb step");
        pprev = pnext;
        p_dt = dt;
    }
    return dot_with != nullptr;
}

}  // namespace hip
}  // namespace utopia
