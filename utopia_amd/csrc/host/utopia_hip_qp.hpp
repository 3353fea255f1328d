// Box-constrained QP solvers over the HIP backend:
//   BoxConstraints        solvers/nonlinear/constrained/utopia_BoxConstraints.hpp:19-345
//   MPRGP                 solvers/nonlinear/constrained/quadratic_programming/utopia_MPRGP.hpp:14-421
//   ProjectedGaussSeidel  backend/petsc/solvers/nonlinear/constrained/quadratic_programming/
//                         utopia_petsc_ProjectedGaussSeidel.cpp:24-540 (multicolour sweep order)
//   ProjectedConjugateGradient  .../quadratic_programming/utopia_ProjectedConjugateGradient.hpp:14-143
//   ProjectedGradient     .../quadratic_programming/utopia_ProjectedGradient.hpp:18-230
#pragma once

#include <cfloat>

#include "utopia_hip_solvers.hpp"

namespace utopia {
// the solvers of the HIP backend (the tensor types, HipVector and HipCrsMatrix, are in utopia)
namespace hip {

struct BoxConstraints {
    std::shared_ptr<HipVector> lower, upper;
    bool has_bound() const { return lower || upper; }
    // BoxConstraints::fill_empty_bounds (utopia_BoxConstraints.hpp:281-300)
    void fill_empty_bounds(const HipContextPtr &ctx, int64_t n) {
        if (!lower || lower->size() != n) lower = std::make_shared<HipVector>(ctx, n, -DBL_MAX);
        if (!upper || upper->size() != n) upper = std::make_shared<HipVector>(ctx, n, DBL_MAX);
    }
};

class QPSolverBase : public IterativeSolver<HipCrsMatrix, HipVector> {
public:
    bool set_box_constraints(const BoxConstraints &box) {
        box_ = box;
        return true;
    }
    const BoxConstraints &get_box_constraints() const { return box_; }

protected:
    BoxConstraints box_;
};

// ---------------------------------------------------------------------------
// MPRGP (non-"hardik" variant, no preconditioner), utopia_MPRGP.hpp:57-262
// ---------------------------------------------------------------------------
class MPRGP final : public QPSolverBase {
public:
    void set_eig_comp_tol(double v) { eps_eig_est_ = v; }
    void power_method_max_it(int64_t v) { power_method_max_it_ = v; }
    double alpha_bar() const { return alpha_bar_; }

    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        QPSolverBase::update(op);
        const int64_t n = op->rows();
        if (fi_.size() != n || fi_.empty())
            for (HipVector *v : {&fi_, &beta_, &gp_, &p_, &Ap_, &Abeta_, &g_, &h1_, &h2_})
                *v = HipVector(op->context(), n, 0.0);
    }

    bool apply(const HipVector &rhs, HipVector &x) override {
        const HipCrsMatrix &A = *this->op_;
        const int64_t n = A.rows();
        uhip_ctx c = A.c();
        double *s = A.context()->scalars();
        if (x.empty() || x.size() != n) x.zeros_like(rhs);
        box_.fill_empty_bounds(A.context(), n);
        const double *lb = box_.lower->data(), *ub = box_.upper->data();
        const double gamma = 1.0;
        alpha_bar_ = 1.95 / power_method(A);
        this->status_.clear();
        int64_t it = 0;
        bool converged = false;
        double gnorm = -1, gnorm0 = -1;
        hip_check(uhip_box_project_f64(c, n, x.data(), lb, ub, x.data()), "project");
        hip_check(uhip_spmv_axpbz(c, &A.raw(), 1.0, x.data(), -1.0, rhs.data(), g_.data()), "g = Ax - b");
        hip_check(uhip_mprgp_state_f64(c, n, x.data(), g_.data(), lb, ub, fi_.data(), beta_.data(), gp_.data(),
                                       s + S_BB),
                  "state");
        p_.copy_from(fi_);
        double st[3];
        A.context()->read_scalars(s + S_BB, 2, st);
        double beta_beta = st[0], fi_fi = st[1];
        // A p with p.Ap and g.p, and alpha_f, of the current p: a CG step's first launches. They
        // read x, g and p only, so they are enqueued ahead of the read-back of the state that
        // decides whether the next iteration is a CG step (most are); a proportioning step then
        // leaves their results unused.
        bool ahead = false;
        auto cg_products = [&]() {
            static_assert(S_GP == S_PAP + 1 && S_AF == S_PAP + 2, "p.Ap, g.p and alpha_f in consecutive slots");
            hip_check(uhip_spmv_dot2_f64(c, &A.raw(), p_.data(), Ap_.data(), g_.data(), s + S_PAP), "Ap, g.p");
            hip_check(uhip_mprgp_alpha_f_f64(c, n, x.data(), p_.data(), lb, ub, s + S_AF), "alpha_f");
            A.context()->fetch_scalars(1, s + S_PAP, 3);
        };
        while (!converged) {
            bool state_done = false;
            if (beta_beta <= (gamma * gamma * fi_fi)) {
                if (!ahead) cg_products();
                double h[3];
                A.context()->fetch_wait(1, 3, h);
                const double pAp = h[0], gp_dot = h[1], alpha_f = h[2];
                if (pAp <= 0.0) return true;  // negative curvature
                const double alpha_cg = gp_dot / pAp;
                if (alpha_cg <= alpha_f) {
                    // the state of the new iterate comes with the step (it reads x and g only, so
                    // it may precede the p update)
                    static_assert(S_FIAP == S_BB + 3, "state dots and fi.Ap in consecutive slots");
                    hip_check(uhip_mprgp_cg_step_state_f64(c, n, x.data(), g_.data(), p_.data(), Ap_.data(), lb, ub,
                                                           fi_.data(), beta_.data(), gp_.data(), alpha_cg, s + S_BB),
                              "cg step");
                    hip_check(uhip_mprgp_pupdate_f64(c, n, p_.data(), fi_.data(), s, S_FIAP, S_PAP), "p update");
                    state_done = true;
                } else {
                    hip_check(uhip_mprgp_expand_f64(c, n, x.data(), g_.data(), p_.data(), Ap_.data(), lb, ub, alpha_f,
                                                    alpha_bar_),
                              "expansion");
                    hip_check(uhip_spmv_axpbz(c, &A.raw(), 1.0, x.data(), -1.0, rhs.data(), g_.data()), "g");
                    hip_check(uhip_mprgp_free_grad_f64(c, n, x.data(), g_.data(), lb, ub, p_.data()), "p = fi");
                }
            } else {
                static_assert(S_GB == S_BAB + 1, "beta.Abeta and g.beta in consecutive slots");
                hip_check(uhip_spmv_dot2_f64(c, &A.raw(), beta_.data(), Abeta_.data(), g_.data(), s + S_BAB),
                          "Abeta, g.beta");
                double h[2];
                A.context()->read_scalars(s + S_BAB, 2, h);
                const double beta_Abeta = h[0], g_betta = h[1];
                if (beta_Abeta <= 0.0) return true;
                const double alpha_cg = g_betta / beta_Abeta;
                hip_check(uhip_mprgp_proportion_f64(c, n, x.data(), g_.data(), beta_.data(), Abeta_.data(), lb, ub,
                                                    p_.data(), alpha_cg),
                          "proportioning");
            }
            if (!state_done)
                hip_check(uhip_mprgp_state_f64(c, n, x.data(), g_.data(), lb, ub, fi_.data(), beta_.data(),
                                               gp_.data(), s + S_BB),
                          "state");
            A.context()->fetch_scalars(0, s + S_BB, 3);
            cg_products();
            ahead = true;
            A.context()->fetch_wait(0, 3, st);
            beta_beta = st[0];
            fi_fi = st[1];
            gnorm = std::sqrt(st[2]);
            if (it == 0) gnorm0 = gnorm;
            it++;
            this->print_iter(it, gnorm);
            converged = this->check_convergence(it, gnorm, gnorm / gnorm0, 1);
        }
        return converged;
    }

private:
    enum { S_BB = 8, S_FF = 9, S_GG = 10, S_FIAP = 11, S_PAP = 12, S_GP = 13, S_AF = 14, S_BAB = 15, S_GB = 16 };

    // MPRGP::power_method (utopia_MPRGP.hpp:351-387)
    double power_method(const HipCrsMatrix &A) {
        h2_.set(1.0);
        int64_t it = 0;
        bool converged = false;
        double gnorm, lambda = 0.0, lambda_old;
        while (!converged) {
            h1_.copy_from(h2_);
            A.multiply(h1_, h2_);
            h2_.scale(1.0 / h2_.norm2());
            lambda_old = lambda;
            A.multiply(h2_, h1_);
            lambda = h2_.dot(h1_);
            hip_check(uhip_vec_waxpby_f64(A.c(), fi_.size(), 1.0, h2_.data(), -1.0, h1_.data(), fi_.data()), "fi");
            gnorm = fi_.norm2();
            converged = ((gnorm < eps_eig_est_) || (std::abs(lambda_old - lambda) < eps_eig_est_) ||
                         it > power_method_max_it_);
            it = it + 1;
        }
        return lambda;
    }

    double eps_eig_est_ = 1e-1, alpha_bar_ = 0;
    int64_t power_method_max_it_ = 10;
    HipVector fi_, beta_, gp_, p_, Ap_, Abeta_, g_, h1_, h2_;
};

// ---------------------------------------------------------------------------
// ProjectedGaussSeidel: PETSc-specialisation semantics (apply/step, check of
// |x - x_old| every check_s_norm_each iterations, n_local_sweeps = 1, no line
// search), with the lexicographic sweep replaced by a multicolour one: colours
// forward, then backward for the symmetric sweep. Rows of one colour are not
// coupled, so each colour updates in parallel.
// ---------------------------------------------------------------------------
class ProjectedGaussSeidel final : public QPSolverBase {
public:
    void use_symmetric_sweep(bool v) { symmetric_ = v; }
    void check_s_norm_each(int64_t v) { check_each_ = v; }
    int32_t ncolors() const { return ncolors_; }

    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        QPSolverBase::update(op);
        const int64_t n = op->rows();
        auto ctx = op->context();
        for (HipVector *v : {&dinv_, &r_, &c_, &x_old_}) *v = HipVector(ctx, n, 0.0);
        hip_check(uhip_pgs_dinv(op->c(), &op->raw(), dinv_.data()), "pgs dinv");
        void *p = nullptr;
        free_groups();
        ctx_ = ctx;
        hip_check(uhip_malloc(op->c(), &p, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)), "alloc colors");
        color_ = (int32_t *)p;
        hip_check(uhip_malloc(op->c(), &p, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)), "alloc rows");
        rows_ = (int32_t *)p;
        hip_check(uhip_graph_color(op->c(), &op->raw(), 0x5eedull, color_, &ncolors_), "coloring");
        offsets_.assign((size_t)ncolors_ + 1, 0);
        hip_check(uhip_color_groups(op->c(), (int32_t)n, color_, ncolors_, rows_, offsets_.data()), "groups");
    }
    ~ProjectedGaussSeidel() override { free_groups(); }

    // ProjectedGaussSeidel::smooth (utopia_petsc_ProjectedGaussSeidel.cpp:111-132): with bounds,
    // `sweeps` projected steps; without, the reference loop
    // `while (unconstrained_step(...) && it++ < sweeps)` runs sweeps + 1 steps.
    bool smooth(const HipVector &b, HipVector &x, int64_t sweeps) {
        const HipCrsMatrix &A = *this->op_;
        if (x.empty() || x.size() != A.rows()) x.zeros_like(b);
        const bool constrained = box_.has_bound();
        if (constrained) box_.fill_empty_bounds(A.context(), A.rows());
        const int64_t steps = constrained ? sweeps : sweeps + 1;
        for (int64_t s = 0; s < steps; ++s) step(A, b, x, constrained);
        return true;
    }

    // apply (utopia_petsc_ProjectedGaussSeidel.cpp:134-178)
    bool apply(const HipVector &b, HipVector &x) override {
        const HipCrsMatrix &A = *this->op_;
        const int64_t n = A.rows();
        if (x.empty() || x.size() != n) x.zeros_like(b);
        const bool constrained = box_.has_bound();
        if (constrained) box_.fill_empty_bounds(A.context(), n);
        this->status_.clear();
        x_old_.copy_from(x);
        bool converged = false;
        int64_t iteration = 0;
        while (!converged) {
            step(A, b, x, constrained);
            if (iteration % check_each_ == 0) {
                c_.copy_from(x);
                c_.axpy(-1.0, x_old_);
                const double diff = c_.norm2();
                this->print_iter(iteration, diff);
                converged = this->check_convergence(iteration, 1, 1, diff);
            }
            ++iteration;
            if (converged) break;
            x_old_.copy_from(x);
        }
        return converged;
    }

private:
    // step / unconstrained_step (utopia_petsc_ProjectedGaussSeidel.cpp:331-455)
    void step(const HipCrsMatrix &A, const HipVector &b, HipVector &x, bool constrained) {
        uhip_ctx c = A.c();
        A.residual(x, b, r_);
        c_.set(0.0);
        const double *lb = constrained ? box_.lower->data() : nullptr;
        const double *ub = constrained ? box_.upper->data() : nullptr;
        for (int pass = 0; pass < (symmetric_ ? 2 : 1); ++pass) {
            for (int32_t t = 0; t < ncolors_; ++t) {
                const int32_t k = pass == 0 ? t : ncolors_ - 1 - t;
                const int32_t o = offsets_[(size_t)k], cnt = offsets_[(size_t)k + 1] - o;
                hip_check(uhip_pgs_color_sweep(c, &A.raw(), cnt, rows_ + o, dinv_.data(), r_.data(), x.data(), lb, ub,
                                               c_.data(), constrained ? 1 : 0),
                          "pgs sweep");
            }
        }
        x.axpy(1.0, c_);
    }
    void free_groups() {
        if (ctx_) {
            if (color_) uhip_free(ctx_->get(), color_);
            if (rows_) uhip_free(ctx_->get(), rows_);
        }
        color_ = rows_ = nullptr;
    }

    bool symmetric_ = true;
    int64_t check_each_ = 10;
    int32_t ncolors_ = 0;
    int32_t *color_ = nullptr, *rows_ = nullptr;
    std::vector<int32_t> offsets_;
    HipContextPtr ctx_;
    HipVector dinv_, r_, c_, x_old_;
};

// ---------------------------------------------------------------------------
// ProjectedConjugateGradient::apply (utopia_ProjectedConjugateGradient.hpp:33-123),
// step for step; the second A*pk of the reference equals the first (pk is
// unchanged in between), so it is reused.
// ---------------------------------------------------------------------------
class ProjectedConjugateGradient final : public QPSolverBase {
public:
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        QPSolverBase::update(op);
        const int64_t n = op->rows();
        for (HipVector *v : {&x_old_, &x_half_, &uk_, &wk_, &zk_, &pk_, &Apk_, &d_}) *v = HipVector(op->context(), n, 0.0);
    }
    bool apply(const HipVector &b, HipVector &x) override {
        const HipCrsMatrix &A = *this->op_;
        const int64_t n = A.rows();
        uhip_ctx c = A.c();
        double *s = A.context()->scalars();
        if (x.empty() || x.size() != n) x.zeros_like(b);
        box_.fill_empty_bounds(A.context(), n);
        const double *lb = box_.lower->data(), *ub = box_.upper->data();
        this->status_.clear();
        x_old_.copy_from(x);
        A.residual(x, b, uk_);  // uk = b - A x
        zk_.set(0.0);           // init_memory: zk.zeros
        pk_.copy_from(uk_);
        pk_.scale(-1.0);        // pk = -uk
        bool converged = false;
        const int64_t check_each = 10;
        int64_t iteration = 0;
        while (!converged) {
            hip_check(uhip_spmv_dot(c, &A.raw(), pk_.data(), Apk_.data(), s + S_PAP), "Apk, pk.Apk");
            const double *u1[1] = {uk_.data()}, *p1[1] = {pk_.data()};
            hip_check(uhip_vec_dots_f64(c, n, 1, u1, p1, s + S_UP), "uk.pk");
            double h[2];
            A.context()->read_scalars(s + S_PAP, 2, h);  // S_PAP, S_UP adjacent
            const double alpha = h[1] / h[0];
            if (alpha == 0. || std::isinf(alpha) || std::isnan(alpha)) break;
            hip_check(uhip_vec_waxpby_f64(c, n, alpha, pk_.data(), 1.0, x_old_.data(), x_half_.data()), "x_half");
            hip_check(uhip_box_project_f64(c, n, x_half_.data(), lb, ub, x.data()), "project");
            A.residual(x, b, uk_);
            hip_check(uhip_pcg_wz_f64(c, n, uk_.data(), lb, ub, pk_.data(), zk_.data(), wk_.data()), "w, z");
            const double *w1[1] = {wk_.data()}, *a1[1] = {Apk_.data()};
            hip_check(uhip_vec_dots_f64(c, n, 1, w1, a1, s + S_WAP), "wk.Apk");
            double h3[3];
            A.context()->read_scalars(s + S_PAP, 3, h3);  // pk.Apk, uk.pk, wk.Apk
            const double beta = h3[2] / h3[0];
            hip_check(uhip_vec_waxpby_f64(c, n, 1.0, wk_.data(), beta, zk_.data(), pk_.data()), "pk = wk + beta zk");
            if (iteration % check_each == 0 || std::isinf(beta) || std::isnan(beta)) {
                d_.copy_from(x_old_);
                d_.axpy(-1.0, x);
                const double diff = d_.norm2();
                this->print_iter(iteration, diff);
                converged = this->check_convergence(iteration, 1, 1, diff);
            }
            ++iteration;
            if (converged) break;
            x_old_.copy_from(x);
        }
        return converged;
    }

private:
    enum { S_PAP = 20, S_UP = 21, S_WAP = 22 };
    HipVector x_old_, x_half_, uk_, wk_, zk_, pk_, Apk_, d_;
};

// ---------------------------------------------------------------------------
// ProjectedGradient::solve / solve_unconstrained (utopia_ProjectedGradient.hpp:54-213)
// ---------------------------------------------------------------------------
class ProjectedGradient final : public QPSolverBase {
public:
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        QPSolverBase::update(op);
        const int64_t n = op->rows();
        for (HipVector *v : {&x_old_, &x_half_, &p_, &u_, &Ap_, &d_}) *v = HipVector(op->context(), n, 0.0);
    }
    bool apply(const HipVector &b, HipVector &x) override {
        const HipCrsMatrix &A = *this->op_;
        const int64_t n = A.rows();
        uhip_ctx c = A.c();
        double *s = A.context()->scalars();
        if (x.empty() || x.size() != n) x.zeros_like(b);
        this->status_.clear();
        const bool bounded = box_.has_bound();
        if (bounded) box_.fill_empty_bounds(A.context(), n);
        const double *lb = bounded ? box_.lower->data() : nullptr, *ub = bounded ? box_.upper->data() : nullptr;
        x_old_.copy_from(x);
        A.residual(x, b, u_);  // u = b - A x
        p_.copy_from(u_);
        double alpha = 1.;
        bool converged = false;
        const int64_t check_each = 20;
        int64_t iteration = 0;
        while (!converged) {
            if (bounded) {
                hip_check(uhip_vec_waxpby_f64(c, n, 1.0, x.data(), alpha, p_.data(), x_half_.data()), "x_half");
                hip_check(uhip_box_project_f64(c, n, x_half_.data(), lb, ub, x.data()), "project");
                A.residual(x, b, u_);
                hip_check(uhip_pg_direction_f64(c, n, x.data(), u_.data(), lb, ub, p_.data()), "direction");
            } else {
                x.axpy(alpha, p_);
                A.residual(x, b, p_);  // p = b - A x
            }
            if (iteration % check_each == 0) {
                d_.copy_from(x_old_);
                d_.axpy(-1.0, x);
                const double diff = d_.norm2();
                this->print_iter(iteration, diff);
                converged = this->check_convergence(iteration, 1, 1, diff);
            }
            ++iteration;
            if (converged) break;
            x_old_.copy_from(x);
            hip_check(uhip_spmv_dot(c, &A.raw(), p_.data(), Ap_.data(), s + S_PAP), "Ap, p.Ap");
            const double *a1[1] = {bounded ? u_.data() : p_.data()}, *p1[1] = {p_.data()};
            hip_check(uhip_vec_dots_f64(c, n, 1, a1, p1, s + S_NUM), bounded ? "u.p" : "p.p");
            double h[2];
            A.context()->read_scalars(s + S_PAP, 2, h);
            alpha = h[1] / h[0];
            if (std::isinf(alpha) || alpha == 0. || std::isnan(alpha)) {
                d_.copy_from(x_old_);
                d_.axpy(-1.0, x);
                const double diff = d_.norm2();
                converged = this->check_convergence(iteration, 1, 1, diff);
                break;
            }
        }
        return converged;
    }

private:
    enum { S_PAP = 24, S_NUM = 25 };
    HipVector x_old_, x_half_, p_, u_, Ap_, d_;
};

}  // namespace hip
}  // namespace utopia
