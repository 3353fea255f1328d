// MonotoneMultigrid over the HIP backend (solvers/multilevel/utopia_MonotoneMultigrid.hpp:32-378):
// a projected Gauss-Seidel QP smoother on the finest level, and a coarse-space
// correction through the finest interpolation truncated on the active set
// (IPTruncatedTransfer) into a Multigrid (GaussSeidel smoothing, no masks,
// semidefinite fix) over the remaining levels.
#pragma once

#include "utopia_hip_multigrid.hpp"
#include "utopia_hip_qp.hpp"

namespace utopia {
// the solvers of the HIP backend (the tensor types, HipVector and HipCrsMatrix, are in utopia)
namespace hip {

class MonotoneMultigrid final : public QPSolverBase {
public:
    MonotoneMultigrid() : pgs_(std::make_shared<ProjectedGaussSeidel>()) {}

    void pre_smoothing_steps(int v) { pre_ = v; }
    void post_smoothing_steps(int v) { post_ = v; }
    void coarse_space_it(int64_t v) { coarse_space_it_ = v; }
    // ActiveSet::tol (utopia_ActiveSet.hpp:69): active where x <= lb + tol or x >= ub - tol
    void active_set_tol(double v) { as_tol_ = v; }
    int64_t n_levels() const { return (int64_t)I_.size() + 1; }
    const HipVector &indicator() const { return ind_; }
    int64_t active_set_changes() const { return changes_; }
    const std::shared_ptr<Multigrid> &coarse_multigrid() const { return algo_; }

    // Interpolations coarse to fine; the finest one is truncated on the active set.
    void set_transfer_operators(const std::vector<std::shared_ptr<HipCrsMatrix>> &I) {
        if (I.empty()) throw HipError(UHIP_ERR_INVALID, "MonotoneMultigrid: at least one interpolation");
        I_ = I;
    }

    // MonotoneMultigrid::update (:134-155) with init_memory (:209-215)
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        QPSolverBase::update(op);
        auto ctx = op->context();
        const int64_t n = op->rows();
        ind_ = HipVector(ctx, n, 0.0);
        for (HipVector *v : {&help_, &x_old_, &grad_, &r_}) *v = HipVector(ctx, n, 0.0);
        pgs_->use_symmetric_sweep(true);
        pgs_->update(op);
        It_ = std::make_shared<HipCrsMatrix>(I_.back()->copy());
        galerkin_assembly();
    }

    // MonotoneMultigrid::apply (:158-207)
    bool apply(const HipVector &rhs, HipVector &x) override {
        const HipCrsMatrix &A = *this->op_;
        if (x.empty() || x.size() != A.rows()) x.zeros_like(rhs);
        // the smoother sees the caller's bounds (none: unconstrained sweeps); the active
        // set and the criticality measure use them filled with +-DBL_MAX
        filled_ = box_;
        filled_.fill_empty_bounds(A.context(), A.rows());
        this->status_.clear();
        bool converged = false;
        x_old_.copy_from(x);
        if (this->verbose_) this->print_iter(0, grad_norm(A, rhs, x));
        for (int64_t it = 1; it < this->max_it_ && !converged; ++it) {
            step(rhs, x);
            help_.copy_from(x);
            help_.axpy(-1.0, x_old_);
            const double x_diff = help_.norm2();
            const double g_norm = grad_norm(A, rhs, x);
            this->print_iter(it, g_norm);
            converged = this->check_convergence(it, g_norm, 1, x_diff);
            if (!converged) x_old_.copy_from(x);
        }
        return converged;
    }

private:
    // grad = A x - rhs; |P(x - grad) - x|_2 (criticality_measure_inf, :349-365)
    double grad_norm(const HipCrsMatrix &A, const HipVector &rhs, const HipVector &x) {
        hip_check(uhip_spmv_axpbz(A.c(), &A.raw(), 1.0, x.data(), -1.0, rhs.data(), grad_.data()), "grad");
        help_.copy_from(x);
        help_.axpy(-1.0, grad_);
        hip_check(uhip_box_project_f64(A.c(), A.rows(), help_.data(), filled_.lower->data(), filled_.upper->data(),
                                       help_.data()),
                  "project");
        help_.axpy(-1.0, x);
        return help_.norm2();
    }

    // galerkin_assembly (:217-231): A_H = I_t^T A I_t, semidefinite fix, then the
    // coarse Multigrid (or, with two levels, the coarse solver) on A_H
    void galerkin_assembly() {
        Rt_ = std::make_shared<HipCrsMatrix>(It_->transpose());
        auto AH = std::make_shared<HipCrsMatrix>(ptap(*this->op_, *It_));
        hip_check(uhip_csr_fix_semidefinite(AH->c(), &AH->raw()), "fix_semidefinite");
        AH->replan();
        AH_ = AH;
        if (n_levels() > 2) {
            algo_ = std::make_shared<Multigrid>(std::make_shared<GaussSeidel>(), std::make_shared<Factorization>());
            algo_->must_generate_masks(false);
            algo_->fix_semidefinite_operators(true);
            algo_->pre_smoothing_steps(pre_);
            algo_->post_smoothing_steps(post_);
            // the coarse-space Multigrid keeps the IterativeSolver default tolerances
            algo_->set_transfer_operators(std::vector<std::shared_ptr<HipCrsMatrix>>(I_.begin(), I_.end() - 1));
            algo_->update(AH_);
        } else {
            coarse_ = std::make_shared<Factorization>();
            coarse_->factorize(*AH_);
        }
    }

    // step (:243-318)
    void step(const HipVector &rhs, HipVector &x) {
        const HipCrsMatrix &A = *this->op_;
        uhip_ctx c = A.c();
        double *s = A.context()->scalars();
        pgs_->set_box_constraints(box_);
        pgs_->smooth(rhs, x, pre_);
        // ActiveSet::determine; truncate and reassemble when it changed
        hip_check(uhip_active_set_f64(c, A.rows(), x.data(), filled_.lower->data(), filled_.upper->data(), as_tol_,
                                      ind_.data(), s + S_CHANGED),
                  "active set");
        double changed = 0;
        A.context()->read_scalars(s + S_CHANGED, 1, &changed);
        if (changed > 0) {
            ++changes_;
            It_ = std::make_shared<HipCrsMatrix>(I_.back()->copy());
            hip_check(uhip_csr_zero_rows(c, &It_->raw(), ind_.data()), "truncate");
            It_->replan();
            galerkin_assembly();
        }
        A.residual(x, rhs, r_);  // r = rhs - A x
        HipVector rc(A.context(), Rt_->rows(), 0.0), cc(A.context(), Rt_->rows(), 0.0), cf;
        Rt_->multiply(r_, rc);
        if (n_levels() > 2) {
            algo_->max_it(coarse_space_it_);
            algo_->apply(rc, cc);
        } else {
            coarse_->apply(rc, cc);
        }
        It_->multiply(cc, cf);
        x.axpy(1.0, cf);
        pgs_->smooth(rhs, x, post_);
    }

    enum { S_CHANGED = 30 };
    BoxConstraints filled_;
    std::shared_ptr<ProjectedGaussSeidel> pgs_;
    std::vector<std::shared_ptr<HipCrsMatrix>> I_;
    std::shared_ptr<HipCrsMatrix> It_, Rt_;
    std::shared_ptr<const HipCrsMatrix> AH_;
    std::shared_ptr<Multigrid> algo_;
    std::shared_ptr<Factorization> coarse_;
    HipVector ind_, help_, x_old_, grad_, r_;
    int pre_ = 3, post_ = 3;
    int64_t coarse_space_it_ = 1, changes_ = 0;
    double as_tol_ = 0.0;
};

}  // namespace hip
}  // namespace utopia
