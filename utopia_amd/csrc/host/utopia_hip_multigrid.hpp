// Multigrid over the HIP backend: Galerkin hierarchy, masks, smoothers, dense
// coarse solve, and the fused V-cycle engine (FP64 or FP32 levels).
#pragma once

#include <type_traits>

#include <chrono>

#include "utopia_hip_solvers.hpp"

namespace utopia {
// the solvers of the HIP backend (the tensor types, HipVector and HipCrsMatrix, are in utopia)
namespace hip {

// Factorization<M,V> on the coarsest level: dense FP64 Cholesky, then the
// explicit inverse so that every coarse solve is a single GEMV.
class Factorization final : public LinearSolver<HipCrsMatrix, HipVector> {
public:
    ~Factorization() override { release(); }
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        LinearSolver::update(op);
        factorize(*op);
    }
    void factorize(const HipCrsMatrix &A) {
        release();
        ctx_ = A.context();
        n_ = (int32_t)A.rows();
        if (A.rows() != A.cols()) throw HipError(UHIP_ERR_INVALID, "Factorization: square matrix required");
        void *p = nullptr, *q = nullptr;
        hip_check(uhip_malloc(ctx_->get(), &p, sizeof(double) * (size_t)n_ * n_), "alloc dense");
        hip_check(uhip_malloc(ctx_->get(), &q, sizeof(double) * (size_t)n_ * n_), "alloc inverse");
        double *L = (double *)p;
        ainv_ = (double *)q;
        hip_check(uhip_csr_to_dense(ctx_->get(), &A.raw(), L), "to_dense");
        int info = 0;
        hip_check(uhip_dense_cholesky(ctx_->get(), n_, L, n_, &info), "cholesky");
        if (info != 0) {
            uhip_free(ctx_->get(), L);
            throw HipError(UHIP_ERR_NUMERIC, "Factorization: coarse operator not SPD (pivot " + std::to_string(info) + ")");
        }
        hip_check(uhip_dense_cholesky_inverse(ctx_->get(), n_, L, n_, ainv_), "cholesky inverse");
        hip_check(uhip_free(ctx_->get(), L), "free");
    }
    bool apply(const HipVector &rhs, HipVector &sol) override {
        if (sol.size() != rhs.size() || sol.empty()) sol.zeros_like(rhs);
        return solve_raw(rhs.data(), UHIP_F64, sol.data(), UHIP_F64);
    }
    // the coarsest solve stays FP64 in every form: the FP64 inverse, FP64 sums; an
    // FP32 V-cycle hands FP32 vectors in and out (converted exactly on load)
    bool solve_raw(const void *rhs, int rhs_dt, void *sol, int sol_dt) const {
        hip_check(uhip_dense_gemv(ctx_->get(), n_, ainv_, rhs, rhs_dt, sol, sol_dt), "coarse gemv");
        return true;
    }
    int32_t n() const { return n_; }
    const double *inverse() const { return ainv_; }

private:
    void release() {
        if (ainv_ && ctx_) uhip_free(ctx_->get(), ainv_);
        ainv_ = nullptr;
    }
    HipContextPtr ctx_;
    double *ainv_ = nullptr;
    int32_t n_ = 0;
};

// Smoother prototypes handed to Multigrid (cloned per level by the reference).
template <class Matrix, class Vector>
class Chebyshev3level final : public IterativeSolver<Matrix, Vector> {
public:
    using T = typename Vector::Scalar;
    Chebyshev3level() { this->max_it_ = 1; }
    void scale_max_eig(double v) { scale_max_eig_ = v; }
    void scale_min_eig(double v) { scale_min_eig_ = v; }
    double scale_max_eig() const { return scale_max_eig_; }
    double scale_min_eig() const { return scale_min_eig_; }
    void power_method_tol(double t) { pm_.tol(t); }
    void power_method_max_it(int64_t m) { pm_.max_it(m); }
    const PowerMethod<Matrix, Vector> &power_method() const { return pm_; }
    void set_eigs(double emax, double emin) {
        eig_max_ = emax;
        eig_min_ = emin;
    }
    double eig_max() const { return eig_max_; }
    double eig_min() const { return eig_min_; }

    // update -> init_eigs (utopia_Chebyshev3level.hpp:90-121)
    void update(const std::shared_ptr<const Matrix> &op) override {
        IterativeSolver<Matrix, Vector>::update(op);
        eig_max_ = scale_max_eig_ * pm_.get_max_eig(*op);
        eig_min_ = scale_min_eig_ * eig_max_;
        const int64_t n = op->rows();
        for (Vector *v : {&rk_, &pa_, &pb_}) *v = Vector(op->context(), n, 0);
    }
    // smooth: max_it(sweeps) steps from the given x (utopia_Chebyshev3level.hpp:109-116)
    bool apply(const Vector &b, Vector &x) override {
        if (x.empty() || x.size() != b.size()) x.zeros_like(b);
        std::vector<double> a, be;
        chebyshev_coefficients(eig_max_, eig_min_, (int)this->max_it_, a, be);
        chebyshev_smooth_fused<T>(*this->op_, b.data(), Vector::dtype, x.data(), false, x.data(), Vector::dtype,
                                  (int)this->max_it_, a, be, rk_.data(), pa_.data(), pb_.data());
        return true;
    }

private:
    double scale_max_eig_ = 1.2, scale_min_eig_ = 0.06, eig_max_ = 0, eig_min_ = 0;
    PowerMethod<Matrix, Vector> pm_;
    Vector rk_, pa_, pb_;
};

// Jacobi (utopia_Jacobi.hpp:89-140): r = rhs - A x; x += d_inv .* r per sweep, d_inv = pseudo_inv(diag, 0)
template <class Matrix, class Vector>
class Jacobi final : public IterativeSolver<Matrix, Vector> {
public:
    using T = typename Vector::Scalar;
    Jacobi() { this->max_it_ = 1; }
    void update(const std::shared_ptr<const Matrix> &op) override {
        IterativeSolver<Matrix, Vector>::update(op);
        op->build_diag(d_inv_);
        InvDiagPreconditioner<Matrix, Vector>::pseudo_inv(d_inv_, T(0));
        r_ = Vector(op->context(), op->rows(), 0);
    }
    bool apply(const Vector &b, Vector &x) override {
        if (x.empty() || x.size() != b.size()) x.zeros_like(b);
        for (int64_t s = 0; s < this->max_it_; ++s) {
            this->op_->residual(x, b, r_);
            r_.e_mul(d_inv_);
            x.axpy(T(1), r_);
        }
        return true;
    }
    const Vector &d_inv() const { return d_inv_; }

private:
    Vector d_inv_, r_;
};

// PointJacobi (solvers/smoothers/utopia_PointJacobi.hpp:70-145): LU = A - D kept as a matrix
// (the diagonal entries stay in the pattern as zeros), d_inv = 1./diag(A) in the form of
// VecReciprocal (zeros kept) with check_indef turning zeros into ones; a sweep is
// x = d_inv .* (rhs - LU x).
template <class Matrix, class Vector>
class PointJacobi final : public IterativeSolver<Matrix, Vector> {
public:
    using T = typename Vector::Scalar;
    PointJacobi() { this->max_it_ = 1; }
    void update(const std::shared_ptr<const Matrix> &op) override {
        IterativeSolver<Matrix, Vector>::update(op);
        make_point_jacobi(*op, LU_, d_inv_);
        r_ = Vector(op->context(), op->rows(), 0);
    }
    // LU = A - diag(A) (pattern kept) and d_inv = check_indef(1 ./ diag(A))
    static void make_point_jacobi(const Matrix &A, std::shared_ptr<const Matrix> &LU, Vector &d_inv) {
        A.build_diag(d_inv);
        auto m = std::make_shared<Matrix>(A.copy());
        m->shift_diag(d_inv, T(-1));
        LU = m;
        d_inv.reciprocal(T(1));
        d_inv.replace(T(0), T(1));
    }
    // smooth: sweeps from the given x
    bool apply(const Vector &b, Vector &x) override {
        if (x.empty() || x.size() != b.size()) x.zeros_like(b);
        for (int64_t s = 0; s < this->max_it_; ++s) {
            hip_check(uhip_spmv_jacobi(x.ctx(), &LU_->raw(), x.data(), b.data(), Vector::dtype, d_inv_.data(),
                                       r_.data(), Vector::dtype, 0),
                      "point jacobi sweep");
            x.copy_from(r_);
        }
        return true;
    }
    const Vector &d_inv() const { return d_inv_; }
    const Matrix &lu() const { return *LU_; }

private:
    std::shared_ptr<const Matrix> LU_;
    Vector d_inv_, r_;
};

// ---------------------------------------------------------------------------
// Rows of a matrix grouped by colour (Jones-Plassmann colouring with a fixed
// seed): rows of one colour are not coupled, so a Gauss-Seidel sweep updates
// a whole colour in parallel.
// ---------------------------------------------------------------------------
class ColorGroups {
public:
    static constexpr uint64_t kSeed = 0x5eedull;
    ColorGroups() = default;
    ColorGroups(const ColorGroups &) = delete;
    ColorGroups &operator=(const ColorGroups &) = delete;
    ~ColorGroups() { release(); }
    void build(const HipCrsMatrix &A) {
        release();
        ctx_ = A.context();
        const int64_t n = A.rows();
        void *p = nullptr;
        hip_check(uhip_malloc(A.c(), &p, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)), "alloc colors");
        color_ = (int32_t *)p;
        hip_check(uhip_malloc(A.c(), &p, sizeof(int32_t) * (size_t)(n > 0 ? n : 1)), "alloc rows");
        rows_ = (int32_t *)p;
        hip_check(uhip_graph_color(A.c(), &A.raw(), kSeed, color_, &ncolors_), "coloring");
        offsets_.assign((size_t)ncolors_ + 1, 0);
        hip_check(uhip_color_groups(A.c(), (int32_t)n, color_, ncolors_, rows_, offsets_.data()), "colour groups");
    }
    int32_t ncolors() const { return ncolors_; }
    const int32_t *color() const { return color_; }
    // c = 0 on entry; forward over the colours, then (symmetric) backward:
    // c_i = d_inv_i (r_i - sum_{j != i} a_ij c_j)
    void sweep(const HipCrsMatrix &A, const double *dinv, const double *r, double *c, bool symmetric) const {
        for (int pass = 0; pass < (symmetric ? 2 : 1); ++pass)
            for (int32_t t = 0; t < ncolors_; ++t) {
                const int32_t k = pass == 0 ? t : ncolors_ - 1 - t;
                const int32_t o = offsets_[(size_t)k], cnt = offsets_[(size_t)k + 1] - o;
                hip_check(uhip_pgs_color_sweep(A.c(), &A.raw(), cnt, rows_ + o, dinv, r, nullptr, nullptr, nullptr,
                                               c, 0),
                          "gauss-seidel colour sweep");
            }
    }

private:
    void release() {
        if (ctx_) {
            if (color_) uhip_free(ctx_->get(), color_);
            if (rows_) uhip_free(ctx_->get(), rows_);
        }
        color_ = rows_ = nullptr;
    }
    HipContextPtr ctx_;
    int32_t ncolors_ = 0;
    int32_t *color_ = nullptr, *rows_ = nullptr;
    std::vector<int32_t> offsets_;
};

// GaussSeidel (solvers/smoothers/utopia_GaussSeidel.hpp:45-57, defaults :25 and
// :232-234: one local sweep, symmetric, no l1, no line search): per sweep
// r = rhs - A x; c = 0; symmetric Gauss-Seidel on D c = r - (L+U) c; x += c;
// d_inv = 1/diag. The lexicographic row order of the reference sweep becomes a
// colour order.
class GaussSeidel final : public IterativeSolver<HipCrsMatrix, HipVector> {
public:
    GaussSeidel() { this->max_it_ = 1; }
    void use_symmetric_sweep(bool v) { symmetric_ = v; }
    bool symmetric() const { return symmetric_; }
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        IterativeSolver::update(op);
        d_inv_ = HipVector(op->context(), op->rows(), 0.0);
        hip_check(uhip_pgs_dinv(op->c(), &op->raw(), d_inv_.data()), "gauss-seidel d_inv");
        groups_ = std::make_shared<ColorGroups>();
        groups_->build(*op);
        r_ = HipVector(op->context(), op->rows(), 0.0);
        c_ = HipVector(op->context(), op->rows(), 0.0);
    }
    bool apply(const HipVector &b, HipVector &x) override {
        if (x.empty() || x.size() != b.size()) x.zeros_like(b);
        for (int64_t s = 0; s < this->max_it_; ++s) {
            this->op_->residual(x, b, r_);
            c_.set(0.0);
            groups_->sweep(*this->op_, d_inv_.data(), r_.data(), c_.data(), symmetric_);
            x.axpy(1.0, c_);
        }
        return true;
    }

private:
    bool symmetric_ = true;
    HipVector d_inv_, r_, c_;
    std::shared_ptr<ColorGroups> groups_;
};

// ---------------------------------------------------------------------------
// The V-cycle engine on levels with values T. Level 0 is the coarsest.
// ---------------------------------------------------------------------------
template <typename T>
struct MGLevel {
    std::shared_ptr<const HipCrsMatrixT<T>> A;  // level operator
    std::shared_ptr<const HipCrsMatrixT<T>> R;  // restriction to level l-1 (l >= 1)
    std::shared_ptr<const HipCrsMatrixT<T>> P;  // interpolation from level l-1 (l >= 1)
    std::shared_ptr<const HipCrsMatrixT<T>> LU; // PointJacobi: A - D (l >= 1)
    HipVectorT<T> mask;                        // mask of this level (empty: masks off)
    HipVectorT<T> r, c, rR, rk, pa, pb, dinv;
    HipVectorT<T> cI, f;                       // line search: I c_{l-1}; full cycle: restricted rhs
    std::shared_ptr<ColorGroups> groups;        // Gauss-Seidel colour groups (FP64 levels)
    std::vector<double> alpha, beta;            // Chebyshev coefficients for pre / post
    std::vector<double> alpha_post, beta_post;
    double eig_max = 0, eig_min = 0;
    bool pair = false;  // Chebyshev smoothing pairs fused with their residuals (uhip_cheb_pre_residual)
};

enum MGSmoother { MG_CHEBYSHEV = 0, MG_JACOBI = 1, MG_GAUSS_SEIDEL = 2, MG_POINT_JACOBI = 3 };
// MultiLevelBase cycle types (solvers/multilevel/utopia_MultiLevelBase.hpp:12-14)
enum MGCycle { MG_MULTIPLICATIVE_CYCLE = 1, MG_ADDITIVE_CYCLE = 2, MG_FULL_CYCLE = 3 };

template <typename T>
class VCycle {
public:
    static constexpr int dt = HipDtype<T>::value;
    std::vector<MGLevel<T>> levels;
    const Factorization *coarse = nullptr;
    int pre = 3, post = 3, mg_type = 1, smoother = MG_CHEBYSHEV;
    bool masks = true, gs_symmetric = true, line_search = false;
    // smoothing pairs in one pass: 0 never, 1 on levels whose operator allows it and whose pair
    // launch gives every CU a workgroup (kPairFillGroups), 2 wherever the operator allows it
    int fuse_pairs = 1;
    static constexpr int64_t kPairFillGroups = 256;  // the CUs of an MI355X

    void allocate_work() {
        for (size_t l = 0; l < levels.size(); ++l) {
            auto &L = levels[l];
            const int64_t n = L.A->rows();
            auto ctx = L.A->context();
            L.pair = fuse_pairs > 0 && smoother == MG_CHEBYSHEV && l > 0 && uhip_cheb_pair_supported(&L.A->raw()) &&
                     (fuse_pairs == 2 || uhip_cheb_pair_workgroups(&L.A->raw()) >= kPairFillGroups);
            for (HipVectorT<T> *v : {&L.r, &L.c, &L.rR, &L.rk, &L.pa, &L.pb}) *v = HipVectorT<T>(ctx, n, T(0));
            if (line_search && l > 0) L.cI = HipVectorT<T>(ctx, n, T(0));
            if (smoother == MG_CHEBYSHEV && l > 0) {
                chebyshev_coefficients(L.eig_max, L.eig_min, pre, L.alpha, L.beta);
                chebyshev_coefficients(L.eig_max, L.eig_min, post, L.alpha_post, L.beta_post);
            }
        }
        if (line_search && !levels.empty()) ls_ = HipVectorT<double>(levels.back().A->context(), 2, 0.0);
    }

    // Multigrid::standard_cycle (utopia_Multigrid.hpp:252-367). r: level rhs
    // (dtype r_dt); the correction c_l is accumulated in levels[l].c and the
    // final post-smoothed value is written to out (dtype out_dt).
    // With dot_with, also dot_d[0] = out . dot_with when the last kernel can
    // fuse it (returns whether it did).
    bool cycle(int l, const void *r, int r_dt, void *out, int out_dt, const double *dot_with = nullptr,
               double *dot_d = nullptr) {
        auto &L = levels[(size_t)l];
        const int64_t n = L.A->rows();
        uhip_ctx c = L.A->c();
        bool dot_done = false;
        if (l == 0) {
            coarse->solve_raw(r, r_dt, out, out_dt);
            return false;
        }
        auto &Lc = levels[(size_t)l - 1];
        for (int k = 0; k < mg_type; ++k) {
            const bool first = (k == 0);
            const bool final_pass = (k == mg_type - 1);
            if (L.pair && first && pre == 2 && r_dt == dt) {
                // the two-step smoothing from c = 0 and r_R = r - A c in one pass
                hip_check(uhip_cheb_pre_residual(c, &L.A->raw(), r, L.alpha[0], L.beta[1], L.alpha[1], L.c.data(),
                                                 L.rR.data()),
                          "mg pre-smoothing and residual");
            } else {
                smooth(L, r, r_dt, /*zero_guess=*/first, L.c.data(), dt, pre, L.alpha, L.beta);
                // r_R = r - A c
                hip_check(uhip_spmv_residual(c, &L.A->raw(), L.c.data(), r, r_dt, L.rR.data()), "mg residual");
            }
            // r_{l-1} = mask .* (R r_R)
            hip_check(uhip_spmv_masked(c, &L.R->raw(), L.rR.data(), masks && !Lc.mask.empty() ? Lc.mask.data() : nullptr,
                                       Lc.r.data()),
                      "mg restrict");
            cycle(l - 1, Lc.r.data(), dt, Lc.c.data(), dt);
            // c += I c_{l-1}
            if (line_search) line_search_correct(L, Lc);
            else hip_check(uhip_spmv_add(c, &L.P->raw(), Lc.c.data(), L.c.data()), "mg interpolate");
            void *o = final_pass ? out : (void *)L.c.data();
            const int odt = final_pass ? out_dt : dt;
            if (L.pair && post == 2 && r_dt == dt && odt == dt && o != (void *)L.c.data() && o != r) {
                // p0 = r - A c and the last smoothing step on it in one pass
                const double *dw = final_pass ? dot_with : nullptr;
                hip_check(uhip_cheb_residual_post(c, &L.A->raw(), L.c.data(), r, L.alpha_post[0], L.beta_post[1],
                                                  L.alpha_post[1], o, dw, dot_d),
                          "mg residual and post-smoothing");
                dot_done = dw != nullptr;
            } else {
                dot_done = smooth(L, r, r_dt, false, o, odt, post, L.alpha_post, L.beta_post,
                                  final_pass ? dot_with : nullptr, dot_d);
            }
        }
        (void)n;
        return dot_done;
    }

    // Multigrid::full_cycle (utopia_Multigrid.hpp:381-415), FP64 levels: the right-hand side
    // r of level l restricted (and masked) to every level, the coarse solution interpolated
    // upwards with v_cycle_repetition standard cycles on each intermediate level, then
    // v_cycle_repetition standard cycles on level l into out. The steps are the reference's;
    // every standard cycle starts from c = 0 (:264-268), so the result equals one standard
    // cycle of r.
    void full_cycle(int l, const double *r, double *out, int64_t repetitions) {
        if constexpr (!std::is_same<T, double>::value) {
            throw HipError(UHIP_ERR_UNSUPPORTED, "full cycle: FP64 levels only");
        } else {
            if (l == 0) {
                coarse->solve_raw(r, UHIP_F64, out, UHIP_F64);
                return;
            }
            uhip_ctx c = levels[(size_t)l].A->c();
            for (int i = l - 1; i >= 0; --i) {
                auto &Li = levels[(size_t)i];
                auto &Lf = levels[(size_t)i + 1];
                if (Li.f.size() != Li.A->rows() || Li.f.empty()) Li.f = HipVectorT<T>(Li.A->context(), Li.A->rows());
                hip_check(uhip_spmv_masked(c, &Lf.R->raw(), i == l - 1 ? (const void *)r : (const void *)Lf.f.data(),
                                           masks && !Li.mask.empty() ? Li.mask.data() : nullptr, Li.f.data()),
                          "full cycle restrict");
            }
            coarse->solve_raw(levels[0].f.data(), UHIP_F64, levels[0].c.data(), UHIP_F64);
            hip_check(uhip_spmv(c, &levels[1].P->raw(), levels[0].c.data(), levels[1].c.data()), "full cycle interpolate");
            for (int i = 1; i < l; ++i) {
                auto &Li = levels[(size_t)i];
                for (int64_t j = 0; j < repetitions; ++j) cycle(i, Li.f.data(), dt, Li.c.data(), dt);
                hip_check(uhip_spmv(c, &levels[(size_t)i + 1].P->raw(), Li.c.data(), levels[(size_t)i + 1].c.data()),
                          "full cycle interpolate");
            }
            for (int64_t j = 0; j < repetitions; ++j) cycle(l, r, UHIP_F64, out, UHIP_F64);
        }
    }

private:
    HipVectorT<double> ls_;  // line search: c_I.r_R, c_I.A c_I

    static int vdot(uhip_ctx c, int64_t n, const T *x, const T *y, double *d) {
        if constexpr (std::is_same<T, double>::value) return uhip_vec_dot_f64(c, n, x, y, d);
        else return uhip_vec_dot_f32(c, n, x, y, d);
    }
    static int vaxpy(uhip_ctx c, int64_t n, T a, const T *x, T *y) {
        if constexpr (std::is_same<T, double>::value) return uhip_vec_axpy_f64(c, n, a, x, y);
        else return uhip_vec_axpy_f32(c, n, a, x, y);
    }

    // Multigrid::standard_cycle with use_line_search (utopia_Multigrid.hpp:324-341, the
    // symmetric form): c_I = I c_{l-1}; alpha = (c_I . r_R) / ((A c_I) . c_I);
    // c += alpha c_I, or c += c_I when alpha <= 0
    void line_search_correct(MGLevel<T> &L, MGLevel<T> &Lc) {
        uhip_ctx c = L.A->c();
        const int64_t n = L.A->rows();
        hip_check(uhip_spmv(c, &L.P->raw(), Lc.c.data(), L.cI.data()), "mg interpolate");
        hip_check(vdot(c, n, L.cI.data(), L.rR.data(), ls_.data()), "line search c_I.r_R");
        hip_check(uhip_spmv_dot(c, &L.A->raw(), L.cI.data(), L.pb.data(), ls_.data() + 1), "line search A c_I");
        double h[2] = {0, 0};
        L.A->context()->read_scalars(ls_.data(), 2, h);
        const double alpha = h[0] / h[1];
        hip_check(vaxpy(c, n, alpha <= 0. ? T(1) : T(alpha), L.cI.data(), L.c.data()), "c += alpha c_I");
    }

    bool smooth(MGLevel<T> &L, const void *b, int b_dt, bool zero_guess, void *out, int out_dt, int nu,
                const std::vector<double> &a, const std::vector<double> &be, const double *dot_with = nullptr,
                double *dot_d = nullptr) {
        if (nu <= 0 && smoother != MG_CHEBYSHEV) {
            // no sweeps: the result is the guess itself, zero on a zero guess (the
            // standard cycle starts every pass from c = 0). Chebyshev3level takes one
            // step even for sweeps(0) (chebyshev_coefficients).
            if (zero_guess) L.c.set(T(0));
            if (out != (void *)L.c.data())
                hip_check(uhip_vec_axpy_cast(L.A->c(), L.A->rows(), 0.0, L.c.data(), dt, L.c.data(), dt, out, out_dt),
                          "copy");
            return false;
        }
        if (smoother == MG_CHEBYSHEV) {
            return chebyshev_smooth_fused<T>(*L.A, b, b_dt, L.c.data(), zero_guess, out, out_dt, nu, a, be,
                                             L.rk.data(), L.pa.data(), L.pb.data(), dot_with, dot_d);
        }
        uhip_ctx c = L.A->c();
        const int64_t n = L.A->rows();
        if (smoother == MG_GAUSS_SEIDEL) {
            if constexpr (std::is_same<T, double>::value) {
                // per sweep: r = b - A x (= b on a zero guess); c = 0; colour sweeps; x += c
                if (b_dt != UHIP_F64 || out_dt != UHIP_F64)
                    throw HipError(UHIP_ERR_UNSUPPORTED, "Gauss-Seidel smoothing runs on FP64 levels");
                if (zero_guess) L.c.set(0.0);
                for (int s = 0; s < nu; ++s) {
                    const double *r = (const double *)b;
                    if (!(zero_guess && s == 0)) {
                        hip_check(uhip_spmv_residual(c, &L.A->raw(), L.c.data(), b, b_dt, L.rk.data()), "gs r");
                        r = L.rk.data();
                    }
                    L.pa.set(0.0);
                    L.groups->sweep(*L.A, L.dinv.data(), r, L.pa.data(), gs_symmetric);
                    const bool last = (s == nu - 1);
                    hip_check(uhip_vec_axpy_cast(c, n, 1.0, L.pa.data(), dt, L.c.data(), dt,
                                                 last ? out : (void *)L.c.data(), last ? out_dt : dt),
                              "gs x");
                }
                return false;
            } else {
                throw HipError(UHIP_ERR_UNSUPPORTED, "Gauss-Seidel smoothing runs on FP64 levels");
            }
        }
        // Jacobi sweeps x = x + d_inv .* (b - A x), or PointJacobi sweeps x = d_inv .* (b - LU x),
        // each one fused SpMV pass, alternating between c and rk (x = 0 on a zero guess)
        const bool point = smoother == MG_POINT_JACOBI;
        const HipCrsMatrixT<T> &M = point ? *L.LU : *L.A;
        if (zero_guess) L.c.set(T(0));
        T *cur = L.c.data(), *other = L.rk.data();
        for (int s = 0; s < nu; ++s) {
            const bool last = (s == nu - 1);
            if (last && out != (void *)cur) {
                hip_check(uhip_spmv_jacobi(c, &M.raw(), cur, b, b_dt, L.dinv.data(), out, out_dt, point ? 0 : 1),
                          "jacobi sweep");
                break;
            }
            hip_check(uhip_spmv_jacobi(c, &M.raw(), cur, b, b_dt, L.dinv.data(), other, dt, point ? 0 : 1),
                      "jacobi sweep");
            std::swap(cur, other);
            if (last)  // out is the buffer the sweep read: the result is in the other one
                hip_check(uhip_vec_axpy_cast(c, n, 0.0, cur, dt, cur, dt, out, out_dt), "copy");
        }
        return false;
    }
};

// z = M rhs for a V-cycle engine, plus dot_d = z . rhs (fused into the last
// kernel when possible).
template <typename T>
inline void vcycle_apply_dot(VCycle<T> &eng, const HipVector &rhs, HipVector &z, double *dot_d,
                             const float *rhs32 = nullptr) {
    // rhs32 = (float)rhs: an FP32 cycle reads it in place of rhs (same values after
    // the conversion every FP32 kernel applies), the dot still uses the FP64 rhs
    const void *r = rhs32 ? (const void *)rhs32 : (const void *)rhs.data();
    const bool fused = eng.cycle((int)eng.levels.size() - 1, r, rhs32 ? UHIP_F32 : UHIP_F64, z.data(), UHIP_F64,
                                 rhs.data(), dot_d);
    if (!fused) {
        const double *a[1] = {z.data()}, *b[1] = {rhs.data()};
        hip_check(uhip_vec_dots_f64(z.ctx(), z.size(), 1, a, b, dot_d), "dot(z, r)");
    }
}

// ---------------------------------------------------------------------------
// Multigrid<HipCrsMatrix, HipVector>
// ---------------------------------------------------------------------------
class Multigrid : public IterativeSolver<HipCrsMatrix, HipVector> {
public:
    using Smoother = IterativeSolver<HipCrsMatrix, HipVector>;
    static const int V_CYCLE = 1;
    static const int W_CYCLE = 2;

    Multigrid(const std::shared_ptr<Smoother> &smoother, const std::shared_ptr<Factorization> &coarse_solver)
        : smoother_(smoother), coarse_(coarse_solver) {}

    // Interpolation operators ordered coarse to fine; R = transpose(I) kept
    // explicitly (LinearMultiLevel::set_transfer_operators + IPRTransfer).
    bool set_transfer_operators(const std::vector<std::shared_ptr<HipCrsMatrix>> &interpolations) {
        I_.clear();
        R_.clear();
        for (auto &I : interpolations) {
            I_.push_back(I);
            R_.push_back(std::make_shared<HipCrsMatrix>(I->transpose()));
        }
        n_levels_ = (int)interpolations.size() + 1;
        return true;
    }

    void pre_smoothing_steps(int n) { pre_ = n; }
    void post_smoothing_steps(int n) { post_ = n; }
    int pre_smoothing_steps() const { return pre_; }
    int post_smoothing_steps() const { return post_; }
    void mg_type(int t) { mg_type_ = t; }
    // MultiLevelBase::cycle_type / v_cycle_repetition (utopia_MultiLevelBase.hpp:79-90),
    // Multigrid::use_line_search (utopia_Multigrid.hpp:480)
    void cycle_type(int t) { cycle_type_ = t; }
    int cycle_type() const { return cycle_type_; }
    void v_cycle_repetition(int64_t n) { v_cycle_repetition_ = n; }
    int64_t v_cycle_repetition() const { return v_cycle_repetition_; }
    void use_line_search(bool v) {
        line_search_ = v;
        if (eng_.line_search != v && !eng_.levels.empty()) {
            eng_.line_search = v;
            eng_.allocate_work();
        }
    }
    bool use_line_search() const { return line_search_; }
    // the pairs of products where the pre-smoothing ends and the post-smoothing starts run in one
    // pass each on operators that allow it (uhip_cheb_pair_supported): 1 where the pair launch
    // fills the GPU, 2 wherever allowed; 0: one launch per product
    void fused_smoothing(int v) {
        fused_ = v;
        if (eng_.fuse_pairs != v && !eng_.levels.empty()) {
            eng_.fuse_pairs = v;
            eng_.allocate_work();
        }
    }
    int fused_smoothing() const { return fused_; }
    void must_generate_masks(bool v) { masks_ = v; }
    bool must_generate_masks() const { return masks_; }
    void fix_semidefinite_operators(bool v) { fix_semi_ = v; }
    // LinearMultiLevel::set_linear_operators (utopia_LinearMultiLevel.hpp:138-150), used with
    // set_perform_galerkin_assembly(false) (utopia_Multigrid.hpp:129-153): the level operators
    // are taken as given (coarse to fine); no masks are generated and no semidefinite fix applied.
    void set_linear_operators(const std::vector<std::shared_ptr<const HipCrsMatrix>> &A) { given_ = A; }
    void set_perform_galerkin_assembly(bool v) { galerkin_ = v; }
    int n_levels() const { return n_levels_; }
    const HipCrsMatrix &level_operator(int l) const { return *A_[(size_t)l]; }
    const HipVector &mask(int l) const { return mask_[(size_t)l]; }
    double setup_seconds() const { return setup_s_; }
    const VCycle<double> &engine() const { return eng_; }
    const std::shared_ptr<Factorization> &coarse_solver() const { return coarse_; }
    // Chebyshev bounds of level l (for tests that compare against fixed bounds).
    void set_level_eigs(int l, double emax, double emin) {
        eng_.levels[(size_t)l].eig_max = emax;
        eng_.levels[(size_t)l].eig_min = emin;
        eng_.allocate_work();
    }
    double level_eig_max(int l) const { return eng_.levels[(size_t)l].eig_max; }
    double level_eig_min(int l) const { return eng_.levels[(size_t)l].eig_min; }

    // Multigrid::update (utopia_Multigrid.hpp:129-153)
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        const auto t0 = std::chrono::steady_clock::now();
        IterativeSolver::update(op);
        const int L = n_levels_;
        if (L < 2 || (int)I_.size() != L - 1) throw HipError(UHIP_ERR_INVALID, "Multigrid: set_transfer_operators first");
        auto ctx = op->context();
        A_.assign((size_t)L, nullptr);
        A_[(size_t)L - 1] = op;
        const bool masks = masks_ && galerkin_;
        // masks (MultiLevelMask::generate_masks, before the Galerkin products)
        mask_.assign((size_t)L, HipVector());
        if (masks) {
            HipVector &mf = mask_[(size_t)L - 1];
            mf = HipVector(ctx, op->rows());
            hip_check(uhip_csr_offdiag_indicator(op->c(), &op->raw(), std::numeric_limits<double>::epsilon() * 1e6,
                                                 mf.data()),
                      "mask");
            mf.scale(-1.0);  // off_value 1, on_value 0 where an off-diagonal exists
            mf.shift(1.0);
            for (int l = L - 1; l > 0; --l) {
                HipVector &mc = mask_[(size_t)l - 1];
                mc = HipVector(ctx, R_[(size_t)l - 1]->rows());
                hip_check(uhip_csr_boolean_restrict_or(op->c(), &R_[(size_t)l - 1]->raw(),
                                                       std::numeric_limits<double>::epsilon() * 1e6,
                                                       mask_[(size_t)l].data(), mc.data()),
                          "boolean_restrict_or");
            }
            for (auto &m : mask_) {
                m.shift(-1.0);
                m.abs();
            }
        }
        // Galerkin assembly (LinearMultiLevel::galerkin_assembly)
        if (galerkin_) {
            for (int l = L - 1; l > 0; --l) {
                auto Ac = std::make_shared<HipCrsMatrix>(ptap(*A_[(size_t)l], *I_[(size_t)l - 1]));
                if (fix_semi_) hip_check(uhip_csr_fix_semidefinite(Ac->c(), &Ac->raw()), "fix_semidefinite");
                A_[(size_t)l - 1] = Ac;
            }
        } else {
            if ((int)given_.size() != L) throw HipError(UHIP_ERR_INVALID, "Multigrid: set_linear_operators first");
            for (int l = 0; l < L - 1; ++l) A_[(size_t)l] = given_[(size_t)l];
        }
        // smoothers, levels 1..L-1 in this order (rand() draws of the power method)
        auto cheb = std::dynamic_pointer_cast<Chebyshev3level<HipCrsMatrix, HipVector>>(smoother_);
        auto jac = std::dynamic_pointer_cast<Jacobi<HipCrsMatrix, HipVector>>(smoother_);
        auto gs = std::dynamic_pointer_cast<GaussSeidel>(smoother_);
        auto pj = std::dynamic_pointer_cast<PointJacobi<HipCrsMatrix, HipVector>>(smoother_);
        if (!cheb && !jac && !gs && !pj)
            throw HipError(UHIP_ERR_UNSUPPORTED,
                           "Multigrid: smoother must be Chebyshev3level, Jacobi, PointJacobi or GaussSeidel");
        if (cycle_type_ != MG_MULTIPLICATIVE_CYCLE && cycle_type_ != MG_FULL_CYCLE)
            throw HipError(UHIP_ERR_UNSUPPORTED, "Multigrid: unknown MG type (cycle_type 1 or 3)");
        eng_.levels.clear();
        eng_.levels.resize((size_t)L);
        eng_.pre = pre_;
        eng_.post = post_;
        eng_.mg_type = mg_type_;
        eng_.masks = masks;
        eng_.smoother = cheb ? MG_CHEBYSHEV : (gs ? MG_GAUSS_SEIDEL : (pj ? MG_POINT_JACOBI : MG_JACOBI));
        eng_.gs_symmetric = gs ? gs->symmetric() : true;
        eng_.line_search = line_search_;
        eng_.fuse_pairs = fused_;
        for (int l = 0; l < L; ++l) {
            auto &lv = eng_.levels[(size_t)l];
            lv.A = A_[(size_t)l];
            if (l > 0) {
                lv.R = R_[(size_t)l - 1];
                lv.P = I_[(size_t)l - 1];
            }
            if (masks) lv.mask = mask_[(size_t)l];
        }
        for (int l = 1; l < L; ++l) {
            auto &lv = eng_.levels[(size_t)l];
            if (cheb) {
                PowerMethod<HipCrsMatrix, HipVector> pm;
                pm.tol(cheb->power_method().tol());
                pm.max_it(cheb->power_method().max_it());
                lv.eig_max = cheb->scale_max_eig() * pm.get_max_eig(*A_[(size_t)l]);
                lv.eig_min = cheb->scale_min_eig() * lv.eig_max;
            } else if (gs) {
                lv.dinv = HipVector(ctx, A_[(size_t)l]->rows(), 0.0);
                hip_check(uhip_pgs_dinv(A_[(size_t)l]->c(), &A_[(size_t)l]->raw(), lv.dinv.data()), "gs d_inv");
                lv.groups = std::make_shared<ColorGroups>();
                lv.groups->build(*A_[(size_t)l]);
            } else if (pj) {
                PointJacobi<HipCrsMatrix, HipVector>::make_point_jacobi(*A_[(size_t)l], lv.LU, lv.dinv);
            } else {
                A_[(size_t)l]->build_diag(lv.dinv);
                InvDiagPreconditioner<HipCrsMatrix, HipVector>::pseudo_inv(lv.dinv, 0.0);
            }
        }
        coarse_->factorize(*A_[0]);
        eng_.coarse = coarse_.get();
        eng_.allocate_work();
        r_top_ = HipVector(ctx, op->rows(), 0);
        c_top_ = HipVector(ctx, op->rows(), 0);
        op->context()->sync();
        setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // Multigrid::apply (utopia_Multigrid.hpp:162-238), x is the initial guess.
    bool apply(const HipVector &rhs, HipVector &x) override {
        const HipCrsMatrix &A = *A_.back();
        if (x.empty() || x.size() != rhs.size()) x.zeros_like(rhs);
        const int l = n_levels_ - 1;
        A.residual(x, rhs, r_top_);
        double r_norm = r_top_.norm2();
        const double r0_norm = r_norm;
        int64_t it = 1;
        bool converged = false;
        this->status_.clear();
        if (cycle_type_ == MG_FULL_CYCLE) this->max_it(1);  // utopia_Multigrid.hpp:191-193
        while (!converged) {
            if (cycle_type_ == MG_FULL_CYCLE) eng_.full_cycle(l, r_top_.data(), c_top_.data(), v_cycle_repetition_);
            else eng_.cycle(l, r_top_.data(), UHIP_F64, c_top_.data(), UHIP_F64);
            x.axpy(1.0, c_top_);
            A.residual(x, rhs, r_top_);
            r_norm = r_top_.norm2();
            const double rel_norm = r_norm / r0_norm;
            this->print_iter(it, r_norm);
            converged = this->check_convergence(it, r_norm, rel_norm, 1);
            it++;
        }
        return true;
    }

    // Preconditioner use (CG calls it with z = 0): with max_it == 1 the cycle
    // result is the answer: r = rhs - A*0 = rhs, x = 0 + c = c, and the closing
    // residual only feeds check_convergence, which stops at it >= max_it anyway.
    bool apply_zero_initial_guess(const HipVector &rhs, HipVector &z) override {
        if (this->max_it_ != 1 || this->verbose_ || cycle_type_ != MG_MULTIPLICATIVE_CYCLE)
            return LinearSolver::apply_zero_initial_guess(rhs, z);
        if (z.empty() || z.size() != rhs.size()) z.zeros_like(rhs);
        eng_.cycle(n_levels_ - 1, rhs.data(), UHIP_F64, z.data(), UHIP_F64);
        return true;
    }
    bool apply_zero_initial_guess_dot(const HipVector &rhs, HipVector &z, double *dot_d) override {
        if (this->max_it_ != 1 || this->verbose_ || cycle_type_ != MG_MULTIPLICATIVE_CYCLE)
            return LinearSolver::apply_zero_initial_guess_dot(rhs, z, dot_d);
        if (z.empty() || z.size() != rhs.size()) z.zeros_like(rhs);
        vcycle_apply_dot(eng_, rhs, z, dot_d);
        return true;
    }

    const std::vector<std::shared_ptr<HipCrsMatrix>> &interpolations() const { return I_; }
    const std::vector<std::shared_ptr<HipCrsMatrix>> &restrictions() const { return R_; }
    const std::vector<std::shared_ptr<const HipCrsMatrix>> &levels() const { return A_; }

private:
    std::shared_ptr<Smoother> smoother_;
    std::shared_ptr<Factorization> coarse_;
    std::vector<std::shared_ptr<HipCrsMatrix>> I_, R_;
    std::vector<std::shared_ptr<const HipCrsMatrix>> A_, given_;
    std::vector<HipVector> mask_;
    int n_levels_ = 0, pre_ = 3, post_ = 3, mg_type_ = 1, cycle_type_ = MG_MULTIPLICATIVE_CYCLE;
    int64_t v_cycle_repetition_ = 10;
    bool masks_ = true, fix_semi_ = false, galerkin_ = true, line_search_ = false;
    int fused_ = 1;
    VCycle<double> eng_;
    HipVector r_top_, c_top_;
    double setup_s_ = 0;
};

// ---------------------------------------------------------------------------
// Mixed precision: the V-cycle of an FP64-assembled Multigrid run on FP32
// copies of every level operator, transfer and mask (int32 indices shared
// layout), with the FP64 coarse inverse. The FP64 residual enters and the FP64
// correction leaves through the first / last kernels of the cycle.
// ---------------------------------------------------------------------------
class MixedPrecisionMultigrid final : public LinearSolver<HipCrsMatrix, HipVector> {
public:
    explicit MixedPrecisionMultigrid(std::shared_ptr<Multigrid> mg) : mg_(std::move(mg)) {}
    const VCycle<float> &engine() const { return eng_; }
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        LinearSolver::update(op);
        mg_->update(op);
        const auto &src = mg_->engine();
        eng_.levels.clear();
        eng_.levels.resize(src.levels.size());
        eng_.pre = src.pre;
        eng_.post = src.post;
        eng_.mg_type = src.mg_type;
        eng_.masks = src.masks;
        eng_.smoother = src.smoother;
        eng_.line_search = src.line_search;
        eng_.fuse_pairs = src.fuse_pairs;
        eng_.coarse = src.coarse;
        if (eng_.smoother == MG_GAUSS_SEIDEL)
            throw HipError(UHIP_ERR_UNSUPPORTED, "mixed precision: Chebyshev, Jacobi or PointJacobi smoothing");
        if (mg_->cycle_type() != MG_MULTIPLICATIVE_CYCLE)
            throw HipError(UHIP_ERR_UNSUPPORTED, "mixed precision: multiplicative cycle only");
        for (size_t l = 0; l < src.levels.size(); ++l) {
            auto &d = eng_.levels[l];
            const auto &s = src.levels[l];
            d.A = std::make_shared<HipCrsMatrix32>(s.A->cast_f32());
            if (l > 0) {
                d.R = std::make_shared<HipCrsMatrix32>(s.R->cast_f32());
                d.P = std::make_shared<HipCrsMatrix32>(s.P->cast_f32());
            }
            if (!s.mask.empty()) {
                d.mask = HipVector32(op->context(), s.mask.size());
                hip_check(uhip_vec_cast_f64_f32(op->c(), s.mask.size(), s.mask.data(), d.mask.data()), "mask cast");
            }
            // Jacobi / PointJacobi: FP32 copies of the FP64 d_inv and LU
            if (!s.dinv.empty()) {
                d.dinv = HipVector32(op->context(), s.dinv.size());
                hip_check(uhip_vec_cast_f64_f32(op->c(), s.dinv.size(), s.dinv.data(), d.dinv.data()), "d_inv cast");
            }
            if (s.LU) d.LU = std::make_shared<HipCrsMatrix32>(s.LU->cast_f32());
            d.eig_max = s.eig_max;
            d.eig_min = s.eig_min;
        }
        eng_.allocate_work();
    }
    bool apply(const HipVector &rhs, HipVector &sol) override {
        if (sol.empty() || sol.size() != rhs.size()) sol.zeros_like(rhs);
        HipVector c(rhs.context(), rhs.size(), 0.0), r(rhs);
        const HipCrsMatrix &A = *this->op_;
        A.residual(sol, rhs, r);
        eng_.cycle((int)eng_.levels.size() - 1, r.data(), UHIP_F64, c.data(), UHIP_F64);
        sol.axpy(1.0, c);
        return true;
    }
    bool apply_zero_initial_guess(const HipVector &rhs, HipVector &z) override {
        if (z.empty() || z.size() != rhs.size()) z.zeros_like(rhs);
        eng_.cycle((int)eng_.levels.size() - 1, rhs.data(), UHIP_F64, z.data(), UHIP_F64);
        return true;
    }
    bool apply_zero_initial_guess_dot(const HipVector &rhs, HipVector &z, double *dot_d) override {
        if (z.empty() || z.size() != rhs.size()) z.zeros_like(rhs);
        vcycle_apply_dot(eng_, rhs, z, dot_d);
        return true;
    }
    float *rhs_f32_buffer(int64_t n) override {
        if (r32_.size() != n || r32_.empty()) r32_ = HipVector32(this->op_->context(), n);
        return r32_.data();
    }
    bool apply_zero_initial_guess_dot_f32(const HipVector &rhs, const float *rhs32, HipVector &z,
                                          double *dot_d) override {
        if (z.empty() || z.size() != rhs.size()) z.zeros_like(rhs);
        vcycle_apply_dot(eng_, rhs, z, dot_d, rhs32);
        return true;
    }
    // The FP32 cycle's result is an FP32 vector: kept in FP32 storage it converts
    // exactly wherever the FP64 CG reads it.
    const float *apply_zero_initial_guess_dot_z32(const HipVector &rhs, const float *rhs32, double *dot_d) override {
        const int64_t n = rhs.size();
        if (z32_.size() != n || z32_.empty()) z32_ = HipVector32(this->op_->context(), n);
        const bool fused = eng_.cycle((int)eng_.levels.size() - 1, rhs32, UHIP_F32, z32_.data(), UHIP_F32, rhs.data(),
                                      dot_d);
        if (!fused) {
            if (z64_.size() != n || z64_.empty()) z64_ = HipVector(this->op_->context(), n);
            hip_check(uhip_vec_cast_f32_f64(rhs.ctx(), n, z32_.data(), z64_.data()), "z cast");
            const double *a[1] = {z64_.data()}, *b[1] = {rhs.data()};
            hip_check(uhip_vec_dots_f64(rhs.ctx(), n, 1, a, b, dot_d), "dot(z, r)");
        }
        return z32_.data();
    }
    const Multigrid &fp64() const { return *mg_; }
    // Chebyshev bounds of level l on both the FP64 hierarchy and its FP32 copy
    void set_level_eigs(int l, double emax, double emin) {
        mg_->set_level_eigs(l, emax, emin);
        eng_.levels[(size_t)l].eig_max = emax;
        eng_.levels[(size_t)l].eig_min = emin;
        eng_.allocate_work();
    }

private:
    std::shared_ptr<Multigrid> mg_;
    VCycle<float> eng_;
    HipVector32 r32_, z32_;
    HipVector z64_;
};

}  // namespace hip
}  // namespace utopia
