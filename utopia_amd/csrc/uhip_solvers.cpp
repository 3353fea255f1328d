// Solver-level C ABI: opaque solver objects in the shape of utopia's
// isolver_lsolve_* plug-out (plug_in_and_out/isolver/utopia_isolver_lsolve.cpp:42-220):
// create/init, update_crs, apply, set_* parameters, destroy. Each object owns one of
// the templated solvers of the host layer over HipVector / HipCrsMatrix.
#include <chrono>
#include <cstring>
#include <map>
#include <string>

#include "host/utopia_hip_amg.hpp"
#include "host/utopia_hip_mmg.hpp"
#include "host/utopia_hip_multigrid.hpp"
#include "host/utopia_hip_qp.hpp"

using namespace utopia;
using namespace utopia::hip;

struct uhip_solver_s {
    int type = 0;
    HipContextPtr ctx;
    std::map<std::string, double> params;
    std::vector<std::shared_ptr<HipCrsMatrix>> transfers;
    std::shared_ptr<HipCrsMatrix> A;
    std::shared_ptr<HipCrsMatrix> prec;  // update_with_preconditioner: the preconditioner's matrix
    std::shared_ptr<LinearSolver<HipCrsMatrix, HipVector>> solver;
    std::shared_ptr<ConjugateGradient<HipCrsMatrix, HipVector>> cg;
    std::shared_ptr<Multigrid> mg;
    std::shared_ptr<MixedPrecisionMultigrid> mixed;
    std::shared_ptr<MPRGP> mprgp;
    std::shared_ptr<ProjectedGaussSeidel> pgs;
    std::shared_ptr<QPSolverBase> qp;  // the QP solver of the MPRGP / PGS / PCG / PG / MMG types
    std::shared_ptr<MonotoneMultigrid> mmg;
    std::shared_ptr<AlgebraicMultigrid> amg;
    HipVector lb, ub;
    bool has_lb = false, has_ub = false;
    SolutionStatus status;
    double seconds = 0, setup_seconds = 0;
    int64_t total_iterations = 0, refinements = 0;
};

namespace uhip {
int set_error(int code, const char *fmt, ...);  // uhip_runtime.hip
}

namespace {

// The solver's own context on the caller's device and stream (the caller's
// uhip_ctx stays untouched and may be destroyed independently).
HipContextPtr borrow(uhip_ctx c) { return std::make_shared<HipContext>(uhip_ctx_device(c), uhip_ctx_stream(c)); }

double param(const uhip_solver_s *s, const char *k, double def) {
    auto it = s->params.find(k);
    return it == s->params.end() ? def : it->second;
}

// uhip_last_error() reports the exception's message (for kernel-layer failures it
// already embeds that layer's message)
int fail(const std::exception &e) {
    const auto *he = dynamic_cast<const HipError *>(&e);
    return uhip::set_error(he ? he->code() : UHIP_ERR_INVALID, "%s", e.what());
}

template <class S>
void apply_iterative_params(uhip_solver_s *s, S &solver_) {
    IterativeSolver<HipCrsMatrix, HipVector> &solver = solver_;
    solver.atol(param(s, "atol", solver.atol()));
    solver.rtol(param(s, "rtol", solver.rtol()));
    solver.stol(param(s, "stol", solver.stol()));
    solver.max_it((int64_t)param(s, "max_it", (double)solver.max_it()));
    solver.verbose(param(s, "verbose", 0) != 0);
}

std::shared_ptr<Multigrid> make_multigrid(uhip_solver_s *s, bool with_transfers = true) {
    std::shared_ptr<IterativeSolver<HipCrsMatrix, HipVector>> sm;
    if ((int)param(s, "smoother", 0) == 1) {
        sm = std::make_shared<Jacobi<HipCrsMatrix, HipVector>>();
    } else if ((int)param(s, "smoother", 0) == 3) {
        sm = std::make_shared<PointJacobi<HipCrsMatrix, HipVector>>();
    } else if ((int)param(s, "smoother", 0) == 2) {
        auto gs = std::make_shared<GaussSeidel>();
        gs->use_symmetric_sweep(param(s, "use_symmetric_sweep", 1) != 0);
        sm = gs;
    } else {
        auto ch = std::make_shared<Chebyshev3level<HipCrsMatrix, HipVector>>();
        ch->scale_max_eig(param(s, "scale_max_eig", 1.2));
        ch->scale_min_eig(param(s, "scale_min_eig", 0.06));
        ch->power_method_tol(param(s, "power_method_tol", 1e-2));
        ch->power_method_max_it((int64_t)param(s, "power_method_max_it", 30));
        sm = ch;
    }
    auto mg = std::make_shared<Multigrid>(sm, std::make_shared<Factorization>());
    mg->pre_smoothing_steps((int)param(s, "pre_smoothing_steps", 3));
    mg->post_smoothing_steps((int)param(s, "post_smoothing_steps", 3));
    mg->mg_type((int)param(s, "mg_type", 1));
    mg->cycle_type((int)param(s, "cycle_type", MG_MULTIPLICATIVE_CYCLE));
    mg->v_cycle_repetition((int64_t)param(s, "v_cycle_repetition", 10));
    mg->use_line_search(param(s, "use_line_search", 0) != 0);
    mg->fused_smoothing((int)param(s, "fused_smoothing", 1));
    mg->must_generate_masks(param(s, "must_generate_masks", 1) != 0);
    mg->fix_semidefinite_operators(param(s, "fix_semidefinite_operators", 0) != 0);
    if (with_transfers) mg->set_transfer_operators(s->transfers);
    return mg;
}

// AlgebraicMultigrid (n_levels, block_size, Agglomerate bmax / weight / component)
// over a Multigrid with the smoother and cycle parameters of the MG types
std::shared_ptr<AlgebraicMultigrid> make_amg(uhip_solver_s *s) {
    auto agg = std::make_shared<Agglomerate>();
    agg->bmax((int64_t)param(s, "bmax", 3));
    agg->weight(param(s, "weight", 1.0 / 3));
    agg->block_size((int)param(s, "block_size", 1));
    agg->component((int)param(s, "component", -1));
    auto amg = std::make_shared<AlgebraicMultigrid>(make_multigrid(s, false), agg);
    amg->n_levels((int)param(s, "n_levels", 2));
    return amg;
}

void build(uhip_solver_s *s) {
    s->solver.reset();
    s->cg.reset();
    s->mg.reset();
    s->mprgp.reset();
    s->pgs.reset();
    s->qp.reset();
    s->mmg.reset();
    s->amg.reset();
    switch (s->type) {
        case UHIP_SOLVER_CG:
        case UHIP_SOLVER_CG_JACOBI:
        case UHIP_SOLVER_CG_MG:
        case UHIP_SOLVER_CG_MG_MIXED: {
            auto cg = std::make_shared<ConjugateGradient<HipCrsMatrix, HipVector>>();
            apply_iterative_params(s, *cg);
            cg->reset_initial_guess(param(s, "reset_initial_guess", 0) != 0);
            if (s->type == UHIP_SOLVER_CG_JACOBI) {
                cg->set_preconditioner(std::make_shared<InvDiagPreconditioner<HipCrsMatrix, HipVector>>());
            } else if (s->type == UHIP_SOLVER_CG_MG || s->type == UHIP_SOLVER_CG_MG_MIXED) {
                auto mg = make_multigrid(s);
                mg->max_it(1);
                s->mg = mg;
                if (s->type == UHIP_SOLVER_CG_MG_MIXED) {
                    s->mixed = std::make_shared<MixedPrecisionMultigrid>(mg);
                    cg->set_preconditioner(s->mixed);
                    cg->refine(param(s, "refine", 1) != 0);
                } else {
                    cg->set_preconditioner(mg);
                    cg->refine(param(s, "refine", 0) != 0);
                }
            }
            s->cg = cg;
            s->solver = cg;
            break;
        }
        case UHIP_SOLVER_MG: {
            auto mg = make_multigrid(s);
            apply_iterative_params(s, *mg);
            s->mg = mg;
            s->solver = mg;
            break;
        }
        case UHIP_SOLVER_AMG: {
            auto amg = make_amg(s);
            apply_iterative_params(s, *amg);
            apply_iterative_params(s, *amg->algorithm());  // read(in) reaches the Multigrid too
            s->amg = amg;
            s->mg = amg->algorithm();
            s->solver = amg;
            break;
        }
        case UHIP_SOLVER_CG_AMG: {
            auto cg = std::make_shared<ConjugateGradient<HipCrsMatrix, HipVector>>();
            apply_iterative_params(s, *cg);
            cg->reset_initial_guess(param(s, "reset_initial_guess", 0) != 0);
            auto amg = make_amg(s);
            amg->max_it(1);
            cg->set_preconditioner(amg);
            cg->refine(param(s, "refine", 0) != 0);
            s->amg = amg;
            s->mg = amg->algorithm();
            s->cg = cg;
            s->solver = cg;
            break;
        }
        case UHIP_SOLVER_MPRGP: {
            auto q = std::make_shared<MPRGP>();
            apply_iterative_params(s, *q);
            q->set_eig_comp_tol(param(s, "eig_comp_tol", 1e-1));
            q->power_method_max_it((int64_t)param(s, "power_method_max_it", 10));
            s->mprgp = q;
            s->qp = q;
            s->solver = q;
            break;
        }
        case UHIP_SOLVER_PGS: {
            auto q = std::make_shared<ProjectedGaussSeidel>();
            apply_iterative_params(s, *q);
            q->use_symmetric_sweep(param(s, "use_symmetric_sweep", 1) != 0);
            q->check_s_norm_each((int64_t)param(s, "check_s_norm_each", 10));
            s->pgs = q;
            s->qp = q;
            s->solver = q;
            break;
        }
        case UHIP_SOLVER_MMG: {
            auto q = std::make_shared<MonotoneMultigrid>();
            apply_iterative_params(s, *q);
            q->pre_smoothing_steps((int)param(s, "pre_smoothing_steps", 3));
            q->post_smoothing_steps((int)param(s, "post_smoothing_steps", 3));
            q->coarse_space_it((int64_t)param(s, "coarse_space_it", 1));
            q->active_set_tol(param(s, "active_set_tol", 0.0));
            if (s->transfers.empty()) throw HipError(UHIP_ERR_INVALID, "MonotoneMultigrid: set_transfer_operators first");
            q->set_transfer_operators(s->transfers);
            s->mmg = q;
            s->qp = q;
            s->solver = q;
            break;
        }
        case UHIP_SOLVER_PCG: {
            auto q = std::make_shared<ProjectedConjugateGradient>();
            apply_iterative_params(s, *q);
            s->qp = q;
            s->solver = q;
            break;
        }
        case UHIP_SOLVER_PG: {
            auto q = std::make_shared<ProjectedGradient>();
            apply_iterative_params(s, *q);
            s->qp = q;
            s->solver = q;
            break;
        }
        default:
            throw HipError(UHIP_ERR_INVALID, "unknown solver type");
    }
}

void set_bounds(uhip_solver_s *s) {
    BoxConstraints box;
    if (s->has_lb) box.lower = std::make_shared<HipVector>(s->lb);
    if (s->has_ub) box.upper = std::make_shared<HipVector>(s->ub);
    if (s->qp) s->qp->set_box_constraints(box);
}

int do_update(uhip_solver_s *s) {
    const auto t0 = std::chrono::steady_clock::now();
    build(s);
    set_bounds(s);
    if (s->prec) {
        // PreconditionedSolver::update(op, prec): only the CG types carry a preconditioner
        if (!s->cg) throw HipError(UHIP_ERR_UNSUPPORTED, "update_with_preconditioner: not a preconditioned solver");
        s->cg->update(s->A, s->prec);
    } else {
        s->solver->update(s->A);
    }
    s->ctx->sync();
    s->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return UHIP_OK;
}

}  // namespace

extern "C" {

int uhip_solver_create(uhip_ctx ctx, int type, uhip_solver *out) {
    if (!ctx || !out) return UHIP_ERR_INVALID;
    if (type < UHIP_SOLVER_CG || type > UHIP_SOLVER_CG_AMG) return UHIP_ERR_INVALID;
    try {
        auto *s = new uhip_solver_s();
        s->type = type;
        s->ctx = borrow(ctx);
        *out = s;
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_create_by_name(uhip_ctx ctx, const char *backend, const char *type, uhip_solver *out) {
    if (!ctx || !out || !backend || !type) return UHIP_ERR_INVALID;
    static const std::map<std::string, int> registry = {
        {"cg", UHIP_SOLVER_CG},          {"utopia_cg", UHIP_SOLVER_CG}, {"cg_jacobi", UHIP_SOLVER_CG_JACOBI},
        {"cg_mg", UHIP_SOLVER_CG_MG},    {"cg_mg_mixed", UHIP_SOLVER_CG_MG_MIXED},
        {"mg", UHIP_SOLVER_MG},          {"mprgp", UHIP_SOLVER_MPRGP},  {"pgs", UHIP_SOLVER_PGS},
        {"pcg", UHIP_SOLVER_PCG},        {"pg", UHIP_SOLVER_PG},        {"mmg", UHIP_SOLVER_MMG},
        {"amg", UHIP_SOLVER_AMG},        {"cg_amg", UHIP_SOLVER_CG_AMG}};
    const std::string b(backend), t(type);
    if (b != "any" && b != "hip") return uhip::set_error(UHIP_ERR_INVALID, "[Error] backend %s not found", backend);
    auto it = registry.find(t);
    if (it == registry.end()) return uhip::set_error(UHIP_ERR_INVALID, "[Error] solver %s not found", type);
    return uhip_solver_create(ctx, it->second, out);
}

int uhip_solver_destroy(uhip_solver s) {
    delete s;
    return UHIP_OK;
}

int uhip_solver_set_param(uhip_solver s, const char *key, double value) {
    if (!s || !key) return UHIP_ERR_INVALID;
    s->params[key] = value;
    // SpMV plans of matrices built from now on are chosen by timing (reps > 0)
    if (std::strcmp(key, "autotune") == 0) return uhip_ctx_set_autotune(s->ctx->get(), (int)value);
    // tolerances / iteration limits also act on an already built solver
    if (s->cg) apply_iterative_params(s, *s->cg);
    else if (s->mg && s->type == UHIP_SOLVER_MG) apply_iterative_params(s, *s->mg);
    else if (s->amg) {
        apply_iterative_params(s, *s->amg);
        apply_iterative_params(s, *s->amg->algorithm());
    }
    else if (s->qp) apply_iterative_params(s, *s->qp);
    return UHIP_OK;
}

int uhip_solver_set_transfer_operators(uhip_solver s, int n, const int32_t *nrows, const int32_t *ncols,
                                       const int32_t *const *rowptr, const int32_t *const *col,
                                       const double *const *val) {
    if (!s || n < 1 || !nrows || !ncols || !rowptr || !col || !val) return UHIP_ERR_INVALID;
    try {
        s->transfers.clear();
        for (int i = 0; i < n; ++i) {
            auto I = std::make_shared<HipCrsMatrix>();
            I->from_host(s->ctx, nrows[i], ncols[i], rowptr[i], col[i], val[i], (int32_t)param(s, "block_nnz", 0));
            s->transfers.push_back(I);
        }
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_update_crs(uhip_solver s, int32_t n, const int32_t *rowptr, const int32_t *col, const double *val) {
    if (!s || n < 0 || !rowptr || !col || !val) return UHIP_ERR_INVALID;
    try {
        s->A = std::make_shared<HipCrsMatrix>();
        s->A->from_host(s->ctx, n, n, rowptr, col, val, (int32_t)param(s, "block_nnz", 0));
        s->prec.reset();
        return do_update(s);
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_update_crs_device(uhip_solver s, int32_t n, int64_t nnz, int32_t *rowptr, int32_t *col, double *val) {
    if (!s || n < 0 || !rowptr || !col || !val) return UHIP_ERR_INVALID;
    try {
        s->A = std::make_shared<HipCrsMatrix>();
        s->A->wrap(s->ctx, n, n, nnz, rowptr, col, val, (int32_t)param(s, "block_nnz", 0));
        s->prec.reset();
        return do_update(s);
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_update_with_preconditioner_crs(uhip_solver s, int32_t n, const int32_t *rowptr, const int32_t *col,
                                               const double *val, const int32_t *prec_rowptr,
                                               const int32_t *prec_col, const double *prec_val, int on_device) {
    if (!s || n < 0 || !rowptr || !col || !val || !prec_rowptr || !prec_col || !prec_val) return UHIP_ERR_INVALID;
    try {
        const int32_t bn = (int32_t)param(s, "block_nnz", 0);
        auto A = std::make_shared<HipCrsMatrix>(), P = std::make_shared<HipCrsMatrix>();
        if (on_device) {
            int32_t nz[2] = {0, 0};
            hip_check(uhip_memcpy_dtoh(s->ctx->get(), &nz[0], rowptr + n, sizeof(int32_t)), "nnz");
            hip_check(uhip_memcpy_dtoh(s->ctx->get(), &nz[1], prec_rowptr + n, sizeof(int32_t)), "prec nnz");
            A->wrap(s->ctx, n, n, nz[0], (int32_t *)rowptr, (int32_t *)col, (double *)val, bn);
            P->wrap(s->ctx, n, n, nz[1], (int32_t *)prec_rowptr, (int32_t *)prec_col, (double *)prec_val, bn);
        } else {
            A->from_host(s->ctx, n, n, rowptr, col, val, bn);
            P->from_host(s->ctx, n, n, prec_rowptr, prec_col, prec_val, bn);
        }
        s->A = A;
        s->prec = P;
        return do_update(s);
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_set_box_constraints(uhip_solver s, const double *lb, const double *ub) {
    if (!s || !s->A) return UHIP_ERR_INVALID;
    try {
        const int64_t n = s->A->rows();
        s->has_lb = lb != nullptr;
        s->has_ub = ub != nullptr;
        if (lb) {
            s->lb = HipVector(s->ctx, n);
            s->lb.upload(lb);
        }
        if (ub) {
            s->ub = HipVector(s->ctx, n);
            s->ub.upload(ub);
        }
        set_bounds(s);
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

static int run_apply(uhip_solver s, const HipVector &b, HipVector &x) {
    const auto t0 = std::chrono::steady_clock::now();
    bool ok = s->solver->apply(b, x);
    s->ctx->sync();
    s->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    auto *it = dynamic_cast<IterativeSolver<HipCrsMatrix, HipVector> *>(s->solver.get());
    if (it) s->status = it->solution_status();
    s->total_iterations = s->cg ? s->cg->total_iterations() : s->status.iterates;
    s->refinements = s->cg ? s->cg->refinements() : 0;
    return ok ? 1 : 0;
}

int uhip_solver_apply(uhip_solver s, const double *rhs, double *x) {
    if (!s || !s->solver || !rhs || !x) return UHIP_ERR_INVALID;
    try {
        const int64_t n = s->A->rows();
        HipVector b(s->ctx, n), xv(s->ctx, n);
        b.upload(rhs);
        xv.upload(x);
        const int conv = run_apply(s, b, xv);
        xv.download(x);
        return conv ? UHIP_OK : UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_apply_device(uhip_solver s, const double *rhs_d, double *x_d) {
    if (!s || !s->solver || !rhs_d || !x_d) return UHIP_ERR_INVALID;
    try {
        const int64_t n = s->A->rows();
        HipVector b, xv;
        b.wrap(s->ctx, const_cast<double *>(rhs_d), n);
        xv.wrap(s->ctx, x_d, n);
        run_apply(s, b, xv);
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

// z = M r with the preconditioner of a CG solver (z starts at zero, as in the CG loop).
int uhip_solver_precondition(uhip_solver s, const double *rhs, double *z) {
    if (!s || !s->cg || !s->cg->get_preconditioner() || !rhs || !z) return UHIP_ERR_INVALID;
    try {
        const int64_t n = s->A->rows();
        HipVector b(s->ctx, n), zv(s->ctx, n, 0.0);
        b.upload(rhs);
        s->cg->get_preconditioner()->apply_zero_initial_guess(b, zv);
        zv.download(z);
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_get_status(uhip_solver s, int64_t *iterates, double *gradient_norm, int *reason, double *seconds) {
    if (!s) return UHIP_ERR_INVALID;
    if (iterates) *iterates = s->status.iterates;
    if (gradient_norm) *gradient_norm = s->status.gradient_norm;
    if (reason) *reason = s->status.reason;
    if (seconds) *seconds = s->seconds;
    return UHIP_OK;
}

static int get_info(uhip_solver s, const std::string &k, double *value);

int uhip_solver_get_info(uhip_solver s, const char *key, double *value) {
    if (!s || !key || !value) return UHIP_ERR_INVALID;
    try {
        return get_info(s, std::string(key), value);
    } catch (const std::exception &e) {  // malformed level index, transfer index out of range
        return fail(e);
    }
}

static int get_info(uhip_solver s, const std::string &k, double *value) {
    auto level = [&](size_t at) {
        const int l = std::stoi(k.substr(at));
        if (!s->mg || l < 0 || l >= s->mg->n_levels())
            throw HipError(UHIP_ERR_INVALID, "get_info " + k + ": level out of range");
        return l;
    };
    if (k == "setup_seconds") *value = s->setup_seconds;
    else if (k == "solve_seconds") *value = s->seconds;
    else if (k == "total_iterations") *value = (double)s->total_iterations;
    else if (k == "refinements") *value = (double)s->refinements;
    else if (k == "n_levels") *value = s->mg ? s->mg->n_levels() : 0;
    else if (k == "coarse_n") *value = s->mg ? s->mg->level_operator(0).rows() : 0;
    else if (k == "ncolors" && s->pgs) *value = s->pgs->ncolors();
    else if (k == "alpha_bar" && s->mprgp) *value = s->mprgp->alpha_bar();
    else if (k == "active_set_changes" && s->mmg) *value = (double)s->mmg->active_set_changes();
    else if (k.rfind("transfer_nnz_", 0) == 0 && s->mg) *value = (double)s->mg->interpolations().at((size_t)std::stoi(k.substr(13)))->nnz();
    else if (k.rfind("transfer_rows_", 0) == 0 && s->mg) *value = (double)s->mg->interpolations().at((size_t)std::stoi(k.substr(14)))->rows();
    else if (k.rfind("transfer_cols_", 0) == 0 && s->mg) *value = (double)s->mg->interpolations().at((size_t)std::stoi(k.substr(14)))->cols();
    else if (k.rfind("eig_max_", 0) == 0 && s->mg) *value = s->mg->level_eig_max(level(8));
    else if (k.rfind("eig_min_", 0) == 0 && s->mg) *value = s->mg->level_eig_min(level(8));
    else if (k.rfind("level_nnz_", 0) == 0 && s->mg) *value = (double)s->mg->level_operator(level(10)).nnz();
    else if (k.rfind("level_rows_", 0) == 0 && s->mg) *value = (double)s->mg->level_operator(level(11)).rows();
    else if ((k.rfind("format_", 0) == 0 || k.rfind("format32_", 0) == 0) && s->mg && k.size() >= 10) {
        // format_{A,R,P}_<l> of the FP64 V-cycle, format32_{A,R,P}_<l> of the FP32 one (mixed)
        const bool f32 = k.rfind("format32_", 0) == 0;
        const size_t at = f32 ? 9 : 7;
        const char which = k[at];
        const int l = std::stoi(k.substr(at + 2));
        auto pick = [&](const auto &lv) -> const uhip_csr * {
            if (which == 'A' && lv.A) return &lv.A->raw();
            if (which == 'R' && lv.R) return &lv.R->raw();
            if (which == 'P' && lv.P) return &lv.P->raw();
            return nullptr;
        };
        const uhip_csr *m = nullptr;
        if (f32) {
            if (!s->mixed || l < 0 || l >= (int)s->mixed->engine().levels.size()) return UHIP_ERR_INVALID;
            m = pick(s->mixed->engine().levels[(size_t)l]);
        } else {
            if (l < 0 || l >= (int)s->mg->engine().levels.size()) return UHIP_ERR_INVALID;
            m = pick(s->mg->engine().levels[(size_t)l]);
        }
        if (!m) return UHIP_ERR_INVALID;
        *value = (double)m->format;
    } else return UHIP_ERR_INVALID;
    return UHIP_OK;
}

// Copy level l of the Galerkin hierarchy (l = 0 coarsest) to host arrays;
// rowptr must hold rows+1, col/val nnz entries (query with level_rows_/level_nnz_).
int uhip_solver_get_level(uhip_solver s, int l, int32_t *rowptr, int32_t *col, double *val) {
    if (!s || !s->mg || l < 0 || l >= s->mg->n_levels()) return UHIP_ERR_INVALID;
    try {
        const auto &m = s->mg->level_operator(l).raw();
        uhip_ctx c = s->ctx->get();
        hip_check(uhip_memcpy_dtoh(c, rowptr, m.rowptr, sizeof(int32_t) * ((size_t)m.nrows + 1)), "rowptr");
        hip_check(uhip_memcpy_dtoh(c, col, m.col, sizeof(int32_t) * (size_t)m.nnz), "col");
        hip_check(uhip_memcpy_dtoh(c, val, m.val, sizeof(double) * (size_t)m.nnz), "val");
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_get_transfer(uhip_solver s, int l, int32_t *rowptr, int32_t *col, double *val) {
    if (!s || !s->mg || l < 0 || l >= (int)s->mg->interpolations().size()) return UHIP_ERR_INVALID;
    try {
        const auto &m = s->mg->interpolations()[(size_t)l]->raw();
        uhip_ctx c = s->ctx->get();
        hip_check(uhip_memcpy_dtoh(c, rowptr, m.rowptr, sizeof(int32_t) * ((size_t)m.nrows + 1)), "rowptr");
        hip_check(uhip_memcpy_dtoh(c, col, m.col, sizeof(int32_t) * (size_t)m.nnz), "col");
        hip_check(uhip_memcpy_dtoh(c, val, m.val, sizeof(double) * (size_t)m.nnz), "val");
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_get_mask(uhip_solver s, int l, double *mask) {
    if (!s || !s->mg || l < 0 || l >= s->mg->n_levels() || !mask) return UHIP_ERR_INVALID;
    try {
        const HipVector &m = s->mg->mask(l);
        if (m.empty()) return UHIP_ERR_INVALID;
        m.download(mask);
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

// Override the Chebyshev bounds of level l after update (tests pinning the cycle).
int uhip_solver_get_colors(uhip_solver s, int l, int32_t *color, int32_t *ncolors) {
    if (!s || !s->mg || !color || !ncolors || l < 1 || l >= s->mg->n_levels())
        return uhip::set_error(UHIP_ERR_INVALID, "uhip_solver_get_colors: no smoothed level %d (levels 1..n_levels-1)", l);
    try {
        const auto &lv = s->mg->engine().levels[(size_t)l];
        if (!lv.groups)
            return uhip::set_error(UHIP_ERR_INVALID, "uhip_solver_get_colors: level %d has no colouring (GaussSeidel smoother only)", l);
        *ncolors = lv.groups->ncolors();
        hip_check(uhip_memcpy_dtoh(s->ctx->get(), color, lv.groups->color(), sizeof(int32_t) * (size_t)lv.A->rows()),
                  "colors");
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_solver_set_level_eigs(uhip_solver s, int l, double eig_max, double eig_min) {
    if (!s || !s->mg || l < 1 || l >= s->mg->n_levels()) return UHIP_ERR_INVALID;
    try {
        if (s->mixed) s->mixed->set_level_eigs(l, eig_max, eig_min);
        else s->mg->set_level_eigs(l, eig_max, eig_min);
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

// isolver_lsolve_* setters
int uhip_lsolve_set_max_iterations(uhip_solver s, int max_it) { return uhip_solver_set_param(s, "max_it", max_it); }
int uhip_lsolve_set_atol(uhip_solver s, double v) { return uhip_solver_set_param(s, "atol", v); }
int uhip_lsolve_set_rtol(uhip_solver s, double v) { return uhip_solver_set_param(s, "rtol", v); }
int uhip_lsolve_set_stol(uhip_solver s, double v) { return uhip_solver_set_param(s, "stol", v); }
int uhip_lsolve_set_verbosity(uhip_solver s, int level) { return uhip_solver_set_param(s, "verbose", level > 0); }

}  // extern "C"
