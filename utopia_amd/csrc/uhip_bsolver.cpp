// Solver-level C ABI of the batched MG-preconditioned CG: one opaque object per
// batch, in the create / setup / solve / destroy shape of the isolver_lsolve_*
// plug-out (plug_in_and_out/isolver/utopia_isolver_lsolve.cpp:42-220). The
// batched solver itself has no reference analogue.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "host/utopia_hip_batched.hpp"

using namespace utopia;
using namespace utopia::hip;

struct uhip_bsolver_s {
    HipContextPtr ctx;
    std::unique_ptr<BatchedMultigridCG> solver;
    int32_t nseg = 0;
};

namespace {

int fail(const std::exception &e) {
    const auto *he = dynamic_cast<const HipError *>(&e);
    return he ? he->code() : UHIP_ERR_INVALID;
}

std::shared_ptr<HipCrsMatrix> wrap_device(const HipContextPtr &ctx, const uhip_csr *m) {
    if (!m || m->dtype != UHIP_F64) throw HipError(UHIP_ERR_INVALID, "batched: FP64 CSR expected");
    auto w = std::make_shared<HipCrsMatrix>();
    w->wrap(ctx, m->nrows, m->ncols, m->nnz, m->rowptr, m->col, (double *)m->val);
    return w;
}

}  // namespace

extern "C" {

int uhip_bsolver_create(uhip_ctx ctx, int k, int mixed, uhip_bsolver *out) {
    if (!ctx || !out) return UHIP_ERR_INVALID;
    if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return UHIP_ERR_INVALID;
    try {
        auto *s = new uhip_bsolver_s();
        s->ctx = std::make_shared<HipContext>(uhip_ctx_device(ctx), uhip_ctx_stream(ctx));
        s->solver = std::make_unique<BatchedMultigridCG>(s->ctx, k, mixed != 0);
        *out = s;
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_bsolver_destroy(uhip_bsolver s) {
    delete s;
    return UHIP_OK;
}

int uhip_bsolver_set_param(uhip_bsolver s, const char *key, double value) {
    if (!s || !key) return UHIP_ERR_INVALID;
    try {
        s->solver->set_param(key, value);
    } catch (const std::exception &e) {
        return fail(e);
    }
    return UHIP_OK;
}

int uhip_bsolver_setup(uhip_bsolver s, int32_t nseg, const int32_t *seg, const uhip_csr *A, int nlev,
                       const uhip_csr *const *interps, const int32_t *const *coarse_seg) {
    if (!s || nseg < 1 || !seg || !A || nlev < 1 || (nlev > 1 && (!interps || !coarse_seg))) return UHIP_ERR_INVALID;
    try {
        std::vector<int32_t> top(seg, seg + nseg + 1);
        if (top.back() != A->nrows || A->nrows != A->ncols) throw HipError(UHIP_ERR_INVALID, "batched: seg/A size");
        std::vector<std::shared_ptr<HipCrsMatrix>> P;
        std::vector<std::vector<int32_t>> cs;
        int32_t fine_rows = A->nrows;
        for (int l = nlev - 2; l >= 0; --l) {  // check the chain from the finest down
            const uhip_csr *I = interps[l];
            if (!I || I->nrows != fine_rows) throw HipError(UHIP_ERR_INVALID, "batched: interpolation rows");
            if (coarse_seg[l][nseg] != I->ncols) throw HipError(UHIP_ERR_INVALID, "batched: coarse seg size");
            fine_rows = I->ncols;
        }
        for (int l = 0; l < nlev - 1; ++l) {
            P.push_back(wrap_device(s->ctx, interps[l]));
            cs.emplace_back(coarse_seg[l], coarse_seg[l] + nseg + 1);
        }
        s->solver->setup(top, wrap_device(s->ctx, A), P, cs);
        s->nseg = nseg;
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_bsolver_solve(uhip_bsolver s, const double *B_d, double *X_d, int64_t *iterations) {
    if (!s || !B_d || !X_d || s->nseg == 0) return UHIP_ERR_INVALID;
    try {
        const int64_t it = s->solver->solve(B_d, X_d);
        if (iterations) *iterations = it;
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_bsolver_get_conv(uhip_bsolver s, int32_t *conv_it, int32_t *flags) {
    if (!s) return UHIP_ERR_INVALID;
    try {
        if (conv_it) {
            auto v = s->solver->conv_iterations();
            std::memcpy(conv_it, v.data(), sizeof(int32_t) * v.size());
        }
        if (flags) {
            auto v = s->solver->flags();
            std::memcpy(flags, v.data(), sizeof(int32_t) * v.size());
        }
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_bsolver_pattern_grad(uhip_bsolver s, const double *lambda_d, const double *x_d, double *grad_d) {
    if (!s || !lambda_d || !x_d || !grad_d || s->nseg == 0) return UHIP_ERR_INVALID;
    try {
        s->solver->gradient(lambda_d, x_d, grad_d);
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

int uhip_bsolver_get_info(uhip_bsolver s, const char *key, double *value) {
    if (!s || !key || !value) return UHIP_ERR_INVALID;
    const std::string k(key);
    const auto &b = *s->solver;
    if (k == "setup_seconds") *value = b.setup_seconds();
    else if (k == "solve_seconds") *value = b.solve_seconds();
    else if (k == "iterations") *value = (double)b.iterations();
    else if (k == "graph_replays") *value = (double)b.graph_replays();
    else if (k == "graph_captures") *value = (double)b.graph_captures();
    else if (k == "graph_iterations") *value = std::max(1.0, std::floor(b.param("graph_iterations", 16)));
    else if (k == "n_levels") *value = (double)b.levels();
    else if (k.rfind("eig_max_", 0) == 0) {
        int l = -1, sys = -1;
        if (std::sscanf(key, "eig_max_%d_%d", &l, &sys) != 2 || l < 1 || l >= b.levels() || sys < 0 ||
            sys >= s->nseg)
            return UHIP_ERR_INVALID;
        *value = b.eig_max(l)[(size_t)sys];
    } else if (k.rfind("level_nnz_", 0) == 0) {
        int l = -1;
        if (std::sscanf(key, "level_nnz_%d", &l) != 1 || l < 0 || l >= b.levels()) return UHIP_ERR_INVALID;
        *value = (double)b.level_operator(l).nnz();
    } else {
        return UHIP_ERR_INVALID;
    }
    return UHIP_OK;
}

int uhip_bsolver_set_level_eig(uhip_bsolver s, int l, int32_t sys, double eig_max) {
    if (!s || l < 1 || l >= s->solver->levels() || sys < 0 || sys >= s->nseg) return UHIP_ERR_INVALID;
    try {
        auto v = s->solver->eig_max(l);
        v[(size_t)sys] = eig_max;
        s->solver->set_eig_max(l, v);
        return UHIP_OK;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

}  // extern "C"
