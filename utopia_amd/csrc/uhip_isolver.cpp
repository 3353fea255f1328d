// utopia's isolver linear-solver plug-out (plug_in_and_out/isolver/utopia_isolver_lsolve.cpp:42-220)
// over the solver-level C ABI: init / update_crs / update_with_preconditioner_crs / apply / set_* /
// destroy with the reference's names and signatures (include/utopia_hip_isolver.h).
#include <cstdint>
#include <cstdlib>
#include <string>

#include "utopia_hip.h"
#include "utopia_hip_isolver.h"

namespace uhip {
int set_error(int code, const char *fmt, ...);  // uhip_runtime.hip
}

namespace {

struct Plug {
    uhip_ctx ctx = nullptr;
    uhip_solver solver = nullptr;
    int32_t n = -1;
};

Plug *plug_of(const isolver_lsolve_t *info) { return info ? (Plug *)info->private_data : nullptr; }

// one GPU, one process: the local rows are all rows, and they fit the int32 indices
int check_sizes(ptrdiff_t n_local, ptrdiff_t n_global) {
    if (n_local != n_global)
        return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve: n_local %lld != n_global %lld (one GPU, one process)",
                               (long long)n_local, (long long)n_global);
    if (n_local < 0 || n_local > INT32_MAX)
        return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve: n %lld outside the int32 index range",
                               (long long)n_local);
    return UHIP_OK;
}

}  // namespace

extern "C" {

int isolver_lsolve_init(isolver_lsolve_t *info) {
    if (!info) return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve_init: null info");
    const char *dev = std::getenv("UTOPIA_HIP_DEVICE");
    const char *type = std::getenv("UTOPIA_HIP_LINEAR_SOLVER");
    auto *p = new Plug();
    int st = uhip_ctx_create(dev ? std::atoi(dev) : 0, nullptr, &p->ctx);
    if (st == UHIP_OK) st = uhip_solver_create_by_name(p->ctx, "hip", type && *type ? type : "cg_amg", &p->solver);
    if (st != UHIP_OK) {
        if (p->ctx) uhip_ctx_destroy(p->ctx);
        delete p;
        info->private_data = nullptr;
        return st;
    }
    info->private_data = p;
    return UHIP_OK;
}

int isolver_lsolve_update_crs(const isolver_lsolve_t *info, const ptrdiff_t n_local, const ptrdiff_t n_global,
                              const isolver_idx_t *const rowptr, const isolver_idx_t *const colidx,
                              const isolver_scalar_t *const values) {
    Plug *p = plug_of(info);
    if (!p) return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve_update_crs: not initialised");
    if (const int st = check_sizes(n_local, n_global)) return st;
    if (!rowptr || !colidx || !values) return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve_update_crs: null array");
    const int32_t n = (int32_t)n_local;
    int st;
    if (uhip_pointer_is_device(rowptr)) {
        int32_t nnz = 0;
        st = uhip_memcpy_dtoh(p->ctx, &nnz, rowptr + n, sizeof(int32_t));
        if (st == UHIP_OK)
            st = uhip_solver_update_crs_device(p->solver, n, nnz, (int32_t *)rowptr, (int32_t *)colidx,
                                               (double *)values);
    } else {
        st = uhip_solver_update_crs(p->solver, n, rowptr, colidx, values);
    }
    if (st == UHIP_OK) p->n = n;
    return st;
}

int isolver_lsolve_update_with_preconditioner_crs(const isolver_lsolve_t *info, const ptrdiff_t n_local,
                                                  const ptrdiff_t n_global, const isolver_idx_t *const rowptr,
                                                  const isolver_idx_t *const colidx,
                                                  const isolver_scalar_t *const values,
                                                  const isolver_idx_t *const prec_rowptr,
                                                  const isolver_idx_t *const prec_colidx,
                                                  const isolver_scalar_t *const prec_values) {
    Plug *p = plug_of(info);
    if (!p) return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve_update_with_preconditioner_crs: not initialised");
    if (const int st = check_sizes(n_local, n_global)) return st;
    const int32_t n = (int32_t)n_local;
    const int on_device = uhip_pointer_is_device(rowptr);
    if (on_device != uhip_pointer_is_device(prec_rowptr))
        return uhip::set_error(UHIP_ERR_INVALID,
                               "isolver_lsolve_update_with_preconditioner_crs: both matrices on the host or both on "
                               "the device");
    const int st = uhip_solver_update_with_preconditioner_crs(p->solver, n, rowptr, colidx, values, prec_rowptr,
                                                              prec_colidx, prec_values, on_device);
    if (st == UHIP_OK) p->n = n;
    return st;
}

int isolver_lsolve_apply(const isolver_lsolve_t *info, const isolver_scalar_t *const rhs, isolver_scalar_t *const x) {
    Plug *p = plug_of(info);
    if (!p || p->n < 0) return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve_apply: no operator (update first)");
    if (!rhs || !x) return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve_apply: null vector");
    const int dr = uhip_pointer_is_device(rhs), dx = uhip_pointer_is_device(x);
    if (dr != dx)
        return uhip::set_error(UHIP_ERR_INVALID, "isolver_lsolve_apply: rhs and x both on the host or both on the device");
    return dr ? uhip_solver_apply_device(p->solver, rhs, x) : uhip_solver_apply(p->solver, rhs, x);
}

int isolver_lsolve_set_max_iterations(const isolver_lsolve_t *info, const int max_it) {
    Plug *p = plug_of(info);
    return p ? uhip_lsolve_set_max_iterations(p->solver, max_it) : UHIP_ERR_INVALID;
}

int isolver_lsolve_set_atol(const isolver_lsolve_t *info, const isolver_scalar_t atol) {
    Plug *p = plug_of(info);
    return p ? uhip_lsolve_set_atol(p->solver, atol) : UHIP_ERR_INVALID;
}

int isolver_lsolve_set_stol(const isolver_lsolve_t *info, const isolver_scalar_t stol) {
    Plug *p = plug_of(info);
    return p ? uhip_lsolve_set_stol(p->solver, stol) : UHIP_ERR_INVALID;
}

int isolver_lsolve_set_rtol(const isolver_lsolve_t *info, const isolver_scalar_t rtol) {
    Plug *p = plug_of(info);
    return p ? uhip_lsolve_set_rtol(p->solver, rtol) : UHIP_ERR_INVALID;
}

int isolver_lsolve_set_verbosity(const isolver_lsolve_t *info, const int verbosity_level) {
    Plug *p = plug_of(info);
    return p ? uhip_lsolve_set_verbosity(p->solver, verbosity_level) : UHIP_ERR_INVALID;
}

int isolver_lsolve_destroy(isolver_lsolve_t *info) {
    Plug *p = plug_of(info);
    if (!p) return UHIP_OK;
    uhip_solver_destroy(p->solver);
    uhip_ctx_destroy(p->ctx);
    delete p;
    info->private_data = nullptr;
    return UHIP_OK;
}

}  // extern "C"
