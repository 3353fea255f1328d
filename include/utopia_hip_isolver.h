/*
 * utopia_hip_isolver.h - utopia's isolver linear-solver plug-out, exported by libutopia_hip.so
 * with the reference's names and signatures
 * (plug_in_and_out/isolver/utopia_isolver_lsolve.cpp:42-220), so that a code that loads utopia's
 * isolver_lsolve plug-in can load this library in its place.
 *
 * The isolver headers (isolver_lsolve.h, an external project that the reference does not
 * vendor) are not available; the types below restate what utopia_isolver_lsolve.cpp uses of
 * them: isolver_idx_t (PETSc's 32-bit PetscInt), isolver_scalar_t (double), and isolver_lsolve_t
 * with a communicator and the private_data slot the plug-out owns.
 *
 * Differences from the PETSc plug-out, all on purpose:
 *  - one GPU, one process: n_local must equal n_global (else UHIP_ERR_INVALID);
 *  - the solver is chosen by the environment variable UTOPIA_HIP_LINEAR_SOLVER (a registry name
 *    of uhip_solver_create_by_name: "cg_amg" by default, the CG + AlgebraicMultigrid counterpart
 *    of the reference's default GMRES + hypre; "cg_jacobi", "cg", ...) instead of
 *    UTOPIA_LINEAR_SOLVER_CONFIG, a utopia configuration file this library does not parse;
 *    UTOPIA_HIP_DEVICE selects the GPU (default 0);
 *  - arrays may be host memory (copied to the GPU) or device memory (used in place, never
 *    freed, as PetscMatrix::wrap does with the caller's arrays);
 *  - every function returns 0 or a negative UHIP_ERR_* code (the reference returns 0 always and
 *    reports PETSc errors by printing); uhip_last_error() has the message.
 */
#ifndef UTOPIA_HIP_ISOLVER_H
#define UTOPIA_HIP_ISOLVER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int isolver_idx_t;
typedef double isolver_scalar_t;

typedef struct isolver_lsolve {
    void *comm;         /* unused: one process drives one GPU */
    void *private_data; /* owned by the plug-out between init and destroy */
} isolver_lsolve_t;

/* isolver_lsolve_init (utopia_isolver_lsolve.cpp:42-65) */
int isolver_lsolve_init(isolver_lsolve_t *info);
/* isolver_lsolve_update_crs (:67-94) */
int isolver_lsolve_update_crs(const isolver_lsolve_t *info, const ptrdiff_t n_local, const ptrdiff_t n_global,
                              const isolver_idx_t *const rowptr, const isolver_idx_t *const colidx,
                              const isolver_scalar_t *const values);
/* isolver_lsolve_update_with_preconditioner_crs (:96-133) */
int isolver_lsolve_update_with_preconditioner_crs(const isolver_lsolve_t *info, const ptrdiff_t n_local,
                                                  const ptrdiff_t n_global, const isolver_idx_t *const rowptr,
                                                  const isolver_idx_t *const colidx,
                                                  const isolver_scalar_t *const values,
                                                  const isolver_idx_t *const prec_rowptr,
                                                  const isolver_idx_t *const prec_colidx,
                                                  const isolver_scalar_t *const prec_values);
/* isolver_lsolve_apply (:135-157): x is the initial guess and the result */
int isolver_lsolve_apply(const isolver_lsolve_t *info, const isolver_scalar_t *const rhs, isolver_scalar_t *const x);
/* isolver_lsolve_set_* (:159-205) */
int isolver_lsolve_set_max_iterations(const isolver_lsolve_t *info, const int max_it);
int isolver_lsolve_set_atol(const isolver_lsolve_t *info, const isolver_scalar_t atol);
int isolver_lsolve_set_stol(const isolver_lsolve_t *info, const isolver_scalar_t stol);
int isolver_lsolve_set_rtol(const isolver_lsolve_t *info, const isolver_scalar_t rtol);
int isolver_lsolve_set_verbosity(const isolver_lsolve_t *info, const int verbosity_level);
/* isolver_lsolve_destroy (:207-218) */
int isolver_lsolve_destroy(isolver_lsolve_t *info);

#ifdef __cplusplus
}
#endif

#endif /* UTOPIA_HIP_ISOLVER_H */
