/*
 * utopia_hip.h - C ABI of the MI355X (gfx950) sparse-solver backend.
 *
 * Two layers, both plain C (no C++ or torch types in any signature):
 *
 *  (i)  kernel level  - the operations utopia's PETSc backend hands to PETSc
 *       (MatMult, MatMultTranspose, MatPtAP, VecDot, VecNorm, VecAXPY, ...),
 *       here over device pointers and a row-block CSR descriptor;
 *  (ii) solver level  - opaque solver objects mirroring utopia's
 *       isolver_lsolve_* plug-out (init / update_crs / apply / set_* /
 *       destroy) for the templated CG, Multigrid, MPRGP and
 *       ProjectedGaussSeidel solvers.
 *
 * Every function returns an int status: 0 (UHIP_OK) or a negative error code;
 * uhip_last_error() returns the message of the last failure on the calling
 * thread. Reference citations are to zulianp/utopia (paths relative to its
 * repository root).
 */
#ifndef UTOPIA_HIP_H
#define UTOPIA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UHIP_OK 0
#define UHIP_ERR_INVALID (-1)
#define UHIP_ERR_HIP (-2)
#define UHIP_ERR_ALLOC (-3)
#define UHIP_ERR_NUMERIC (-4)
#define UHIP_ERR_UNSUPPORTED (-5)

#define UHIP_F64 0
#define UHIP_F32 1

#define UHIP_NUM_SCALARS 64

/* ------------------------------------------------------------------------
 * Runtime: context (device + stream + reduction workspace), memory, errors.
 * Replaces PETSc's object lifetime + error handling
 * (backend/petsc/utopia_petsc_Error.hpp:12-17: PETSc errors only printed;
 * here every call reports a status).
 * ---------------------------------------------------------------------- */
typedef struct uhip_context_s *uhip_ctx;

const char *uhip_last_error(void);
/* sizeof of the ABI structs (0: uhip_csr, 1: uhip_cheb_step) for binding checks */
int64_t uhip_abi_sizeof(int which);
int uhip_device_count(int *count);
/* stream == NULL: the context creates its own (default-priority) stream. */
int uhip_ctx_create(int device, void *stream, uhip_ctx *out);
int uhip_ctx_destroy(uhip_ctx ctx);
void *uhip_ctx_stream(uhip_ctx ctx);
int uhip_ctx_device(uhip_ctx ctx);
/* UHIP_NUM_SCALARS device doubles owned by the context (scratch for results). */
double *uhip_ctx_scalars(uhip_ctx ctx);
int uhip_sync(uhip_ctx ctx);
int uhip_malloc(uhip_ctx ctx, void **ptr, size_t bytes);
int uhip_free(uhip_ctx ctx, void *ptr);
int uhip_memcpy_htod(uhip_ctx ctx, void *dst, const void *src, size_t bytes);
/* 1 if p is device (or managed) memory, e.g. from hipMalloc or torch, 0 for host memory */
int uhip_pointer_is_device(const void *p);
int uhip_memcpy_dtoh(uhip_ctx ctx, void *dst, const void *src, size_t bytes);
int uhip_memcpy_dtod(uhip_ctx ctx, void *dst, const void *src, size_t bytes);
int uhip_read_scalars(uhip_ctx ctx, const double *src_d, int n, double *dst_h);
/* Asynchronous read-back of n device scalars into one of two slots (0, 1): fetch enqueues the copy
 * on the context stream (the host goes on enqueuing work), wait blocks until that slot's copy is
 * done and returns the values. One fetch in flight per slot: a second fetch into a slot replaces
 * the first. No reference analogue (it lets the CG enqueue its preconditioner before it reads
 * alpha and |r|, and MPRGP its next product before it reads the state of the new iterate). */
int uhip_fetch_scalars(uhip_ctx ctx, int slot, const double *src_d, int n);
int uhip_fetch_wait(uhip_ctx ctx, int slot, int n, double *dst_h);
int uhip_write_scalars(uhip_ctx ctx, double *dst_d, int n, const double *src_h);
int uhip_event_timer_start(uhip_ctx ctx, void **ev_pair);
int uhip_event_timer_stop(uhip_ctx ctx, void *ev_pair, float *ms);
/* Pre-size the reduction workspace (per_slot partial sums per dot) so that a
 * later captured sequence allocates nothing. */
int uhip_ctx_reserve(uhip_ctx ctx, int64_t per_slot);
/* Generation of the reduction workspace: changes whenever it is reallocated (a reservation
 * or an operation needing more partial sums). A graph captured at one generation holds the
 * workspace address of that generation and must be captured again after a change. */
int64_t uhip_ctx_workspace_generation(uhip_ctx ctx);
/* gate_d (device int32, NULL = none): while set, the multi-RHS / batched kernels launched on
 * this context (uhip_spmm, uhip_spmm_cheb_step, uhip_axpy_seg, uhip_segdot, uhip_bcg_step)
 * return without work whenever *gate_d reads 0 when they start. The batched CG points it at
 * its active-pair count, so a captured run of iterations past convergence costs only its
 * launches. No reference analogue. */
int uhip_ctx_set_gate(uhip_ctx ctx, const int32_t *gate_d);
/* hipGraph capture of the work issued on the context stream between begin and
 * end; launch replays it on the same stream. No reference analogue (the
 * reference drives every iteration from the host). */
int uhip_graph_begin(uhip_ctx ctx);
int uhip_graph_end(uhip_ctx ctx, void **exec);
int uhip_graph_launch(uhip_ctx ctx, void *exec);
int uhip_graph_destroy(void *exec);

/* ------------------------------------------------------------------------
 * BLAS1 on device vectors (n entries). Suffix _f64 / _f32.
 * Replaces PetscVector methods (backend/petsc/utopia_petsc_Vector.hpp:39-930):
 *   set      VecSet            (.hpp:260)
 *   copy     VecCopy           (.cpp:636-660)
 *   axpy     VecAXPY  y+=a*x   (.cpp:671)
 *   axpby    VecAXPBY y=a*x+b*y(.hpp:675-680)
 *   waxpby   VecWAXPY-like w=a*x+b*y
 *   scale    VecScale          (.cpp:624)
 *   shift    VecShift          (.cpp:18)
 *   abs      VecAbs (transform(Abs))
 *   emul     VecPointwiseMult  (.cpp:588)
 *   ediv     VecPointwiseDivide
 *   pseudo_inv  ElementWisePseudoInverse (utopia_core/utopia_ElementWisePseudoInverse.hpp:20-56)
 *   emin/emax   VecPointwiseMin/Max z = a < b ? a : b / a < b ? b : a  (e_min/e_max, .cpp:597-603)
 *   reciprocal  VecReciprocal (1/x where x != 0, zeros kept), then *num unless num == 1 (.hpp:711-720)
 *   replace     x_i = to where x_i == from (PointJacobi::check_indef, solvers/smoothers/utopia_PointJacobi.hpp:116-127)
 *   dot      VecDot            (.cpp:59)     result in device memory
 *   norm2    VecNorm NORM_2    (.hpp:301-305) result (already sqrt) in device memory
 *   dots     EvalDots split-phase VecDotBegin/End (backend/petsc/eval/utopia_petsc_EvalDotVecVecs.cpp:57-92),
 *            k <= 4 dots xs[j].ys[j] in one pass
 *   reduce   op 0 sum, 1 min, 2 max, 3 max|x| (norm_infty), 4 #non-finite (has_nan_or_inf)
 *            (utopia_petsc_Vector.hpp:381-397)
 * Reductions accumulate in double and finish in a fixed order (deterministic).
 * ---------------------------------------------------------------------- */
#define UHIP_DECLARE_BLAS1(SUF, T)                                                                    \
    int uhip_vec_set_##SUF(uhip_ctx ctx, int64_t n, T v, T *x);                                       \
    int uhip_vec_copy_##SUF(uhip_ctx ctx, int64_t n, const T *x, T *y);                               \
    int uhip_vec_axpy_##SUF(uhip_ctx ctx, int64_t n, T a, const T *x, T *y);                          \
    int uhip_vec_axpby_##SUF(uhip_ctx ctx, int64_t n, T a, const T *x, T b, T *y);                    \
    int uhip_vec_waxpby_##SUF(uhip_ctx ctx, int64_t n, T a, const T *x, T b, const T *y, T *w);       \
    int uhip_vec_scale_##SUF(uhip_ctx ctx, int64_t n, T a, T *x);                                     \
    int uhip_vec_shift_##SUF(uhip_ctx ctx, int64_t n, T a, T *x);                                     \
    int uhip_vec_abs_##SUF(uhip_ctx ctx, int64_t n, const T *x, T *y);                                \
    int uhip_vec_emul_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z);                   \
    int uhip_vec_ediv_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z);                   \
    int uhip_vec_pseudo_inv_##SUF(uhip_ctx ctx, int64_t n, const T *x, T *y, T tol);                  \
    int uhip_vec_emin_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z);                   \
    int uhip_vec_emax_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z);                   \
    int uhip_vec_reciprocal_##SUF(uhip_ctx ctx, int64_t n, T num, const T *x, T *y);                  \
    int uhip_vec_replace_##SUF(uhip_ctx ctx, int64_t n, T from, T to, T *x);                          \
    int uhip_vec_dots_##SUF(uhip_ctx ctx, int64_t n, int k, const T *const *xs, const T *const *ys,  \
                            double *result_d);                                                        \
    int uhip_vec_dot_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, double *result_d);       \
    int uhip_vec_norm2_##SUF(uhip_ctx ctx, int64_t n, const T *x, double *result_d);                 \
    int uhip_vec_reduce_##SUF(uhip_ctx ctx, int64_t n, const T *x, int op, double *result_d);

UHIP_DECLARE_BLAS1(f64, double)
UHIP_DECLARE_BLAS1(f32, float)
int uhip_vec_cast_f64_f32(uhip_ctx ctx, int64_t n, const double *x, float *y);
int uhip_vec_cast_f32_f64(uhip_ctx ctx, int64_t n, const float *x, double *y);
/* out = xin + a*p with dtype conversion (xin NULL = 0; a == 0 adds nothing, as VecAXPY with
 * alpha 0, so the call copies / casts xin exactly): the first Chebyshev step
 * x = x + alpha*p without a following SpMV (Chebyshev3level.hpp:154), and the
 * FP32<->FP64 boundary of the mixed-precision V-cycle. */
int uhip_vec_axpy_cast(uhip_ctx ctx, int64_t n, double a, const void *p, int p_dtype, const void *xin, int x_dtype,
                       void *out, int out_dtype);

/* Fused CG steps on device scalars s[] (solvers/smoothers/utopia_ConjugateGradient_impl.hpp:238-276):
 * cg_update:  alpha = s[i_num]/s[i_den]; x += alpha p; r_new = r - alpha Ap;
 *             s[i_alpha] = alpha; s[i_rnorm] = |r_new|_2
 * cg_pupdate: beta = s[i_num]/s[i_den]; p = z + beta p; then s[i_den] = s[i_num] */
int uhip_cg_update_f64(uhip_ctx ctx, int64_t n, double *x, const double *p, const double *r, const double *Ap,
                       double *r_new, double *scalars, int i_num, int i_den, int i_alpha, int i_rnorm);
int uhip_cg_pupdate_f64(uhip_ctx ctx, int64_t n, double *p, const double *z, double *scalars, int i_num, int i_den);
/* The same iteration with the x update moved into the p update's pass
 * (ConjugateGradient::preconditioned_solve, solvers/smoothers/utopia_ConjugateGradient_impl.hpp:237-266;
 * identical arithmetic, x_{k+1} = x_k + alpha_k p_k is formed before p_k is replaced):
 * cg_residual:  alpha = s[i_num]/s[i_den]; r_new = r - alpha Ap (r_new32 = (float)r_new if given);
 *               s[i_alpha] = alpha; s[i_rnorm] = |r_new|_2  (non-finite alpha: nothing written but s[i_alpha])
 * cg_xpupdate:  x += s[i_alpha] p; if z: p = z + (s[i_num]/s[i_den]) p, then s[i_den] = s[i_num]
 *               (non-finite s[i_alpha]: x and p are left as they are) */
int uhip_cg_residual_f64(uhip_ctx ctx, int64_t n, const double *r, const double *Ap, double *r_new, float *r_new32,
                         double *scalars, int i_num, int i_den, int i_alpha, int i_rnorm);
/* x += s[i_alpha] p; p = z + (s[i_num]/s[i_den]) p with z FP64 or FP32 (an FP32 preconditioner
 * output converts exactly); the scalar slots are left as they are (a caller alternating the
 * two r.z slots needs no copy); non-finite s[i_alpha]: x and p are left as they are */
int uhip_cg_xpupdate2(uhip_ctx ctx, int64_t n, double *x, double *p, const void *z, int z_dtype, double *scalars,
                      int i_alpha, int i_num, int i_den);
int uhip_cg_xpupdate_f64(uhip_ctx ctx, int64_t n, double *x, double *p, const double *z, double *scalars, int i_alpha,
                         int i_num, int i_den);

/* ------------------------------------------------------------------------
 * CSR matrices. Replaces PetscMatrix (AIJ) (backend/petsc/utopia_petsc_Matrix.hpp:103-1028).
 * int32 rowptr[nrows+1] / col[nnz] (sorted within each row) and values
 * (double or float) live on the device. The row-block plan (blocks[nblocks+1])
 * is made by uhip_csr_plan and owned by it.
 * ---------------------------------------------------------------------- */
typedef struct uhip_csr {
    int32_t nrows;
    int32_t ncols;
    int64_t nnz;
    int32_t *rowptr;
    int32_t *col;
    void *val;
    int32_t dtype;      /* UHIP_F64 or UHIP_F32 */
    /* ---- SpMV plan (made by uhip_csr_plan*, owned by it) ---- */
    int32_t format;     /* UHIP_FMT_ROWBLOCK or UHIP_FMT_SELL64 */
    int32_t *blocks;    /* row-block plan, device, nblocks+1 entries */
    int32_t nblocks;
    int32_t block_capk; /* LDS budget: 256*block_capk products per block */
    int32_t nslices;    /* SELL-64: slices of 64 consecutive rows */
    int32_t *slice_off; /* SELL-64: nslices+1 offsets (entries) */
    int32_t *sell_col;  /* SELL-64: column-major within a slice, padded with zeros */
    void *sell_val;
    int32_t nparts;     /* workgroups of one SpMV launch (partial sums of fused dots) */
    int32_t ndiag;      /* DIA: distinct column offsets (col - row), <= 32 */
    int32_t *dia_off;   /* DIA: ndiag offsets, ascending (device) */
    void *dia_val;      /* DIA: ndiag x ld values (ld = nrows rounded up to 64), diagonal-major, zero (or NaN,
                         * see dia_nan) where absent */
    void *dia_mask;     /* DIA: per-row bit d set iff the row stores offset d (uint8 / uint16 / uint32
                         * for ndiag <= 8 / 16 / 32) */
    int32_t dia_sym;    /* DIA of a bitwise-symmetric matrix: dia_val holds only the offsets >= 0
                         * (ndiag/2 + 1 arrays); offset -o of row r is read as offset +o of row r - o.
                         * 2: the same storage of a 7-point operator whose largest offset is a multiple
                         * of 64, run by the plane-walking 7-point kernel in both precisions */
    int32_t spmm_kernel; /* multi-RHS products (uhip_spmm*): 0 by row length (wavefront per row when
                          * nnz >= 32 rows), 1 one thread per row, 2 one wavefront per row */
    int32_t *vi_base;   /* VI: per-row base column (the row's first column; 0 for an empty row) */
    void *vi_table;     /* VI: the matrix's distinct values (dtype), at most 256 */
    int32_t vi_nvals;   /* VI: number of distinct values */
    struct uhip_csr *tr; /* A^T with its own plan (uhip_csr_plan_transpose), freed with A's plan */
    int32_t dia_nan;    /* DIA: 1 when the matrix holds no NaN; the entries a row does not hold are
                         * then NaN in dia_val (else zero) and kernels may select on the value */
    int32_t dia_s1, dia_s2; /* dia_sym 2 or dia_w27: the offsets S1 and S2 (host copies of dia_off[5],
                             * dia_off[6], or dia_off[16], dia_off[22] of a 27-point plan) */
    int32_t vi_nsig;    /* VI: > 0 when the rows' (slot mask, codes) pairs take at most 256 distinct
                         * values: dia_mask then holds one uint8 signature id per row and dia_val
                         * the vi_nsig signatures (mask word, then the code words) */
    int32_t dia_w27;    /* DIA: 1 when the offsets are the 27 of dz S2 + dy S1 + dx (dz, dy, dx in
                         * {-1, 0, 1}, S2 a multiple of 64) and the plane-walking 27-point kernel
                         * runs (full or symmetric layout) */
    void *dia_delta;    /* DIA, 27-point walk, dia_sym 1: values symmetric only up to rounding
                         * (a Galerkin product). dia_val holds the upper half; dia_delta the 13
                         * lower diagonals as differences of each value's bit pattern from its
                         * mirror's (int32 for FP64, int8 for FP32; dia_ld(n) per diagonal), so
                         * every value is exact. NULL otherwise */
} uhip_csr;

#define UHIP_FMT_AUTO (-1)
#define UHIP_FMT_ROWBLOCK 0
#define UHIP_FMT_SELL64 1
#define UHIP_FMT_DIA 2
#define UHIP_FMT_VI 3 /* value-indexed, row-relative offsets (CSR-VI style, see uhip_csr_plan_format) */

/* Build the row-block plan. block_nnz selects the LDS budget (1024, 2048 or 4096
 * nonzeros per block; 0 = automatic format choice, see uhip_csr_plan_format).
 * Rows longer than block_nnz/4 get a block of their own. */
int uhip_csr_plan(uhip_ctx ctx, uhip_csr *A, int32_t block_nnz);
/* Plan with an explicit format. UHIP_FMT_SELL64 copies the matrix into slices
 * of 64 rows stored column-major (one wavefront per slice, every load
 * coalesced). UHIP_FMT_DIA stores a square matrix with sorted rows and at most
 * 32 distinct offsets col - row as one value array per offset plus a per-row
 * bitmask (no column indices; fails with UHIP_ERR_UNSUPPORTED otherwise); when
 * the pattern and the values are symmetric bit for bit, only the offsets >= 0 are
 * stored (dia_sym = 1).
 * UHIP_FMT_VI stores a matrix (rectangular too) whose rows are at most 32 distinct
 * offsets from their own first column and whose values take at most 256 distinct
 * bit patterns (geometric and aggregation transfer operators) as a per-row base
 * column, a per-row slot mask and one 8-bit value code per slot (the value index
 * of CSR-VI, Kourtis et al. 2008), the distinct values in a table. dia_val then
 * holds the codes row by row: 1, 2 or 4 64-bit words per row (ndiag <= 8, 16,
 * 32), the code of slot d in byte d % 8 of word d / 8.
 * DIA value arrays hold NaN for the entries a row does not hold when the matrix holds no NaN
 * (dia_nan = 1), zero otherwise.
 * UHIP_FMT_AUTO takes DIA when it applies with at most 25% padding, else VI when
 * it applies with at most 3 slots per stored entry, else SELL-64 when its padding
 * adds at most 25% to nnz, else the row-block CSR kernel. Each y_i is the sum of exactly the row's CSR entries in column
 * order in every format. */
int uhip_csr_plan_format(uhip_ctx ctx, uhip_csr *A, int32_t format, int32_t block_nnz);
int uhip_csr_plan_destroy(uhip_ctx ctx, uhip_csr *A);
/* Autotuner: times y = A x for every applicable plan (DIA, VI, SELL-64, row-block CSR
 * with 1024 / 2048 / 4096 nonzeros per workgroup), reps launches each, and keeps
 * the fastest (format and best average ms returned). No reference analogue
 * (PETSc's MatMult has one fixed AIJ kernel). */
int uhip_csr_autotune(uhip_ctx ctx, uhip_csr *A, int reps, int32_t *format, float *ms);
/* reps > 0: every UHIP_FMT_AUTO plan made through this context is chosen by the
 * autotuner; 0 (default): by the padding rules of uhip_csr_plan_format. */
int uhip_ctx_set_autotune(uhip_ctx ctx, int reps);

/* y = A x                         MatMult (utopia_petsc_Matrix.cpp:1326-1358) */
int uhip_spmv(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y);
/* Build (once) A's transpose and its automatic plan in A->tr, for uhip_spmv_transpose. */
int uhip_csr_plan_transpose(uhip_ctx ctx, uhip_csr *A);
/* y = A^T x                       MatMultTranspose (PetscMatrix::transpose_multiply, utopia_petsc_Matrix.cpp:1389),
 * used by IPTransfer::restrict (solvers/multilevel/transfer/utopia_IPTransfer.hpp:65-74). Row c of A^T
 * lists A's entries of column c by increasing row, so y_c adds a_ic x_i in the order of MatMultTranspose's
 * scatter over the rows of A: the same sums bit for bit. Plans the transpose on first use. */
int uhip_spmv_transpose(uhip_ctx ctx, uhip_csr *A, const void *x, void *y);
/* y = alpha A x + beta y          MatMult + VecAXPBY (Eval_Binary, utopia_core/eval/utopia_Eval_Binary.hpp:88-94) */
int uhip_spmv_axpby(uhip_ctx ctx, const uhip_csr *A, double alpha, const void *x, double beta, void *y);
/* y = alpha A x + beta z          g = A x - rhs of MPRGP
 * (solvers/nonlinear/constrained/quadratic_programming/utopia_MPRGP.hpp:114-115,205-206) */
int uhip_spmv_axpbz(uhip_ctx ctx, const uhip_csr *A, double alpha, const void *x, double beta, const void *z,
                    void *y);
/* r = b - A x                     Eval<ResidualExpr> (utopia_core/eval/utopia_Eval_Residual.hpp:19-40)
 * b_dtype may differ from A->dtype (FP64 right-hand side of an FP32 level);
 * x and r have A->dtype. */
int uhip_spmv_residual(uhip_ctx ctx, const uhip_csr *A, const void *x, const void *b, int b_dtype, void *r);
/* y = A x and dot_d[0] = x.y      A.apply(p,Ap) + dot(p,Ap) of ConjugateGradient
 * (solvers/smoothers/utopia_ConjugateGradient_impl.hpp:237-238) */
int uhip_spmv_dot(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y, double *dot_d);
/* uhip_spmv_dot allocating nothing (no reference counterpart): for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops). Same kernels and results
 * bit for bit; the per-workgroup partial sums go to a file-scope device array of 2^18 doubles
 * (2 MiB, one per GPU, held until the process ends) instead of the context's workspace, which
 * uhip_spmv_dot grows with hipMalloc on demand. A must already carry its plan; a plan with more
 * than 2^18 workgroups is rejected (UHIP_ERR_INVALID). Calls on one GPU share the array:
 * overlapping calls (other streams, other contexts) race. */
int uhip_spmv_dot_noalloc(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y, double *dot_d);
/* y = A x, dots_d[0] = x.y, dots_d[1] = z.x: MPRGP's A p, p.Ap and g.p (or A beta, beta.A beta,
 * g.beta) in one pass (solvers/nonlinear/constrained/quadratic_programming/utopia_MPRGP.hpp:126-128,
 * 211-213) */
int uhip_spmv_dot2_f64(uhip_ctx ctx, const uhip_csr *A, const double *x, double *y, const double *z,
                       double *dots_d);
/* y = mask .* (A x), mask may be NULL   IPRTransfer::restrict + MultiLevelMask::apply
 * (solvers/multilevel/transfer/utopia_IPRTransfer.hpp:130-134, utopia_MultiLevelMask.hpp:34-37) */
int uhip_spmv_masked(uhip_ctx ctx, const uhip_csr *A, const void *x, const void *mask, void *y);
/* y += A x                         IPRTransfer::interpolate + c += c_I (utopia_Multigrid.hpp:317,343) */
int uhip_spmv_add(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y);

/* One Jacobi sweep fused with its SpMV, r = b - A xin rounded, then
 * add != 0: out = xin + d .* r   Jacobi::sweep (solvers/smoothers/utopia_Jacobi.hpp:132-138), d = pseudo_inv(diag A, 0)
 * add == 0: out = d .* r         PointJacobi::sweep (solvers/smoothers/utopia_PointJacobi.hpp:132-142), A = LU
 * xin and d have A->dtype; b (b_dtype) and out (out_dtype) may be FP64 on an FP32 level. out must
 * not alias xin, b or d (other rows still read xin). */
int uhip_spmv_jacobi(uhip_ctx ctx, const uhip_csr *A, const void *xin, const void *b, int b_dtype, const void *d,
                     void *out, int out_dtype, int add);

/* One fused step of Chebyshev3level::solve (solvers/smoothers/utopia_Chebyshev3level.hpp:138-175):
 * with s = A pprev:  r_k = rprev + a_prev*s ; p_k = -r_k + beta*pprev ;
 * xout = xin + a_prev*pprev (+ a_cur*p_k when last). rprev NULL: rprev = -pprev.
 * xin NULL: zero initial guess. rout/pout written only when last == 0.
 * pprev has p_dtype, xout has x_dtype (the FP32 level may hand its result to an
 * FP64 vector); xin, rprev, rout and pout have A->dtype. */
typedef struct uhip_cheb_step {
    const void *pprev;
    int32_t p_dtype;
    const void *rprev;
    const void *xin;
    void *xout;
    int32_t x_dtype;
    void *rout;
    void *pout;
    double a_prev, beta, a_cur;
    int32_t last;
    /* optional (last step only): dot_d[0] = xout . dot_with (FP64 vector), the
     * dot(z_new, r_new) of ConjugateGradient (utopia_ConjugateGradient_impl.hpp:270) */
    const double *dot_with;
    double *dot_d;
} uhip_cheb_step;
int uhip_cheb_step_apply(uhip_ctx ctx, const uhip_csr *A, const uhip_cheb_step *step);
/* Two Chebyshev products in one pass over the matrix (temporal blocking of a smoothing pair, in
 * A's precision for every vector; dot_with FP64). A must be a plane-walking 7-point operator
 * (dia_sym 2, dia_nan 1) whose planes split into whole chunks of 512 rows, with S1 <= 256 and
 * nrows a multiple of S2; uhip_cheb_pair_supported says whether it is, and the other entry
 * points return UHIP_ERR_UNSUPPORTED when it is not (the caller then runs the two launches).
 *  uhip_cheb_pre_residual: the last step of a two-step smoothing from a zero guess, c = a_prev b +
 *      a_cur (-(-b + a_prev A b) + beta b), written to c, then r_R = b - A c written to rR: the
 *      values of uhip_cheb_step_apply (pprev = b, no rprev, no xin, last) followed by
 *      uhip_spmv_residual, bit for bit.
 *  uhip_cheb_residual_post: pa = b - A x (not stored), then the last step with pprev = pa,
 *      xin = x: out = x + a_prev pa + a_cur (-(-pa + a_prev A pa) + beta pa), and dot_d[0] =
 *      out . dot_with when given: the values of uhip_spmv_residual followed by
 *      uhip_cheb_step_apply, bit for bit except the dot, summed in another order.
 * The reference applies the smoother's products one at a time (ChebyshevSmoother::smooth,
 * solvers/smoothers/utopia_ChebyshevSmoother.hpp); these fuse the two that meet where the
 * pre-smoothing ends and the post-smoothing starts (Multigrid::standard_cycle,
 * utopia_Multigrid.hpp:280-300, 346-360). */
int uhip_cheb_pair_supported(const uhip_csr *A);
/* workgroups of a pair launch on A (one 1024- or 512-thread group per CU at a time; 0 when not
 * supported): a grid that gives fewer than the GPU's CUs leaves CUs idle, and the multigrid then
 * runs the two launches instead. No reference analogue. */
int64_t uhip_cheb_pair_workgroups(const uhip_csr *A);
int uhip_cheb_pre_residual(uhip_ctx ctx, const uhip_csr *A, const void *b, double a_prev, double beta,
                           double a_cur, void *c, void *rR);
int uhip_cheb_residual_post(uhip_ctx ctx, const uhip_csr *A, const void *x, const void *b, double a_prev,
                            double beta, double a_cur, void *out, const double *dot_with, double *dot_d);

/* Diagonal:  d_i = A_ii (0 if absent)   MatGetDiagonal (utopia_petsc_Matrix.cpp:744-756) */
int uhip_csr_diag(uhip_ctx ctx, const uhip_csr *A, void *d);
/* A_ii += alpha * d_i where row i stores its diagonal (values only; re-plan a DIA / SELL-64 / VI
 * matrix after the call)   PetscMatrix::shift_diag -> MatDiagonalSet ADD_VALUES
 * (backend/petsc/utopia_petsc_Matrix.cpp:1783-1788); with alpha = -1 and d = diag(A) it forms
 * PointJacobi's LU = A - D (solvers/smoothers/utopia_PointJacobi.hpp:89-96) */
int uhip_csr_shift_diag(uhip_ctx ctx, uhip_csr *A, const void *d, double alpha);
/* A_ii += 1 where |A_ii| < 1e-12  LinearMultiLevel::fix_semidefinite_operator
 * (solvers/multilevel/utopia_LinearMultiLevel.hpp:89-113) */
int uhip_csr_fix_semidefinite(uhip_ctx ctx, uhip_csr *A);
/* Values cast f64 -> f32 of the same pattern (for the FP32 V-cycle). */
int uhip_csr_cast_values_f64_f32(uhip_ctx ctx, int64_t nnz, const double *src, float *dst);
/* mask for the finest level: 1 where the row has an off-diagonal |a_ij| > 1e6*eps,
 * MultiLevelMask::generate_mask_from_matrix with on=0/off=1 followed by shift(-1), abs
 * (solvers/multilevel/utopia_MultiLevelMask.hpp:39-115): output is the final mask. */
int uhip_csr_offdiag_indicator(uhip_ctx ctx, const uhip_csr *A, double tol, void *ind);
/* ind_c = min(1, Rbool * ind_f) where Rbool_ij = |R_ij| > tol
 * IPRTransfer::boolean_restrict_or (solvers/multilevel/transfer/utopia_IPRTransfer.hpp:144-173) */
int uhip_csr_boolean_restrict_or(uhip_ctx ctx, const uhip_csr *R, double tol, const void *ind_f, void *ind_c);

/* ------------------------------------------------------------------------
 * Sparse matrix products (setup). Pattern arithmetic is integer-exact.
 * ---------------------------------------------------------------------- */
/* B = A^T (structure and values, FP64 or FP32; cols sorted)   MatTranspose (utopia_petsc_Matrix.cpp:462),
 * used by IPRTransfer's R = transpose(I) (utopia_IPRTransfer.hpp:28-39).
 * Allocates B->rowptr/col/val (free with uhip_csr_free). */
int uhip_csr_transpose(uhip_ctx ctx, const uhip_csr *A, uhip_csr *B);
/* C = A * B (Gustavson, structural pattern, sorted cols). Allocates C arrays. */
int uhip_csr_spgemm(uhip_ctx ctx, const uhip_csr *A, const uhip_csr *B, uhip_csr *C);
/* Ac = P^T A P  (Galerkin triple product)   MatPtAP
 * (backend/petsc/eval/utopia_petsc_Eval_Multiply.cpp:12-61; IPRTransfer::restrict(M) :183-187).
 * Allocates Ac arrays. */
int uhip_csr_ptap(uhip_ctx ctx, const uhip_csr *A, const uhip_csr *P, uhip_csr *Ac);
/* Free arrays allocated by transpose/spgemm/ptap/alloc (and the plan). */
int uhip_csr_free(uhip_ctx ctx, uhip_csr *A);
/* Dense row-major copy (nrows x ncols doubles) for the coarsest level. */
int uhip_csr_to_dense(uhip_ctx ctx, const uhip_csr *A, double *dense);

/* ------------------------------------------------------------------------
 * Dense coarsest-level solve. Replaces Factorization<PetscMatrix,PetscVector>
 * (backend/petsc/solvers/linear/utopia_petsc_Factorization.hpp:13-84, KSPPREONLY+PCLU):
 * A = L L^T (FP64, MFMA trailing updates), then A^-1 = L^-T L^-1 kept on the
 * device so that each apply is one GEMV.
 * ---------------------------------------------------------------------- */
int uhip_dense_cholesky(uhip_ctx ctx, int32_t n, double *A, int32_t lda, int *info_h);
int uhip_dense_cholesky_inverse(uhip_ctx ctx, int32_t n, const double *L, int32_t lda, double *Ainv);
/* y = M x (row-major n x n, FP64), x/y of dtype x_dtype / y_dtype */
int uhip_dense_gemv(uhip_ctx ctx, int32_t n, const double *M, const void *x, int x_dtype, void *y, int y_dtype);

/* ------------------------------------------------------------------------
 * Box-constrained QP kernels (FP64).
 * ---------------------------------------------------------------------- */
/* out_i = (lb_i >= x_i) ? lb_i : ((ub_i <= x_i) ? ub_i : x_i)
 * VariableBoundSolverInterface::project (solvers/nonlinear/constrained/utopia_VariableBoundSolverInterface.hpp:107-155) */
int uhip_box_project_f64(uhip_ctx ctx, int64_t n, const double *x, const double *lb, const double *ub, double *out);
/* MPRGP::get_fi / get_beta (solvers/nonlinear/constrained/quadratic_programming/utopia_MPRGP.hpp:266-287,324-348),
 * gp = fi + beta and dots_d[0..2] = beta.beta, fi.fi, gp.gp in one pass (:238-243). fi/beta/gp may be NULL. */
int uhip_mprgp_state_f64(uhip_ctx ctx, int64_t n, const double *x, const double *g, const double *lb, const double *ub,
                         double *fi, double *beta, double *gp, double *dots_d);
int uhip_mprgp_free_grad_f64(uhip_ctx ctx, int64_t n, const double *x, const double *g, const double *lb,
                             const double *ub, double *fi);
/* MPRGP::get_alpha_f + multi_min (utopia_MPRGP.hpp:289-322; utopia_core/eval/utopia_MultiReduce.hpp:10-25) */
int uhip_mprgp_alpha_f_f64(uhip_ctx ctx, int64_t n, const double *x, const double *p, const double *lb,
                           const double *ub, double *result_d);
/* CG step (utopia_MPRGP.hpp:167-182): x -= alpha p; g -= alpha Ap; fi = get_fi(x,g); s[i_fiAp] = fi.Ap */
int uhip_mprgp_cg_step_f64(uhip_ctx ctx, int64_t n, double *x, double *g, const double *p, const double *Ap,
                           const double *lb, const double *ub, double *fi, double alpha, double *scalars, int i_fiAp);
/* The CG step followed by the state of the new iterate in one pass (utopia_MPRGP.hpp:167-182
 * then :238-243): x -= alpha p; g -= alpha Ap; fi, beta, gp = fi + beta of the new x, g;
 * dots_d[0..3] = beta.beta, fi.fi, gp.gp, fi.Ap. The sums equal those of
 * uhip_mprgp_cg_step_f64 and uhip_mprgp_state_f64 bit for bit. */
int uhip_mprgp_cg_step_state_f64(uhip_ctx ctx, int64_t n, double *x, double *g, const double *p, const double *Ap,
                                 const double *lb, const double *ub, double *fi, double *beta, double *gp,
                                 double alpha, double *dots_d);
/* p = fi - (s[i_num]/s[i_den]) p */
int uhip_mprgp_pupdate_f64(uhip_ctx ctx, int64_t n, double *p, const double *fi, const double *scalars, int i_num,
                           int i_den);
/* expansion step (utopia_MPRGP.hpp:184-198): x = P(x - af p - abar get_fi(x - af p, g - af Ap)) */
int uhip_mprgp_expand_f64(uhip_ctx ctx, int64_t n, double *x, const double *g, const double *p, const double *Ap,
                          const double *lb, const double *ub, double alpha_f, double alpha_bar);
/* proportioning step (utopia_MPRGP.hpp:223-235): x -= a beta; g -= a Abeta; p = get_fi(x, g) */
int uhip_mprgp_proportion_f64(uhip_ctx ctx, int64_t n, double *x, double *g, const double *beta, const double *Abeta,
                              const double *lb, const double *ub, double *p, double alpha);
/* ActiveSet::determine (solvers/nonlinear/constrained/utopia_ActiveSet.hpp:41-79): indicator_i = 1 where
 * x_i <= lb_i + tol or x_i >= ub_i - tol, else 0; changed_d[0] = number of entries that changed */
int uhip_active_set_f64(uhip_ctx ctx, int64_t n, const double *x, const double *lb, const double *ub, double tol,
                        double *indicator, double *changed_d);
/* IPTruncatedTransfer::truncate_interpolation (transfer/utopia_IPTruncatedTransfer.hpp:203-223):
 * zero the values of the rows with indicator == 1 (FP64, pattern kept). DIA and SELL-64
 * plans hold their own copy of the values: re-plan after the call. */
int uhip_csr_zero_rows(uhip_ctx ctx, uhip_csr *A, const double *indicator);
/* ProjectedConjugateGradient element step (solvers/nonlinear/constrained/quadratic_programming/
 * utopia_ProjectedConjugateGradient.hpp:71-97): with elem = z_i (the reference's view reads zk),
 * w_i = val = (|elem-ub_i| <= eps or |elem-lb_i| <= eps) ? max(u_i,0) : u_i; z_i = val == 0 ? max(p_i,0) : p_i */
int uhip_pcg_wz_f64(uhip_ctx ctx, int64_t n, const double *u, const double *lb, const double *ub, const double *p,
                    double *z, double *w);
/* ProjectedGradient direction (utopia_ProjectedGradient.hpp:157-170): p_i = 0 where x_i is within
 * 10 eps of lb_i or ub_i, else u_i */
int uhip_pg_direction_f64(uhip_ctx ctx, int64_t n, const double *x, const double *u, const double *lb,
                          const double *ub, double *p);
/* Jones-Plassmann colouring of a structurally symmetric pattern: color[i] in [0, ncolors),
 * no two coupled rows share a colour. */
int uhip_graph_color(uhip_ctx ctx, const uhip_csr *A, uint64_t seed, int32_t *color, int32_t *ncolors);
/* rows grouped by colour; offsets_h (host, ncolors+1) */
int uhip_color_groups(uhip_ctx ctx, int32_t n, const int32_t *color, int32_t ncolors, int32_t *rows,
                      int32_t *offsets_h);
/* d_inv_i = 1/a_ii where |a_ii| > 0 else 0 (ProjectedScalarGaussSeidelSweep::init_from_local_matrix,
 * solvers/nonlinear/constrained/quadratic_programming/utopia_ProjectedGaussSeidelSweep.hpp:95-104) */
int uhip_pgs_dinv(uhip_ctx ctx, const uhip_csr *A, double *dinv);
/* One colour of the projected (constrained != 0) or plain Gauss-Seidel sweep
 * (utopia_ProjectedGaussSeidelSweep.hpp:161-235), bounds lb - x, ub - x. */
int uhip_pgs_color_sweep(uhip_ctx ctx, const uhip_csr *A, int32_t cnt, const int32_t *rows, const double *dinv,
                         const double *r, const double *x, const double *lb, const double *ub, double *c,
                         int constrained);

/* ------------------------------------------------------------------------
 * Solver level: isolver_lsolve_* shaped objects
 * (plug_in_and_out/isolver/utopia_isolver_lsolve.cpp:42-220).
 * ---------------------------------------------------------------------- */
typedef struct uhip_solver_s *uhip_solver;
#define UHIP_SOLVER_CG 0          /* ConjugateGradient<..., HOMEMADE>, no preconditioner */
#define UHIP_SOLVER_CG_JACOBI 1   /* + InvDiagPreconditioner */
#define UHIP_SOLVER_CG_MG 2       /* + Multigrid (FP64 V-cycle), max_it 1 */
#define UHIP_SOLVER_CG_MG_MIXED 3 /* + Multigrid on FP32 levels (FP64 CG, residual, coarse solve) */
#define UHIP_SOLVER_MG 4          /* Multigrid as a solver */
#define UHIP_SOLVER_MPRGP 5       /* MPRGP box-constrained QP */
#define UHIP_SOLVER_PGS 6         /* ProjectedGaussSeidel (multicolour) */
#define UHIP_SOLVER_PCG 7         /* ProjectedConjugateGradient */
#define UHIP_SOLVER_PG 8          /* ProjectedGradient */
#define UHIP_SOLVER_MMG 9         /* MonotoneMultigrid (PGS smoother, truncated finest transfer) */
#define UHIP_SOLVER_AMG 10        /* AlgebraicMultigrid (Agglomerate transfers) as a solver */
#define UHIP_SOLVER_CG_AMG 11     /* ConjugateGradient + AlgebraicMultigrid preconditioner, max_it 1 */

/* isolver_lsolve_init */
int uhip_solver_create(uhip_ctx ctx, int type, uhip_solver *out);
/* QPSolverRegistry::find(backend, type) (backend/polymorphic/utopia_polymorphic_QPSolver_impl.hpp:81-97,
 * names registered at :119-125) and OmniLinearSolver's "type" (utopia_polymorphic_LinearSolver_impl.hpp:20-23,
 * SolverType names solvers/utopia_SolverType.hpp:33-34). backend: "any" or "hip"; type: "cg" / "utopia_cg"
 * (CG), "cg_jacobi", "cg_mg", "cg_mg_mixed", "mg", "mprgp", "pgs", "pcg", "pg", "mmg", "amg" (the
 * PETSc LinearSolverFactory's name, backend/petsc/solvers/linear/utopia_petsc_LinearSolverFactory.cpp:73),
 * "cg_amg". An unknown backend
 * or type is UHIP_ERR_INVALID with the reference's message ("solver X not found") in uhip_last_error();
 * the reference falls back to its default solver instead, which this library does not carry. */
int uhip_solver_create_by_name(uhip_ctx ctx, const char *backend, const char *type, uhip_solver *out);
/* isolver_lsolve_destroy */
int uhip_solver_destroy(uhip_solver s);
/* read(Input&)-style parameters: atol rtol stol max_it verbose reset_initial_guess refine
 * pre_smoothing_steps post_smoothing_steps mg_type must_generate_masks fix_semidefinite_operators
 * cycle_type (1 MULTIPLICATIVE_CYCLE, 3 FULL_CYCLE; utopia_MultiLevelBase.hpp:12-14) v_cycle_repetition
 * use_line_search (utopia_Multigrid.hpp:324-341) fused_smoothing (1, the default: on plane-walking
 * 7-point levels whose pair launch fills the GPU, uhip_cheb_pair_workgroups >= 256, the
 * pre-smoothing's last step and the residual, and the post-smoothing's residual and step, run as
 * one pass each, uhip_cheb_pre_residual / uhip_cheb_residual_post; 2: on every level that allows
 * it; 0: one launch per product)
 * smoother(0 Chebyshev3level, 1 Jacobi, 2 GaussSeidel, multicolour, 3 PointJacobi) scale_max_eig scale_min_eig power_method_tol
 * power_method_max_it block_nnz eig_comp_tol use_symmetric_sweep check_s_norm_each coarse_space_it
 * active_set_tol; AlgebraicMultigrid: n_levels bmax weight block_size (2-4: BlockAgglomerate,
 * utopia_AlgebraicMultigrid.hpp:87-103) component (block weight entry, default block_size - 1).
 * Read at the next update.
 * autotune (reps; 0 = off) applies to the SpMV plans of matrices built after it is set. */
int uhip_solver_set_param(uhip_solver s, const char *key, double value);
/* Multigrid::set_transfer_operators (utopia_LinearMultiLevel.hpp:65-79), host CSR, coarse to fine */
int uhip_solver_set_transfer_operators(uhip_solver s, int n, const int32_t *nrows, const int32_t *ncols,
                                       const int32_t *const *rowptr, const int32_t *const *col,
                                       const double *const *val);
/* isolver_lsolve_update_crs: host CSR arrays (copied to the device) */
int uhip_solver_update_crs(uhip_solver s, int32_t n, const int32_t *rowptr, const int32_t *col, const double *val);
/* same on caller-owned device arrays (PetscMatrix::wrap semantics: not freed) */
int uhip_solver_update_crs_device(uhip_solver s, int32_t n, int64_t nnz, int32_t *rowptr, int32_t *col, double *val);
/* isolver_lsolve_update_with_preconditioner_crs (plug_in_and_out/isolver/utopia_isolver_lsolve.cpp:96-133)
 * -> PreconditionedSolver::update(op, prec) (solvers/linear/utopia_PreconditionedSolver.hpp:48-61): the CG
 * iterates on A, its preconditioner (Jacobi, Multigrid, AMG) is built from the second matrix. Host arrays
 * (copied) or, with on_device, caller-owned device arrays (wrapped, not freed). UHIP_ERR_UNSUPPORTED for
 * solver types without a preconditioner. */
int uhip_solver_update_with_preconditioner_crs(uhip_solver s, int32_t n, const int32_t *rowptr, const int32_t *col,
                                               const double *val, const int32_t *prec_rowptr,
                                               const int32_t *prec_col, const double *prec_val, int on_device);
/* QPSolver::set_box_constraints (solvers/nonlinear/constrained/quadratic_programming/utopia_QPSolver.hpp:34-44);
 * NULL bound = none (filled with -/+DBL_MAX, BoxConstraints::fill_empty_bounds). Call after update. */
int uhip_solver_set_box_constraints(uhip_solver s, const double *lb, const double *ub);
/* isolver_lsolve_apply: host arrays, x is the initial guess and the result */
int uhip_solver_apply(uhip_solver s, const double *rhs, double *x);
int uhip_solver_apply_device(uhip_solver s, const double *rhs_d, double *x_d);
/* z = M rhs with the preconditioner of a CG solver (host arrays) */
int uhip_solver_precondition(uhip_solver s, const double *rhs, double *z);
/* SolutionStatus (iterates, gradient_norm, reason) and solve wall time */
int uhip_solver_get_status(uhip_solver s, int64_t *iterates, double *gradient_norm, int *reason, double *seconds);
/* setup_seconds solve_seconds total_iterations refinements n_levels coarse_n eig_max_<l> eig_min_<l>
 * level_nnz_<l> level_rows_<l> ncolors transfer_{nnz,rows,cols}_<i> alpha_bar active_set_changes;
 * format_{A,R,P}_<l>: the SpMV plan (UHIP_FMT_*) of level l's operator, restriction or
 * interpolation in the FP64 V-cycle, format32_{A,R,P}_<l> in the FP32 one of a mixed solver */
int uhip_solver_get_info(uhip_solver s, const char *key, double *value);
int uhip_solver_get_level(uhip_solver s, int l, int32_t *rowptr, int32_t *col, double *val);
int uhip_solver_get_mask(uhip_solver s, int l, double *mask);
/* interpolation l (coarse to fine; sizes from get_info transfer_rows_<l> transfer_cols_<l>
 * transfer_nnz_<l>): the given ones, or those AlgebraicMultigrid built (Agglomerate::create_transfer,
 * solvers/multilevel/agglomerate/utopia_Agglomerate.hpp:37-39) */
int uhip_solver_get_transfer(uhip_solver s, int l, int32_t *rowptr, int32_t *col, double *val);
int uhip_solver_set_level_eigs(uhip_solver s, int l, double eig_max, double eig_min);
/* colour of every row of level l (GaussSeidel smoother) and the number of colours */
int uhip_solver_get_colors(uhip_solver s, int l, int32_t *color, int32_t *ncolors);
int uhip_lsolve_set_max_iterations(uhip_solver s, int max_it);
int uhip_lsolve_set_atol(uhip_solver s, double v);
int uhip_lsolve_set_rtol(uhip_solver s, double v);
int uhip_lsolve_set_stol(uhip_solver s, double v);
int uhip_lsolve_set_verbosity(uhip_solver s, int level);

/* ------------------------------------------------------------------------
 * Multi-RHS and batched kernels (no reference analogue: utopia vectors are
 * single-column and every solve is one system). Blocks of k right-hand sides
 * are row-major [n x k]; k in {1, 2, 4, 8, 16}. A batch of independent systems
 * is one block-diagonal CSR; rowseg[i] is the system of row i.
 * ---------------------------------------------------------------------- */
/* op 0: Y = A X; op 1: Y += A X; op 2: Y = B - A X (B dtype b_dtype).
 * Each column is the single-vector SpMV of PetscMatrix::multiply
 * (backend/petsc/utopia_petsc_Matrix.cpp:1137-1152) with per-row sums in entry order,
 * except for matrices averaging >= 32 entries per row (dense blocks), whose rows
 * are summed by one wavefront each in a fixed tree order. */
/* Autotuner of the multi-RHS product: times Y = A X (width k in {1,2,4,8,16}, A's
 * precision) with one thread per row and with one wavefront per row, reps launches
 * each, and keeps the faster in A->spmm_kernel (choice and its average ms returned).
 * The thread-per-row kernel sums each row in column order, the wavefront kernel in a
 * fixed tree. No reference analogue (utopia vectors are single-column). */
int uhip_csr_autotune_spmm(uhip_ctx ctx, uhip_csr *A, int k, int reps, int32_t *choice, float *ms);
int uhip_spmm(uhip_ctx ctx, const uhip_csr *A, int k, const void *X, int x_dtype, void *Y, int y_dtype, int op,
              const void *B, int b_dtype);
/* One Chebyshev3level step (solvers/smoothers/utopia_Chebyshev3level.hpp:150-175)
 * on k columns with per-system coefficients coef[3*s + {0,1,2}] = (alpha_{k-1}, beta_k, alpha_k).
 * rprev NULL: r = -pprev; xin NULL: x = 0; last: x += alpha_k p_k, no rout/pout. */
int uhip_spmm_cheb_step(uhip_ctx ctx, const uhip_csr *A, int k, const void *pprev, int p_dtype, const void *rprev,
                        const void *xin, void *xout, int x_dtype, void *rout, void *pout, const int32_t *rowseg,
                        const double *coef, int last);
/* out = xin + coef[3*rowseg[i]] * p (one-step Chebyshev); xin may be NULL */
int uhip_axpy_seg(uhip_ctx ctx, int32_t n, int k, const int32_t *rowseg, const double *coef, const void *p,
                  int p_dtype, const void *xin, void *out, int out_dtype, int level_dtype);
/* out[s*k + j] = sum over rows of system s of x[i,j]*y[i,j] (sqrt if take_sqrt).
 * Row chunks [c_row0[c], c_row1[c]) never straddle systems; system s owns chunks
 * seg_c0[s] .. seg_c0[s+1]-1. Fixed summation order (deterministic). */
int uhip_segdot(uhip_ctx ctx, int k, int32_t nchunks, const int32_t *c_row0, const int32_t *c_row1, int32_t nseg,
                const int32_t *seg_c0, const void *x, const void *y, int dtype, double *out, int take_sqrt);
/* Batched CG stages for m = nseg*k (system, rhs) pairs, arithmetic of
 * ConjugateGradient::preconditioned_solve (solvers/smoothers/utopia_ConjugateGradient_impl.hpp:195-294).
 * state_d = {rz, pap, alpha, beta, rznew, rr, atol} (device, m each); flags 0 active,
 * 1 converged, 2 stopped on a non-finite alpha. stage 0: alpha = rz/pap, x += alpha p,
 * r -= alpha Ap; stage 1: rr (= |r|) <= atol converges, *iter += 1, conv_it = *iter,
 * *active_count = active pairs; stage 2: beta = rznew/rz, rz = rznew, p = z + beta p. */
int uhip_bcg_step(uhip_ctx ctx, int stage, int32_t n, int k, int32_t m, const int32_t *rowseg, double *x, double *p,
                  double *r, const double *Ap, const double *z, double *const *state_d, int32_t *flags,
                  int32_t *conv_it, int32_t *iter, int32_t *active_count);
/* v = w / norm[rowseg[i]] for systems with done == NULL or done[s] == 0 (power method) */
int uhip_seg_scale(uhip_ctx ctx, int32_t n, const int32_t *rowseg, const double *norm, const int32_t *done,
                   const double *w, double *v);
/* Value gradient on the pattern of A: grad[e] = -sum_j lambda[row(e), j] * x[col(e), j]
 * for L(x(A)), with lambda = A^-T dL/dx (the adjoint solve). */
int uhip_pattern_grad(uhip_ctx ctx, const uhip_csr *A, int k, const double *lambda, const double *x, double *grad);

/* ------------------------------------------------------------------------
 * Batched MG-preconditioned CG (no reference analogue): nseg independent SPD
 * systems, k right-hand sides each, FP64 CG with an FP64 or FP32 V-cycle
 * (Chebyshev smoothing, Galerkin coarse operators, dense coarse inverses per
 * system). One CG iteration of the whole batch is a captured hipGraph.
 * ---------------------------------------------------------------------- */
typedef struct uhip_bsolver_s *uhip_bsolver;
int uhip_bsolver_create(uhip_ctx ctx, int k, int mixed, uhip_bsolver *out);
int uhip_bsolver_destroy(uhip_bsolver s);
/* rtol max_it pre_smoothing_steps post_smoothing_steps scale_max_eig scale_min_eig
 * power_method_tol power_method_max_it use_graph graph_iterations (CG iterations per captured graph,
 * default 16; kernels gated on the active count, see uhip_ctx_set_gate) workspace_reserve (pre-size
 * the reduction workspace; a captured graph is captured again after it grows); autotune (reps > 0: at setup
 * the multi-RHS kernel of every level operator, transfer and coarse block inverse is chosen by
 * uhip_csr_autotune_spmm) */
int uhip_bsolver_set_param(uhip_bsolver s, const char *key, double value);
/* Block-diagonal operator (device CSR, FP64) with seg[0..nseg] row offsets;
 * nlev-1 interpolations coarse to fine (device CSR, finest last), coarse_seg
 * holds the nseg+1 row offsets of each coarse level (coarsest first). */
int uhip_bsolver_setup(uhip_bsolver s, int32_t nseg, const int32_t *seg, const uhip_csr *A, int nlev,
                       const uhip_csr *const *interps, const int32_t *const *coarse_seg);
/* Device arrays [n x k]: X holds the initial guess and receives the solution. */
int uhip_bsolver_solve(uhip_bsolver s, const double *B_d, double *X_d, int64_t *iterations);
/* per (system, rhs): iterations to convergence (-1: not converged) and flags */
int uhip_bsolver_get_conv(uhip_bsolver s, int32_t *conv_it, int32_t *flags);
/* dL/dA on the pattern of the finest operator from lambda and x (device, [n x k]) */
int uhip_bsolver_pattern_grad(uhip_bsolver s, const double *lambda_d, const double *x_d, double *grad_d);
/* setup_seconds solve_seconds iterations graph_replays graph_captures graph_iterations n_levels eig_max_<l>_<s> */
int uhip_bsolver_get_info(uhip_bsolver s, const char *key, double *value);
int uhip_bsolver_set_level_eig(uhip_bsolver s, int l, int32_t sys, double eig_max);

/* ------------------------------------------------------------------------
 * Problem generators (host, seeded): the inputs every path consumes.
 * splitmix64 stream; uniform in [0,1) = (u >> 11) * 2^-53; normals by
 * Box-Muller. Arrays are caller-allocated host memory.
 * ---------------------------------------------------------------------- */
/* nnz of the stencil operators below (for allocation). */
int64_t uhip_gen_nnz(int kind, int32_t nx, int32_t ny, int32_t nz);
/* kind 0: 2D 5-point Laplacian [-1,-1,4,-1,-1] on nx*ny interior (Dirichlet eliminated)
 * kind 1: 3D 7-point Laplacian on nx*ny*nz (diag 6)
 * kind 2: 3D 7-point variable coefficient: cell kappa = exp(N(0,1)) from seed,
 *         face coefficient 2 ka kb/(ka+kb), boundary faces use the cell value
 * kind 3: 2D 9-point Laplacian (diag 8, all eight neighbours -1) + shift*I
 * kind 4: 2D 5-point with per-edge coefficients exp(N(0, 0.5)) from seed
 * kind 5: kind 2 renumbered, P A P^T: the full 64-row blocks are permuted by a seeded
 *         Fisher-Yates shuffle (a tail block of fewer than 64 rows keeps its place), rows keep
 *         sorted columns. A symmetric operator with no diagonal structure, for the general
 *         sparse formats. */
int uhip_gen_stencil(int kind, int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double shift,
                     int32_t *rowptr, int32_t *col, double *val);
/* Trilinear (bilinear / linear for nz==1 / ny==1) prolongation from the coarse
 * grid (nx/2, ny/2, nz/2) (integer halves; a side of 1 stays 1) to the fine grid (nx, ny, nz).
 * Fine point f (0-based) takes coarse point (f-1)/2 with weight 1 when f is odd,
 * coarse points f/2-1 and f/2 with weight 1/2 each (those inside the grid) when
 * f is even. */
int64_t uhip_gen_prolongation_nnz(int32_t nx, int32_t ny, int32_t nz);
int uhip_gen_prolongation(int32_t nx, int32_t ny, int32_t nz, int32_t *rowptr, int32_t *col, double *val);
/* x_i = lo + (hi-lo)*u_i */
int uhip_gen_uniform(uint64_t seed, int64_t n, double lo, double hi, double *x);
/* x_i ~ N(mu, sigma) */
int uhip_gen_normal(uint64_t seed, int64_t n, double mu, double sigma, double *x);

#ifdef __cplusplus
}
#endif

#endif /* UTOPIA_HIP_H */
