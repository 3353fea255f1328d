// Fused vector kernels of the Krylov and smoother loops. Step lengths are read
// from device scalars so a CG iteration needs one host read-back (the residual
// norm) instead of one per dot product.
#include "uhip_common.hpp"

namespace uhip {

// out = xin + a * p   (xin may be null = 0), with dtype conversion. a == 0 adds nothing
// (VecAXPY with alpha 0 is a no-op), so the call is an exact copy / cast of xin.
template <typename TP, typename TX, typename TO, typename TAcc>
__global__ __launch_bounds__(kBlock) void k_axpy_cast(int64_t n, TAcc a, const TP *__restrict__ p,
                                                      const TX *__restrict__ xin, TO *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        TAcc v = xin ? (TAcc)xin[i] : TAcc(0);
        if (a != TAcc(0)) v = v + a * (TAcc)p[i];
        out[i] = (TO)v;
    }
}

// alpha = s[i_num]/s[i_den]; x += alpha*p; rn = r - alpha*Ap; partial |rn|^2.
// Block 0 stores alpha to s[i_alpha].
__global__ __launch_bounds__(kBlock) void k_cg_update(int64_t n, double *__restrict__ x,
                                                      const double *__restrict__ p, const double *__restrict__ r,
                                                      const double *__restrict__ Ap, double *__restrict__ rn,
                                                      double *__restrict__ s, int i_num, int i_den, int i_alpha,
                                                      double *__restrict__ partials) {
    const double alpha = s[i_num] / s[i_den];
    if (!isfinite(alpha)) {  // CG stops before touching x (utopia_ConjugateGradient_impl.hpp:241-244)
        if (blockIdx.x == 0 && threadIdx.x == 0) s[i_alpha] = alpha;
        return;
    }
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        x[i] = x[i] + alpha * p[i];
        const double v = r[i] - alpha * Ap[i];
        rn[i] = v;
        acc += v * v;
    }
    acc = block_sum<kBlock>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc;
        if (blockIdx.x == 0) s[i_alpha] = alpha;
    }
}

// beta = s[i_num]/s[i_den]; p = z + beta*p; afterwards s[i_den] = s[i_num] (the
// next iteration's dot(z, r) is this iteration's dot(z_new, r_new)).
__global__ __launch_bounds__(kBlock) void k_cg_pupdate(int64_t n, double *__restrict__ p, const double *__restrict__ z,
                                                       const double *__restrict__ s, int i_num, int i_den) {
    const double beta = s[i_num] / s[i_den];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        p[i] = z[i] + beta * p[i];
}

// alpha = s[i_num]/s[i_den]; rn = r - alpha*Ap (and rn32 = (float)rn when
// given); partial |rn|^2. x is left to k_cg_xpupdate. Block 0 stores alpha.
__global__ __launch_bounds__(kBlock) void k_cg_residual(int64_t n, const double *__restrict__ r,
                                                        const double *__restrict__ Ap, double *__restrict__ rn,
                                                        float *__restrict__ rn32, double *__restrict__ s, int i_num,
                                                        int i_den, int i_alpha, double *__restrict__ partials) {
    const double alpha = s[i_num] / s[i_den];
    if (!isfinite(alpha)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) s[i_alpha] = alpha;
        return;
    }
    // four grid-strided entries per step, their loads issued together (the grid is capped at
    // kMaxRedBlocks, so each thread covers many entries)
    double acc = 0.0;
    const int64_t st = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + 3 * st < n; i += 4 * st) {
        double rv[4], av[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            rv[u] = __builtin_nontemporal_load(r + i + u * st);
            av[u] = __builtin_nontemporal_load(Ap + i + u * st);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double v = rv[u] - alpha * av[u];
            __builtin_nontemporal_store(v, rn + i + u * st);
            // the FP32 copy is the preconditioner's next input: a cached store keeps it in the MALL
            if (rn32) rn32[i + u * st] = (float)v;
            acc += v * v;
        }
    }
    for (; i < n; i += st) {
        const double v = r[i] - alpha * Ap[i];
        __builtin_nontemporal_store(v, rn + i);
        if (rn32) rn32[i] = (float)v;
        acc += v * v;
    }
    acc = block_sum<kBlock>(acc);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = acc;
        if (blockIdx.x == 0) s[i_alpha] = alpha;
    }
}

// x += s[i_alpha]*p, then (z given) p = z + beta*p with beta = s[i_num]/s[i_den]:
// the x update of iteration k uses p_k before it is replaced by p_{k+1}.
template <typename TZ>
__global__ __launch_bounds__(kBlock) void k_cg_xpupdate(int64_t n, double *__restrict__ x, double *__restrict__ p,
                                                        const TZ *__restrict__ z, const double *__restrict__ s,
                                                        int i_alpha, int i_num, int i_den) {
    const double alpha = s[i_alpha];
    if (!isfinite(alpha)) return;  // the CG stops before touching x
    const double beta = z ? s[i_num] / s[i_den] : 0.0;
    // x streams (loaded and stored non-temporally: it is next read by the next iteration's update,
    // so caching it only evicts the next product's operands); p and z go through the caches (z was
    // just written by the preconditioner, p is read next by A p). Measured at 256^3 against
    // cached x loads: MG-PCG mixed -1.4 to -1.8 ms, FP64 -2 to -3 ms; loading p and z
    // non-temporally as well gave back 1.2 - 1.5 ms of it.
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double pi = p[i];
        __builtin_nontemporal_store(__builtin_nontemporal_load(x + i) + alpha * pi, x + i);
        if (z) p[i] = (double)z[i] + beta * pi;
    }
}

__global__ void k_scalar_copy(double *s, int dst, int src) { s[dst] = s[src]; }

}  // namespace uhip

using namespace uhip;

extern "C" {

int uhip_vec_axpy_cast(uhip_ctx ctx, int64_t n, double a, const void *p, int p_dtype, const void *xin, int x_dtype,
                       void *out, int out_dtype) {
    UHIP_REQUIRE(ctx && p && out, "uhip_vec_axpy_cast: null argument");
    if (n <= 0) return UHIP_OK;
    dim3 g(grid_for(n, 1, 1 << 16)), b(kBlock);
    hipStream_t s = ctx->stream;
    // working precision: float only if every operand is float
    const bool f32 = p_dtype == UHIP_F32 && (xin == nullptr || x_dtype == UHIP_F32) && out_dtype == UHIP_F32;
#define UHIP_AC(TP, TX, TO, TA)                                                                     \
    hipLaunchKernelGGL((k_axpy_cast<TP, TX, TO, TA>), g, b, 0, s, n, (TA)a, (const TP *)p, (const TX *)xin, \
                       (TO *)out)
    if (f32) {
        UHIP_AC(float, float, float, float);
    } else if (p_dtype == UHIP_F64 && x_dtype != UHIP_F32 && out_dtype == UHIP_F64) {
        UHIP_AC(double, double, double, double);
    } else if (p_dtype == UHIP_F32 && out_dtype == UHIP_F64) {
        if (xin && x_dtype == UHIP_F64) UHIP_AC(float, double, double, double);
        else UHIP_AC(float, float, double, float);
    } else if (p_dtype == UHIP_F64 && out_dtype == UHIP_F32) {
        if (xin && x_dtype == UHIP_F64) UHIP_AC(double, double, float, double);
        else UHIP_AC(double, float, float, float);
    } else {
        UHIP_AC(double, float, double, double);
    }
#undef UHIP_AC
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_cg_update_f64(uhip_ctx ctx, int64_t n, double *x, const double *p, const double *r, const double *Ap,
                       double *r_new, double *scalars, int i_num, int i_den, int i_alpha, int i_rnorm) {
    UHIP_REQUIRE(ctx && x && p && r && Ap && r_new && scalars, "uhip_cg_update_f64: null argument");
    const int g = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, g));
    hipLaunchKernelGGL(k_cg_update, dim3(g), dim3(kBlock), 0, ctx->stream, n, x, p, r, Ap, r_new, scalars, i_num,
                       i_den, i_alpha, ctx->partials);
    UHIP_CHECK_LAUNCH();
    return finish_sums(ctx, 1, g, ctx->max_partials, scalars + i_rnorm, 1u);
}

int uhip_cg_pupdate_f64(uhip_ctx ctx, int64_t n, double *p, const double *z, double *scalars, int i_num, int i_den) {
    UHIP_REQUIRE(ctx && p && z && scalars, "uhip_cg_pupdate_f64: null argument");
    if (n > 0) {
        hipLaunchKernelGGL(k_cg_pupdate, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, p, z,
                           scalars, i_num, i_den);
        UHIP_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_scalar_copy, dim3(1), dim3(1), 0, ctx->stream, scalars, i_den, i_num);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_cg_residual_f64(uhip_ctx ctx, int64_t n, const double *r, const double *Ap, double *r_new, float *r_new32,
                         double *scalars, int i_num, int i_den, int i_alpha, int i_rnorm) {
    UHIP_REQUIRE(ctx && r && Ap && r_new && scalars, "uhip_cg_residual_f64: null argument");
    const int g = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, g));
    hipLaunchKernelGGL(k_cg_residual, dim3(g), dim3(kBlock), 0, ctx->stream, n, r, Ap, r_new, r_new32, scalars, i_num,
                       i_den, i_alpha, ctx->partials);
    UHIP_CHECK_LAUNCH();
    return finish_sums(ctx, 1, g, ctx->max_partials, scalars + i_rnorm, 1u);
}

int uhip_cg_xpupdate_f64(uhip_ctx ctx, int64_t n, double *x, double *p, const double *z, double *scalars, int i_alpha,
                         int i_num, int i_den) {
    UHIP_REQUIRE(ctx && x && p && scalars, "uhip_cg_xpupdate_f64: null argument");
    if (n > 0) {
        hipLaunchKernelGGL(k_cg_xpupdate<double>, dim3(grid_for(n, 4, 1 << 16)), dim3(kBlock), 0, ctx->stream, n, x,
                           p, z, scalars, i_alpha, i_num, i_den);
        UHIP_CHECK_LAUNCH();
    }
    if (z) {
        hipLaunchKernelGGL(k_scalar_copy, dim3(1), dim3(1), 0, ctx->stream, scalars, i_den, i_num);
        UHIP_CHECK_LAUNCH();
    }
    return UHIP_OK;
}

int uhip_cg_xpupdate2(uhip_ctx ctx, int64_t n, double *x, double *p, const void *z, int z_dtype, double *scalars,
                      int i_alpha, int i_num, int i_den) {
    UHIP_REQUIRE(ctx && x && p && z && scalars, "uhip_cg_xpupdate2: null argument");
    UHIP_REQUIRE(z_dtype == UHIP_F64 || z_dtype == UHIP_F32, "uhip_cg_xpupdate2: z dtype");
    if (n <= 0) return UHIP_OK;
    const dim3 g(grid_for(n, 4, 1 << 16)), b(kBlock);
    if (z_dtype == UHIP_F64)
        hipLaunchKernelGGL(k_cg_xpupdate<double>, g, b, 0, ctx->stream, n, x, p, (const double *)z, scalars, i_alpha,
                           i_num, i_den);
    else
        hipLaunchKernelGGL(k_cg_xpupdate<float>, g, b, 0, ctx->stream, n, x, p, (const float *)z, scalars, i_alpha,
                           i_num, i_den);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

}  // extern "C"
