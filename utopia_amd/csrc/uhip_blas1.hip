// Streaming vector kernels (BLAS1) and fused reductions: the operations the
// solver templates call on PetscVector (VecSet, VecCopy, VecAXPY, VecAXPBY,
// VecScale, VecPointwiseMult/Divide, VecShift, VecAbs, VecDot, VecNorm,
// VecMin/VecMax, VecDotBegin/End) replaced by HBM-bound HIP kernels.
#include "uhip_common.hpp"

namespace uhip {

// Grid-stride loop with four independent elements in flight per lane. Results
// are written with streaming (non-temporal) stores: each is written once.
#define UHIP_FOR4(i, n)                                                                    \
    for (int64_t base_ = (int64_t)blockIdx.x * kBlock + threadIdx.x,                      \
                 stride_ = (int64_t)gridDim.x * kBlock;                                    \
         base_ < (n); base_ += 4 * stride_)                                                \
        _Pragma("unroll") for (int u_ = 0; u_ < 4; ++u_) if (int64_t i = base_ + u_ * stride_; i < (n))

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t n, T v, T *__restrict__ x) {
    UHIP_FOR4(i, n) { __builtin_nontemporal_store(v, x + i); }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const TI *__restrict__ x, TO *__restrict__ y) {
    UHIP_FOR4(i, n) { __builtin_nontemporal_store((TO)x[i], y + i); }
}

// y = y + a*x
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy(int64_t n, T a, const T *__restrict__ x, T *__restrict__ y) {
    UHIP_FOR4(i, n) { __builtin_nontemporal_store(y[i] + a * x[i], y + i); }
}

// y = a*x + b*y
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, T a, const T *__restrict__ x, T b, T *__restrict__ y) {
    UHIP_FOR4(i, n) { __builtin_nontemporal_store(a * x[i] + b * y[i], y + i); }
}

// w = a*x + b*y
template <typename T>
__global__ __launch_bounds__(kBlock) void k_waxpby(int64_t n, T a, const T *__restrict__ x, T b,
                                                   const T *__restrict__ y, T *__restrict__ w) {
    UHIP_FOR4(i, n) { __builtin_nontemporal_store(a * x[i] + b * y[i], w + i); }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(int64_t n, T a, T *__restrict__ x) {
    UHIP_FOR4(i, n) { __builtin_nontemporal_store(a * x[i], x + i); }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_shift(int64_t n, T a, T *__restrict__ x) {
    UHIP_FOR4(i, n) { __builtin_nontemporal_store(x[i] + a, x + i); }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_abs(int64_t n, const T *__restrict__ x, T *__restrict__ y) {
    UHIP_FOR4(i, n) { y[i] = fabs(x[i]); }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_emul(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                 T *__restrict__ z) {
    UHIP_FOR4(i, n) { z[i] = x[i] * y[i]; }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ediv(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                 T *__restrict__ z) {
    UHIP_FOR4(i, n) { z[i] = x[i] / y[i]; }
}

// z = min(x, y) / max(x, y) entry by entry in the form of PETSc's VecPointwiseMin /
// VecPointwiseMax (PetscMin(a, b) = a < b ? a : b, PetscMax(a, b) = a < b ? b : a),
// called by PetscVector::e_min / e_max (backend/petsc/utopia_petsc_Vector.cpp:597-603).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_emin(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                 T *__restrict__ z) {
    UHIP_FOR4(i, n) {
        const T a = x[i], b = y[i];
        z[i] = a < b ? a : b;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_emax(int64_t n, const T *__restrict__ x, const T *__restrict__ y,
                                                 T *__restrict__ z) {
    UHIP_FOR4(i, n) {
        const T a = x[i], b = y[i];
        z[i] = a < b ? b : a;
    }
}

// PetscVector::reciprocal(num) (backend/petsc/utopia_petsc_Vector.hpp:711-720): VecReciprocal
// (1/x where x != 0, zeros kept), then a scale by num unless num == 1.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reciprocal(int64_t n, T num, const T *__restrict__ x,
                                                       T *__restrict__ y) {
    UHIP_FOR4(i, n) {
        const T v = x[i];
        const T r = v != T(0) ? T(1) / v : v;
        y[i] = num == T(1) ? r : num * r;
    }
}

// x_i = to where x_i == from (PointJacobi::check_indef turns zeros into ones,
// solvers/smoothers/utopia_PointJacobi.hpp:116-127)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_replace(int64_t n, T from, T to, T *__restrict__ x) {
    UHIP_FOR4(i, n) {
        if (x[i] == from) x[i] = to;
    }
}

// y_i = 1/x_i where |x_i| > tol, else 0 (utopia_ElementWisePseudoInverse.hpp:20-56).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pseudo_inv(int64_t n, const T *__restrict__ x, T *__restrict__ y,
                                                       T tol) {
    UHIP_FOR4(i, n) {
        const T v = x[i];
        y[i] = (fabs(v) > tol) ? T(1) / v : T(0);
    }
}

// Up to kMaxDots dot products in one pass; partial sums in double.
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_dots(int64_t n, const T *__restrict__ x0, const T *__restrict__ y0,
                                                 const T *__restrict__ x1, const T *__restrict__ y1,
                                                 const T *__restrict__ x2, const T *__restrict__ y2,
                                                 const T *__restrict__ x3, const T *__restrict__ y3,
                                                 double *__restrict__ partials, int64_t stride) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    UHIP_FOR4(i, n) {
        s0 += (double)x0[i] * (double)y0[i];
        if (K > 1) s1 += (double)x1[i] * (double)y1[i];
        if (K > 2) s2 += (double)x2[i] * (double)y2[i];
        if (K > 3) s3 += (double)x3[i] * (double)y3[i];
    }
    s0 = block_sum<kBlock>(s0);
    if (K > 1) s1 = block_sum<kBlock>(s1);
    if (K > 2) s2 = block_sum<kBlock>(s2);
    if (K > 3) s3 = block_sum<kBlock>(s3);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s0;
        if (K > 1) partials[stride + blockIdx.x] = s1;
        if (K > 2) partials[2 * stride + blockIdx.x] = s2;
        if (K > 3) partials[3 * stride + blockIdx.x] = s3;
    }
}

// op 0: sum, 1: min, 2: max, 3: max |x|, 4: count of non-finite entries.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reduce(int64_t n, const T *__restrict__ x, int op,
                                                   double *__restrict__ partials) {
    double acc = (op == 1) ? INFINITY : (op == 2 ? -INFINITY : 0.0);
    UHIP_FOR4(i, n) {
        const double v = (double)x[i];
        if (op == 0) acc += v;
        else if (op == 1) acc = fmin(acc, v);
        else if (op == 2) acc = fmax(acc, v);
        else if (op == 3) acc = fmax(acc, fabs(v));
        else acc += isfinite(v) ? 0.0 : 1.0;
    }
    __shared__ double red[kBlock];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double a = red[threadIdx.x], b = red[threadIdx.x + s];
            red[threadIdx.x] = (op == 0 || op == 4) ? a + b : (op == 1 ? fmin(a, b) : fmax(a, b));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_reduce_finish(const double *__restrict__ partials, int64_t np, int op,
                                                          double *__restrict__ out) {
    double acc = (op == 1) ? INFINITY : (op == 2 ? -INFINITY : 0.0);
    for (int64_t b = threadIdx.x; b < np; b += kBlock) {
        const double v = partials[b];
        if (op == 0 || op == 4) acc += v;
        else if (op == 1) acc = fmin(acc, v);
        else acc = fmax(acc, v);
    }
    __shared__ double red[kBlock];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double a = red[threadIdx.x], b = red[threadIdx.x + s];
            red[threadIdx.x] = (op == 0 || op == 4) ? a + b : (op == 1 ? fmin(a, b) : fmax(a, b));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <typename T>
static int launch_dots(uhip_ctx ctx, int64_t n, int k, const T *const *xs, const T *const *ys, double *result_d,
                       unsigned sqrt_mask) {
    UHIP_REQUIRE(ctx && result_d, "dots: null argument");
    UHIP_REQUIRE(k >= 1 && k <= kMaxDots, "dots: k=%d out of range", k);
    for (int j = 0; j < k; ++j) UHIP_REQUIRE(n == 0 || (xs[j] && ys[j]), "dots: null vector %d", j);
    const int g = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, g));
    const int64_t st = ctx->max_partials;
    const T *a[4] = {nullptr, nullptr, nullptr, nullptr}, *b[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int j = 0; j < k; ++j) a[j] = xs[j], b[j] = ys[j];
    hipStream_t s = stream_of(ctx);
#define UHIP_DOTS_CASE(K)                                                                                    \
    case K:                                                                                                  \
        hipLaunchKernelGGL((k_dots<T, K>), dim3(g), dim3(kBlock), 0, s, n, a[0], b[0], a[1], b[1], a[2], b[2], \
                           a[3], b[3], ctx->partials, st);                                                    \
        break;
    switch (k) {
        UHIP_DOTS_CASE(1)
        UHIP_DOTS_CASE(2)
        UHIP_DOTS_CASE(3)
        UHIP_DOTS_CASE(4)
    }
#undef UHIP_DOTS_CASE
    UHIP_CHECK_LAUNCH();
    return finish_sums(ctx, k, g, st, result_d, sqrt_mask);
}

template <typename T>
static int launch_reduce(uhip_ctx ctx, int64_t n, const T *x, int op, double *result_d) {
    UHIP_REQUIRE(ctx && result_d && (x || n == 0), "reduce: null argument");
    UHIP_REQUIRE(op >= 0 && op <= 4, "reduce: op %d", op);
    const int g = grid_for(n, 4, kMaxRedBlocks);
    UHIP_TRY(ensure_partials(ctx, g));
    hipStream_t s = stream_of(ctx);
    hipLaunchKernelGGL((k_reduce<T>), dim3(g), dim3(kBlock), 0, s, n, x, op, ctx->partials);
    UHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(kBlock), 0, s, ctx->partials, (int64_t)g, op, result_d);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

}  // namespace uhip

using namespace uhip;

#define UHIP_LAUNCH1(kern, n, ...)                                                                   \
    do {                                                                                             \
        UHIP_REQUIRE(ctx, "null context");                                                           \
        if ((n) <= 0) return UHIP_OK;                                                                \
        hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(kBlock), 0, stream_of(ctx), n, __VA_ARGS__); \
        UHIP_CHECK_LAUNCH();                                                                         \
        return UHIP_OK;                                                                              \
    } while (0)

#define UHIP_BLAS1_API(SUF, T)                                                                                   \
    int uhip_vec_set_##SUF(uhip_ctx ctx, int64_t n, T v, T *x) { UHIP_LAUNCH1(k_fill<T>, n, v, x); }             \
    int uhip_vec_copy_##SUF(uhip_ctx ctx, int64_t n, const T *x, T *y) {                                        \
        if (x == y) return UHIP_OK;                                                                              \
        UHIP_LAUNCH1((k_copy<T, T>), n, x, y);                                                                   \
    }                                                                                                            \
    int uhip_vec_axpy_##SUF(uhip_ctx ctx, int64_t n, T a, const T *x, T *y) { UHIP_LAUNCH1(k_axpy<T>, n, a, x, y); } \
    int uhip_vec_axpby_##SUF(uhip_ctx ctx, int64_t n, T a, const T *x, T b, T *y) {                              \
        UHIP_LAUNCH1(k_axpby<T>, n, a, x, b, y);                                                                 \
    }                                                                                                            \
    int uhip_vec_waxpby_##SUF(uhip_ctx ctx, int64_t n, T a, const T *x, T b, const T *y, T *w) {                 \
        UHIP_LAUNCH1(k_waxpby<T>, n, a, x, b, y, w);                                                             \
    }                                                                                                            \
    int uhip_vec_scale_##SUF(uhip_ctx ctx, int64_t n, T a, T *x) { UHIP_LAUNCH1(k_scale<T>, n, a, x); }          \
    int uhip_vec_shift_##SUF(uhip_ctx ctx, int64_t n, T a, T *x) { UHIP_LAUNCH1(k_shift<T>, n, a, x); }          \
    int uhip_vec_abs_##SUF(uhip_ctx ctx, int64_t n, const T *x, T *y) { UHIP_LAUNCH1(k_abs<T>, n, x, y); }       \
    int uhip_vec_emul_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z) {                             \
        UHIP_LAUNCH1(k_emul<T>, n, x, y, z);                                                                     \
    }                                                                                                            \
    int uhip_vec_ediv_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z) {                             \
        UHIP_LAUNCH1(k_ediv<T>, n, x, y, z);                                                                     \
    }                                                                                                            \
    int uhip_vec_emin_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z) {                             \
        UHIP_LAUNCH1(k_emin<T>, n, x, y, z);                                                                     \
    }                                                                                                            \
    int uhip_vec_emax_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, T *z) {                             \
        UHIP_LAUNCH1(k_emax<T>, n, x, y, z);                                                                     \
    }                                                                                                            \
    int uhip_vec_reciprocal_##SUF(uhip_ctx ctx, int64_t n, T num, const T *x, T *y) {                            \
        UHIP_LAUNCH1(k_reciprocal<T>, n, num, x, y);                                                             \
    }                                                                                                            \
    int uhip_vec_replace_##SUF(uhip_ctx ctx, int64_t n, T from, T to, T *x) {                                    \
        UHIP_LAUNCH1(k_replace<T>, n, from, to, x);                                                              \
    }                                                                                                            \
    int uhip_vec_pseudo_inv_##SUF(uhip_ctx ctx, int64_t n, const T *x, T *y, T tol) {                            \
        UHIP_LAUNCH1(k_pseudo_inv<T>, n, x, y, tol);                                                             \
    }                                                                                                            \
    int uhip_vec_dots_##SUF(uhip_ctx ctx, int64_t n, int k, const T *const *xs, const T *const *ys,             \
                            double *result_d) {                                                                  \
        return launch_dots<T>(ctx, n, k, xs, ys, result_d, 0u);                                                  \
    }                                                                                                            \
    int uhip_vec_dot_##SUF(uhip_ctx ctx, int64_t n, const T *x, const T *y, double *result_d) {                 \
        return launch_dots<T>(ctx, n, 1, &x, &y, result_d, 0u);                                                  \
    }                                                                                                            \
    int uhip_vec_norm2_##SUF(uhip_ctx ctx, int64_t n, const T *x, double *result_d) {                           \
        return launch_dots<T>(ctx, n, 1, &x, &x, result_d, 1u);                                                  \
    }                                                                                                            \
    int uhip_vec_reduce_##SUF(uhip_ctx ctx, int64_t n, const T *x, int op, double *result_d) {                  \
        return launch_reduce<T>(ctx, n, x, op, result_d);                                                        \
    }

extern "C" {
UHIP_BLAS1_API(f64, double)
UHIP_BLAS1_API(f32, float)

int uhip_vec_cast_f64_f32(uhip_ctx ctx, int64_t n, const double *x, float *y) {
    UHIP_LAUNCH1((k_copy<double, float>), n, x, y);
}
int uhip_vec_cast_f32_f64(uhip_ctx ctx, int64_t n, const float *x, double *y) {
    UHIP_LAUNCH1((k_copy<float, double>), n, x, y);
}
}
