// Batched, multi-right-hand-side kernels: many independent SPD systems stored
// as one block-diagonal CSR matrix (segments of rows), K right-hand sides per
// system stored row-major (entry (i, j) at i*K + j).
//
//   SpMM      one thread per row owns all K right-hand sides; the row's
//             values/columns are read once and the gathered x entries of a
//             column index are K contiguous values.
//   segdot    per (segment, column) dot products, two fixed-order stages.
//   batched CG / Chebyshev steps with per-(segment, column) step lengths.
#include <algorithm>

#include "uhip_common.hpp"

namespace uhip {

// A kernel of the batched family returns at once when the context's gate (a device
// counter, uhip_ctx_set_gate) reads 0: a captured run of CG iterations then costs
// only its launches once every (system, rhs) pair has converged.
#define UHIP_GATE() \
    if (gate && *gate == 0) return

// SpMM: one thread per row holds all K sums; each stored entry's value and
// column are read once and the K x values of that column are one contiguous
// K*sizeof(TX)-byte segment. EB entries' loads are issued before their
// products are added (in entry order, so each column's sum is the sequential
// CSR sum).
template <int K, typename TX>
__device__ __forceinline__ const TX *xrow(const TX *X, int32_t c) {
    const TX *p = X + (int64_t)c * K;
    if constexpr (K * sizeof(TX) >= 16) return (const TX *)__builtin_assume_aligned(p, 16);
    return p;
}

template <int K, typename TAcc, class Epi, typename TA, typename TX>
__global__ __launch_bounds__(256) void k_spmm_csr(const int32_t *__restrict__ gate, int32_t n, const int32_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ col, const TA *__restrict__ val,
                                                  const TX *__restrict__ X, Epi epi) {
    UHIP_GATE();
    constexpr int EB = K >= 16 ? 2 : 4;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int32_t r = b * 256 + (int32_t)threadIdx.x;
    if (r >= n) return;
    TAcc acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = TAcc(0);
    int32_t e = rowptr[r];
    const int32_t e1 = rowptr[r + 1];
    for (; e < e1; e += EB) {
        int32_t c[EB];
        TA v[EB];
#pragma unroll
        for (int u = 0; u < EB; ++u) {
            c[u] = e + u < e1 ? col[e + u] : 0;
            v[u] = e + u < e1 ? val[e + u] : TA(0);
        }
        TX xv[EB][K];
#pragma unroll
        for (int u = 0; u < EB; ++u) {
            const TX *xr = xrow<K>(X, c[u]);
#pragma unroll
            for (int j = 0; j < K; ++j) xv[u][j] = xr[j];
        }
#pragma unroll
        for (int u = 0; u < EB; ++u)
            if (e + u < e1) {
#pragma unroll
                for (int j = 0; j < K; ++j) acc[j] += (TAcc)v[u] * (TAcc)xv[u][j];
            }
    }
    epi(r, acc);
}

// Row r's K entries of a row-major [n x K] array, 16-byte aligned when the row
// spans at least 16 bytes (hipMalloc'd base).
template <int K, typename T>
__device__ __forceinline__ T *krow(T *base, int32_t r) {
    T *p = base + (int64_t)r * K;
    if constexpr (K * sizeof(T) >= 16) return (T *)__builtin_assume_aligned(p, 16);
    return p;
}

// Epilogues read all K inputs of the row before writing any output, so an
// output may alias an input of the same row (in-place updates).
template <int K, typename TAcc, typename TY>
struct EpiKStore {
    TY *y;
    __device__ void operator()(int32_t r, const TAcc *s) const {
        TY *yr = krow<K>(y, r);
#pragma unroll
        for (int j = 0; j < K; ++j) yr[j] = (TY)s[j];
    }
};

template <int K, typename TAcc, typename TY>
struct EpiKAdd {
    TY *y;
    __device__ void operator()(int32_t r, const TAcc *s) const {
        TY *yr = krow<K>(y, r);
        TY t[K];
#pragma unroll
        for (int j = 0; j < K; ++j) t[j] = yr[j];
#pragma unroll
        for (int j = 0; j < K; ++j) yr[j] = (TY)((TAcc)t[j] + s[j]);
    }
};

template <int K, typename TAcc, typename TB, typename TY>
struct EpiKResidual {
    const TB *b;
    TY *y;
    __device__ void operator()(int32_t r, const TAcc *s) const {
        const TB *br = krow<K>(b, r);
        TB t[K];
#pragma unroll
        for (int j = 0; j < K; ++j) t[j] = br[j];
        TY *yr = krow<K>(y, r);
#pragma unroll
        for (int j = 0; j < K; ++j) yr[j] = (TY)((TAcc)t[j] - s[j]);
    }
};

// Chebyshev step (same algebra as the single-vector EpiCheb) with per-segment
// coefficients coef[3*seg + {0,1,2}] = (a_prev, beta, a_cur).
template <int K, typename TAcc, typename TP, typename TXO>
struct EpiKCheb {
    const TP *pprev;
    const TAcc *rprev;
    const TAcc *xin;
    TXO *xout;
    TAcc *rout;
    TAcc *pout;
    const int32_t *rowseg;
    const double *coef;
    int last;
    __device__ void operator()(int32_t r, const TAcc *s) const {
        const int32_t g = rowseg[r];
        const TAcc a_prev = (TAcc)coef[3 * g], beta = (TAcc)coef[3 * g + 1], a_cur = (TAcc)coef[3 * g + 2];
        TAcc pp[K], rp[K], xv[K];
        const TP *ppr = krow<K>(pprev, r);
#pragma unroll
        for (int j = 0; j < K; ++j) pp[j] = (TAcc)ppr[j];
        if (rprev) {
            const TAcc *rpr = krow<K>(rprev, r);
#pragma unroll
            for (int j = 0; j < K; ++j) rp[j] = rpr[j];
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j) rp[j] = -pp[j];
        }
        if (xin) {
            const TAcc *xr = krow<K>(xin, r);
#pragma unroll
            for (int j = 0; j < K; ++j) xv[j] = xr[j];
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j) xv[j] = TAcc(0);
        }
        TAcc rk[K], pk[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            rk[j] = rp[j] + a_prev * s[j];
            pk[j] = -rk[j] + beta * pp[j];
            xv[j] = xv[j] + a_prev * pp[j];
        }
        if (last) {
#pragma unroll
            for (int j = 0; j < K; ++j) xv[j] = xv[j] + a_cur * pk[j];
        } else {
            TAcc *ro = krow<K>(rout, r), *po = krow<K>(pout, r);
#pragma unroll
            for (int j = 0; j < K; ++j) ro[j] = rk[j];
#pragma unroll
            for (int j = 0; j < K; ++j) po[j] = pk[j];
        }
        TXO *xo = krow<K>(xout, r);
#pragma unroll
        for (int j = 0; j < K; ++j) xo[j] = (TXO)xv[j];
    }
};

// Long rows (dense coarse blocks): one wavefront per row, lane l takes entries
// l, l+64, ... in order, and a fixed butterfly combines the lanes.
template <int K, typename TAcc, class Epi, typename TA, typename TX>
__global__ __launch_bounds__(256) void k_spmm_wave(const int32_t *__restrict__ gate, int32_t n, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col, const TA *__restrict__ val,
                                                   const TX *__restrict__ X, Epi epi) {
    UHIP_GATE();
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int32_t r = b * 4 + (int32_t)(threadIdx.x >> 6);
    const int l = threadIdx.x & 63;
    if (r >= n) return;
    TAcc acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = TAcc(0);
    const int32_t e1 = rowptr[r + 1];
    for (int32_t e = rowptr[r] + l; e < e1; e += 64) {
        const TAcc v = (TAcc)val[e];
        const TX *xr = xrow<K>(X, col[e]);
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] += v * (TAcc)xr[j];
    }
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = wave_sum(acc[j]);
    if (l == 0) epi(r, acc);
}

// Rows averaging >= 32 entries take the wavefront-per-row kernel (its sums are
// tree-ordered); all others the row-per-thread kernel (sequential CSR order),
// unless the matrix carries an autotuned choice (spmm_kernel 1 or 2).
template <int K, typename TAcc, class Epi, typename TA, typename TX>
int launch_spmm(uhip_ctx ctx, const uhip_csr *A, const TX *X, const Epi &e) {
    const int32_t n = A->nrows;
    if (n == 0) return UHIP_OK;
    const bool wave = A->spmm_kernel == 2 || (A->spmm_kernel != 1 && A->nnz >= 32LL * n);
    if (wave)
        hipLaunchKernelGGL((k_spmm_wave<K, TAcc, Epi, TA, TX>), dim3((n + 3) / 4), dim3(256), 0, ctx->stream,
                           ctx->gate, n, A->rowptr, A->col, (const TA *)A->val, X, e);
    else
        hipLaunchKernelGGL((k_spmm_csr<K, TAcc, Epi, TA, TX>), dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                           ctx->gate, n, A->rowptr, A->col, (const TA *)A->val, X, e);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

// ---- segmented dots: out[s*K + j] = sum_{i in seg s} x[i,j] * y[i,j] ----
// Chunks of rows never straddle a segment; chunk c covers [c_row0[c], c_row1[c]).
template <int K, typename TX, typename TY>
__global__ __launch_bounds__(256) void k_segdot_chunk(const int32_t *__restrict__ gate, const int32_t *__restrict__ c_row0,
                                                      const int32_t *__restrict__ c_row1, const TX *__restrict__ x,
                                                      const TY *__restrict__ y, double *__restrict__ partials) {
    UHIP_GATE();
    constexpr int RPB = 256 / K;
    __shared__ double red[256];
    const int c = blockIdx.x;
    const int32_t r0 = c_row0[c], r1 = c_row1[c];
    const int j = threadIdx.x % K, ro = threadIdx.x / K;
    double acc = 0.0;
    for (int32_t r = r0 + ro; r < r1; r += RPB) {
        const int64_t i = (int64_t)r * K + j;
        acc += (double)x[i] * (double)y[i];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = RPB / 2; s > 0; s >>= 1) {  // fixed tree over row offsets, per column
        if (ro < s) red[threadIdx.x] += red[threadIdx.x + s * K];
        __syncthreads();
    }
    if (ro == 0) partials[(int64_t)c * K + j] = red[j];
}

template <int K>
// One wavefront per segment: lane l sums the segment's chunks l, l+64, ... in
// order, then a fixed butterfly combines the lanes (deterministic).
__global__ __launch_bounds__(64) void k_segdot_finish(const int32_t *__restrict__ gate, int32_t nseg, const int32_t *__restrict__ seg_c0,
                                                      const double *__restrict__ partials, double *__restrict__ out,
                                                      int take_sqrt) {
    UHIP_GATE();
    const int s = blockIdx.x;
    const int l = threadIdx.x;
    const int32_t c0 = seg_c0[s], c1 = seg_c0[s + 1];
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = 0.0;
    for (int32_t c = c0 + l; c < c1; c += 64) {
#pragma unroll
        for (int j = 0; j < K; ++j) acc[j] += partials[(int64_t)c * K + j];
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const double t = wave_sum(acc[j]);
        if (l == 0) out[(int64_t)s * K + j] = take_sqrt ? sqrt(t) : t;
    }
}

// ---- batched CG on per-(segment, column) scalars ----
// state: rz, pap, alpha, beta, rznew, rr, atol, flags (0 active, 1 converged, 2 stopped on a
// non-finite step), conv_it.
struct BCGState {
    double *rz, *pap, *alpha, *beta, *rznew, *rr, *atol;
    int32_t *flags, *conv_it;
};

__global__ __launch_bounds__(256) void k_bcg_alpha(const int32_t *__restrict__ gate, int32_t m, BCGState st) {
    UHIP_GATE();
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= m) return;
    double a = st.rz[t] / st.pap[t];
    if (st.flags[t] != 0) {
        a = 0.0;
    } else if (!isfinite(a)) {  // utopia_ConjugateGradient_impl.hpp:241-244: stop without updating x
        st.flags[t] = 2;
        a = 0.0;
    }
    st.alpha[t] = a;
}

template <int K>
__global__ __launch_bounds__(256) void k_bcg_update(const int32_t *__restrict__ gate, int32_t n, const int32_t *__restrict__ rowseg,
                                                    const double *__restrict__ alpha, double *__restrict__ x,
                                                    const double *__restrict__ p, double *__restrict__ r,
                                                    const double *__restrict__ Ap) {
    UHIP_GATE();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)n * K) return;
    const int32_t row = (int32_t)(i / K);
    const double a = alpha[rowseg[row] * K + (int)(i % K)];
    if (a == 0.0) return;
    x[i] = x[i] + a * p[i];
    r[i] = r[i] - a * Ap[i];
}

// after |r| is known: converge (flag 1) where |r| <= atol; count the still active ones
// iter (device) counts the iterations done; it is advanced here so that a
// captured iteration replays with the right count.
__global__ __launch_bounds__(256) void k_bcg_check(const int32_t *__restrict__ gate, int32_t m, BCGState st, int32_t *__restrict__ iter,
                                                   int32_t *__restrict__ active_count) {
    UHIP_GATE();
    __shared__ int32_t cnt[256];
    const int32_t it = *iter + 1;
    int32_t c = 0;
    for (int t = threadIdx.x; t < m; t += 256) {
        if (st.flags[t] == 0) {
            if (st.rr[t] <= st.atol[t]) {
                st.flags[t] = 1;
                st.conv_it[t] = it;
            } else {
                ++c;
            }
        }
    }
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) cnt[threadIdx.x] += cnt[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *active_count = cnt[0];
        *iter = it;
    }
}

__global__ __launch_bounds__(256) void k_bcg_beta(const int32_t *__restrict__ gate, int32_t m, BCGState st) {
    UHIP_GATE();
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= m) return;
    st.beta[t] = (st.flags[t] == 0) ? st.rznew[t] / st.rz[t] : 0.0;
    st.rz[t] = st.rznew[t];
}

template <int K>
__global__ __launch_bounds__(256) void k_bcg_pupdate(const int32_t *__restrict__ gate, int32_t n, const int32_t *__restrict__ rowseg,
                                                     const double *__restrict__ beta, double *__restrict__ p,
                                                     const double *__restrict__ z) {
    UHIP_GATE();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)n * K) return;
    const int32_t row = (int32_t)(i / K);
    p[i] = z[i] + beta[rowseg[row] * K + (int)(i % K)] * p[i];
}

// x_i = xin_i + a[seg]*p_i (Chebyshev with one sweep), dtype conversions
template <int K, typename TP, typename TX, typename TO, typename TAcc>
__global__ __launch_bounds__(256) void k_axpy_seg(const int32_t *__restrict__ gate, int32_t n, const int32_t *__restrict__ rowseg,
                                                  const double *__restrict__ coef, const TP *__restrict__ p,
                                                  const TX *__restrict__ xin, TO *__restrict__ out) {
    UHIP_GATE();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)n * K) return;
    const TAcc a = (TAcc)coef[3 * rowseg[i / K]];
    TAcc v = xin ? (TAcc)xin[i] : TAcc(0);
    out[i] = (TO)(v + a * (TAcc)p[i]);
}

// per segment power-method helpers (single column)
__global__ __launch_bounds__(256) void k_seg_scale(int32_t n, const int32_t *__restrict__ rowseg,
                                                   const double *__restrict__ norm, const int32_t *__restrict__ done,
                                                   const double *__restrict__ w, double *__restrict__ v) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int32_t s = rowseg[i];
    if (done && done[s]) return;
    v[i] = (1. / norm[s]) * w[i];
}

// grad_e = -sum_j lambda[i, j] * x[col_e, j] over the pattern (dL/dA_ij = -lambda_i x_j)
template <int K>
__global__ __launch_bounds__(256) void k_pattern_grad(int32_t n, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ lam, const double *__restrict__ x,
                                                      double *__restrict__ grad) {
    const int32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    double l[K];
#pragma unroll
    for (int j = 0; j < K; ++j) l[j] = lam[(int64_t)r * K + j];
    for (int32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
        const int64_t c = (int64_t)col[e] * K;
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) s += l[j] * x[c + j];
        grad[e] = -s;
    }
}

}  // namespace uhip

using namespace uhip;

// ---- dispatch helpers (K in {1, 2, 4, 8, 16}) ----
#define UHIP_K_SWITCH(k, ...)                                         \
    switch (k) {                                                      \
        case 1: { constexpr int KK = 1; __VA_ARGS__; } break;         \
        case 2: { constexpr int KK = 2; __VA_ARGS__; } break;         \
        case 4: { constexpr int KK = 4; __VA_ARGS__; } break;         \
        case 8: { constexpr int KK = 8; __VA_ARGS__; } break;         \
        case 16: { constexpr int KK = 16; __VA_ARGS__; } break;       \
        default: return set_error(UHIP_ERR_INVALID, "K=%d not in {1,2,4,8,16}", k); \
    }

extern "C" {

int uhip_spmm(uhip_ctx ctx, const uhip_csr *A, int k, const void *X, int x_dtype, void *Y, int y_dtype, int op,
              const void *B, int b_dtype) {
    UHIP_REQUIRE(ctx && A && X && Y, "uhip_spmm: null argument");
    UHIP_REQUIRE(op >= 0 && op <= 2, "uhip_spmm: op %d", op);
    UHIP_REQUIRE(op != 2 || B, "uhip_spmm: residual needs B");
    // op 0: Y = A X; 1: Y += A X; 2: Y = B - A X. Accumulation in the matrix precision.
#define UHIP_SPMM_T(TA, TX, TY, TB)                                                                          \
    do {                                                                                                     \
        if (op == 0) {                                                                                       \
            UHIP_K_SWITCH(k, UHIP_TRY((launch_spmm<KK, TA, EpiKStore<KK, TA, TY>, TA, TX>(                 \
                                 ctx, A, (const TX *)X, EpiKStore<KK, TA, TY>{(TY *)Y}))));                  \
        } else if (op == 1) {                                                                                \
            UHIP_K_SWITCH(k, UHIP_TRY((launch_spmm<KK, TA, EpiKAdd<KK, TA, TY>, TA, TX>(                   \
                                 ctx, A, (const TX *)X, EpiKAdd<KK, TA, TY>{(TY *)Y}))));                    \
        } else {                                                                                             \
            UHIP_K_SWITCH(k, UHIP_TRY((launch_spmm<KK, TA, EpiKResidual<KK, TA, TB, TY>, TA, TX>(          \
                                 ctx, A, (const TX *)X, EpiKResidual<KK, TA, TB, TY>{(const TB *)B, (TY *)Y})))); \
        }                                                                                                    \
        return UHIP_OK;                                                                                      \
    } while (0)
    if (A->dtype == UHIP_F64) {
        if (x_dtype == UHIP_F64 && y_dtype == UHIP_F64) UHIP_SPMM_T(double, double, double, double);
        if (x_dtype == UHIP_F32 && y_dtype == UHIP_F32) UHIP_SPMM_T(double, float, float, double);
        if (x_dtype == UHIP_F32 && y_dtype == UHIP_F64) UHIP_SPMM_T(double, float, double, double);
        if (x_dtype == UHIP_F64 && y_dtype == UHIP_F32) UHIP_SPMM_T(double, double, float, double);
    } else {
        UHIP_REQUIRE(y_dtype == UHIP_F32, "uhip_spmm: FP32 matrix writes FP32");
        if (x_dtype == UHIP_F32) {
            if (op != 2 || b_dtype == UHIP_F32) UHIP_SPMM_T(float, float, float, float);
            UHIP_SPMM_T(float, float, float, double);
        }
    }
#undef UHIP_SPMM_T
    return set_error(UHIP_ERR_INVALID, "uhip_spmm: dtype combination");
}

int uhip_csr_autotune_spmm(uhip_ctx ctx, uhip_csr *A, int k, int reps, int32_t *choice, float *ms_out) {
    UHIP_REQUIRE(ctx && A && reps > 0, "uhip_csr_autotune_spmm: bad argument");
    UHIP_REQUIRE(k == 1 || k == 2 || k == 4 || k == 8 || k == 16, "uhip_csr_autotune_spmm: K=%d", k);
    const size_t vb = A->dtype == UHIP_F64 ? sizeof(double) : sizeof(float);
    void *x = nullptr, *y = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&x, vb * (size_t)k * (size_t)std::max(A->ncols, 1)));
    UHIP_CHECK_HIP(hipMalloc(&y, vb * (size_t)k * (size_t)std::max(A->nrows, 1)));
    UHIP_CHECK_HIP(hipMemsetAsync(x, 0, vb * (size_t)k * (size_t)std::max(A->ncols, 1), ctx->stream));
    hipEvent_t e0, e1;
    UHIP_CHECK_HIP(hipEventCreate(&e0));
    UHIP_CHECK_HIP(hipEventCreate(&e1));
    const int32_t saved = A->spmm_kernel;
    float best = 1e30f;
    int32_t win = 1;
    int st = UHIP_OK;
    const int dt = A->dtype;
    for (int32_t cand = 1; cand <= 2 && st == UHIP_OK; ++cand) {
        A->spmm_kernel = cand;
        if ((st = uhip_spmm(ctx, A, k, x, dt, y, dt, 0, nullptr, 0)) != UHIP_OK) break;  // warm-up
        hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && st == UHIP_OK; ++r) st = uhip_spmm(ctx, A, k, x, dt, y, dt, 0, nullptr, 0);
        hipEventRecord(e1, ctx->stream);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)reps;
        if (ms < best) best = ms, win = cand;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamSynchronize(ctx->stream);
    hipFree(x);
    hipFree(y);
    if (st != UHIP_OK) {
        A->spmm_kernel = saved;
        return st;
    }
    A->spmm_kernel = win;
    if (choice) *choice = win;
    if (ms_out) *ms_out = best;
    return UHIP_OK;
}

int uhip_spmm_cheb_step(uhip_ctx ctx, const uhip_csr *A, int k, const void *pprev, int p_dtype, const void *rprev,
                        const void *xin, void *xout, int x_dtype, void *rout, void *pout, const int32_t *rowseg,
                        const double *coef, int last) {
    UHIP_REQUIRE(ctx && A && pprev && xout && rowseg && coef, "uhip_spmm_cheb_step: null argument");
    UHIP_REQUIRE(last || (rout && pout), "uhip_spmm_cheb_step: rout/pout needed");
#define UHIP_SPMM_CHEB(TA, TP, TX)                                                                              \
    do {                                                                                                        \
        UHIP_K_SWITCH(k,                                                                                      \
                          EpiKCheb<KK, TA, TP, TX> e{(const TP *)pprev, (const TA *)rprev, (const TA *)xin,     \
                                                     (TX *)xout,         (TA *)rout,         (TA *)pout,         \
                                                     rowseg,             coef,               last};              \
                          UHIP_TRY((launch_spmm<KK, TA, EpiKCheb<KK, TA, TP, TX>, TA, TP>(ctx, A, (const TP *)pprev, \
                                                                                           e)));                \
                      );                                                                                      \
        return UHIP_OK;                                                                                         \
    } while (0)
    if (A->dtype == UHIP_F64) {
        UHIP_REQUIRE(p_dtype == UHIP_F64 && x_dtype == UHIP_F64, "uhip_spmm_cheb_step: FP64 level");
        UHIP_SPMM_CHEB(double, double, double);
    }
    if (p_dtype == UHIP_F32 && x_dtype == UHIP_F32) UHIP_SPMM_CHEB(float, float, float);
    if (p_dtype == UHIP_F64 && x_dtype == UHIP_F32) UHIP_SPMM_CHEB(float, double, float);
    if (p_dtype == UHIP_F32 && x_dtype == UHIP_F64) UHIP_SPMM_CHEB(float, float, double);
    if (p_dtype == UHIP_F64 && x_dtype == UHIP_F64) UHIP_SPMM_CHEB(float, double, double);
#undef UHIP_SPMM_CHEB
    return set_error(UHIP_ERR_INVALID, "uhip_spmm_cheb_step: dtypes");
}

int uhip_axpy_seg(uhip_ctx ctx, int32_t n, int k, const int32_t *rowseg, const double *coef, const void *p,
                  int p_dtype, const void *xin, void *out, int out_dtype, int level_dtype) {
    UHIP_REQUIRE(ctx && rowseg && coef && p && out, "uhip_axpy_seg: null argument");
    if (n == 0) return UHIP_OK;
    const int64_t m = (int64_t)n * k;
    dim3 g((unsigned)((m + 255) / 256)), b(256);
#define UHIP_AXS(TP, TX, TO, TA)                                                                              \
    UHIP_K_SWITCH(k, hipLaunchKernelGGL((k_axpy_seg<KK, TP, TX, TO, TA>), g, b, 0, ctx->stream, ctx->gate, n, rowseg, coef, \
                                         (const TP *)p, (const TX *)xin, (TO *)out))
    if (level_dtype == UHIP_F64) {
        UHIP_AXS(double, double, double, double);
    } else if (p_dtype == UHIP_F64 && out_dtype == UHIP_F32) {
        UHIP_AXS(double, float, float, float);
    } else if (p_dtype == UHIP_F32 && out_dtype == UHIP_F64) {
        UHIP_AXS(float, float, double, float);
    } else if (p_dtype == UHIP_F64 && out_dtype == UHIP_F64) {
        UHIP_AXS(double, float, double, float);
    } else {
        UHIP_AXS(float, float, float, float);
    }
#undef UHIP_AXS
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_segdot(uhip_ctx ctx, int k, int32_t nchunks, const int32_t *c_row0, const int32_t *c_row1, int32_t nseg,
                const int32_t *seg_c0, const void *x, const void *y, int dtype, double *out, int take_sqrt) {
    UHIP_REQUIRE(ctx && c_row0 && c_row1 && seg_c0 && x && y && out, "uhip_segdot: null argument");
    if (nchunks == 0) return UHIP_OK;
    UHIP_TRY(ensure_partials(ctx, (int64_t)nchunks * k));
    if (dtype == UHIP_F64) {
        UHIP_K_SWITCH(k, hipLaunchKernelGGL((k_segdot_chunk<KK, double, double>), dim3(nchunks), dim3(256), 0,
                                             ctx->stream, ctx->gate, c_row0, c_row1, (const double *)x, (const double *)y,
                                             ctx->partials));
    } else {
        UHIP_K_SWITCH(k, hipLaunchKernelGGL((k_segdot_chunk<KK, float, float>), dim3(nchunks), dim3(256), 0,
                                             ctx->stream, ctx->gate, c_row0, c_row1, (const float *)x, (const float *)y,
                                             ctx->partials));
    }
    UHIP_CHECK_LAUNCH();
    const int m = nseg * k;
    UHIP_K_SWITCH(k, hipLaunchKernelGGL((k_segdot_finish<KK>), dim3(nseg), dim3(64), 0, ctx->stream, ctx->gate, nseg, seg_c0,
                                         ctx->partials, out, take_sqrt));
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_bcg_step(uhip_ctx ctx, int stage, int32_t n, int k, int32_t m, const int32_t *rowseg, double *x, double *p,
                  double *r, const double *Ap, const double *z, double *const *state_d, int32_t *flags,
                  int32_t *conv_it, int32_t *iter, int32_t *active_count) {
    // state_d: rz, pap, alpha, beta, rznew, rr, atol
    UHIP_REQUIRE(ctx && state_d && flags && conv_it && iter && active_count, "uhip_bcg_step: null argument");
    BCGState st{state_d[0], state_d[1], state_d[2], state_d[3], state_d[4], state_d[5], state_d[6], flags, conv_it};
    const int64_t nk = (int64_t)n * k;
    dim3 gm((m + 255) / 256), gv((unsigned)((nk + 255) / 256)), b(256);
    switch (stage) {
        case 0:  // alpha and x, r update
            hipLaunchKernelGGL(k_bcg_alpha, gm, b, 0, ctx->stream, ctx->gate, m, st);
            UHIP_CHECK_LAUNCH();
            if (nk > 0) {
                UHIP_K_SWITCH(k, hipLaunchKernelGGL((k_bcg_update<KK>), gv, b, 0, ctx->stream, ctx->gate, n, rowseg, st.alpha, x,
                                                     (const double *)p, r, Ap));
                UHIP_CHECK_LAUNCH();
            }
            return UHIP_OK;
        case 1:  // convergence check (rr holds |r|)
            hipLaunchKernelGGL(k_bcg_check, dim3(1), b, 0, ctx->stream, ctx->gate, m, st, iter, active_count);
            UHIP_CHECK_LAUNCH();
            return UHIP_OK;
        case 2:  // beta and p update
            hipLaunchKernelGGL(k_bcg_beta, gm, b, 0, ctx->stream, ctx->gate, m, st);
            UHIP_CHECK_LAUNCH();
            if (nk > 0) {
                UHIP_K_SWITCH(k, hipLaunchKernelGGL((k_bcg_pupdate<KK>), gv, b, 0, ctx->stream, ctx->gate, n, rowseg, st.beta, p,
                                                     z));
                UHIP_CHECK_LAUNCH();
            }
            return UHIP_OK;
        default:
            return set_error(UHIP_ERR_INVALID, "uhip_bcg_step: stage %d", stage);
    }
}

int uhip_seg_scale(uhip_ctx ctx, int32_t n, const int32_t *rowseg, const double *norm, const int32_t *done,
                   const double *w, double *v) {
    UHIP_REQUIRE(ctx && rowseg && norm && w && v, "uhip_seg_scale: null argument");
    if (n == 0) return UHIP_OK;
    hipLaunchKernelGGL(k_seg_scale, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, rowseg, norm, done, w, v);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_pattern_grad(uhip_ctx ctx, const uhip_csr *A, int k, const double *lambda, const double *x, double *grad) {
    UHIP_REQUIRE(ctx && A && lambda && x && grad, "uhip_pattern_grad: null argument");
    const int32_t n = A->nrows;
    if (n == 0) return UHIP_OK;
    UHIP_K_SWITCH(k, hipLaunchKernelGGL((k_pattern_grad<KK>), dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n,
                                         A->rowptr, A->col, lambda, x, grad));
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

}  // extern "C"
