// Row-block CSR SpMV skeleton shared by every sparse kernel of the library.
//
// A plan splits the rows into blocks whose nonzeros fit in LDS. One workgroup
// per block:
//   phase 1  the block's (val, col) stream is read fully coalesced (lane t
//            takes entries e0+t, e0+t+WG, ...), x[col] is gathered through the
//            L2/MALL, and the products val*x go to LDS;
//   phase 2  lane t owns row r0+t (+WG ...) and sums its products from LDS in
//            column order, then hands (row, sum) to an epilogue that fuses the
//            per-row vector work (residual, Chebyshev step, correction, dots).
// The sequential per-row sum makes each y_i bit-identical to a textbook CSR
// loop with the products rounded before they are added.
#pragma once

#include "uhip_common.hpp"

namespace uhip {

constexpr int kSpmvWG = 256;

template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
    return __builtin_nontemporal_load(p);
}

// Epilogues: operator()(row, sum, part) with part[kDots] per-thread partial sums
// that the kernel reduces per block into partials[d*stride + block].
template <typename TAcc, int CAPK, class Epi, typename TA, typename TX>
__global__ __launch_bounds__(kSpmvWG) void k_spmv_rowblock(const int32_t *__restrict__ blocks,
                                                           const int32_t *__restrict__ rowptr,
                                                           const int32_t *__restrict__ col,
                                                           const TA *__restrict__ val, const TX *__restrict__ x,
                                                           Epi epi, double *__restrict__ partials, int64_t stride) {
    constexpr int CAP = kSpmvWG * CAPK;
    constexpr int ND = Epi::kDots > 0 ? Epi::kDots : 1;
    __shared__ TAcc prod[CAP];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int r0 = blocks[b], r1 = blocks[b + 1];
    const int e0 = rowptr[r0], e1 = rowptr[r1];
    const int tid = threadIdx.x;
    double part[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) part[d] = 0.0;

    if (e1 - e0 <= CAP) {
        int c[CAPK];
        TA v[CAPK];
#pragma unroll
        for (int k = 0; k < CAPK; ++k) {
            const int e = e0 + tid + k * kSpmvWG;
            if (e < e1) {
                c[k] = ld_stream(col + e);
                v[k] = ld_stream(val + e);
            }
        }
#pragma unroll
        for (int k = 0; k < CAPK; ++k) {
            const int e = e0 + tid + k * kSpmvWG;
            if (e < e1) prod[e - e0] = (TAcc)v[k] * (TAcc)x[c[k]];
        }
        __syncthreads();
        for (int r = r0 + tid; r < r1; r += kSpmvWG) {
            const int a = rowptr[r] - e0, z = rowptr[r + 1] - e0;
            TAcc s = TAcc(0);
            for (int j = a; j < z; ++j) s += prod[j];
            epi(r, s, part);
        }
    } else {
        // A single row longer than the LDS budget (the plan isolates such rows).
        TAcc s = TAcc(0);
        for (int e = e0 + tid; e < e1; e += kSpmvWG) s += (TAcc)val[e] * (TAcc)x[col[e]];
        s = block_sum<kSpmvWG>(s);
        if (tid == 0) epi(r0, s, part);
    }
    if constexpr (Epi::kDots > 0) {
#pragma unroll
        for (int d = 0; d < Epi::kDots; ++d) {
            const double t = block_sum<kSpmvWG>(part[d]);
            if (tid == 0) partials[d * stride + b] = t;
        }
    }
}

// Dispatch over the LDS budget (block_nnz = 4/5 * CAPK * WG). Matrix values
// have the accumulation type TAcc.
template <typename TAcc, typename TX, class Epi>
int launch_rowblock(uhip_ctx ctx, const uhip_csr *A, const TX *x, const Epi &epi) {
    using TA = TAcc;
    UHIP_REQUIRE(A && A->blocks && A->nblocks >= 0, "spmv: matrix has no row-block plan");
    if (A->nblocks == 0) return UHIP_OK;
    if (Epi::kDots > 0) UHIP_TRY(ensure_partials(ctx, A->nblocks));
    const int capk = A->block_capk;
    const TA *val = (const TA *)A->val;
    hipStream_t s = stream_of(ctx);
    dim3 g(A->nblocks), blk(kSpmvWG);
    double *parts = ctx->partials;
    const int64_t st = ctx->max_partials;
    switch (capk) {
        case 5:
            hipLaunchKernelGGL((k_spmv_rowblock<TAcc, 5, Epi, TA, TX>), g, blk, 0, s, A->blocks, A->rowptr, A->col,
                               val, x, epi, parts, st);
            break;
        case 10:
            hipLaunchKernelGGL((k_spmv_rowblock<TAcc, 10, Epi, TA, TX>), g, blk, 0, s, A->blocks, A->rowptr, A->col,
                               val, x, epi, parts, st);
            break;
        case 20:
            hipLaunchKernelGGL((k_spmv_rowblock<TAcc, 20, Epi, TA, TX>), g, blk, 0, s, A->blocks, A->rowptr, A->col,
                               val, x, epi, parts, st);
            break;
        default:
            return set_error(UHIP_ERR_INVALID, "spmv: unsupported plan capk=%d", capk);
    }
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

// ---- epilogues ------------------------------------------------------------

// y = s
template <typename TAcc, typename TY>
struct EpiStore {
    static constexpr int kDots = 0;
    TY *y;
    __device__ void operator()(int r, TAcc s, double *) const { y[r] = (TY)s; }
};

// y = alpha*s + beta*y   (beta == 0 never reads y)
template <typename TAcc, typename TY>
struct EpiAxpby {
    static constexpr int kDots = 0;
    TY *y;
    TAcc alpha, beta;
    __device__ void operator()(int r, TAcc s, double *) const {
        y[r] = (beta == TAcc(0)) ? (TY)(alpha * s) : (TY)(alpha * s + beta * (TAcc)y[r]);
    }
};

// y = alpha*s + beta*z
template <typename TAcc, typename TY>
struct EpiAxpbz {
    static constexpr int kDots = 0;
    const TY *z;
    TY *y;
    TAcc alpha, beta;
    __device__ void operator()(int r, TAcc s, double *) const { y[r] = (TY)(alpha * s + beta * (TAcc)z[r]); }
};

// r = b - s
template <typename TAcc, typename TB, typename TY>
struct EpiResidual {
    static constexpr int kDots = 0;
    const TB *b;
    TY *y;
    __device__ void operator()(int r, TAcc s, double *) const { y[r] = (TY)((TAcc)b[r] - s); }
};

// y = s, part += x_r * s  (p.Ap of CG)
template <typename TAcc, typename TX, typename TY>
struct EpiStoreDot {
    static constexpr int kDots = 1;
    const TX *x;
    TY *y;
    __device__ void operator()(int r, TAcc s, double *part) const {
        y[r] = (TY)s;
        part[0] += (double)x[r] * (double)s;
    }
};

// y = mask .* s  (restriction followed by MultiLevelMask::apply); mask may be null
template <typename TAcc, typename TY>
struct EpiMaskStore {
    static constexpr int kDots = 0;
    const TY *mask;
    TY *y;
    __device__ void operator()(int r, TAcc s, double *) const {
        y[r] = mask ? (TY)((TAcc)mask[r] * s) : (TY)s;
    }
};

// y = y + s  (coarse-grid correction c += I c_c)
template <typename TAcc, typename TY>
struct EpiAddTo {
    static constexpr int kDots = 0;
    TY *y;
    __device__ void operator()(int r, TAcc s, double *) const { y[r] = (TY)((TAcc)y[r] + s); }
};

// One step of Chebyshev3level::solve (utopia_Chebyshev3level.hpp:138-175) fused
// with the SpMV that feeds it. With s = (A p_{k-1})_r:
//   r_k = r_{k-1} + a_prev*s ; p_k = -r_k + beta*p_{k-1} ; x += a_prev*p_{k-1}
// and on the last step also x += a_cur*p_k. rprev == null means r_{k-1} = -p_{k-1}
// (first step), xin == null means x = 0 (zero initial guess).
template <typename TAcc, typename TP, typename TV, typename TXI, typename TXO>
struct EpiCheb {
    static constexpr int kDots = 0;
    const TP *pprev;
    const TV *rprev;
    const TXI *xin;
    TXO *xout;
    TV *rout;
    TV *pout;
    TAcc a_prev, beta, a_cur;
    int last;
    __device__ void operator()(int r, TAcc s, double *) const {
        const TAcc pp = (TAcc)pprev[r];
        const TAcc rp = rprev ? (TAcc)rprev[r] : -pp;
        const TAcc rk = rp + a_prev * s;
        const TAcc pk = -rk + beta * pp;
        TAcc xv = xin ? (TAcc)xin[r] : TAcc(0);
        xv = xv + a_prev * pp;
        if (last) {
            xv = xv + a_cur * pk;
        } else {
            rout[r] = (TV)rk;
            pout[r] = (TV)pk;
        }
        xout[r] = (TXO)xv;
    }
};

}  // namespace uhip
