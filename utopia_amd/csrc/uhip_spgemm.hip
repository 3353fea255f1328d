// Setup-time sparse products: transpose, C = A*B and the Galerkin triple
// product Ac = P^T A P. All pattern work is integer-exact and the value of
// each C_ij is summed in the order a row-wise (Gustavson) product visits it:
// over k in A's row order, so results are deterministic and equal to a
// sequential dense-accumulator loop.
#include <hipcub/hipcub.hpp>

#include "uhip_common.hpp"

namespace uhip {

__global__ __launch_bounds__(kBlock) void k_col_hist(int64_t nnz, const int32_t *__restrict__ col,
                                                     int32_t *__restrict__ cnt) {
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kBlock)
        atomicAdd(&cnt[col[e]], 1);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_transpose_keys(int32_t n, const int32_t *__restrict__ rowptr,
                                                           const int32_t *__restrict__ col,
                                                           const T *__restrict__ val,
                                                           uint64_t *__restrict__ keys, T *__restrict__ vals) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    for (int32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) {
        keys[e] = ((uint64_t)(uint32_t)col[e] << 32) | (uint32_t)r;
        vals[e] = val[e];
    }
}

__global__ __launch_bounds__(kBlock) void k_keys_low(int64_t nnz, const uint64_t *__restrict__ keys,
                                                     int32_t *__restrict__ col) {
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kBlock)
        col[e] = (int32_t)(keys[e] & 0xffffffffu);
}

// Upper bound of row i of A*B: sum of |B_k| over k in A_i.
__global__ __launch_bounds__(kBlock) void k_row_ub(int32_t n, const int32_t *__restrict__ arp,
                                                   const int32_t *__restrict__ acol,
                                                   const int32_t *__restrict__ brp, int32_t *__restrict__ ub) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    int64_t s = 0;
    for (int32_t e = arp[r]; e < arp[r + 1]; ++e) {
        const int32_t k = acol[e];
        s += brp[k + 1] - brp[k];
    }
    ub[r] = (int32_t)(s > 0x7fffffff ? 0x7fffffff : s);
}

// One row of C = A*B per block iteration. PASS 0 writes the number of distinct
// columns to cnt[i]; PASS 1 writes sorted columns and values at crp[i].
// Candidates get key (col << 32 | position) so equal columns stay in visiting
// order through the (unstable) bitonic sort.
template <int NT, int CAP, int PASS>
__global__ __launch_bounds__(NT) void k_spgemm_rows(int32_t n, const int32_t *__restrict__ arp,
                                                    const int32_t *__restrict__ acol,
                                                    const double *__restrict__ aval,
                                                    const int32_t *__restrict__ brp,
                                                    const int32_t *__restrict__ bcol,
                                                    const double *__restrict__ bval, int32_t *__restrict__ cnt,
                                                    const int32_t *__restrict__ crp, int32_t *__restrict__ ccol,
                                                    double *__restrict__ cval) {
    __shared__ uint64_t key[CAP];
    __shared__ double val[CAP];
    __shared__ int32_t off[NT + 1];
    __shared__ int32_t wsum[NT / 64 + 1];
    const int tid = threadIdx.x;
    for (int32_t row = blockIdx.x; row < n; row += gridDim.x) {
        const int32_t a0 = arp[row], a1 = arp[row + 1];
        int32_t m = 0;  // candidates so far
        for (int32_t c0 = a0; c0 < a1; c0 += NT) {
            const int32_t t = c0 + tid;
            int32_t k = 0, len = 0;
            double av = 0.0;
            if (t < a1) {
                k = acol[t];
                av = aval[t];
                len = brp[k + 1] - brp[k];
            }
            // block exclusive scan of len
            int32_t incl = len;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int32_t y = __shfl_up(incl, d, 64);
                if ((tid & 63) >= d) incl += y;
            }
            if ((tid & 63) == 63) wsum[tid >> 6] = incl;
            __syncthreads();
            int32_t wbase = 0;
            for (int w = 0; w < (tid >> 6); ++w) wbase += wsum[w];
            const int32_t excl = wbase + incl - len;
            int32_t total = 0;
            for (int w = 0; w < NT / 64; ++w) total += wsum[w];
            if (t < a1) {
                const int32_t b0 = brp[k];
                for (int32_t q = 0; q < len; ++q) {
                    const int32_t p = m + excl + q;
                    key[p] = ((uint64_t)(uint32_t)bcol[b0 + q] << 32) | (uint32_t)p;
                    if (PASS == 1) val[p] = av * bval[b0 + q];
                }
            }
            m += total;
            __syncthreads();
        }
        // pad to a power of two and sort
        int32_t P = 1;
        while (P < m) P <<= 1;
        for (int32_t p = m + tid; p < P; p += NT) key[p] = ~0ull;
        __syncthreads();
        for (int32_t size = 2; size <= P; size <<= 1) {
            for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (int32_t i = tid; i < P; i += NT) {
                    const int32_t j = i ^ stride;
                    if (j > i) {
                        const bool up = (i & size) == 0;
                        const uint64_t ki = key[i], kj = key[j];
                        if ((ki > kj) == up) {
                            key[i] = kj;
                            key[j] = ki;
                            if (PASS == 1) {
                                const double vi = val[i];
                                val[i] = val[j];
                                val[j] = vi;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        // run heads: chunk [tid*per, (tid+1)*per) per thread, then a block scan of head counts
        const int32_t per = (m + NT - 1) / NT;
        const int32_t s0 = tid * per, s1 = min(m, s0 + per);
        int32_t heads = 0;
        for (int32_t i = s0; i < s1; ++i)
            if (i == 0 || (key[i] >> 32) != (key[i - 1] >> 32)) ++heads;
        int32_t incl = heads;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t y = __shfl_up(incl, d, 64);
            if ((tid & 63) >= d) incl += y;
        }
        if ((tid & 63) == 63) wsum[tid >> 6] = incl;
        __syncthreads();
        int32_t wbase = 0;
        for (int w = 0; w < (tid >> 6); ++w) wbase += wsum[w];
        int32_t total = 0;
        for (int w = 0; w < NT / 64; ++w) total += wsum[w];
        if (PASS == 0) {
            if (tid == 0) cnt[row] = total;
        } else {
            int32_t rank = wbase + incl - heads;
            const int32_t base = crp[row];
            for (int32_t i = s0; i < s1; ++i) {
                if (i == 0 || (key[i] >> 32) != (key[i - 1] >> 32)) {
                    const uint64_t c = key[i] >> 32;
                    double s = val[i];
                    for (int32_t j = i + 1; j < m && (key[j] >> 32) == c; ++j) s += val[j];
                    ccol[base + rank] = (int32_t)c;
                    cval[base + rank] = s;
                    ++rank;
                }
            }
        }
        __syncthreads();
    }
}

template <int NT, int CAP>
static int spgemm_passes(uhip_ctx ctx, const uhip_csr *A, const uhip_csr *B, uhip_csr *C) {
    hipStream_t s = ctx->stream;
    const int32_t n = A->nrows;
    int32_t *cnt = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&cnt, sizeof(int32_t) * ((size_t)n + 1)));
    const int grid = (int)std::min<int64_t>(n, 1 << 16);
    hipLaunchKernelGGL((k_spgemm_rows<NT, CAP, 0>), dim3(grid), dim3(NT), 0, s, n, A->rowptr, A->col,
                       (const double *)A->val, B->rowptr, B->col, (const double *)B->val, cnt, nullptr, nullptr,
                       nullptr);
    UHIP_CHECK_LAUNCH();
    UHIP_CHECK_HIP(hipMalloc(&C->rowptr, sizeof(int32_t) * ((size_t)n + 1)));
    void *tmp = nullptr;
    size_t tb = 0;
    UHIP_CHECK_HIP(hipMemsetAsync(cnt + n, 0, sizeof(int32_t), s));
    UHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, C->rowptr, n + 1, s));
    UHIP_CHECK_HIP(hipMalloc(&tmp, tb));
    UHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, C->rowptr, n + 1, s));
    int32_t nnz = 0;
    UHIP_CHECK_HIP(hipMemcpyAsync(&nnz, C->rowptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    C->nnz = nnz;
    UHIP_CHECK_HIP(hipMalloc(&C->col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    UHIP_CHECK_HIP(hipMalloc(&C->val, sizeof(double) * std::max<int64_t>(nnz, 1)));
    hipLaunchKernelGGL((k_spgemm_rows<NT, CAP, 1>), dim3(grid), dim3(NT), 0, s, n, A->rowptr, A->col,
                       (const double *)A->val, B->rowptr, B->col, (const double *)B->val, nullptr, C->rowptr,
                       C->col, (double *)C->val);
    UHIP_CHECK_LAUNCH();
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    hipFree(cnt);
    hipFree(tmp);
    return UHIP_OK;
}

__global__ __launch_bounds__(kBlock) void k_to_dense(int32_t n, int32_t ncols, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col,
                                                     const double *__restrict__ val, double *__restrict__ dense) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    for (int32_t e = rowptr[r]; e < rowptr[r + 1]; ++e) dense[(int64_t)r * ncols + col[e]] += val[e];
}

}  // namespace uhip

using namespace uhip;

extern "C" {

int uhip_csr_free(uhip_ctx ctx, uhip_csr *A) {
    if (!A) return UHIP_OK;
    if (ctx) UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (A->rowptr) hipFree(A->rowptr);
    if (A->col) hipFree(A->col);
    if (A->val) hipFree(A->val);
    A->rowptr = nullptr;
    A->col = nullptr;
    A->val = nullptr;
    return uhip_csr_plan_destroy(ctx, A);
}

}  // extern "C"

template <typename T>
static int transpose_typed(uhip_ctx ctx, const uhip_csr *A, uhip_csr *B) {
    hipStream_t s = ctx->stream;
    const int64_t nnz = A->nnz;
    const int32_t n = A->nrows, m = A->ncols;
    memset(B, 0, sizeof(*B));
    B->nrows = m;
    B->ncols = n;
    B->nnz = nnz;
    B->dtype = A->dtype;
    UHIP_CHECK_HIP(hipMalloc(&B->rowptr, sizeof(int32_t) * ((size_t)m + 1)));
    UHIP_CHECK_HIP(hipMalloc(&B->col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    UHIP_CHECK_HIP(hipMalloc(&B->val, sizeof(T) * std::max<int64_t>(nnz, 1)));
    int32_t *cnt = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&cnt, sizeof(int32_t) * ((size_t)m + 1)));
    UHIP_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * ((size_t)m + 1), s));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_col_hist, dim3(grid_for(nnz, 4, 1 << 16)), dim3(kBlock), 0, s, nnz, A->col, cnt);
        UHIP_CHECK_LAUNCH();
    }
    void *tmp = nullptr;
    size_t tb = 0;
    UHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, B->rowptr, m + 1, s));
    UHIP_CHECK_HIP(hipMalloc(&tmp, tb));
    UHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, B->rowptr, m + 1, s));
    hipFree(tmp);
    tmp = nullptr;
    if (nnz > 0) {
        uint64_t *k0 = nullptr, *k1 = nullptr;
        T *v0 = nullptr;
        UHIP_CHECK_HIP(hipMalloc(&k0, sizeof(uint64_t) * nnz));
        UHIP_CHECK_HIP(hipMalloc(&k1, sizeof(uint64_t) * nnz));
        UHIP_CHECK_HIP(hipMalloc(&v0, sizeof(T) * nnz));
        hipLaunchKernelGGL(k_transpose_keys<T>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                           A->rowptr, A->col, (const T *)A->val, k0, v0);
        UHIP_CHECK_LAUNCH();
        int hi_bits = 1;
        while ((1ll << hi_bits) < (int64_t)m) ++hi_bits;
        const int end_bit = 32 + hi_bits;
        UHIP_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, (T *)B->val, (int)nnz, 0,
                                                          end_bit, s));
        UHIP_CHECK_HIP(hipMalloc(&tmp, tb));
        UHIP_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, (T *)B->val, (int)nnz, 0,
                                                          end_bit, s));
        hipLaunchKernelGGL(k_keys_low, dim3(grid_for(nnz, 4, 1 << 16)), dim3(kBlock), 0, s, nnz, k1, B->col);
        UHIP_CHECK_LAUNCH();
        UHIP_CHECK_HIP(hipStreamSynchronize(s));
        hipFree(k0);
        hipFree(k1);
        hipFree(v0);
        hipFree(tmp);
    }
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    hipFree(cnt);
    return UHIP_OK;
}

extern "C" {

int uhip_csr_transpose(uhip_ctx ctx, const uhip_csr *A, uhip_csr *B) {
    UHIP_REQUIRE(ctx && A && B, "uhip_csr_transpose: null argument");
    if (A->dtype == UHIP_F64) return transpose_typed<double>(ctx, A, B);
    if (A->dtype == UHIP_F32) return transpose_typed<float>(ctx, A, B);
    return set_error(UHIP_ERR_INVALID, "uhip_csr_transpose: dtype %d", A->dtype);
}

int uhip_csr_spgemm(uhip_ctx ctx, const uhip_csr *A, const uhip_csr *B, uhip_csr *C) {
    UHIP_REQUIRE(ctx && A && B && C, "uhip_csr_spgemm: null argument");
    UHIP_REQUIRE(A->dtype == UHIP_F64 && B->dtype == UHIP_F64, "uhip_csr_spgemm: FP64 matrices only");
    UHIP_REQUIRE(A->ncols == B->nrows, "uhip_csr_spgemm: inner dimensions %d != %d", A->ncols, B->nrows);
    hipStream_t s = ctx->stream;
    const int32_t n = A->nrows;
    memset(C, 0, sizeof(*C));
    C->nrows = n;
    C->ncols = B->ncols;
    C->dtype = UHIP_F64;
    if (n == 0) {
        UHIP_CHECK_HIP(hipMalloc(&C->rowptr, sizeof(int32_t)));
        UHIP_CHECK_HIP(hipMemsetAsync(C->rowptr, 0, sizeof(int32_t), s));
        UHIP_CHECK_HIP(hipMalloc(&C->col, sizeof(int32_t)));
        UHIP_CHECK_HIP(hipMalloc(&C->val, sizeof(double)));
        return UHIP_OK;
    }
    int32_t *ub = nullptr, *mx = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&ub, sizeof(int32_t) * n));
    UHIP_CHECK_HIP(hipMalloc(&mx, sizeof(int32_t)));
    hipLaunchKernelGGL(k_row_ub, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->rowptr, A->col,
                       B->rowptr, ub);
    UHIP_CHECK_LAUNCH();
    void *tmp = nullptr;
    size_t tb = 0;
    UHIP_CHECK_HIP(hipcub::DeviceReduce::Max(nullptr, tb, ub, mx, n, s));
    UHIP_CHECK_HIP(hipMalloc(&tmp, tb));
    UHIP_CHECK_HIP(hipcub::DeviceReduce::Max(tmp, tb, ub, mx, n, s));
    int32_t maxub = 0;
    UHIP_CHECK_HIP(hipMemcpyAsync(&maxub, mx, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    hipFree(tmp);
    hipFree(ub);
    hipFree(mx);
    if (maxub <= 1024) return spgemm_passes<64, 1024>(ctx, A, B, C);
    if (maxub <= 4096) return spgemm_passes<256, 4096>(ctx, A, B, C);
    if (maxub <= 8192) return spgemm_passes<256, 8192>(ctx, A, B, C);
    return set_error(UHIP_ERR_UNSUPPORTED, "uhip_csr_spgemm: a row has %d candidate products (> 8192)", maxub);
}

int uhip_csr_ptap(uhip_ctx ctx, const uhip_csr *A, const uhip_csr *P, uhip_csr *Ac) {
    UHIP_REQUIRE(ctx && A && P && Ac, "uhip_csr_ptap: null argument");
    UHIP_REQUIRE(A->nrows == A->ncols && P->nrows == A->ncols, "uhip_csr_ptap: shapes A %dx%d P %dx%d", A->nrows,
                 A->ncols, P->nrows, P->ncols);
    uhip_csr R, AP;
    memset(&R, 0, sizeof(R));
    memset(&AP, 0, sizeof(AP));
    int st = uhip_csr_transpose(ctx, P, &R);
    if (st == UHIP_OK) st = uhip_csr_spgemm(ctx, A, P, &AP);
    if (st == UHIP_OK) st = uhip_csr_spgemm(ctx, &R, &AP, Ac);
    uhip_csr_free(ctx, &R);
    uhip_csr_free(ctx, &AP);
    return st;
}

int uhip_csr_to_dense(uhip_ctx ctx, const uhip_csr *A, double *dense) {
    UHIP_REQUIRE(ctx && A && dense, "uhip_csr_to_dense: null argument");
    UHIP_REQUIRE(A->dtype == UHIP_F64, "uhip_csr_to_dense: FP64 only");
    const int32_t n = A->nrows;
    UHIP_CHECK_HIP(hipMemsetAsync(dense, 0, sizeof(double) * (size_t)n * A->ncols, ctx->stream));
    if (n == 0) return UHIP_OK;
    hipLaunchKernelGGL(k_to_dense, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, n, A->ncols,
                       A->rowptr, A->col, (const double *)A->val, dense);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

}  // extern "C"
