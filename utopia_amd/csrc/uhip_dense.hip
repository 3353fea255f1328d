// Dense FP64 kernels for the coarsest multigrid level: blocked Cholesky
// (A = L L^T) whose GEMM-shaped trailing updates run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), the explicit inverse A^-1 = L^-T L^-1, and the GEMV
// that applies it once per V-cycle.
#include "uhip_common.hpp"

namespace uhip {

constexpr int kNB = 64;  // panel width

typedef double v4d __attribute__((ext_vector_type(4)));

// C(i,j) = beta*C(i,j) + alpha * sum_k A(i,k) * B(j,k) for a 64x64 tile of C per
// workgroup (4 waves, each a 32x32 quarter = 2x2 MFMA 16x16 tiles).
// Operand lanes of v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and
// B[l/16][l%16] (k = l/16).
// Element (i,k) of A is A[i*sai + k*sak], of B is B[j*sbj + k*sbk], C is row-major ldc.
// lower_only: skip tiles strictly above the diagonal.
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, double alpha, const double *__restrict__ A,
                                                  int64_t sai, int64_t sak, const double *__restrict__ B,
                                                  int64_t sbj, int64_t sbk, double beta, double *__restrict__ C,
                                                  int64_t ldc, int lower_only) {
    const int ti = blockIdx.y * 64, tj = blockIdx.x * 64;
    if (lower_only && tj > ti + 63) return;
    __shared__ double As[64][17];
    __shared__ double Bs[64][17];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    v4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = tid; e < 64 * 16; e += 256) {
            const int r = e >> 4, kk = e & 15;
            const int gi = ti + r, gj = tj + r, gk = k0 + kk;
            As[r][kk] = (gi < M && gk < K) ? A[gi * sai + gk * sak] : 0.0;
            Bs[r][kk] = (gj < N && gk < K) ? B[gj * sbj + gk * sbk] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            const int kl = ks + (lane >> 4);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const double av = As[wi + a * 16 + (lane & 15)][kl];
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const double bv = Bs[wj + b * 16 + (lane & 15)][kl];
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[a][b], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // v_mfma_f64_16x16x4f64 result: item r of lane l is D[4*r + l/16][l%16]
                const int gi = ti + wi + a * 16 + 4 * r + (lane >> 4);
                const int gj = tj + wj + b * 16 + (lane & 15);
                if (gi < M && gj < N && (!lower_only || gj <= gi)) {
                    double *c = C + (int64_t)gi * ldc + gj;
                    *c = (beta == 0.0 ? 0.0 : beta * *c) + alpha * acc[a][b][r];
                }
            }
}

// Unblocked Cholesky of the nb x nb diagonal block at (k,k), in place (lower).
__global__ __launch_bounds__(256) void k_potrf_diag(int n, double *__restrict__ A, int64_t lda, int k,
                                                    int *__restrict__ info) {
    __shared__ double T[kNB][kNB + 1];
    const int nb = min(kNB, n - k);
    const int tid = threadIdx.x;
    for (int e = tid; e < nb * nb; e += 256) T[e / nb][e % nb] = A[(int64_t)(k + e / nb) * lda + k + e % nb];
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (tid == 0) {
            const double d = T[j][j];
            if (!(d > 0.0)) {
                if (*info == 0) *info = k + j + 1;
                T[j][j] = 1.0;
            } else {
                T[j][j] = sqrt(d);
            }
        }
        __syncthreads();
        const double djj = T[j][j];
        for (int i = j + 1 + tid; i < nb; i += 256) T[i][j] = T[i][j] / djj;
        __syncthreads();
        for (int e = tid; e < (nb - j - 1) * (nb - j - 1); e += 256) {
            const int i = j + 1 + e / (nb - j - 1), c = j + 1 + e % (nb - j - 1);
            if (c <= i) T[i][c] = T[i][c] - T[i][j] * T[c][j];
        }
        __syncthreads();
    }
    for (int e = tid; e < nb * nb; e += 256) {
        const int i = e / nb, c = e % nb;
        A[(int64_t)(k + i) * lda + k + c] = (c <= i) ? T[i][c] : 0.0;
    }
}

// Rows i >= k+nb of the panel: L(i, k:k+nb) = A(i, k:k+nb) * Lkk^-T (one row per thread).
__global__ __launch_bounds__(256) void k_trsm_panel(int n, double *__restrict__ A, int64_t lda, int k) {
    __shared__ double L[kNB][kNB + 1];
    const int nb = min(kNB, n - k);
    for (int e = threadIdx.x; e < nb * nb; e += 256) L[e / nb][e % nb] = A[(int64_t)(k + e / nb) * lda + k + e % nb];
    __syncthreads();
    const int i = k + nb + blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double *row = A + (int64_t)i * lda + k;
    double x[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) x[c] = (c < nb) ? row[c] : 0.0;
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
        if (c < nb) {
            double s = x[c];
            for (int q = 0; q < c; ++q) s -= x[q] * L[c][q];
            x[c] = s / L[c][c];
        }
    }
#pragma unroll
    for (int c = 0; c < kNB; ++c)
        if (c < nb) row[c] = x[c];
}

// Forward substitution for the block row k of Y in L Y = B (B overwritten by Y):
// Y(k:k+nb, :) = Lkk^-1 * B(k:k+nb, :), one column per thread.
__global__ __launch_bounds__(256) void k_trsm_rows(int n, const double *__restrict__ Lm, int64_t lda,
                                                   double *__restrict__ B, int64_t ldb, int k) {
    __shared__ double L[kNB][kNB + 1];
    const int nb = min(kNB, n - k);
    for (int e = threadIdx.x; e < nb * nb; e += 256) L[e / nb][e % nb] = Lm[(int64_t)(k + e / nb) * lda + k + e % nb];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    double y[kNB];
#pragma unroll
    for (int r = 0; r < kNB; ++r) y[r] = (r < nb) ? B[(int64_t)(k + r) * ldb + c] : 0.0;
#pragma unroll
    for (int r = 0; r < kNB; ++r) {
        if (r < nb) {
            double s = y[r];
            for (int q = 0; q < r; ++q) s -= L[r][q] * y[q];
            y[r] = s / L[r][r];
        }
    }
#pragma unroll
    for (int r = 0; r < kNB; ++r)
        if (r < nb) B[(int64_t)(k + r) * ldb + c] = y[r];
}

__global__ __launch_bounds__(256) void k_identity(int n, double *__restrict__ B) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < (int64_t)n * n; e += (int64_t)gridDim.x * 256)
        B[e] = (e / n == e % n) ? 1.0 : 0.0;
}

// y = M x, one wavefront per row. Lane l keeps eight chains, chain u summing columns
// l + 64 u + 512 j in order; the chains are combined in a fixed tree and the lanes by the
// butterfly, so the result is the same on every run. M is streamed (non-temporal loads). With
// an FP32 x the main loop issues two steps' loads (sixteen per lane) before adding them to the
// chains in the same order as one step at a time; an FP64 x keeps one step's eight. Measured
// with the 4096-row coarse inverse of a 256^3 MG-PCG hierarchy: FP32 x 30.5 -> 23.8 us per
// product; sixteen loads with an FP64 x took 35.7 us against 29.8 (the streamed M alone: FP64
// MG-PCG -2 ms).
template <typename TX, typename TY, typename TM = double>
__global__ __launch_bounds__(256) void k_gemv(int n, const TM *__restrict__ M, const TX *__restrict__ x,
                                              TY *__restrict__ y) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const TM *m = M + (int64_t)row * n;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int c = lane;
    if constexpr (sizeof(TX) == 4) {
        for (; c + 15 * 64 < n; c += 16 * 64) {
            double mv[16], xv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                mv[u] = (double)__builtin_nontemporal_load(m + c + u * 64);
                xv[u] = (double)x[c + u * 64];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] += mv[u] * xv[u];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] += mv[8 + u] * xv[8 + u];
        }
    }
    for (; c + 7 * 64 < n; c += 8 * 64) {
        double mv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            mv[u] = (double)__builtin_nontemporal_load(m + c + u * 64);
            xv[u] = (double)x[c + u * 64];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += mv[u] * xv[u];
    }
    for (int u = 0; c < n; c += 64, ++u) a[u] += (double)m[c] * (double)x[c];
    double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    s = wave_sum(s);
    if (lane == 0) y[row] = (TY)s;
}

static int gemm(uhip_ctx ctx, int M, int N, int K, double alpha, const double *A, int64_t sai, int64_t sak,
                const double *B, int64_t sbj, int64_t sbk, double beta, double *C, int64_t ldc, int lower) {
    if (M <= 0 || N <= 0) return UHIP_OK;
    dim3 g((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL(k_gemm_f64, g, dim3(256), 0, ctx->stream, M, N, K, alpha, A, sai, sak, B, sbj, sbk, beta, C,
                       ldc, lower);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

}  // namespace uhip

using namespace uhip;

extern "C" {

int uhip_dense_cholesky(uhip_ctx ctx, int32_t n, double *A, int32_t lda, int *info_h) {
    UHIP_REQUIRE(ctx && A && info_h && n >= 0 && lda >= n, "uhip_dense_cholesky: bad argument");
    int *info = (int *)ctx->counters;
    UHIP_CHECK_HIP(hipMemsetAsync(info, 0, sizeof(int), ctx->stream));
    for (int k = 0; k < n; k += kNB) {
        const int nb = std::min(kNB, n - k);
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, ctx->stream, n, A, (int64_t)lda, k, info);
        UHIP_CHECK_LAUNCH();
        const int rest = n - k - nb;
        if (rest > 0) {
            hipLaunchKernelGGL(k_trsm_panel, dim3((rest + 255) / 256), dim3(256), 0, ctx->stream, n, A,
                               (int64_t)lda, k);
            UHIP_CHECK_LAUNCH();
            // A22 -= L21 L21^T (lower triangle)
            double *L21 = A + (int64_t)(k + nb) * lda + k;
            double *A22 = A + (int64_t)(k + nb) * lda + k + nb;
            UHIP_TRY(gemm(ctx, rest, rest, nb, -1.0, L21, lda, 1, L21, lda, 1, 1.0, A22, lda, 1));
        }
    }
    UHIP_CHECK_HIP(hipMemcpyAsync(info_h, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return UHIP_OK;
}

int uhip_dense_cholesky_inverse(uhip_ctx ctx, int32_t n, const double *L, int32_t lda, double *Ainv) {
    UHIP_REQUIRE(ctx && L && Ainv && n >= 0 && lda >= n, "uhip_dense_cholesky_inverse: bad argument");
    if (n == 0) return UHIP_OK;
    double *Y = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&Y, sizeof(double) * (size_t)n * n));
    hipLaunchKernelGGL(k_identity, dim3(grid_for((int64_t)n * n)), dim3(256), 0, ctx->stream, n, Y);
    UHIP_CHECK_LAUNCH();
    // Y = L^-1 (blocked forward substitution on the identity)
    for (int k = 0; k < n; k += kNB) {
        const int nb = std::min(kNB, n - k);
        hipLaunchKernelGGL(k_trsm_rows, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, L, (int64_t)lda, Y,
                           (int64_t)n, k);
        UHIP_CHECK_LAUNCH();
        const int rest = n - k - nb;
        if (rest > 0) {
            // Y(k+nb:, :) -= L(k+nb:, k:k+nb) * Y(k:k+nb, :)
            const double *L21 = L + (int64_t)(k + nb) * lda + k;
            const double *Yk = Y + (int64_t)k * n;
            int st = gemm(ctx, rest, n, nb, -1.0, L21, lda, 1, Yk, 1, n, 1.0, Y + (int64_t)(k + nb) * n, n, 0);
            if (st != UHIP_OK) {
                hipFree(Y);
                return st;
            }
        }
    }
    // A^-1 = Y^T Y:  Ainv(i,j) = sum_k Y(k,i) Y(k,j)
    int st = gemm(ctx, n, n, n, 1.0, Y, 1, n, Y, 1, n, 0.0, Ainv, n, 0);
    hipStreamSynchronize(ctx->stream);
    hipFree(Y);
    return st;
}

int uhip_dense_gemv(uhip_ctx ctx, int32_t n, const double *M, const void *x, int x_dtype, void *y, int y_dtype) {
    UHIP_REQUIRE(ctx && M && x && y && n >= 0, "uhip_dense_gemv: bad argument");
    if (n == 0) return UHIP_OK;
    dim3 g((n + 3) / 4), b(256);
    hipStream_t s = ctx->stream;
    if (x_dtype == UHIP_F64 && y_dtype == UHIP_F64)
        hipLaunchKernelGGL((k_gemv<double, double>), g, b, 0, s, n, M, (const double *)x, (double *)y);
    else if (x_dtype == UHIP_F32 && y_dtype == UHIP_F32)
        hipLaunchKernelGGL((k_gemv<float, float>), g, b, 0, s, n, M, (const float *)x, (float *)y);
    else if (x_dtype == UHIP_F32 && y_dtype == UHIP_F64)
        hipLaunchKernelGGL((k_gemv<float, double>), g, b, 0, s, n, M, (const float *)x, (double *)y);
    else
        hipLaunchKernelGGL((k_gemv<double, float>), g, b, 0, s, n, M, (const double *)x, (float *)y);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

}  // extern "C"
