// Row-block plan, SpMV family (MatMult and its fused callers), and the
// per-row matrix utilities (diagonal, masks, semidefinite fix).
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include <mutex>

#include "uhip_spmv.hpp"

#include <limits>

namespace uhip {

// ---- partial-sum scratch of uhip_spmv_dot_noalloc ---------------------------------
// File-scope device array (each GPU the library runs on holds its own copy until the
// process ends): kNoAllocParts per-workgroup partial sums of one dot. In host code its
// name is a stand-in, not a device address: the address is looked up once per device
// (spmv_scratch_init, from uhip_ctx_create with the device current) and only that
// address is handed to kernels.
constexpr int64_t kNoAllocParts = 1 << 18;  // 2 MiB
static __device__ __align__(256) double scratch[kNoAllocParts];
constexpr int kMaxScratchDevices = 64;
static double *g_scratch_ptr[kMaxScratchDevices] = {};
static std::mutex g_scratch_mu;

int spmv_scratch_init(int device) {
    UHIP_REQUIRE(device >= 0 && device < kMaxScratchDevices, "spmv scratch: device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return UHIP_OK;
    double *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(UHIP_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return UHIP_OK;
}

static double *scratch_of(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Row r starts a block when its cost window (rowptr[r] + r) / B differs from the
// previous row's, or when r or r-1 is a long row (> B/4 nonzeros). A multi-row
// block then holds < B + B/4 <= 256*capk nonzeros.
__global__ __launch_bounds__(kBlock) void k_plan_flags(int32_t n, const int32_t *__restrict__ rowptr, int32_t B,
                                                       uint8_t *__restrict__ flags) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t lim = B / 4;
    const bool long_r = (rowptr[r + 1] - rowptr[r]) > lim;
    uint8_t f = 1;
    if (r > 0) {
        const int64_t w0 = ((int64_t)rowptr[r - 1] + (r - 1)) / B;
        const int64_t w1 = ((int64_t)rowptr[r] + r) / B;
        const bool long_prev = (rowptr[r] - rowptr[r - 1]) > lim;
        f = (w0 != w1) || long_r || long_prev;
    }
    flags[r] = f;
}

__global__ void k_plan_tail(int32_t *__restrict__ blocks, const int32_t *__restrict__ count, int32_t n) {
    blocks[*count] = n;
}

// entries of slice s = 64 * (longest row of the slice)
__global__ __launch_bounds__(kBlock) void k_slice_width(int32_t n, int32_t nslices, const int32_t *__restrict__ rowptr,
                                                        int32_t *__restrict__ width) {
    const int32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s > nslices) return;
    if (s == nslices) {
        width[s] = 0;
        return;
    }
    int32_t w = 0;
    const int32_t r1 = min(n, (s + 1) * 64);
    for (int32_t r = s * 64; r < r1; ++r) w = max(w, rowptr[r + 1] - rowptr[r]);
    width[s] = 64 * w;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_fill(int32_t n, const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col, const T *__restrict__ val,
                                                      const int32_t *__restrict__ slice_off,
                                                      int32_t *__restrict__ scol, T *__restrict__ sval) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int32_t s = r >> 6, lane = r & 63;
    const int32_t o = slice_off[s], w = (slice_off[s + 1] - o) >> 6;
    const int32_t a = rowptr[r], len = rowptr[r + 1] - a;
    for (int32_t j = 0; j < w; ++j) {
        const int64_t d = (int64_t)o + 64 * j + lane;
        scol[d] = j < len ? col[a + j] : 0;
        sval[d] = j < len ? val[a + j] : T(0);
    }
}

// Plan SELL-64 arrays; returns UHIP_ERR_UNSUPPORTED (nothing allocated) when the
// padding exceeds max_ratio * nnz.
static int plan_sell(uhip_ctx ctx, uhip_csr *A, double max_ratio) {
    hipStream_t s = ctx->stream;
    const int32_t n = A->nrows;
    const int32_t ns = (n + 63) / 64;
    int32_t *w = nullptr, *off = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&w, sizeof(int32_t) * ((size_t)ns + 1)));
    UHIP_CHECK_HIP(hipMalloc(&off, sizeof(int32_t) * ((size_t)ns + 1)));
    hipLaunchKernelGGL(k_slice_width, dim3(ns / kBlock + 1), dim3(kBlock), 0, s, n, ns, A->rowptr, w);
    UHIP_CHECK_LAUNCH();
    void *tmp = nullptr;
    size_t tb = 0;
    UHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, w, off, ns + 1, s));
    UHIP_CHECK_HIP(hipMalloc(&tmp, tb));
    UHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, w, off, ns + 1, s));
    int32_t total = 0;
    UHIP_CHECK_HIP(hipMemcpyAsync(&total, off + ns, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    hipFree(tmp);
    hipFree(w);
    if ((double)total > max_ratio * (double)A->nnz || total < 0) {
        hipFree(off);
        return UHIP_ERR_UNSUPPORTED;
    }
    const size_t vb = A->dtype == UHIP_F64 ? sizeof(double) : sizeof(float);
    UHIP_CHECK_HIP(hipMalloc(&A->sell_col, sizeof(int32_t) * (size_t)std::max(total, 1)));
    UHIP_CHECK_HIP(hipMalloc(&A->sell_val, vb * (size_t)std::max(total, 1)));
    if (A->dtype == UHIP_F64)
        hipLaunchKernelGGL(k_sell_fill<double>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->rowptr,
                           A->col, (const double *)A->val, off, A->sell_col, (double *)A->sell_val);
    else
        hipLaunchKernelGGL(k_sell_fill<float>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->rowptr,
                           A->col, (const float *)A->val, off, A->sell_col, (float *)A->sell_val);
    UHIP_CHECK_LAUNCH();
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    A->slice_off = off;
    A->nslices = ns;
    A->format = UHIP_FMT_SELL64;
    A->nparts = (ns + kSellWaves - 1) / kSellWaves;
    return UHIP_OK;
}

// ---- DIA plan ---------------------------------------------------------------
// Distinct offsets col - row are collected in a 64-slot open-addressing set per
// workgroup (LDS), then merged into one global 64-slot set. flags bit 0: a row
// is not strictly increasing in column; bit 1: more than 64 distinct offsets.
constexpr int kDiaSlots = 64;
constexpr int32_t kDiaEmpty = INT32_MIN;

__device__ __forceinline__ int dia_hash(int32_t key) { return (int)(((uint32_t)key * 2654435761u) >> 26); }

__global__ __launch_bounds__(kBlock) void k_dia_init(int32_t *__restrict__ table, int32_t *__restrict__ flags) {
    if (threadIdx.x < kDiaSlots) table[threadIdx.x] = kDiaEmpty;
    if (threadIdx.x == 0) *flags = 0;
}

// ROWBASE: offsets are taken from each row's first column instead of its index.
template <bool ROWBASE>
__global__ __launch_bounds__(kBlock) void k_dia_scan(int32_t n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col, int32_t *__restrict__ gtable,
                                                     int32_t *__restrict__ gflags) {
    __shared__ int32_t t[kDiaSlots];
    __shared__ int32_t bad;
    if (threadIdx.x < kDiaSlots) t[threadIdx.x] = kDiaEmpty;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r < n) {
        const int32_t a = rowptr[r], e1 = rowptr[r + 1];
        int32_t prev = INT32_MIN;
        for (int32_t e = a; e < e1; ++e) {
            const int32_t c = col[e];
            if (e > a && c <= prev) atomicOr(&bad, 1);
            prev = c;
            const int32_t key = c - (ROWBASE ? col[a] : r);
            const int h = dia_hash(key);
            bool placed = false;
            for (int p = 0; p < kDiaSlots && !placed; ++p) {
                const int slot = (h + p) & (kDiaSlots - 1);
                const int32_t cur = t[slot];
                if (cur == key) {
                    placed = true;
                } else if (cur == kDiaEmpty) {
                    const int32_t old = atomicCAS(&t[slot], kDiaEmpty, key);
                    placed = (old == kDiaEmpty || old == key);
                }
            }
            if (!placed) atomicOr(&bad, 2);
        }
    }
    __syncthreads();
    if (threadIdx.x < kDiaSlots && t[threadIdx.x] != kDiaEmpty) {
        const int32_t key = t[threadIdx.x];
        const int h = dia_hash(key);
        bool placed = false;
        for (int p = 0; p < kDiaSlots && !placed; ++p) {
            const int slot = (h + p) & (kDiaSlots - 1);
            const int32_t cur = __hip_atomic_load(&gtable[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == key) {
                placed = true;
            } else if (cur == kDiaEmpty) {
                const int32_t old = atomicCAS(&gtable[slot], kDiaEmpty, key);
                placed = (old == kDiaEmpty || old == key);
            }
        }
        if (!placed) atomicOr(gflags, 2);
    }
    if (threadIdx.x == 0 && bad) atomicOr(gflags, bad);
}

// Row r writes all ndiag slots of its diagonals (its value, or NaN where the row does not hold
// the diagonal) and its mask; has_nan[0] |= 1 when a stored value is NaN (the plan then zeroes
// the absent slots instead).
template <typename T, typename TM>
__global__ __launch_bounds__(kBlock) void k_dia_fill(int32_t n, int32_t ndiag, const int32_t *__restrict__ off,
                                                     const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col, const T *__restrict__ val,
                                                     T *__restrict__ dval, TM *__restrict__ mask,
                                                     int32_t *__restrict__ has_nan) {
    __shared__ int32_t so[32];
    if (threadIdx.x < ndiag) so[threadIdx.x] = off[threadIdx.x];
    __syncthreads();
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    int32_t e = rowptr[r];
    const int32_t e1 = rowptr[r + 1];
    uint32_t m = 0;
    bool nan_seen = false;
    const int64_t ld = dia_ld(n);
    for (int d = 0; d < ndiag; ++d) {
        T v = (T)__builtin_nan("");
        if (e < e1 && col[e] - r == so[d]) {
            v = val[e];
            nan_seen |= (v != v);
            m |= 1u << d;
            ++e;
        }
        dval[(int64_t)d * ld + r] = v;
    }
    mask[r] = (TM)m;
    if (nan_seen) atomicOr(has_nan, 1);
}

// absent slots to zero (a matrix that holds NaN values keeps the mask as the only record)
template <typename T, typename TM>
__global__ __launch_bounds__(kBlock) void k_dia_zero_absent(int32_t n, int32_t ndiag, T *__restrict__ dval,
                                                            const TM *__restrict__ mask) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t ld = dia_ld(n);
    const uint32_t m = (uint32_t)mask[r];
    for (int d = 0; d < ndiag; ++d)
        if (!((m >> d) & 1u)) dval[(int64_t)d * ld + r] = T(0);
}

__device__ __forceinline__ bool same_bits(double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); }
__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }

// flag[0] = 1 unless every stored entry (r, r + o), o > 0, has a stored mirror
// (r + o, r) with the same bits, and the pattern is symmetric.
template <typename T, typename TM>
__global__ __launch_bounds__(kBlock) void k_dia_symcheck(int32_t n, int32_t ndiag, const int32_t *__restrict__ off,
                                                         const T *__restrict__ dval, const TM *__restrict__ mask,
                                                         int32_t *__restrict__ flag) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t ld = dia_ld(n);
    const int32_t k = ndiag >> 1;
    const uint32_t m = (uint32_t)mask[r];
    bool ok = true;
    for (int d = k + 1; d < ndiag; ++d) {
        const int32_t o = off[d], md = ndiag - 1 - d;
        const bool up = (m >> d) & 1u;
        const bool lo = (r + o < n) && (((uint32_t)mask[r + o] >> md) & 1u);
        if (up != lo) ok = false;
        if (up && lo) {
            const T a = dval[(int64_t)d * ld + r], b = dval[(int64_t)md * ld + r + o];
            if (!same_bits(a, b)) ok = false;
        }
    }
    if (!ok) flag[0] = 1;
}

// Bit patterns of FP64 / FP32 values as signed integers, and the delta type of each: int32 for
// FP64, int8 for FP32 (the FP32 copy of a near-symmetric operator differs by an ulp or two).
__device__ __forceinline__ long long value_bits(double v) { return __double_as_longlong(v); }
__device__ __forceinline__ long long value_bits(float v) { return (long long)__float_as_int(v); }
template <typename T>
struct DeltaOf {
    typedef int32_t t;
};
template <>
struct DeltaOf<float> {
    typedef int8_t t;
};

// Near-symmetric values (a Galerkin product, symmetric up to rounding): flag[0] = 1 unless every
// stored lower entry (r, r - o) has a stored mirror (r - o, r) of the same sign whose bit pattern
// differs from its own by a value of the delta type.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dia_deltacheck(int32_t n, int32_t ndiag, const int32_t *__restrict__ off,
                                                           const T *__restrict__ dval,
                                                           const uint32_t *__restrict__ mask,
                                                           int32_t *__restrict__ flag) {
    typedef typename DeltaOf<T>::t TD;
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t ld = dia_ld(n);
    const int32_t k = ndiag >> 1;
    const uint32_t m = mask[r];
    bool ok = true;
    for (int d = 0; d < k; ++d) {
        if (!((m >> d) & 1u)) continue;
        const int32_t rm = r + off[d], md = ndiag - 1 - d;
        if (rm < 0 || !((mask[rm] >> md) & 1u)) {
            ok = false;
            continue;
        }
        const long long a = value_bits(dval[(int64_t)d * ld + r]);
        const long long b = value_bits(dval[(int64_t)md * ld + rm]);
        const long long e = a - b;
        if ((a ^ b) < 0 || e < std::numeric_limits<TD>::min() || e > std::numeric_limits<TD>::max()) ok = false;
    }
    if (!ok) flag[0] = 1;
}

// delta[d][r] = bits(value of lower entry d of row r) - bits(its mirror), 0 where absent
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dia_deltafill(int32_t n, int32_t ndiag, const int32_t *__restrict__ off,
                                                          const T *__restrict__ dval,
                                                          const uint32_t *__restrict__ mask,
                                                          typename DeltaOf<T>::t *__restrict__ delta) {
    typedef typename DeltaOf<T>::t TD;
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const int64_t ld = dia_ld(n);
    const int32_t k = ndiag >> 1;
    const uint32_t m = mask[r];
    for (int d = 0; d < k; ++d) {
        TD e = 0;
        if ((m >> d) & 1u) {
            const int32_t rm = r + off[d], md = ndiag - 1 - d;
            e = (TD)(value_bits(dval[(int64_t)d * ld + r]) - value_bits(dval[(int64_t)md * ld + rm]));
        }
        delta[(int64_t)d * ld + r] = e;
    }
}

// Plan DIA arrays; UHIP_ERR_UNSUPPORTED (nothing allocated) when the matrix is
// not square, has unsorted or duplicate columns in a row, more than 32 distinct
// offsets, or ndiag * n > max_ratio * nnz.
// walk: the plane-walking 27-point kernel (and with it the near-symmetric layout) 0 never, 1 when
// the grid gives it enough waves (automatic plans), 2 whenever the offsets allow (an explicit DIA
// plan; the autotuner times 2 against 0).
static int plan_dia(uhip_ctx ctx, uhip_csr *A, double max_ratio, int walk) {
    hipStream_t s = ctx->stream;
    const int32_t n = A->nrows;
    if (n != A->ncols || n == 0) return UHIP_ERR_UNSUPPORTED;
    int32_t *table = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&table, sizeof(int32_t) * (kDiaSlots + 1)));
    int32_t *flags = table + kDiaSlots;
    hipLaunchKernelGGL(k_dia_init, dim3(1), dim3(kBlock), 0, s, table, flags);
    UHIP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_dia_scan<false>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->rowptr, A->col,
                       table, flags);
    UHIP_CHECK_LAUNCH();
    int32_t h[kDiaSlots + 1];
    UHIP_CHECK_HIP(hipMemcpyAsync(h, table, sizeof(h), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> offs;
    for (int i = 0; i < kDiaSlots; ++i)
        if (h[i] != kDiaEmpty) offs.push_back(h[i]);
    std::sort(offs.begin(), offs.end());
    const int32_t nd = (int32_t)offs.size();
    if (h[kDiaSlots] != 0 || nd < 1 || nd > 32 || (double)nd * (double)n > max_ratio * (double)A->nnz) {
        hipFree(table);
        return UHIP_ERR_UNSUPPORTED;
    }
    const size_t vb = A->dtype == UHIP_F64 ? sizeof(double) : sizeof(float);
    const size_t mb = nd <= 8 ? 1 : (nd <= 16 ? 2 : 4);
    UHIP_CHECK_HIP(hipMemcpyAsync(table, offs.data(), sizeof(int32_t) * (size_t)nd, hipMemcpyHostToDevice, s));
    UHIP_CHECK_HIP(hipMalloc(&A->dia_val, vb * (size_t)nd * (size_t)dia_ld(n)));
    UHIP_CHECK_HIP(hipMalloc(&A->dia_mask, mb * (size_t)n));
    const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
    UHIP_CHECK_HIP(hipMemsetAsync(flags, 0, sizeof(int32_t), s));
#define UHIP_DIA_FILL(T, TM)                                                                                      \
    hipLaunchKernelGGL((k_dia_fill<T, TM>), g, b, 0, s, n, nd, table, A->rowptr, A->col, (const T *)A->val,      \
                       (T *)A->dia_val, (TM *)A->dia_mask, flags)
    if (A->dtype == UHIP_F64) {
        if (mb == 1) UHIP_DIA_FILL(double, uint8_t);
        else if (mb == 2) UHIP_DIA_FILL(double, uint16_t);
        else UHIP_DIA_FILL(double, uint32_t);
    } else {
        if (mb == 1) UHIP_DIA_FILL(float, uint8_t);
        else if (mb == 2) UHIP_DIA_FILL(float, uint16_t);
        else UHIP_DIA_FILL(float, uint32_t);
    }
#undef UHIP_DIA_FILL
    UHIP_CHECK_LAUNCH();
    int32_t has_nan = 0;
    UHIP_CHECK_HIP(hipMemcpyAsync(&has_nan, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    if (has_nan) {
#define UHIP_DIA_ZERO(T, TM) \
    hipLaunchKernelGGL((k_dia_zero_absent<T, TM>), g, b, 0, s, n, nd, (T *)A->dia_val, (const TM *)A->dia_mask)
        if (A->dtype == UHIP_F64) {
            if (mb == 1) UHIP_DIA_ZERO(double, uint8_t);
            else if (mb == 2) UHIP_DIA_ZERO(double, uint16_t);
            else UHIP_DIA_ZERO(double, uint32_t);
        } else {
            if (mb == 1) UHIP_DIA_ZERO(float, uint8_t);
            else if (mb == 2) UHIP_DIA_ZERO(float, uint16_t);
            else UHIP_DIA_ZERO(float, uint32_t);
        }
#undef UHIP_DIA_ZERO
        UHIP_CHECK_LAUNCH();
    }
    // the 27 offsets dz S2 + dy S1 + dx in this (sorted) order, S2 a multiple of 64 (and for an
    // automatic plan enough waves to fill the GPU, kDia27MinWaves: a small or flat grid keeps the
    // generic kernels): the plane-walking 27-point kernel
    bool w27 = false;
    if (nd == 27 && walk > 0) {
        const int32_t S1 = offs[16], S2 = offs[22];
        w27 = S1 > 2 && S2 > 2 * S1 + 2 && S2 % 64 == 0 &&
              (int64_t)n + (int64_t)(dia27_planes(A->dtype) + 2) * S2 < INT32_MAX &&
              (walk == 2 ||
               (int64_t)dia7z_workgroups(n, S2, dia27_planes(A->dtype)) * (kDiaWG / 64) >= kDia27MinWaves);
        for (int i = 0; i < 27 && w27; ++i)
            w27 = offs[(size_t)i] == (i / 9 - 1) * S2 + (i / 3 % 3 - 1) * S1 + (i % 3 - 1);
    }
    // symmetric offsets -o..0..+o: keep only 0..+o when the values are symmetric bit for bit
    bool sym_offs = (nd & 1) && offs[(size_t)nd / 2] == 0;
    for (int32_t d = 0; d < nd && sym_offs; ++d) sym_offs = offs[(size_t)d] == -offs[(size_t)(nd - 1 - d)];
    int32_t sym = 0;
    if (sym_offs && nd > 1) {
        UHIP_CHECK_HIP(hipMemsetAsync(flags, 0, sizeof(int32_t), s));
#define UHIP_DIA_SYM(T, TM)                                                                                      \
    hipLaunchKernelGGL((k_dia_symcheck<T, TM>), g, b, 0, s, n, nd, table, (const T *)A->dia_val,              \
                       (const TM *)A->dia_mask, flags)
        if (A->dtype == UHIP_F64) {
            if (mb == 1) UHIP_DIA_SYM(double, uint8_t);
            else if (mb == 2) UHIP_DIA_SYM(double, uint16_t);
            else UHIP_DIA_SYM(double, uint32_t);
        } else {
            if (mb == 1) UHIP_DIA_SYM(float, uint8_t);
            else if (mb == 2) UHIP_DIA_SYM(float, uint16_t);
            else UHIP_DIA_SYM(float, uint32_t);
        }
#undef UHIP_DIA_SYM
        UHIP_CHECK_LAUNCH();
        int32_t asym = 1;
        UHIP_CHECK_HIP(hipMemcpyAsync(&asym, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        UHIP_CHECK_HIP(hipStreamSynchronize(s));
        if (asym == 0) {
            const size_t ld = (size_t)dia_ld(n), keep = (size_t)(nd / 2 + 1);
            void *half = nullptr;
            UHIP_CHECK_HIP(hipMalloc(&half, vb * keep * ld));
            UHIP_CHECK_HIP(hipMemcpyAsync(half, (const char *)A->dia_val + vb * (size_t)(nd / 2) * ld, vb * keep * ld,
                                          hipMemcpyDeviceToDevice, s));
            UHIP_CHECK_HIP(hipStreamSynchronize(s));
            UHIP_CHECK_HIP(hipFree(A->dia_val));
            A->dia_val = half;
            sym = 1;
        } else if (w27 && !has_nan) {
            // symmetric up to rounding: the upper half plus each lower value's bit-pattern
            // difference from its mirror (int32 for FP64, int8 for FP32), for the 27-point kernel only
            const bool f64 = A->dtype == UHIP_F64;
            UHIP_CHECK_HIP(hipMemsetAsync(flags, 0, sizeof(int32_t), s));
            if (f64)
                hipLaunchKernelGGL(k_dia_deltacheck<double>, g, b, 0, s, n, nd, table, (const double *)A->dia_val,
                                   (const uint32_t *)A->dia_mask, flags);
            else
                hipLaunchKernelGGL(k_dia_deltacheck<float>, g, b, 0, s, n, nd, table, (const float *)A->dia_val,
                                   (const uint32_t *)A->dia_mask, flags);
            UHIP_CHECK_LAUNCH();
            int32_t far = 1;
            UHIP_CHECK_HIP(hipMemcpyAsync(&far, flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            UHIP_CHECK_HIP(hipStreamSynchronize(s));
            if (far == 0) {
                const size_t ld = (size_t)dia_ld(n), keep = (size_t)(nd / 2 + 1);
                const size_t db = f64 ? sizeof(int32_t) : sizeof(int8_t);
                void *half = nullptr, *delta = nullptr;
                UHIP_CHECK_HIP(hipMalloc(&half, vb * keep * ld));
                UHIP_CHECK_HIP(hipMalloc(&delta, db * (size_t)(nd / 2) * ld));
                if (f64)
                    hipLaunchKernelGGL(k_dia_deltafill<double>, g, b, 0, s, n, nd, table, (const double *)A->dia_val,
                                       (const uint32_t *)A->dia_mask, (int32_t *)delta);
                else
                    hipLaunchKernelGGL(k_dia_deltafill<float>, g, b, 0, s, n, nd, table, (const float *)A->dia_val,
                                       (const uint32_t *)A->dia_mask, (int8_t *)delta);
                UHIP_CHECK_LAUNCH();
                UHIP_CHECK_HIP(hipMemcpyAsync(half, (const char *)A->dia_val + vb * (size_t)(nd / 2) * ld,
                                              vb * keep * ld, hipMemcpyDeviceToDevice, s));
                UHIP_CHECK_HIP(hipStreamSynchronize(s));
                UHIP_CHECK_HIP(hipFree(A->dia_val));
                A->dia_val = half;
                A->dia_delta = delta;
                sym = 1;
            }
        }
    }
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    // symmetric 7-point operators with the largest offset a multiple of 64 take the
    // plane-walking 7-point kernel (the same arrays)
    const int planes = dia7_planes(A->dtype);
    if (sym && nd == 7 && offs[4] == 1 && offs[5] > 1 && offs[6] > offs[5] && offs[6] % 64 == 0 &&
        (int64_t)n + (int64_t)planes * offs[6] < INT32_MAX)
        sym = 2;
    A->dia_off = table;  // first nd slots hold the sorted offsets
    A->ndiag = nd;
    A->dia_sym = sym;
    A->dia_nan = has_nan ? 0 : 1;
    A->dia_w27 = w27 ? 1 : 0;
    A->dia_s1 = sym == 2 ? offs[5] : (w27 ? offs[16] : 0);
    A->dia_s2 = sym == 2 ? offs[6] : (w27 ? offs[22] : 0);
    A->format = UHIP_FMT_DIA;
    A->nparts = sym == 2 ? dia7z_workgroups(n, offs[6], planes)
                         : (w27 ? dia7z_workgroups(n, offs[22], dia27_planes(A->dtype))
                                : (n + dia_rows_per_wg(A->dtype) - 1) / dia_rows_per_wg(A->dtype));
    return UHIP_OK;
}

// ---- VI plan (value-indexed, row-relative offsets) ---------------------------
// Distinct value bit patterns are collected like the offsets: a 512-slot set per
// workgroup in LDS, merged into one global set. flags bit 2: more than 512 distinct
// values, bit 3: a NaN value (not indexed).
constexpr int kViSlots = 512;
constexpr uint64_t kViEmpty = ~0ull;

__device__ __forceinline__ int vi_hash(uint64_t key) {
    return (int)(((key ^ (key >> 29)) * 0x9E3779B97F4A7C15ull) >> 55);  // 9 bits
}
template <typename T>
__device__ __forceinline__ uint64_t vi_bits(T v) {
    if constexpr (sizeof(T) == 8) return (uint64_t)__double_as_longlong((double)v);
    else return (uint64_t)__float_as_uint((float)v);
}

__global__ __launch_bounds__(kBlock) void k_vi_init(unsigned long long *__restrict__ table) {
    for (int i = threadIdx.x; i < kViSlots; i += kBlock) table[i] = kViEmpty;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_vi_scan(int64_t nnz, const T *__restrict__ val,
                                                    unsigned long long *__restrict__ gtable,
                                                    int32_t *__restrict__ gflags) {
    __shared__ unsigned long long t[kViSlots];
    __shared__ int32_t bad, count;
    for (int i = threadIdx.x; i < kViSlots; i += kBlock) t[i] = kViEmpty;
    if (threadIdx.x == 0) bad = count = 0;
    __syncthreads();
    // The block stops at its 257th distinct value (the table holds at most 256), so a
    // matrix with many values costs a few probes per entry, not a walk of the table.
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * kBlock) {
        if (__hip_atomic_load(&bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & 4) break;
        const T v = val[e];
        if (v != v) {
            atomicOr(&bad, 8);
            continue;
        }
        const unsigned long long key = vi_bits(v);
        const int h = vi_hash(key);
        bool placed = false;
        for (int p = 0; p < kViSlots && !placed; ++p) {
            const int slot = (h + p) & (kViSlots - 1);
            const unsigned long long cur = t[slot];
            if (cur == key) {
                placed = true;
            } else if (cur == kViEmpty) {
                const unsigned long long old = atomicCAS(&t[slot], kViEmpty, key);
                if (old == kViEmpty && atomicAdd(&count, 1) >= 256) atomicOr(&bad, 4);
                placed = (old == kViEmpty || old == key);
            }
        }
        if (!placed) atomicOr(&bad, 4);
    }
    __syncthreads();
    if (bad & 4) {
        if (threadIdx.x == 0) atomicOr(gflags, bad);
        return;
    }
    for (int i = threadIdx.x; i < kViSlots; i += kBlock) {
        const unsigned long long key = t[i];
        if (key == kViEmpty) continue;
        const int h = vi_hash(key);
        bool placed = false;
        for (int p = 0; p < kViSlots && !placed; ++p) {
            const int slot = (h + p) & (kViSlots - 1);
            const unsigned long long cur =
                __hip_atomic_load(&gtable[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == key) {
                placed = true;
            } else if (cur == kViEmpty) {
                const unsigned long long old = atomicCAS(&gtable[slot], kViEmpty, key);
                placed = (old == kViEmpty || old == key);
            }
        }
        if (!placed) atomicOr(gflags, 4);
    }
    if (threadIdx.x == 0 && bad) atomicOr(gflags, bad);
}

// Row r: base = its first column; slot d gets the code of the entry at column
// base + off[d] (0 where absent) in byte d % 8 of word r * NW + d / 8 (NW =
// sizeof(TM)), and bit d of the mask says whether it is present.
template <typename T, typename TM>
__global__ __launch_bounds__(kBlock) void k_vi_fill(int32_t n, int32_t ndiag, const int32_t *__restrict__ off,
                                                    int32_t nvals, const unsigned long long *__restrict__ sorted_bits,
                                                    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                    const T *__restrict__ val, int32_t *__restrict__ base,
                                                    uint64_t *__restrict__ codes, TM *__restrict__ mask) {
    constexpr int NW = (int)sizeof(TM);
    __shared__ int32_t so[32];
    __shared__ unsigned long long sv[256];
    if (threadIdx.x < ndiag) so[threadIdx.x] = off[threadIdx.x];
    for (int i = threadIdx.x; i < nvals; i += kBlock) sv[i] = sorted_bits[i];
    __syncthreads();
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    int32_t e = rowptr[r];
    const int32_t e1 = rowptr[r + 1];
    const int32_t b = e < e1 ? col[e] : 0;
    uint32_t m = 0;
    uint64_t word[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) word[w] = 0;
#pragma unroll
    for (int d = 0; d < 8 * NW; ++d) {
        if (d < ndiag && e < e1 && col[e] - b == so[d]) {
            const unsigned long long key = vi_bits(val[e]);
            int lo = 0, hi = nvals - 1;  // the key is in the table (collected by k_vi_scan)
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (sv[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            word[d / 8] |= (uint64_t)lo << (8 * (d % 8));
            m |= 1u << d;
            ++e;
        }
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) codes[(int64_t)r * NW + w] = word[w];
    base[r] = b;
    mask[r] = (TM)m;
}

template <typename T>
__global__ void k_vi_table(int32_t nvals, const unsigned long long *__restrict__ sorted_bits, T *__restrict__ table) {
    const int i = threadIdx.x;
    if (i >= nvals) return;
    if constexpr (sizeof(T) == 8) table[i] = (T)__longlong_as_double((long long)sorted_bits[i]);
    else table[i] = (T)__uint_as_float((unsigned)sorted_bits[i]);
}

// Row signatures of a VI plan: a row's (slot mask, code words) pair. Transfer operators have few
// distinct ones (interior rows, faces, edges, corners), so a plan keeps one uint8 signature id
// per row and a table of at most 256 signatures instead of every row's mask and codes.
__device__ __forceinline__ uint64_t sig_mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 29);
}

template <typename TM>
__global__ __launch_bounds__(kBlock) void k_vi_sig_hash(int32_t n, const uint64_t *__restrict__ codes,
                                                        const TM *__restrict__ mask, uint64_t *__restrict__ key) {
    constexpr int NW = (int)sizeof(TM);
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    uint64_t h = sig_mix(0x243F6A8885A308D3ull, (uint64_t)mask[r]);
#pragma unroll
    for (int w = 0; w < NW; ++w) h = sig_mix(h, codes[(int64_t)r * NW + w]);
    key[r] = h;
}

// signature id of each row (its key's place among the distinct keys) and, per signature, the
// first row that has it. The minimum is taken in LDS per workgroup first (over the rows that
// start a run of equal keys: only those can be a signature's first row), then each workgroup
// offers its minimum to the global one, skipped when that is already smaller (it only decreases,
// so an old value read here only lets an offer through). Transfer operators' signatures alternate
// from row to row (even and odd grid points): one global atomic per row queued the rows of a 16.7M-row
// interpolation operator on 20 addresses for 97 ms.
__global__ __launch_bounds__(kBlock) void k_vi_sig_assign(int32_t n, const uint64_t *__restrict__ key,
                                                          const uint64_t *__restrict__ uniq,
                                                          const int32_t *__restrict__ nuniq, uint8_t *__restrict__ sig,
                                                          int32_t *__restrict__ rep) {
    static_assert(kBlock >= 256, "one thread per signature");
    __shared__ uint64_t u[256];
    __shared__ int32_t first[256];
    const int nu = *nuniq;
    if (threadIdx.x < 256) {
        if ((int)threadIdx.x < nu) u[threadIdx.x] = uniq[threadIdx.x];
        first[threadIdx.x] = INT32_MAX;
    }
    __syncthreads();
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r < n) {
        const uint64_t k = key[r];
        int lo = 0, hi = nu - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (u[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        sig[r] = (uint8_t)lo;
        if (r == 0 || key[r - 1] != k) atomicMin(&first[lo], r);
    }
    __syncthreads();
    if ((int)threadIdx.x < nu) {
        const int32_t f = first[threadIdx.x];
        if (f != INT32_MAX && f < __hip_atomic_load(&rep[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&rep[threadIdx.x], f);
    }
}

template <typename TM>
__global__ void k_vi_sig_table(int32_t nsig, const int32_t *__restrict__ rep, const uint64_t *__restrict__ codes,
                               const TM *__restrict__ mask, uint64_t *__restrict__ tab) {
    constexpr int NW = (int)sizeof(TM);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsig) return;
    const int32_t r = rep[i];
    tab[(int64_t)i * (NW + 1)] = (uint64_t)mask[r];
#pragma unroll
    for (int w = 0; w < NW; ++w) tab[(int64_t)i * (NW + 1) + 1 + w] = codes[(int64_t)r * NW + w];
}

// every row's mask and codes equal its signature's (a hash collision would break this)
template <typename TM>
__global__ __launch_bounds__(kBlock) void k_vi_sig_check(int32_t n, const uint8_t *__restrict__ sig,
                                                         const uint64_t *__restrict__ tab,
                                                         const uint64_t *__restrict__ codes,
                                                         const TM *__restrict__ mask, int32_t *__restrict__ bad) {
    constexpr int NW = (int)sizeof(TM);
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const uint64_t *t = tab + (int64_t)sig[r] * (NW + 1);
    bool same = t[0] == (uint64_t)mask[r];
#pragma unroll
    for (int w = 0; w < NW; ++w) same = same && t[1 + w] == codes[(int64_t)r * NW + w];
    if (!same) atomicOr(bad, 1);
}

// Replace a VI plan's per-row masks and codes by signature ids when at most 256 distinct
// signatures occur. Leaves the plan as it is otherwise.
template <typename TM>
static int vi_signatures(uhip_ctx ctx, uhip_csr *A) {
    hipStream_t s = ctx->stream;
    const int32_t n = A->nrows;
    constexpr int NW = (int)sizeof(TM);
    const uint64_t *codes = (const uint64_t *)A->dia_val;
    const TM *mask = (const TM *)A->dia_mask;
    uint64_t *key = nullptr, *sorted = nullptr, *uniq = nullptr, *tab = nullptr;
    int32_t *nsel = nullptr, *rep = nullptr;
    uint8_t *sig = nullptr;
    void *tmp = nullptr;
    size_t tb = 0, tb2 = 0;
    auto done = [&](int ret) {
        hipFree(key);
        hipFree(sorted);
        hipFree(uniq);
        hipFree(nsel);
        hipFree(rep);
        hipFree(tmp);
        if (ret != UHIP_OK) {
            hipFree(sig);
            hipFree(tab);
        }
        return ret;
    };
    const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
    if (hipMalloc(&key, sizeof(uint64_t) * n) != hipSuccess || hipMalloc(&sorted, sizeof(uint64_t) * n) != hipSuccess ||
        hipMalloc(&uniq, sizeof(uint64_t) * n) != hipSuccess || hipMalloc(&nsel, sizeof(int32_t) * 2) != hipSuccess)
        return done(set_error(UHIP_ERR_ALLOC, "VI signatures: hipMalloc"));
    hipLaunchKernelGGL((k_vi_sig_hash<TM>), g, b, 0, s, n, codes, mask, key);
    hipcub::DeviceRadixSort::SortKeys(nullptr, tb, key, sorted, n, 0, 64, s);
    hipcub::DeviceSelect::Unique(nullptr, tb2, sorted, uniq, nsel, n, s);
    if (hipMalloc(&tmp, std::max(tb, tb2)) != hipSuccess) return done(set_error(UHIP_ERR_ALLOC, "VI signatures: temp"));
    tb = tb2 = std::max(tb, tb2);
    if (hipcub::DeviceRadixSort::SortKeys(tmp, tb, key, sorted, n, 0, 64, s) != hipSuccess ||
        hipcub::DeviceSelect::Unique(tmp, tb2, sorted, uniq, nsel, n, s) != hipSuccess)
        return done(set_error(UHIP_ERR_HIP, "VI signatures: sort / unique"));
    int32_t nu = 0;
    UHIP_CHECK_HIP(hipMemcpyAsync(&nu, nsel, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    if (nu < 1 || nu > 256) return done(UHIP_OK);  // too many: keep the per-row codes
    std::vector<int32_t> big(256, INT32_MAX);
    if (hipMalloc(&rep, sizeof(int32_t) * 256) != hipSuccess || hipMalloc(&sig, (size_t)n) != hipSuccess ||
        hipMalloc(&tab, sizeof(uint64_t) * (NW + 1) * (size_t)nu) != hipSuccess)
        return done(set_error(UHIP_ERR_ALLOC, "VI signatures: hipMalloc"));
    UHIP_CHECK_HIP(hipMemcpyAsync(rep, big.data(), sizeof(int32_t) * 256, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_vi_sig_assign, g, b, 0, s, n, key, uniq, nsel, sig, rep);
    hipLaunchKernelGGL((k_vi_sig_table<TM>), dim3(1), dim3(256), 0, s, nu, rep, codes, mask, tab);
    UHIP_CHECK_HIP(hipMemsetAsync(nsel + 1, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL((k_vi_sig_check<TM>), g, b, 0, s, n, sig, tab, codes, mask, nsel + 1);
    if (hipGetLastError() != hipSuccess) return done(set_error(UHIP_ERR_HIP, "VI signatures: launch"));
    int32_t bad = 0;
    UHIP_CHECK_HIP(hipMemcpyAsync(&bad, nsel + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    if (bad) return done(UHIP_OK);  // a hash collision: keep the per-row codes
    UHIP_CHECK_HIP(hipFree(A->dia_val));
    UHIP_CHECK_HIP(hipFree(A->dia_mask));
    A->dia_val = tab;
    A->dia_mask = sig;
    A->vi_nsig = nu;
    return done(UHIP_OK);
}

// Plan VI arrays; UHIP_ERR_UNSUPPORTED (nothing kept) when a row is unsorted, the
// rows reach more than 32 distinct offsets from their first column, the values
// more than 256 distinct bit patterns (or any NaN), or ndiag * n > max_ratio * nnz.
static int plan_vi(uhip_ctx ctx, uhip_csr *A, double max_ratio) {
    hipStream_t s = ctx->stream;
    const int32_t n = A->nrows;
    if (n == 0 || A->nnz == 0) return UHIP_ERR_UNSUPPORTED;
    int32_t *table = nullptr;
    unsigned long long *vt = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&table, sizeof(int32_t) * (kDiaSlots + 1)));
    UHIP_CHECK_HIP(hipMalloc(&vt, sizeof(unsigned long long) * kViSlots));
    int32_t *flags = table + kDiaSlots;
    hipLaunchKernelGGL(k_dia_init, dim3(1), dim3(kBlock), 0, s, table, flags);
    hipLaunchKernelGGL(k_vi_init, dim3(1), dim3(kBlock), 0, s, vt);
    hipLaunchKernelGGL(k_dia_scan<true>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->rowptr, A->col,
                       table, flags);
    const dim3 gv(grid_for(A->nnz, 8, 4096));
    if (A->dtype == UHIP_F64)
        hipLaunchKernelGGL(k_vi_scan<double>, gv, dim3(kBlock), 0, s, A->nnz, (const double *)A->val, vt, flags);
    else
        hipLaunchKernelGGL(k_vi_scan<float>, gv, dim3(kBlock), 0, s, A->nnz, (const float *)A->val, vt, flags);
    UHIP_CHECK_LAUNCH();
    int32_t h[kDiaSlots + 1];
    std::vector<unsigned long long> hv(kViSlots);
    UHIP_CHECK_HIP(hipMemcpyAsync(h, table, sizeof(h), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipMemcpyAsync(hv.data(), vt, sizeof(unsigned long long) * kViSlots, hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    std::vector<int32_t> offs;
    for (int i = 0; i < kDiaSlots; ++i)
        if (h[i] != kDiaEmpty) offs.push_back(h[i]);
    std::sort(offs.begin(), offs.end());
    std::vector<unsigned long long> bits;
    for (unsigned long long b : hv)
        if (b != kViEmpty) bits.push_back(b);
    std::sort(bits.begin(), bits.end());
    const int32_t nd = (int32_t)offs.size(), nv = (int32_t)bits.size();
    if (h[kDiaSlots] != 0 || nd < 1 || nd > 32 || nv < 1 || nv > 256 ||
        (double)nd * (double)n > max_ratio * (double)A->nnz) {
        hipFree(table);
        hipFree(vt);
        return UHIP_ERR_UNSUPPORTED;
    }
    const size_t vb = A->dtype == UHIP_F64 ? sizeof(double) : sizeof(float);
    const size_t mb = nd <= 8 ? 1 : (nd <= 16 ? 2 : 4);
    UHIP_CHECK_HIP(hipMemcpyAsync(table, offs.data(), sizeof(int32_t) * (size_t)nd, hipMemcpyHostToDevice, s));
    UHIP_CHECK_HIP(hipMemcpyAsync(vt, bits.data(), sizeof(unsigned long long) * (size_t)nv, hipMemcpyHostToDevice, s));
    UHIP_CHECK_HIP(hipMalloc(&A->dia_val, sizeof(uint64_t) * mb * (size_t)n));
    UHIP_CHECK_HIP(hipMalloc(&A->dia_mask, mb * (size_t)n));
    UHIP_CHECK_HIP(hipMalloc(&A->vi_base, sizeof(int32_t) * (size_t)n));
    UHIP_CHECK_HIP(hipMalloc(&A->vi_table, vb * 256));
    const dim3 g((n + kBlock - 1) / kBlock), b(kBlock);
#define UHIP_VI_FILL(T, TM)                                                                                      \
    hipLaunchKernelGGL((k_vi_fill<T, TM>), g, b, 0, s, n, nd, table, nv, vt, A->rowptr, A->col, (const T *)A->val, \
                       A->vi_base, (uint64_t *)A->dia_val, (TM *)A->dia_mask)
    if (A->dtype == UHIP_F64) {
        if (mb == 1) UHIP_VI_FILL(double, uint8_t);
        else if (mb == 2) UHIP_VI_FILL(double, uint16_t);
        else UHIP_VI_FILL(double, uint32_t);
        hipLaunchKernelGGL(k_vi_table<double>, dim3(1), dim3(256), 0, s, nv, vt, (double *)A->vi_table);
    } else {
        if (mb == 1) UHIP_VI_FILL(float, uint8_t);
        else if (mb == 2) UHIP_VI_FILL(float, uint16_t);
        else UHIP_VI_FILL(float, uint32_t);
        hipLaunchKernelGGL(k_vi_table<float>, dim3(1), dim3(256), 0, s, nv, vt, (float *)A->vi_table);
    }
#undef UHIP_VI_FILL
    UHIP_CHECK_LAUNCH();
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    hipFree(vt);
    A->dia_off = table;  // first nd slots hold the sorted offsets
    A->ndiag = nd;
    A->vi_nvals = nv;
    A->vi_nsig = 0;
    A->format = UHIP_FMT_VI;
    if (mb == 1) UHIP_TRY(vi_signatures<uint8_t>(ctx, A));
    else if (mb == 2) UHIP_TRY(vi_signatures<uint16_t>(ctx, A));
    else UHIP_TRY(vi_signatures<uint32_t>(ctx, A));
    const int32_t rows_per_wg = kDiaWG * vi_rows_per_lane(A->dtype, nd);
    A->nparts = (n + rows_per_wg - 1) / rows_per_wg;
    return UHIP_OK;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag(int32_t n, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ col, const T *__restrict__ val,
                                                 T *__restrict__ d) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    T v = T(0);
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
        if (col[j] == r) v = val[j];
    d[r] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fix_semidef(int32_t n, const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col, T *__restrict__ val) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
        if (col[j] == r && fabs((double)val[j]) < 1e-12) val[j] = val[j] + T(1);
}

// A_rr += alpha*d_r where row r stores its diagonal (MatDiagonalSet with ADD_VALUES,
// PetscMatrix::shift_diag, backend/petsc/utopia_petsc_Matrix.cpp:1783-1788)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_shift_diag(int32_t n, const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ col, T *__restrict__ val,
                                                       const T *__restrict__ d, T alpha) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
        if (col[j] == r) val[j] = val[j] + (alpha == T(1) ? d[r] : alpha * d[r]);
}

// ind_r = 1 if row r has an off-diagonal entry with |a| > tol, else 0.
template <typename TA, typename TV>
__global__ __launch_bounds__(kBlock) void k_offdiag_ind(int32_t n, const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col,
                                                        const TA *__restrict__ val, double tol,
                                                        TV *__restrict__ ind) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    TV v = TV(0);
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
        if (col[j] != r && fabs((double)val[j]) > tol) v = TV(1);
    ind[r] = v;
}

// out_r = min(1, sum_j [|R_rj| > tol] * in_j)
template <typename TA, typename TV>
__global__ __launch_bounds__(kBlock) void k_bool_restrict(int32_t n, const int32_t *__restrict__ rowptr,
                                                          const int32_t *__restrict__ col,
                                                          const TA *__restrict__ val, double tol,
                                                          const TV *__restrict__ in, TV *__restrict__ out) {
    const int32_t r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    TV s = TV(0);
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j)
        s += (fabs((double)val[j]) > tol ? TV(1) : TV(0)) * in[col[j]];
    out[r] = s > TV(1) ? TV(1) : s;
}

__global__ __launch_bounds__(kBlock) void k_cast_vals(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

template <typename TA, typename TV, typename TAcc = TV>
static int spmv_typed(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y) {
    EpiStore<TAcc, TV> e{(TV *)y};
    return launch_rowblock<TAcc>(ctx, A, (const TV *)x, e);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_one(int64_t n, T *__restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) x[i] = T(1);
}

static int plan_format(uhip_ctx ctx, uhip_csr *A, int32_t format, int32_t block_nnz, int walk);

// Autotuner: time y = A x (x = 1) for every applicable layout - DIA (27-point plans with and
// without the plane-walking kernel), VI, SELL-64, row-block CSR with 1024 / 2048 / 4096
// nonzeros per workgroup - and keep the fastest plan. Each candidate runs reps times after
// one untimed launch.
static int autotune_plan(uhip_ctx ctx, uhip_csr *A, int reps, float *best_ms_out) {
    struct Cand {
        int32_t format, block_nnz;
        int walk;
    };
    const Cand cands[] = {{UHIP_FMT_DIA, 2048, 2},      {UHIP_FMT_DIA, 2048, 0},      {UHIP_FMT_VI, 2048, 1},
                          {UHIP_FMT_SELL64, 2048, 1},   {UHIP_FMT_ROWBLOCK, 1024, 1}, {UHIP_FMT_ROWBLOCK, 2048, 1},
                          {UHIP_FMT_ROWBLOCK, 4096, 1}};
    const size_t vb = A->dtype == UHIP_F64 ? sizeof(double) : sizeof(float);
    void *x = nullptr, *y = nullptr;
    UHIP_CHECK_HIP(hipMalloc(&x, vb * (size_t)std::max(A->ncols, 1)));
    UHIP_CHECK_HIP(hipMalloc(&y, vb * (size_t)std::max(A->nrows, 1)));
    const dim3 g(grid_for(std::max(A->ncols, 1), 1, 1 << 14)), b(kBlock);
    if (A->dtype == UHIP_F64) hipLaunchKernelGGL(k_fill_one<double>, g, b, 0, ctx->stream, A->ncols, (double *)x);
    else hipLaunchKernelGGL(k_fill_one<float>, g, b, 0, ctx->stream, A->ncols, (float *)x);
    hipEvent_t e0, e1;
    UHIP_CHECK_HIP(hipEventCreate(&e0));
    UHIP_CHECK_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    Cand win{UHIP_FMT_ROWBLOCK, 2048, 1};
    int st = UHIP_OK;
    bool walked = false;  // the DIA plan with the walk allowed took the 27-point kernel
    for (const Cand &c : cands) {
        if (c.format == UHIP_FMT_DIA && c.walk == 0 && !walked) continue;  // the same plan as before
        st = plan_format(ctx, A, c.format, c.block_nnz, c.walk);
        if (st == UHIP_ERR_UNSUPPORTED) continue;  // DIA does not apply
        if (st != UHIP_OK) break;
        if (c.format == UHIP_FMT_DIA && c.walk == 2) walked = A->dia_w27 != 0;
        auto run = [&]() {
            return A->dtype == UHIP_F64 ? spmv_typed<double, double>(ctx, A, x, y) : spmv_typed<float, float>(ctx, A, x, y);
        };
        if ((st = run()) != UHIP_OK) break;
        // Each launch timed on its own, median of reps: in a solve every product waits for the
        // previous kernel, so a launch's own duration is what counts. Back-to-back launches of
        // small kernels all time at the dispatch rate, which ranks candidates at random.
        std::vector<float> t;
        for (int r = 0; r < reps && st == UHIP_OK; ++r) {
            hipEventRecord(e0, ctx->stream);
            st = run();
            hipEventRecord(e1, ctx->stream);
            hipEventSynchronize(e1);
            float one = 0.f;
            hipEventElapsedTime(&one, e0, e1);
            t.push_back(one);
        }
        if (st != UHIP_OK) break;
        std::sort(t.begin(), t.end());
        const float ms = t[t.size() / 2];
        if (ms < best) {
            best = ms;
            win = c;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(x);
    hipFree(y);
    if (st != UHIP_OK && st != UHIP_ERR_UNSUPPORTED) return st;
    if (best_ms_out) *best_ms_out = best;
    return plan_format(ctx, A, win.format, win.block_nnz, win.walk);
}

}  // namespace uhip

using namespace uhip;

extern "C" {

int uhip_csr_plan(uhip_ctx ctx, uhip_csr *A, int32_t block_nnz) {
    return uhip_csr_plan_format(ctx, A, block_nnz == 0 ? UHIP_FMT_AUTO : UHIP_FMT_ROWBLOCK, block_nnz);
}

int uhip_ctx_set_autotune(uhip_ctx ctx, int reps) {
    UHIP_REQUIRE(ctx && reps >= 0, "uhip_ctx_set_autotune: bad argument");
    ctx->autotune_reps = reps;
    return UHIP_OK;
}

int uhip_csr_autotune(uhip_ctx ctx, uhip_csr *A, int reps, int32_t *format, float *ms) {
    UHIP_REQUIRE(ctx && A && A->rowptr && reps > 0, "uhip_csr_autotune: bad argument");
    if (A->nrows == 0 || A->nnz == 0) return uhip_csr_plan_format(ctx, A, UHIP_FMT_ROWBLOCK, 2048);
    UHIP_TRY(autotune_plan(ctx, A, reps, ms));
    if (format) *format = A->format;
    return UHIP_OK;
}

int uhip_csr_plan_format(uhip_ctx ctx, uhip_csr *A, int32_t format, int32_t block_nnz) {
    UHIP_REQUIRE(ctx && A && A->rowptr, "uhip_csr_plan: null argument");
    UHIP_REQUIRE(A->nrows >= 0 && A->ncols >= 0, "uhip_csr_plan: bad shape");
    UHIP_REQUIRE(format == UHIP_FMT_AUTO || format == UHIP_FMT_ROWBLOCK || format == UHIP_FMT_SELL64 ||
                     format == UHIP_FMT_DIA || format == UHIP_FMT_VI,
                 "uhip_csr_plan: format %d", format);
    if (format == UHIP_FMT_AUTO && ctx->autotune_reps > 0 && A->nrows > 0 && A->nnz > 0)
        return autotune_plan(ctx, A, ctx->autotune_reps, nullptr);
    return plan_format(ctx, A, format, block_nnz, format == UHIP_FMT_DIA ? 2 : 1);
}

}  // extern "C"

namespace uhip {

static int plan_format(uhip_ctx ctx, uhip_csr *A, int32_t format, int32_t block_nnz, int walk) {
    if (block_nnz == 0) block_nnz = 2048;
    int capk;
    if (block_nnz == 1024) capk = 5;
    else if (block_nnz == 2048) capk = 10;
    else if (block_nnz == 4096) capk = 20;
    else return set_error(UHIP_ERR_INVALID, "uhip_csr_plan: block_nnz %d not in {1024,2048,4096}", block_nnz);
    UHIP_TRY(uhip_csr_plan_destroy(ctx, A));
    const int32_t n = A->nrows;
    hipStream_t s = ctx->stream;
    A->block_capk = capk;
    A->format = UHIP_FMT_ROWBLOCK;
    if (format == UHIP_FMT_DIA || (format == UHIP_FMT_AUTO && n > 0 && A->nnz > 0)) {
        const int st = plan_dia(ctx, A, format == UHIP_FMT_DIA ? 1e30 : 1.25, walk);
        if (st == UHIP_OK) return UHIP_OK;
        if (st != UHIP_ERR_UNSUPPORTED) return st;
        if (format == UHIP_FMT_DIA)
            return set_error(UHIP_ERR_UNSUPPORTED,
                             "uhip_csr_plan: DIA needs a square matrix with sorted rows and <= 32 distinct offsets");
    }
    if (format == UHIP_FMT_VI || (format == UHIP_FMT_AUTO && n > 0 && A->nnz > 0)) {
        // automatic with <= 3 slots per stored entry (a trilinear interpolation has 2.4-2.7)
        const int st = plan_vi(ctx, A, format == UHIP_FMT_VI ? 1e30 : 3.0);
        if (st == UHIP_OK) return UHIP_OK;
        if (st != UHIP_ERR_UNSUPPORTED) return st;
        if (format == UHIP_FMT_VI)
            return set_error(UHIP_ERR_UNSUPPORTED, "uhip_csr_plan: VI needs sorted rows, <= 32 offsets from each "
                                                   "row's first column and <= 256 distinct values");
    }
    if (n > 0 && A->nnz > 0 && format != UHIP_FMT_ROWBLOCK && format != UHIP_FMT_VI) {
        const int st = plan_sell(ctx, A, format == UHIP_FMT_SELL64 ? 1e30 : 1.25);
        if (st == UHIP_OK) return UHIP_OK;
        if (st != UHIP_ERR_UNSUPPORTED) return st;
    }
    if (n == 0) {
        A->nblocks = 0;
        A->nparts = 0;
        UHIP_CHECK_HIP(hipMalloc(&A->blocks, sizeof(int32_t)));
        UHIP_CHECK_HIP(hipMemsetAsync(A->blocks, 0, sizeof(int32_t), s));
        return UHIP_OK;
    }
    uint8_t *flags = nullptr;
    int32_t *out = nullptr, *count = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    UHIP_CHECK_HIP(hipMalloc(&flags, n));
    UHIP_CHECK_HIP(hipMalloc(&out, sizeof(int32_t) * ((size_t)n + 1)));
    UHIP_CHECK_HIP(hipMalloc(&count, sizeof(int32_t)));
    hipLaunchKernelGGL(k_plan_flags, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, A->rowptr, block_nnz,
                       flags);
    UHIP_CHECK_LAUNCH();
    hipcub::CountingInputIterator<int32_t> it(0);
    UHIP_CHECK_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, it, flags, out, count, n, s));
    UHIP_CHECK_HIP(hipMalloc(&tmp, tmp_bytes));
    UHIP_CHECK_HIP(hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, it, flags, out, count, n, s));
    hipLaunchKernelGGL(k_plan_tail, dim3(1), dim3(1), 0, s, out, count, n);
    UHIP_CHECK_LAUNCH();
    int32_t nb = 0;
    UHIP_CHECK_HIP(hipMemcpyAsync(&nb, count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    UHIP_CHECK_HIP(hipMalloc(&A->blocks, sizeof(int32_t) * ((size_t)nb + 1)));
    UHIP_CHECK_HIP(hipMemcpyAsync(A->blocks, out, sizeof(int32_t) * ((size_t)nb + 1), hipMemcpyDeviceToDevice, s));
    UHIP_CHECK_HIP(hipStreamSynchronize(s));
    hipFree(flags);
    hipFree(out);
    hipFree(count);
    hipFree(tmp);
    A->nblocks = nb;
    A->nparts = nb;
    return UHIP_OK;
}

}  // namespace uhip

extern "C" {

int uhip_csr_plan_destroy(uhip_ctx ctx, uhip_csr *A) {
    UHIP_REQUIRE(A, "uhip_csr_plan_destroy: null matrix");
    if (A->tr) {
        uhip_csr *t = A->tr;
        A->tr = nullptr;
        const int st = uhip_csr_free(ctx, t);
        delete t;
        if (st != UHIP_OK) return st;
    }
    if (A->blocks || A->slice_off || A->sell_col || A->sell_val || A->dia_off || A->dia_val || A->dia_mask ||
        A->dia_delta || A->vi_base || A->vi_table) {
        if (ctx) UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (A->dia_off) UHIP_CHECK_HIP(hipFree(A->dia_off));
    if (A->dia_val) UHIP_CHECK_HIP(hipFree(A->dia_val));
    if (A->dia_mask) UHIP_CHECK_HIP(hipFree(A->dia_mask));
    if (A->vi_base) UHIP_CHECK_HIP(hipFree(A->vi_base));
    if (A->vi_table) UHIP_CHECK_HIP(hipFree(A->vi_table));
    A->vi_base = nullptr;
    A->vi_table = nullptr;
    A->vi_nvals = 0;
    A->dia_off = nullptr;
    A->dia_val = nullptr;
    A->dia_mask = nullptr;
    A->ndiag = 0;
    A->dia_sym = 0;
    A->dia_nan = 0;
    A->dia_s1 = A->dia_s2 = 0;
    A->vi_nsig = 0;
    A->dia_w27 = 0;
    if (A->dia_delta) UHIP_CHECK_HIP(hipFree(A->dia_delta));
    A->dia_delta = nullptr;
    if (A->blocks) UHIP_CHECK_HIP(hipFree(A->blocks));
    if (A->slice_off) UHIP_CHECK_HIP(hipFree(A->slice_off));
    if (A->sell_col) UHIP_CHECK_HIP(hipFree(A->sell_col));
    if (A->sell_val) UHIP_CHECK_HIP(hipFree(A->sell_val));
    A->blocks = nullptr;
    A->nblocks = 0;
    A->slice_off = nullptr;
    A->sell_col = nullptr;
    A->sell_val = nullptr;
    A->nslices = 0;
    A->nparts = 0;
    A->format = UHIP_FMT_ROWBLOCK;
    return UHIP_OK;
}

int uhip_spmv(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y) {
    UHIP_REQUIRE(ctx && A && x && y, "uhip_spmv: null argument");
    if (A->dtype == UHIP_F64) return spmv_typed<double, double>(ctx, A, x, y);
    if (A->dtype == UHIP_F32) return spmv_typed<float, float>(ctx, A, x, y);
    return set_error(UHIP_ERR_INVALID, "uhip_spmv: dtype %d", A->dtype);
}

int uhip_csr_plan_transpose(uhip_ctx ctx, uhip_csr *A) {
    UHIP_REQUIRE(ctx && A, "uhip_csr_plan_transpose: null argument");
    if (A->tr) return UHIP_OK;
    auto *t = new uhip_csr();
    memset(t, 0, sizeof(*t));
    int st = uhip_csr_transpose(ctx, A, t);
    if (st == UHIP_OK) st = uhip_csr_plan(ctx, t, 0);
    if (st != UHIP_OK) {
        uhip_csr_free(ctx, t);
        delete t;
        return st;
    }
    A->tr = t;
    return UHIP_OK;
}

int uhip_spmv_transpose(uhip_ctx ctx, uhip_csr *A, const void *x, void *y) {
    UHIP_REQUIRE(ctx && A && x && y, "uhip_spmv_transpose: null argument");
    UHIP_TRY(uhip_csr_plan_transpose(ctx, A));
    return uhip_spmv(ctx, A->tr, x, y);
}

int uhip_spmv_axpby(uhip_ctx ctx, const uhip_csr *A, double alpha, const void *x, double beta, void *y) {
    UHIP_REQUIRE(ctx && A && x && y, "uhip_spmv_axpby: null argument");
    if (A->dtype == UHIP_F64) {
        EpiAxpby<double, double> e{(double *)y, alpha, beta};
        return launch_rowblock<double>(ctx, A, (const double *)x, e);
    }
    EpiAxpby<float, float> e{(float *)y, (float)alpha, (float)beta};
    return launch_rowblock<float>(ctx, A, (const float *)x, e);
}

int uhip_spmv_axpbz(uhip_ctx ctx, const uhip_csr *A, double alpha, const void *x, double beta, const void *z,
                    void *y) {
    UHIP_REQUIRE(ctx && A && x && z && y, "uhip_spmv_axpbz: null argument");
    if (A->dtype == UHIP_F64) {
        EpiAxpbz<double, double> e{(const double *)z, (double *)y, alpha, beta};
        return launch_rowblock<double>(ctx, A, (const double *)x, e);
    }
    EpiAxpbz<float, float> e{(const float *)z, (float *)y, (float)alpha, (float)beta};
    return launch_rowblock<float>(ctx, A, (const float *)x, e);
}

int uhip_spmv_residual(uhip_ctx ctx, const uhip_csr *A, const void *x, const void *b, int b_dtype, void *r) {
    UHIP_REQUIRE(ctx && A && x && b && r, "uhip_spmv_residual: null argument");
    if (A->dtype == UHIP_F64) {
        UHIP_REQUIRE(b_dtype == UHIP_F64, "uhip_spmv_residual: FP64 matrix needs FP64 b");
        EpiResidual<double, double, double> e{(const double *)b, (double *)r};
        return launch_rowblock<double>(ctx, A, (const double *)x, e);
    }
    if (b_dtype == UHIP_F64) {
        EpiResidual<float, double, float> e{(const double *)b, (float *)r};
        return launch_rowblock<float>(ctx, A, (const float *)x, e);
    }
    EpiResidual<float, float, float> e{(const float *)b, (float *)r};
    return launch_rowblock<float>(ctx, A, (const float *)x, e);
}

int uhip_spmv_dot(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y, double *dot_d) {
    UHIP_REQUIRE(ctx && A && x && y && dot_d, "uhip_spmv_dot: null argument");
    UHIP_REQUIRE(A->nrows == A->ncols, "uhip_spmv_dot: matrix must be square");
    if (A->dtype == UHIP_F64) {
        EpiStoreDot<double, double, double> e{(const double *)x, (double *)y};
        UHIP_TRY(launch_rowblock<double>(ctx, A, (const double *)x, e));
    } else {
        EpiStoreDot<float, float, float> e{(const float *)x, (float *)y};
        UHIP_TRY(launch_rowblock<float>(ctx, A, (const float *)x, e));
    }
    return finish_sums(ctx, 1, A->nparts, ctx->max_partials, dot_d);
}

int uhip_spmv_dot_noalloc(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y, double *dot_d) {
    UHIP_REQUIRE(ctx && A && x && y && dot_d, "uhip_spmv_dot_noalloc: null argument");
    UHIP_REQUIRE(A->nrows == A->ncols, "uhip_spmv_dot_noalloc: matrix must be square");
    const int64_t need = A->format == UHIP_FMT_ROWBLOCK ? A->nblocks : A->nparts;
    UHIP_REQUIRE(need <= kNoAllocParts,
                 "uhip_spmv_dot_noalloc: the plan has %lld workgroups, the scratch holds %lld partial sums "
                 "(at most about 2^18 workgroups; use uhip_spmv_dot)",
                 (long long)need, (long long)kNoAllocParts);
    double *parts = scratch_of(ctx->device);
    UHIP_REQUIRE(parts, "uhip_spmv_dot_noalloc: scratch not initialised on device %d", ctx->device);
    if (A->dtype == UHIP_F64) {
        EpiStoreDot<double, double, double> e{(const double *)x, (double *)y};
        UHIP_TRY(launch_rowblock<double>(ctx, A, (const double *)x, e, parts, kNoAllocParts));
    } else {
        EpiStoreDot<float, float, float> e{(const float *)x, (float *)y};
        UHIP_TRY(launch_rowblock<float>(ctx, A, (const float *)x, e, parts, kNoAllocParts));
    }
    return finish_sums_from(ctx, parts, 1, A->nparts, kNoAllocParts, dot_d);
}

int uhip_spmv_dot2_f64(uhip_ctx ctx, const uhip_csr *A, const double *x, double *y, const double *z,
                       double *dots_d) {
    UHIP_REQUIRE(ctx && A && x && y && z && dots_d, "uhip_spmv_dot2_f64: null argument");
    UHIP_REQUIRE(A->nrows == A->ncols && A->dtype == UHIP_F64, "uhip_spmv_dot2_f64: square FP64 matrix");
    EpiStoreDot2<double, double, double> e{x, z, y};
    UHIP_TRY(launch_rowblock<double>(ctx, A, x, e));
    return finish_sums(ctx, 2, A->nparts, ctx->max_partials, dots_d);
}

int uhip_spmv_masked(uhip_ctx ctx, const uhip_csr *A, const void *x, const void *mask, void *y) {
    UHIP_REQUIRE(ctx && A && x && y, "uhip_spmv_masked: null argument");
    if (A->dtype == UHIP_F64) {
        EpiMaskStore<double, double> e{(const double *)mask, (double *)y};
        return launch_rowblock<double>(ctx, A, (const double *)x, e);
    }
    EpiMaskStore<float, float> e{(const float *)mask, (float *)y};
    return launch_rowblock<float>(ctx, A, (const float *)x, e);
}

int uhip_spmv_add(uhip_ctx ctx, const uhip_csr *A, const void *x, void *y) {
    UHIP_REQUIRE(ctx && A && x && y, "uhip_spmv_add: null argument");
    if (A->dtype == UHIP_F64) {
        EpiAddTo<double, double> e{(double *)y};
        return launch_rowblock<double>(ctx, A, (const double *)x, e);
    }
    EpiAddTo<float, float> e{(float *)y};
    return launch_rowblock<float>(ctx, A, (const float *)x, e);
}

int uhip_spmv_jacobi(uhip_ctx ctx, const uhip_csr *A, const void *xin, const void *b, int b_dtype, const void *d,
                     void *out, int out_dtype, int add) {
    UHIP_REQUIRE(ctx && A && b && d && out, "uhip_spmv_jacobi: null argument");
    UHIP_REQUIRE(A->nrows == A->ncols, "uhip_spmv_jacobi: matrix must be square");
    UHIP_REQUIRE(out != xin && out != b && out != d, "uhip_spmv_jacobi: out must not alias xin, b or d");
    // a zero guess is a zero vector of the caller's: the reference forms A*0 too
    UHIP_REQUIRE(xin, "uhip_spmv_jacobi: xin required (a zero vector for a zero guess)");
    if (A->dtype == UHIP_F64)
        UHIP_REQUIRE(b_dtype == UHIP_F64 && out_dtype == UHIP_F64, "uhip_spmv_jacobi: FP64 matrix needs FP64 b, out");
#define UHIP_JAC(TACC, TB, TXO)                                                                                 \
    do {                                                                                                        \
        if (add) {                                                                                              \
            EpiJacobi<TACC, TB, TXO, true> e{(const TB *)b, (const TACC *)d, (const TACC *)xin, (TXO *)out};   \
            return launch_rowblock<TACC>(ctx, A, (const TACC *)xin, e);                                         \
        }                                                                                                       \
        EpiJacobi<TACC, TB, TXO, false> e{(const TB *)b, (const TACC *)d, (const TACC *)xin, (TXO *)out};      \
        return launch_rowblock<TACC>(ctx, A, (const TACC *)xin, e);                                             \
    } while (0)
    if (A->dtype == UHIP_F64) UHIP_JAC(double, double, double);
    if (b_dtype == UHIP_F64 && out_dtype == UHIP_F64) UHIP_JAC(float, double, double);
    if (b_dtype == UHIP_F64) UHIP_JAC(float, double, float);
    if (out_dtype == UHIP_F64) UHIP_JAC(float, float, double);
    UHIP_JAC(float, float, float);
#undef UHIP_JAC
}

}  // extern "C"

namespace uhip {

template <typename TAcc, typename TP, typename TX, bool DOT>
static int cheb_typed(uhip_ctx ctx, const uhip_csr *A, const uhip_cheb_step *st) {
    EpiCheb<TAcc, TP, TAcc, TAcc, TX, DOT> e{(const TP *)st->pprev, (const TAcc *)st->rprev,
                                             (const TAcc *)st->xin, (TX *)st->xout,
                                             (TAcc *)st->rout,      (TAcc *)st->pout,
                                             (TAcc)st->a_prev,      (TAcc)st->beta,
                                             (TAcc)st->a_cur,       st->last,
                                             st->dot_with};
    UHIP_TRY(launch_rowblock<TAcc>(ctx, A, (const TP *)st->pprev, e));
    if (DOT) return finish_sums(ctx, 1, A->nparts, ctx->max_partials, st->dot_d);
    return UHIP_OK;
}

template <typename TAcc, typename TP, typename TX>
static int cheb_dispatch(uhip_ctx ctx, const uhip_csr *A, const uhip_cheb_step *st) {
    if (st->dot_with) return cheb_typed<TAcc, TP, TX, true>(ctx, A, st);
    return cheb_typed<TAcc, TP, TX, false>(ctx, A, st);
}

}  // namespace uhip

extern "C" {

int uhip_cheb_step_apply(uhip_ctx ctx, const uhip_csr *A, const uhip_cheb_step *st) {
    UHIP_REQUIRE(ctx && A && st && st->pprev && st->xout, "uhip_cheb_step_apply: null argument");
    UHIP_REQUIRE(st->last || (st->rout && st->pout), "uhip_cheb_step_apply: rout/pout needed when last == 0");
    UHIP_REQUIRE(!st->dot_with || (st->last && st->dot_d), "uhip_cheb_step_apply: dot only on the last step");
    if (A->dtype == UHIP_F64) {
        UHIP_REQUIRE(st->p_dtype == UHIP_F64 && st->x_dtype == UHIP_F64, "uhip_cheb_step_apply: FP64 level");
        return cheb_dispatch<double, double, double>(ctx, A, st);
    }
    if (st->p_dtype == UHIP_F32 && st->x_dtype == UHIP_F32) return cheb_dispatch<float, float, float>(ctx, A, st);
    if (st->p_dtype == UHIP_F64 && st->x_dtype == UHIP_F32) return cheb_dispatch<float, double, float>(ctx, A, st);
    if (st->p_dtype == UHIP_F32 && st->x_dtype == UHIP_F64) return cheb_dispatch<float, float, double>(ctx, A, st);
    if (st->p_dtype == UHIP_F64 && st->x_dtype == UHIP_F64) return cheb_dispatch<float, double, double>(ctx, A, st);
    return set_error(UHIP_ERR_INVALID, "uhip_cheb_step_apply: dtypes");
}

int uhip_csr_diag(uhip_ctx ctx, const uhip_csr *A, void *d) {
    UHIP_REQUIRE(ctx && A && d, "uhip_csr_diag: null argument");
    const int32_t n = A->nrows;
    if (n == 0) return UHIP_OK;
    dim3 g((n + kBlock - 1) / kBlock);
    if (A->dtype == UHIP_F64)
        hipLaunchKernelGGL(k_diag<double>, g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (const double *)A->val, (double *)d);
    else
        hipLaunchKernelGGL(k_diag<float>, g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (const float *)A->val, (float *)d);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_csr_shift_diag(uhip_ctx ctx, uhip_csr *A, const void *d, double alpha) {
    UHIP_REQUIRE(ctx && A && d, "uhip_csr_shift_diag: null argument");
    const int32_t n = A->nrows < A->ncols ? A->nrows : A->ncols;
    if (n == 0) return UHIP_OK;
    dim3 g((n + kBlock - 1) / kBlock);
    if (A->dtype == UHIP_F64)
        hipLaunchKernelGGL(k_shift_diag<double>, g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (double *)A->val, (const double *)d, alpha);
    else
        hipLaunchKernelGGL(k_shift_diag<float>, g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (float *)A->val, (const float *)d, (float)alpha);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_csr_fix_semidefinite(uhip_ctx ctx, uhip_csr *A) {
    UHIP_REQUIRE(ctx && A, "uhip_csr_fix_semidefinite: null argument");
    const int32_t n = A->nrows;
    if (n == 0) return UHIP_OK;
    dim3 g((n + kBlock - 1) / kBlock);
    if (A->dtype == UHIP_F64)
        hipLaunchKernelGGL(k_fix_semidef<double>, g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (double *)A->val);
    else
        hipLaunchKernelGGL(k_fix_semidef<float>, g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (float *)A->val);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_csr_cast_values_f64_f32(uhip_ctx ctx, int64_t nnz, const double *src, float *dst) {
    UHIP_REQUIRE(ctx && (nnz == 0 || (src && dst)), "uhip_csr_cast_values_f64_f32: null argument");
    if (nnz == 0) return UHIP_OK;
    hipLaunchKernelGGL(k_cast_vals, dim3(grid_for(nnz)), dim3(kBlock), 0, ctx->stream, nnz, src, dst);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_csr_offdiag_indicator(uhip_ctx ctx, const uhip_csr *A, double tol, void *ind) {
    UHIP_REQUIRE(ctx && A && ind, "uhip_csr_offdiag_indicator: null argument");
    const int32_t n = A->nrows;
    if (n == 0) return UHIP_OK;
    dim3 g((n + kBlock - 1) / kBlock);
    if (A->dtype == UHIP_F64)
        hipLaunchKernelGGL((k_offdiag_ind<double, double>), g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (const double *)A->val, tol, (double *)ind);
    else
        hipLaunchKernelGGL((k_offdiag_ind<float, float>), g, dim3(kBlock), 0, ctx->stream, n, A->rowptr, A->col,
                           (const float *)A->val, tol, (float *)ind);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int uhip_csr_boolean_restrict_or(uhip_ctx ctx, const uhip_csr *R, double tol, const void *ind_f, void *ind_c) {
    UHIP_REQUIRE(ctx && R && ind_f && ind_c, "uhip_csr_boolean_restrict_or: null argument");
    const int32_t n = R->nrows;
    if (n == 0) return UHIP_OK;
    dim3 g((n + kBlock - 1) / kBlock);
    if (R->dtype == UHIP_F64)
        hipLaunchKernelGGL((k_bool_restrict<double, double>), g, dim3(kBlock), 0, ctx->stream, n, R->rowptr, R->col,
                           (const double *)R->val, tol, (const double *)ind_f, (double *)ind_c);
    else
        hipLaunchKernelGGL((k_bool_restrict<float, float>), g, dim3(kBlock), 0, ctx->stream, n, R->rowptr, R->col,
                           (const float *)R->val, tol, (const float *)ind_f, (float *)ind_c);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

}  // extern "C"
