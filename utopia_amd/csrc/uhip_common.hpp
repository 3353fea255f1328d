// Internal header shared by the HIP translation units of libutopia_hip.
// Not part of the public C ABI (that is include/utopia_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/utopia_hip.h"

namespace uhip {

// Launch geometry shared by all streaming kernels.
constexpr int kBlock = 256;          // 4 wavefronts of 64 lanes
constexpr int kMaxRedBlocks = 1024;  // upper bound on blocks of a grid-stride reduction
constexpr int kMaxDots = 4;          // fused dots per reduction launch
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin

// Record a message for uhip_last_error(); returns code.
int set_error(int code, const char *fmt, ...);

#define UHIP_CHECK_HIP(expr)                                                                  \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::uhip::set_error(UHIP_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #expr, \
                                     hipGetErrorString(e_));                                  \
    } while (0)

#define UHIP_CHECK_LAUNCH() UHIP_CHECK_HIP(hipGetLastError())

#define UHIP_REQUIRE(cond, ...)                                                   \
    do {                                                                          \
        if (!(cond)) return ::uhip::set_error(UHIP_ERR_INVALID, __VA_ARGS__);     \
    } while (0)

#define UHIP_TRY(expr)              \
    do {                            \
        int s_ = (expr);            \
        if (s_ != UHIP_OK) return s_; \
    } while (0)

}  // namespace uhip

// The context: one device, one stream, and the small device workspace every
// reduction writes its per-block partial sums into.
struct uhip_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 256;
    double *partials = nullptr;     // kMaxDots * max_partials doubles
    int64_t max_partials = 0;       // capacity per dot slot
    int64_t workspace_gen = 0;      // bumped whenever partials is (re)allocated
    const int32_t *gate = nullptr;  // device counter: batched kernels skip their work while it reads 0
    double *dev_scalars = nullptr;  // UHIP_NUM_SCALARS device scalar slots
    double *host_scalars = nullptr; // pinned mirror used for read-back
    double *host_fetch = nullptr;        // pinned targets of the asynchronous read-backs (2 slots)
    hipEvent_t fetch_done[2] = {nullptr, nullptr}; // recorded after each slot's read-back
    unsigned int *counters = nullptr;
    int autotune_reps = 0;          // > 0: automatic SpMV plans are chosen by timing
};

namespace uhip {

inline hipStream_t stream_of(uhip_ctx ctx) { return ctx ? ctx->stream : nullptr; }

// Workgroup index remap so that each XCD walks a contiguous slice of the
// logical grid (the hardware hands workgroup b to XCD b % 8). Keeps
// neighbouring rows - and the x entries they share - in one XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int xcd = b % kNumXcd;
    const int local = b / kNumXcd;
    const int per = n / kNumXcd;
    const int rem = n % kNumXcd;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// Wavefront (64-lane) sum via cross-lane shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum for blockDim.x == BS (multiple of 64); result valid in thread 0.
template <int BS, typename T>
__device__ __forceinline__ T block_sum(T v) {
    __shared__ T red[BS / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < BS / 64; ++i) s += red[i];
    }
    __syncthreads();
    return s;
}

inline int grid_for(int64_t n, int per_thread = 4, int cap = 1 << 20) {
    int64_t g = (n + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// Deterministic second stage of all reductions: result[k] = sum_b partials[k*stride + b]
// summed in a fixed order (thread-strided, then a fixed tree); slot k is
// replaced by its square root when bit k of sqrt_mask is set.
int finish_sums(uhip_ctx ctx, int k, int64_t nparts, int64_t stride, double *result_d, unsigned sqrt_mask = 0);
// The same second stage over partial sums held elsewhere than ctx->partials.
int finish_sums_from(uhip_ctx ctx, const double *partials, int k, int64_t nparts, int64_t stride, double *result_d,
                     unsigned sqrt_mask = 0);
// Device address of the SpMV library's file-scope partial-sum scratch on `device` (the device
// must be current): looked up once per device at context creation.
int spmv_scratch_init(int device);

// Grow the per-slot capacity of ctx->partials to at least per_slot entries.
int ensure_partials(uhip_ctx ctx, int64_t per_slot);

}  // namespace uhip
