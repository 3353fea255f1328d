// Context, device memory and error plumbing of the C ABI, plus the fixed-order
// second stage shared by every reduction.
#include <cstdarg>
#include <mutex>

#include "uhip_common.hpp"

namespace uhip {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// result[k] = sum_{b < nparts} partials[k*stride + b], one 1024-thread block per k.
// Thread t sums b = t, t+1024, ... in order, then a fixed tree combines the
// threads, so the result is identical from run to run.
__global__ __launch_bounds__(1024) void k_finish_sums(const double *__restrict__ partials, int64_t nparts,
                                                      int64_t stride, double *__restrict__ result,
                                                      unsigned sqrt_mask) {
    const int k = blockIdx.x;
    const double *p = partials + (int64_t)k * stride;
    // eight independent chains keep eight loads in flight; combined in a fixed order
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t b = threadIdx.x;
    for (; b + 7 * 1024 < nparts; b += 8 * 1024) {
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += p[b + u * 1024];
    }
    for (int u = 0; b < nparts; b += 1024, ++u) a[u] += p[b];
    double s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    s = block_sum<1024>(s);
    if (threadIdx.x == 0) result[k] = ((sqrt_mask >> k) & 1u) ? sqrt(s) : s;
}

int finish_sums_from(uhip_ctx ctx, const double *partials, int k, int64_t nparts, int64_t stride, double *result_d,
                     unsigned sqrt_mask) {
    hipLaunchKernelGGL(k_finish_sums, dim3(k), dim3(1024), 0, stream_of(ctx), partials, nparts, stride, result_d,
                       sqrt_mask);
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

int finish_sums(uhip_ctx ctx, int k, int64_t nparts, int64_t stride, double *result_d, unsigned sqrt_mask) {
    return finish_sums_from(ctx, ctx->partials, k, nparts, stride, result_d, sqrt_mask);
}

int ensure_partials(uhip_ctx ctx, int64_t per_slot) {
    if (per_slot <= ctx->max_partials) return UHIP_OK;
    int64_t cap = ctx->max_partials > 0 ? ctx->max_partials : 4096;
    while (cap < per_slot) cap *= 2;
    if (ctx->partials) {
        UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        UHIP_CHECK_HIP(hipFree(ctx->partials));
        ctx->partials = nullptr;
    }
    UHIP_CHECK_HIP(hipMalloc(&ctx->partials, sizeof(double) * kMaxDots * cap));
    ctx->max_partials = cap;
    ++ctx->workspace_gen;
    return UHIP_OK;
}

}  // namespace uhip

using namespace uhip;

extern "C" {

const char *uhip_last_error(void) { return g_last_error.c_str(); }

int64_t uhip_abi_sizeof(int which) {
    if (which == 0) return (int64_t)sizeof(uhip_csr);
    if (which == 1) return (int64_t)sizeof(uhip_cheb_step);
    return -1;
}

int uhip_device_count(int *count) {
    UHIP_REQUIRE(count, "uhip_device_count: null output");
    UHIP_CHECK_HIP(hipGetDeviceCount(count));
    return UHIP_OK;
}

int uhip_ctx_create(int device, void *stream, uhip_ctx *out) {
    UHIP_REQUIRE(out, "uhip_ctx_create: null output");
    *out = nullptr;
    int n = 0;
    UHIP_CHECK_HIP(hipGetDeviceCount(&n));
    UHIP_REQUIRE(device >= 0 && device < n, "uhip_ctx_create: device %d out of range (%d devices)", device, n);
    UHIP_CHECK_HIP(hipSetDevice(device));
    auto *c = new uhip_context_s();
    c->device = device;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        // Default priority on purpose: streams of mixed priority are not used.
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return set_error(UHIP_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
        }
        c->own_stream = true;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    int st = spmv_scratch_init(device);
    if (st == UHIP_OK) st = ensure_partials(c, 1 << 16);
    if (st == UHIP_OK && hipMalloc(&c->dev_scalars, sizeof(double) * UHIP_NUM_SCALARS) != hipSuccess)
        st = set_error(UHIP_ERR_ALLOC, "hipMalloc scalars");
    if (st == UHIP_OK && hipHostMalloc(&c->host_scalars, sizeof(double) * UHIP_NUM_SCALARS) != hipSuccess)
        st = set_error(UHIP_ERR_ALLOC, "hipHostMalloc scalars");
    if (st == UHIP_OK && hipHostMalloc(&c->host_fetch, sizeof(double) * 2 * UHIP_NUM_SCALARS) != hipSuccess)
        st = set_error(UHIP_ERR_ALLOC, "hipHostMalloc fetch scalars");
    for (int k = 0; k < 2; ++k)
        if (st == UHIP_OK && hipEventCreateWithFlags(&c->fetch_done[k], hipEventDisableTiming) != hipSuccess)
            st = set_error(UHIP_ERR_HIP, "hipEventCreate");
    if (st == UHIP_OK && hipMalloc(&c->counters, sizeof(unsigned int) * 64) != hipSuccess)
        st = set_error(UHIP_ERR_ALLOC, "hipMalloc counters");
    if (st == UHIP_OK && hipMemsetAsync(c->dev_scalars, 0, sizeof(double) * UHIP_NUM_SCALARS, c->stream) != hipSuccess)
        st = set_error(UHIP_ERR_HIP, "memset scalars");
    if (st == UHIP_OK && hipMemsetAsync(c->counters, 0, sizeof(unsigned int) * 64, c->stream) != hipSuccess)
        st = set_error(UHIP_ERR_HIP, "memset counters");
    if (st == UHIP_OK && hipStreamSynchronize(c->stream) != hipSuccess) st = set_error(UHIP_ERR_HIP, "sync");
    if (st != UHIP_OK) {
        uhip_ctx_destroy(c);
        return st;
    }
    *out = c;
    return UHIP_OK;
}

int uhip_ctx_destroy(uhip_ctx ctx) {
    if (!ctx) return UHIP_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->partials) hipFree(ctx->partials);
    if (ctx->dev_scalars) hipFree(ctx->dev_scalars);
    if (ctx->host_scalars) hipHostFree(ctx->host_scalars);
    if (ctx->host_fetch) hipHostFree(ctx->host_fetch);
    for (int k = 0; k < 2; ++k)
        if (ctx->fetch_done[k]) hipEventDestroy(ctx->fetch_done[k]);
    if (ctx->counters) hipFree(ctx->counters);
    if (ctx->own_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return UHIP_OK;
}

int64_t uhip_ctx_workspace_generation(uhip_ctx ctx) { return ctx ? ctx->workspace_gen : -1; }

int uhip_pointer_is_device(const void *p) {
    if (!p) return 0;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory: not an error
        return 0;
    }
    return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) ? 1 : 0;
}

int uhip_ctx_set_gate(uhip_ctx ctx, const int32_t *gate_d) {
    UHIP_REQUIRE(ctx, "uhip_ctx_set_gate: null context");
    ctx->gate = gate_d;
    return UHIP_OK;
}

void *uhip_ctx_stream(uhip_ctx ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int uhip_ctx_device(uhip_ctx ctx) { return ctx ? ctx->device : -1; }

double *uhip_ctx_scalars(uhip_ctx ctx) { return ctx ? ctx->dev_scalars : nullptr; }

int uhip_sync(uhip_ctx ctx) {
    UHIP_REQUIRE(ctx, "uhip_sync: null context");
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return UHIP_OK;
}

int uhip_malloc(uhip_ctx ctx, void **ptr, size_t bytes) {
    UHIP_REQUIRE(ctx && ptr, "uhip_malloc: null argument");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    UHIP_CHECK_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return set_error(UHIP_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return UHIP_OK;
}

int uhip_free(uhip_ctx ctx, void *ptr) {
    if (!ptr) return UHIP_OK;
    if (ctx) UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    UHIP_CHECK_HIP(hipFree(ptr));
    return UHIP_OK;
}

int uhip_memcpy_htod(uhip_ctx ctx, void *dst, const void *src, size_t bytes) {
    UHIP_REQUIRE(ctx, "uhip_memcpy_htod: null context");
    if (bytes == 0) return UHIP_OK;
    UHIP_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return UHIP_OK;
}

int uhip_memcpy_dtoh(uhip_ctx ctx, void *dst, const void *src, size_t bytes) {
    UHIP_REQUIRE(ctx, "uhip_memcpy_dtoh: null context");
    if (bytes == 0) return UHIP_OK;
    UHIP_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return UHIP_OK;
}

int uhip_memcpy_dtod(uhip_ctx ctx, void *dst, const void *src, size_t bytes) {
    UHIP_REQUIRE(ctx, "uhip_memcpy_dtod: null context");
    if (bytes == 0 || dst == src) return UHIP_OK;
    UHIP_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return UHIP_OK;
}

int uhip_read_scalars(uhip_ctx ctx, const double *src_d, int n, double *dst_h) {
    UHIP_REQUIRE(ctx && src_d && dst_h, "uhip_read_scalars: null argument");
    UHIP_REQUIRE(n > 0 && n <= UHIP_NUM_SCALARS, "uhip_read_scalars: n=%d out of range", n);
    UHIP_CHECK_HIP(hipMemcpyAsync(ctx->host_scalars, src_d, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(dst_h, ctx->host_scalars, sizeof(double) * n);
    return UHIP_OK;
}

int uhip_fetch_scalars(uhip_ctx ctx, int slot, const double *src_d, int n) {
    UHIP_REQUIRE(ctx && src_d, "uhip_fetch_scalars: null argument");
    UHIP_REQUIRE(slot == 0 || slot == 1, "uhip_fetch_scalars: slot %d (0 or 1)", slot);
    UHIP_REQUIRE(n > 0 && n <= UHIP_NUM_SCALARS, "uhip_fetch_scalars: n=%d out of range", n);
    UHIP_CHECK_HIP(hipMemcpyAsync(ctx->host_fetch + slot * UHIP_NUM_SCALARS, src_d, sizeof(double) * n,
                                  hipMemcpyDeviceToHost, ctx->stream));
    UHIP_CHECK_HIP(hipEventRecord(ctx->fetch_done[slot], ctx->stream));
    return UHIP_OK;
}

int uhip_fetch_wait(uhip_ctx ctx, int slot, int n, double *dst_h) {
    UHIP_REQUIRE(ctx && dst_h, "uhip_fetch_wait: null argument");
    UHIP_REQUIRE(slot == 0 || slot == 1, "uhip_fetch_wait: slot %d (0 or 1)", slot);
    UHIP_REQUIRE(n > 0 && n <= UHIP_NUM_SCALARS, "uhip_fetch_wait: n=%d out of range", n);
    UHIP_CHECK_HIP(hipEventSynchronize(ctx->fetch_done[slot]));
    memcpy(dst_h, ctx->host_fetch + slot * UHIP_NUM_SCALARS, sizeof(double) * n);
    return UHIP_OK;
}

int uhip_write_scalars(uhip_ctx ctx, double *dst_d, int n, const double *src_h) {
    UHIP_REQUIRE(ctx && src_h && dst_d, "uhip_write_scalars: null argument");
    UHIP_REQUIRE(n > 0 && n <= UHIP_NUM_SCALARS, "uhip_write_scalars: n=%d out of range", n);
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(ctx->host_scalars, src_h, sizeof(double) * n);
    UHIP_CHECK_HIP(hipMemcpyAsync(dst_d, ctx->host_scalars, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    UHIP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return UHIP_OK;
}

int uhip_event_timer_start(uhip_ctx ctx, void **ev_pair) {
    UHIP_REQUIRE(ctx && ev_pair, "uhip_event_timer_start: null argument");
    hipEvent_t *e = new hipEvent_t[2];
    UHIP_CHECK_HIP(hipEventCreate(&e[0]));
    UHIP_CHECK_HIP(hipEventCreate(&e[1]));
    UHIP_CHECK_HIP(hipEventRecord(e[0], ctx->stream));
    *ev_pair = e;
    return UHIP_OK;
}

int uhip_event_timer_stop(uhip_ctx ctx, void *ev_pair, float *ms) {
    UHIP_REQUIRE(ctx && ev_pair && ms, "uhip_event_timer_stop: null argument");
    hipEvent_t *e = (hipEvent_t *)ev_pair;
    UHIP_CHECK_HIP(hipEventRecord(e[1], ctx->stream));
    UHIP_CHECK_HIP(hipEventSynchronize(e[1]));
    UHIP_CHECK_HIP(hipEventElapsedTime(ms, e[0], e[1]));
    hipEventDestroy(e[0]);
    hipEventDestroy(e[1]);
    delete[] e;
    return UHIP_OK;
}

int uhip_ctx_reserve(uhip_ctx ctx, int64_t per_slot) {
    UHIP_REQUIRE(ctx && per_slot >= 0, "uhip_ctx_reserve: bad argument");
    return ensure_partials(ctx, per_slot);
}

int uhip_graph_begin(uhip_ctx ctx) {
    UHIP_REQUIRE(ctx && ctx->stream, "uhip_graph_begin: context needs its own (non-null) stream");
    UHIP_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    return UHIP_OK;
}

int uhip_graph_end(uhip_ctx ctx, void **exec) {
    UHIP_REQUIRE(ctx && exec, "uhip_graph_end: null argument");
    *exec = nullptr;
    hipGraph_t g = nullptr;
    UHIP_CHECK_HIP(hipStreamEndCapture(ctx->stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(UHIP_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    *exec = ge;
    return UHIP_OK;
}

int uhip_graph_launch(uhip_ctx ctx, void *exec) {
    UHIP_REQUIRE(ctx && exec, "uhip_graph_launch: null argument");
    UHIP_CHECK_HIP(hipGraphLaunch((hipGraphExec_t)exec, ctx->stream));
    return UHIP_OK;
}

int uhip_graph_destroy(void *exec) {
    if (exec) UHIP_CHECK_HIP(hipGraphExecDestroy((hipGraphExec_t)exec));
    return UHIP_OK;
}

}  // extern "C"
