/*
 * api.hip - context, memory and the volume/points entries of the C ABI.
 */
#include <cmath>
#include <cstring>
#include <mutex>

#include "internal.hpp"

namespace lyap {

int fail(lyap_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

int ctx_alloc(lyap_ctx *ctx, void **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) return LYAP_OK;
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
        return fail(ctx, LYAP_ERR_BUDGET,
                    "allocation of %zu bytes exceeds the budget (%zu of %zu in use)",
                    bytes, ctx->in_use, ctx->budget);
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(ctx, LYAP_ERR_DEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    ctx->allocations[*ptr] = bytes;
    ctx->in_use += bytes;
    return LYAP_OK;
}

void ctx_free(lyap_ctx *ctx, void *ptr)
{
    if (!ptr) return;
    auto it = ctx->allocations.find(ptr);
    if (it != ctx->allocations.end()) {
        ctx->in_use -= it->second;
        ctx->allocations.erase(it);
    }
    (void)hipFree(ptr);
}

int ensure_reeval_list(lyap_ctx *ctx, size_t entries)
{
    if (entries <= ctx->reeval_cap) return LYAP_OK;
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx_free(ctx, ctx->reeval_list);
    ctx->reeval_list = nullptr;
    ctx->reeval_cap = 0;
    void *p = nullptr;
    int rc = ctx_alloc(ctx, &p, entries * sizeof(uint32_t));
    if (rc) return rc;
    ctx->reeval_list = (uint32_t *)p;
    ctx->reeval_cap = entries;
    return LYAP_OK;
}

static int check_params(lyap_ctx *ctx, const LyapParams *prm, const int32_t *seq,
                        EvalDesc *e, uint32_t renorm)
{
    if (!prm || !seq) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null parameters or sequence");
    if (!make_eval_desc(e, *prm, seq, renorm))
        return fail(ctx, LYAP_ERR_BAD_SEQUENCE,
                    "sequence must hold 1..%d letter indices 0..3 before the -1",
                    LYAP_MAX_SEQ);
    return LYAP_OK;
}

int timing_mark(lyap_ctx *ctx)
{
    if (!ctx->timing) return LYAP_OK;
    if (ctx->timing_used == ctx->timing_events.size()) {
        hipEvent_t e;
        LYAP_HIP(ctx, hipEventCreate(&e));
        ctx->timing_events.push_back(e);
    }
    LYAP_HIP(ctx, hipEventRecord(ctx->timing_events[ctx->timing_used], ctx->stream));
    ctx->timing_used++;
    return LYAP_OK;
}

} /* namespace lyap */

using namespace lyap;

extern "C" {

const char *lyap_strerror(int status)
{
    switch (status) {
    case LYAP_OK: return "ok";
    case LYAP_ERR_INVALID_ARGUMENT: return "invalid argument";
    case LYAP_ERR_BAD_SEQUENCE: return "bad sequence";
    case LYAP_ERR_UNSUPPORTED: return "unsupported";
    case LYAP_ERR_NO_DEVICE: return "no device";
    case LYAP_ERR_DEVICE: return "device error";
    case LYAP_ERR_BUDGET: return "HBM budget exceeded";
    case LYAP_ERR_QUEUE_FULL: return "queue full";
    case LYAP_ERR_NOT_FOUND: return "not found";
    case LYAP_ERR_CANCELLED: return "cancelled";
    case LYAP_ERR_PENDING: return "pending";
    case LYAP_ERR_NO_LIBRARY: return "library not loaded";
    default: return "unknown status";
    }
}

int lyap_create(lyap_ctx **out, int device, size_t hbm_budget_bytes)
{
    if (!out) return LYAP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return LYAP_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return LYAP_ERR_INVALID_ARGUMENT;
    lyap_ctx *ctx = new lyap_ctx();
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return LYAP_ERR_DEVICE;
    }
    hipDeviceProp_t props;
    if (hipGetDeviceProperties(&props, device) == hipSuccess) ctx->num_cus = props.multiProcessorCount;
    void *p = nullptr;
    if (ctx_alloc(ctx, &p, 256) != LYAP_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return LYAP_ERR_DEVICE;
    }
    ctx->counters = (uint32_t *)p;
    /* the allocation-free entry's scratch address on this device; a failure
     * is reported by that entry alone */
    (void)noalloc_scratch_setup(device);
    if (launch_zero_u32(ctx, ctx->counters, 64) != LYAP_OK ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        lyap_destroy(ctx);
        return LYAP_ERR_DEVICE;
    }
    *out = ctx;
    return LYAP_OK;
}

void lyap_destroy(lyap_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_render_resources(ctx);
    for (hipEvent_t e : ctx->timing_events) (void)hipEventDestroy(e);
    for (auto &kv : ctx->allocations) (void)hipFree(kv.first);
    ctx->allocations.clear();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int lyap_sync(lyap_ctx *ctx)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LYAP_OK;
}

void *lyap_get_stream(lyap_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

const char *lyap_last_error(lyap_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : "no context";
}

size_t lyap_bytes_in_use(lyap_ctx *ctx) { return ctx ? ctx->in_use : 0; }

int lyap_device_alloc(lyap_ctx *ctx, void **ptr, size_t bytes)
{
    if (!ctx || !ptr) return LYAP_ERR_INVALID_ARGUMENT;
    LYAP_HIP(ctx, hipMalloc(ptr, bytes));
    return LYAP_OK;
}

int lyap_device_free(lyap_ctx *ctx, void *ptr)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LYAP_HIP(ctx, hipFree(ptr));
    return LYAP_OK;
}

int lyap_copy_to_device(lyap_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LYAP_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LYAP_OK;
}

int lyap_copy_to_host(lyap_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LYAP_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LYAP_OK;
}

int lyap_calc_volume(lyap_ctx *ctx, float *d_exps, const LyapParams *prm,
                     const int32_t *seq, const float lo[3], const float hi[3],
                     unsigned nx, unsigned ny, unsigned nz,
                     const LyapVolumeMode *mode_in)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LyapVolumeMode mode{};
    if (mode_in) mode = *mode_in;
    if (mode.precision < LYAP_PREC_PLAIN || mode.precision > LYAP_PREC_FP64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown precision %d", mode.precision);
    if (!lo || !hi) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null box");
    if ((uint64_t)nx * ny * nz == 0) return LYAP_OK;
    if (!d_exps) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null output");
    EvalDesc e;
    int rc = check_params(ctx, prm, seq, &e, mode.renorm);
    if (rc) return rc;
    return launch_volume(ctx, d_exps, e, lo, hi, nx, ny, nz, nullptr, 0, mode);
}

int lyap_calc_volume_noalloc(lyap_ctx *ctx, float *d_exps, const LyapParams *prm,
                             const int32_t *seq, const float lo[3], const float hi[3],
                             unsigned nx, unsigned ny, unsigned nz,
                             const LyapVolumeMode *mode_in)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LyapVolumeMode mode{};
    if (mode_in) mode = *mode_in;
    if (mode.precision < LYAP_PREC_PLAIN || mode.precision > LYAP_PREC_FP64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown precision %d", mode.precision);
    if (!lo || !hi) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null box");
    if ((uint64_t)nx * ny * nz == 0) return LYAP_OK;
    if (!d_exps) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null output");
    EvalDesc e;
    int rc = check_params(ctx, prm, seq, &e, mode.renorm);
    if (rc) return rc;
    return launch_volume_noalloc(ctx, d_exps, e, lo, hi, nx, ny, nz, mode);
}

int lyap_calc_points(lyap_ctx *ctx, float *d_out, const float *d_xyz, size_t n,
                     const LyapParams *prm, const int32_t *seq,
                     const LyapVolumeMode *mode_in)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LyapVolumeMode mode{};
    if (mode_in) mode = *mode_in;
    if (mode.precision < LYAP_PREC_PLAIN || mode.precision > LYAP_PREC_FP64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown precision %d", mode.precision);
    if (n == 0) return LYAP_OK;
    if (!d_out || !d_xyz) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null buffer");
    EvalDesc e;
    int rc = check_params(ctx, prm, seq, &e, mode.renorm);
    if (rc) return rc;
    const float zero[3] = {0, 0, 0};
    return launch_volume(ctx, d_out, e, zero, zero, 0, 0, 0, d_xyz, n, mode);
}

int lyap_calc_points_halves(lyap_ctx *ctx, float *d_lambda, float *d_half1, float *d_half2,
                            const float *d_xyz, size_t n, const LyapParams *prm,
                            const int32_t *seq, unsigned renorm)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    if (n == 0) return LYAP_OK;
    if (!d_lambda || !d_half1 || !d_half2 || !d_xyz)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null buffer");
    EvalDesc e;
    int rc = check_params(ctx, prm, seq, &e, renorm);
    if (rc) return rc;
    return launch_points_halves(ctx, d_lambda, d_half1, d_half2, e, d_xyz, n);
}

int lyap_last_reevaluated(lyap_ctx *ctx, uint64_t *count)
{
    if (!ctx || !count) return LYAP_ERR_INVALID_ARGUMENT;
    uint32_t c = 0;
    LYAP_HIP(ctx, hipMemcpyAsync(&c, ctx->counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = c;
    return LYAP_OK;
}

int lyap_timing_enable(lyap_ctx *ctx, int on)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    /* switching off pauses: the launches recorded so far stay until
     * collected, so a caller can time a sample of its frames */
    ctx->timing = on != 0;
    return LYAP_OK;
}

int lyap_timing_collect(lyap_ctx *ctx, double *kernel_ms, uint64_t *launches)
{
    if (!ctx || !kernel_ms || !launches) return LYAP_ERR_INVALID_ARGUMENT;
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double total = 0.0;
    for (size_t i = 0; i + 1 < ctx->timing_used; i += 2) {
        float ms = 0.0f;
        LYAP_HIP(ctx, hipEventElapsedTime(&ms, ctx->timing_events[i], ctx->timing_events[i + 1]));
        total += ms;
    }
    *kernel_ms = total;
    *launches = ctx->timing_used / 2;
    ctx->timing_used = 0;
    return LYAP_OK;
}

int lyap_last_evals(lyap_ctx *ctx, uint64_t *count)
{
    if (!ctx || !count) return LYAP_ERR_INVALID_ARGUMENT;
    unsigned long long c = 0;
    const uint32_t *src = ctx->last_counters ? ctx->last_counters : ctx->counters;
    LYAP_HIP(ctx, hipMemcpyAsync(&c, src + 2, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *count = c;
    return LYAP_OK;
}

} /* extern "C" */
