/*
 * internal.hpp - host-side state shared by the library's translation units.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "lyap.h"
#include "lyap_device.hpp"

struct lyap_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;          /* 0: no cap */
    size_t in_use = 0;          /* bytes the library allocated */
    std::string last_error;
    std::unordered_map<void *, size_t> allocations;

    /* scratch: [0] FP64 re-evaluation count, [1..2] 64-bit evaluation count */
    uint32_t *counters = nullptr;
    uint32_t *reeval_list = nullptr;   /* voxel indices for the FP64 pass */
    size_t reeval_cap = 0;             /* entries */

    /* the render path's launch resources (render.hip) */
    void *render_res = nullptr;
    uint32_t *last_counters = nullptr; /* counters of the last render launch */

    LyapTuning tuned{};                /* autotuner's choice, zero = unset */
    void *debug_wave_log = nullptr;    /* per-wave render record, when set */
    int num_cus = 0;

    /* kernel timing: event pairs around render launches, reused */
    bool timing = false;
    std::vector<hipEvent_t> timing_events;
    size_t timing_used = 0;            /* events recorded since the last collect */
};

namespace lyap {

int fail(lyap_ctx *ctx, int code, const char *fmt, ...);

/* Allocation through the context's budget. */
int ctx_alloc(lyap_ctx *ctx, void **ptr, size_t bytes);
void ctx_free(lyap_ctx *ctx, void *ptr);

int ensure_reeval_list(lyap_ctx *ctx, size_t entries);

/* Records the next timing event on the context's stream (no-op when timing
 * is off). */
int timing_mark(lyap_ctx *ctx);

/* Launchers in the kernel translation units. */
int launch_zero_u32(lyap_ctx *ctx, uint32_t *p, uint32_t n);

int launch_volume(lyap_ctx *ctx, float *out, const EvalDesc &e,
                  const float lo[3], const float hi[3], unsigned nx,
                  unsigned ny, unsigned nz, const float *xyz, size_t npoints,
                  const LyapVolumeMode &mode);

void free_render_resources(lyap_ctx *ctx);

int launch_volume_slab(lyap_ctx *ctx, float *out, const EvalDesc &e,
                       const float lo[3], const float hi[3], unsigned nx,
                       unsigned ny, unsigned nz, unsigned z_begin, unsigned z_end,
                       const float *xyz, size_t npoints, const LyapVolumeMode &mode);

/* The allocation-free volume path: entries of its scratch list (16 MiB). */
constexpr uint32_t kNoAllocListCap = 1u << 22;
/* Looks up the scratch array's address on `device` (current) once per
 * device; lyap_create calls it. */
int noalloc_scratch_setup(int device);
int launch_volume_noalloc(lyap_ctx *ctx, float *out, const EvalDesc &e, const float lo[3],
                          const float hi[3], unsigned nx, unsigned ny, unsigned nz,
                          const LyapVolumeMode &mode);

int launch_points_halves(lyap_ctx *ctx, float *lam, float *h1, float *h2, const EvalDesc &e,
                         const float *xyz, size_t n);

/* Server launch slots (render.hip). */
struct RenderSlot;
RenderSlot *render_slot_create();
void render_slot_destroy(lyap_ctx *ctx, RenderSlot *slot);
/* The slot's counter block: words [2..3] hold the last launch's lambda
 * evaluations (rays' algorithmic samples plus one per voxel). */
const uint32_t *render_slot_counters(const RenderSlot *slot);
/* FP64 re-evaluations of each item of the last launch, kItemCountStride
 * words apart (views' entries stay 0). */
constexpr uint32_t kItemCountStride = 32;
const uint32_t *render_slot_vox_counts(const RenderSlot *slot);
int render_slot_launch(lyap_ctx *ctx, RenderSlot *slot, const LyapView *views,
                       const uint32_t *ray_begin, const uint32_t *ray_count, unsigned n,
                       const LyapVolumeBlock *blocks, unsigned nb, const LyapVolumeMode &vmode,
                       const LyapRenderMode &mode, const LyapTuning &tune);

int launch_render(lyap_ctx *ctx, LyapRGBA *rgba, LyapPoint *points,
                  const LyapCam &cam, const LyapParams &prm, const EvalDesc &e,
                  const LyapLight *lights, unsigned num_lights, unsigned w,
                  unsigned h, float *depth, float *lambda, uint32_t *evals,
                  const LyapRenderMode &mode);

} /* namespace lyap */

#define LYAP_HIP(ctx, call)                                                   \
    do {                                                                      \
        hipError_t _e = (call);                                               \
        if (_e != hipSuccess)                                                 \
            return lyap::fail((ctx), LYAP_ERR_DEVICE, "%s: %s (%s:%d)", #call, \
                              hipGetErrorString(_e), __FILE__, __LINE__);     \
    } while (0)
