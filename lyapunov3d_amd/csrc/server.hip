/*
 * server.hip - the serving loop's device side: continuous batching of view
 * and volume requests on one GPU.
 *
 * The scheduler (serve_core.cpp) decides what each launch carries; this file
 * runs the launches. Every slot owns a context of its own (stream, launch
 * buffers, counters), so launches of different slots overlap on the device
 * and a launch that is finishing its long rays shares the CUs with the next
 * one. A mixed slot runs its views' ray tiles and its volumes' voxel tiles
 * in one persistent launch (lyap_run_batch's path: the first launch's work
 * queue deals the ray tiles and then the voxel tiles, so the waves whose rays
 * end take voxels). With LYAP_SERVER_SEPARATE_VOLUMES (or when the volume
 * precision differs from the views') render slots take views only and
 * volume slots run slabs in the volume kernels.
 *
 * The reference renders one view per launch from an interactive loop
 * (lyap_interactive.cu:711) and one volume per process (lyap_calculate.cu:72);
 * it has no request queue.
 */
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "internal.hpp"
#include "serve_core.hpp"

using namespace lyap;
using namespace lyap::serve;

namespace {

struct HipSlot {
    lyap_ctx *ctx = nullptr;
    int kind = SLOT_MIXED;
    RenderSlot *rs = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    /* pinned: the launch's lambda evaluations (rays' samples, one per voxel) */
    unsigned long long *host_evals = nullptr;
    /* pinned: FP64 re-evaluations of each slab, kItemCountStride words apart */
    uint32_t *host_reeval = nullptr;
    unsigned nblocks = 0;
};

class HipBackend final : public ServeBackend {
public:
    HipBackend(lyap_ctx *owner, const LyapServerConfig &cfg)
        : owner_(owner), cfg_(cfg), t0_(std::chrono::steady_clock::now())
    {
    }

    /* Creates the slots' contexts, events and pinned result words. */
    int init(const std::vector<int> &kinds)
    {
        slots_.resize(kinds.size());
        for (size_t i = 0; i < kinds.size(); ++i) {
            HipSlot &sl = slots_[i];
            sl.kind = kinds[i];
            int rc = lyap_create(&sl.ctx, owner_->device, 0);
            if (rc) return fail(owner_, rc, "server slot context: %s", lyap_strerror(rc));
            const size_t reeval_bytes = kMaxSlabs * kItemCountStride * sizeof(uint32_t);
            if (hipEventCreate(&sl.start) != hipSuccess || hipEventCreate(&sl.end) != hipSuccess ||
                hipHostMalloc((void **)&sl.host_evals, sizeof(unsigned long long), hipHostMallocDefault) !=
                    hipSuccess ||
                hipHostMalloc((void **)&sl.host_reeval, reeval_bytes, hipHostMallocDefault) != hipSuccess)
                return fail(owner_, LYAP_ERR_DEVICE, "server slot set-up failed");
            *sl.host_evals = 0;
            std::memset(sl.host_reeval, 0, reeval_bytes);
            if (sl.kind != SLOT_VOLUME) {
                sl.rs = render_slot_create();
            } else if (cfg_.vol_mode.precision == LYAP_PREC_PRODUCT &&
                       (cfg_.vol_mode.fp64_margin > 0.0f || cfg_.vol_mode.converge_tau > 0.0f)) {
                /* re-evaluation list for the largest slab (256 x 256 x 256) up
                 * front, so a slab never reallocates while others run */
                rc = ensure_reeval_list(sl.ctx, 256u * 256u * 256u);
                if (rc) {
                    owner_->last_error = sl.ctx->last_error;
                    return rc;
                }
            }
        }
        return LYAP_OK;
    }

    ~HipBackend() override
    {
        sync_all();
        for (auto &sl : slots_) {
            if (sl.rs) render_slot_destroy(sl.ctx, sl.rs);
            if (sl.start) (void)hipEventDestroy(sl.start);
            if (sl.end) (void)hipEventDestroy(sl.end);
            if (sl.host_evals) (void)hipHostFree(sl.host_evals);
            if (sl.host_reeval) (void)hipHostFree(sl.host_reeval);
            if (sl.ctx) lyap_destroy(sl.ctx);
        }
    }

    double now_ms() override
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
    }

    int alloc(size_t bytes, void **out) override { return ctx_alloc(owner_, out, bytes); }

    void release(void *p) override
    {
        sync_all();   /* a free waits for the device anyway */
        ctx_free(owner_, p);
    }

    int launch(unsigned i, const SlotLaunch &l) override
    {
        HipSlot &sl = slots_[i];
        const int rc = launch_on(sl, l);
        if (rc) owner_->last_error = sl.ctx->last_error;
        return rc;
    }

    int query(unsigned i, LaunchResult *res) override
    {
        HipSlot &sl = slots_[i];
        const hipError_t q = hipEventQuery(sl.end);
        if (q == hipErrorNotReady) return LYAP_ERR_PENDING;
        if (q != hipSuccess) return fail(owner_, LYAP_ERR_DEVICE, "server launch: %s", hipGetErrorString(q));
        float ms = 0.0f;
        res->ms = hipEventElapsedTime(&ms, sl.start, sl.end) == hipSuccess ? (double)ms : -1.0;
        res->evals = *sl.host_evals;
        for (unsigned b = 0; b < sl.nblocks && b < kMaxSlabs; ++b)
            res->reeval[b] = sl.host_reeval[b * kItemCountStride];
        return LYAP_OK;
    }

    void idle() override { std::this_thread::sleep_for(std::chrono::microseconds(20)); }

    int error(int code, const char *msg) override { return fail(owner_, code, "%s", msg); }

private:
    void sync_all()
    {
        for (auto &sl : slots_)
            if (sl.ctx) (void)hipStreamSynchronize(sl.ctx->stream);
    }

    int launch_on(HipSlot &sl, const SlotLaunch &l)
    {
        sl.nblocks = (unsigned)l.blocks.size();
        LYAP_HIP(sl.ctx, hipEventRecord(sl.start, sl.ctx->stream));
        if (sl.kind == SLOT_VOLUME) {
            /* the volume kernels, slab by slab; each slab's re-evaluation
             * count (the counter restarts per slab) */
            for (size_t b = 0; b < l.blocks.size(); ++b) {
                const LyapVolumeBlock &k = l.blocks[b];
                EvalDesc e;
                if (!make_eval_desc(&e, k.prm, k.seq, cfg_.vol_mode.renorm))
                    return fail(sl.ctx, LYAP_ERR_BAD_SEQUENCE, "bad sequence");
                int rc = launch_volume_slab(sl.ctx, k.out, e, k.lo, k.hi, k.n[0], k.n[1], k.n[2], k.z_begin,
                                            k.z_end, nullptr, 0, cfg_.vol_mode);
                if (rc) return rc;
                LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_reeval + b * kItemCountStride, sl.ctx->counters, 4,
                                                hipMemcpyDeviceToHost, sl.ctx->stream));
            }
            *sl.host_evals = l.voxels;
        } else {
            int rc = render_slot_launch(sl.ctx, sl.rs, l.views.data(), l.begin.data(), l.count.data(),
                                        (unsigned)l.views.size(), l.blocks.data(), (unsigned)l.blocks.size(),
                                        cfg_.vol_mode, cfg_.mode, owner_->tuned);
            if (rc) return rc;
            LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_evals, render_slot_counters(sl.rs) + 2, 8,
                                            hipMemcpyDeviceToHost, sl.ctx->stream));
            if (!l.blocks.empty())
                LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_reeval,
                                                render_slot_vox_counts(sl.rs) + l.views.size() * kItemCountStride,
                                                l.blocks.size() * kItemCountStride * sizeof(uint32_t),
                                                hipMemcpyDeviceToHost, sl.ctx->stream));
        }
        LYAP_HIP(sl.ctx, hipEventRecord(sl.end, sl.ctx->stream));
        return LYAP_OK;
    }

    lyap_ctx *owner_;                  /* owner of resident outputs */
    LyapServerConfig cfg_;
    std::chrono::steady_clock::time_point t0_;
    std::vector<HipSlot> slots_;
};

} /* namespace */

extern "C" int lyap_server_create(lyap_ctx *ctx, const LyapServerConfig *cfg_in, lyap_server **out)
{
    if (!out) return LYAP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LyapServerConfig cfg;
    std::string why;
    int rc = check_config(cfg_in, &cfg, &why);
    if (rc) return fail(ctx, rc, "%s", why.c_str());
    HipBackend *dev = new HipBackend(ctx, cfg);
    rc = dev->init(slot_kinds(cfg));
    if (rc) {
        delete dev;
        return rc;
    }
    *out = server_new(cfg, dev);
    return LYAP_OK;
}
