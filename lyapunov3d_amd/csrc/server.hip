/*
 * server.hip - continuous batching of view and volume requests on one device.
 *
 * The reference renders one view per launch from an interactive loop
 * (lyap_interactive.cu:711) and one volume per process (lyap_calculate.cu:72);
 * it has no request queue. This is the serving layer over the same kernels:
 *
 *  - a view is a range of 64-ray blocks (8x8 pixels), cut into pieces of a
 *    multiple of 64 rays; a volume is a range of z slices;
 *  - a free slot takes the pending pieces of both kinds with the earliest
 *    deadlines, up to about slice_ms of estimated device time, and runs them
 *    as one persistent launch: the work queue of its first launch deals the
 *    views' ray tiles and then the slabs' voxel tiles (lyap_run_batch's
 *    path), so the waves whose rays end take voxels. With
 *    LYAP_SERVER_SEPARATE_VOLUMES (or when the volume precision differs from
 *    the views') render slots take views only and volume slots run slabs in
 *    the volume kernels;
 *  - every slot owns a context of its own (stream, launch buffers, counters),
 *    so launches of different slots overlap on the device and a launch that
 *    is finishing its long rays shares the CUs with the next one;
 *  - a request is admitted when its first piece is issued: its resident
 *    outputs are taken from a pool of device buffers against the budget;
 *    released buffers go back to the pool, so serving never frees device
 *    memory (a free waits for the whole device).
 *
 * Costs are estimated in "iteration units" (rays or voxels times settle +
 * accum) and converted to milliseconds with a running measure of the slots'
 * launch durations.
 */
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace lyap {
namespace {

enum ReqState { QUEUED = 0, ACTIVE = 1, FINISHED = 2 };

struct Request {
    uint64_t ticket = 0;
    LyapRequest rq{};
    int state = QUEUED;
    int status = LYAP_ERR_PENDING;
    bool cancel = false;
    double submit_ms = 0, start_ms = 0, finish_ms = 0, deadline = 0;
    /* view: rays [0, n_units) in 64-ray blocks; volume: z slices */
    uint32_t n_units = 0, next_unit = 0;
    uint32_t outstanding = 0;          /* pieces on the device */
    double unit_cost = 0;              /* iteration units per ray or per z slice */
    uint64_t evals = 0;
    /* resident outputs the server owns */
    void *own_rgba = nullptr;
    float *own_lambda = nullptr;
    float *own_volume = nullptr;
    size_t rgba_bytes = 0, lambda_bytes = 0, volume_bytes = 0;
    EvalDesc vol_e{};
};

/* z slabs one volume launch may queue */
constexpr size_t kMaxSlabs = 16;

struct Piece {
    uint64_t ticket;
    uint32_t begin, count;
    int kind;                          /* LYAP_REQ_* */
    uint32_t item;                     /* volume: index among the launch's slabs */
};

enum SlotKind { SLOT_MIXED = 0, SLOT_RENDER = 1, SLOT_VOLUME = 2 };

struct Slot {
    lyap_ctx *ctx = nullptr;
    int kind = SLOT_MIXED;
    RenderSlot *rs = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    bool busy = false;
    std::vector<Piece> pieces;
    double est_view_ms = 0, est_vol_ms = 0;    /* estimated device time of each part */
    /* pinned: the launch's lambda evaluations (rays' samples, one per voxel) */
    unsigned long long *host_evals = nullptr;
    /* pinned: FP64 re-evaluations of each slab, kItemCountStride words apart */
    uint32_t *host_reeval = nullptr;
    uint64_t rays = 0, voxels = 0;              /* work of the launch */
};

double now_ms(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

} /* namespace */
} /* namespace lyap */

using namespace lyap;

struct lyap_server {
    lyap_ctx *ctx = nullptr;           /* owner of resident outputs */
    LyapServerConfig cfg{};
    std::chrono::steady_clock::time_point t0;
    uint64_t next_ticket = 1;
    std::map<uint64_t, Request> reqs;
    std::vector<Slot> slots;
    /* device buffers by size, not in use */
    std::multimap<size_t, void *> pool;
    size_t resident = 0, pooled = 0, peak = 0;
    /* measured device milliseconds per iteration unit */
    double render_ms_per_unit = 1.0e-9, volume_ms_per_unit = 1.0e-10;
    LyapServerStats stats{};
    std::vector<uint64_t> finished;    /* not yet handed out by lyap_server_finished */
    size_t finished_head = 0;
};

namespace {

/* ---- resident memory ------------------------------------------------------ */

int pool_take(lyap_server *s, size_t bytes, void **out)
{
    *out = nullptr;
    if (bytes == 0) return LYAP_OK;
    auto it = s->pool.find(bytes);
    if (it != s->pool.end()) {
        *out = it->second;
        s->pool.erase(it);
        s->pooled -= bytes;
    } else {
        /* make room: pooled buffers of other sizes count against the budget */
        const size_t cap = s->cfg.hbm_budget_bytes;
        while (cap && s->resident + s->pooled + bytes > cap && !s->pool.empty()) {
            auto big = std::prev(s->pool.end());
            for (auto &sl : s->slots)   /* a free waits for the device anyway */
                if (sl.busy) (void)hipEventSynchronize(sl.end);
            ctx_free(s->ctx, big->second);
            s->pooled -= big->first;
            s->pool.erase(big);
        }
        int rc = ctx_alloc(s->ctx, out, bytes);
        if (rc) return rc;
    }
    s->resident += bytes;
    if (s->resident > s->peak) s->peak = s->resident;
    return LYAP_OK;
}

void pool_give(lyap_server *s, void *p, size_t bytes)
{
    if (!p) return;
    s->resident -= bytes;
    s->pool.emplace(bytes, p);
    s->pooled += bytes;
}

size_t request_bytes(const Request &r)
{
    if (r.rq.kind == LYAP_REQ_VIEW) {
        const size_t px = (size_t)r.rq.view.width * r.rq.view.height;
        size_t b = 0;
        if (!r.rq.view.rgba) b += px * sizeof(LyapRGBA);
        if (!r.rq.view.rgba && !r.rq.view.lambda) b += px * sizeof(float);
        return b;
    }
    if (r.rq.vol_out) return 0;
    return (size_t)r.rq.vol_n[0] * r.rq.vol_n[1] * r.rq.vol_n[2] * sizeof(float);
}

/* Admits a queued request: takes its resident outputs. LYAP_ERR_BUDGET when
 * they do not fit now (the request waits). */
int admit(lyap_server *s, Request &r)
{
    const size_t need = request_bytes(r);
    if (s->cfg.hbm_budget_bytes && s->resident + need > s->cfg.hbm_budget_bytes)
        return LYAP_ERR_BUDGET;
    if (r.rq.kind == LYAP_REQ_VIEW) {
        const size_t px = (size_t)r.rq.view.width * r.rq.view.height;
        if (!r.rq.view.rgba) {
            void *c = nullptr, *l = nullptr;
            int rc = pool_take(s, px * sizeof(LyapRGBA), &c);
            if (rc) return rc;
            if (!r.rq.view.lambda) {
                rc = pool_take(s, px * sizeof(float), &l);
                if (rc) {
                    pool_give(s, c, px * sizeof(LyapRGBA));
                    return rc;
                }
                r.own_lambda = (float *)l;
                r.lambda_bytes = px * sizeof(float);
                r.rq.view.lambda = r.own_lambda;
            }
            r.own_rgba = c;
            r.rgba_bytes = px * sizeof(LyapRGBA);
            r.rq.view.rgba = (LyapRGBA *)r.own_rgba;
        }
    } else if (!r.rq.vol_out) {
        void *p = nullptr;
        const size_t b = request_bytes(r);
        int rc = pool_take(s, b, &p);
        if (rc) return rc;
        r.own_volume = (float *)p;
        r.volume_bytes = b;
        r.rq.vol_out = r.own_volume;
    }
    r.state = ACTIVE;
    r.start_ms = now_ms(s->t0);
    s->stats.queued--;
    s->stats.active++;
    return LYAP_OK;
}

void release_outputs(lyap_server *s, Request &r)
{
    pool_give(s, r.own_rgba, r.rgba_bytes);
    pool_give(s, r.own_lambda, r.lambda_bytes);
    pool_give(s, r.own_volume, r.volume_bytes);
    r.own_rgba = nullptr;
    r.own_lambda = nullptr;
    r.own_volume = nullptr;
}

void finish(lyap_server *s, Request &r, int status)
{
    if (r.state == FINISHED) return;
    if (r.state == QUEUED) s->stats.queued--;
    else s->stats.active--;
    r.state = FINISHED;
    r.status = status;
    r.finish_ms = now_ms(s->t0);
    s->finished.push_back(r.ticket);
    if (status == LYAP_OK) s->stats.completed++;
    else if (status == LYAP_ERR_CANCELLED) s->stats.cancelled++;
    else s->stats.failed++;
}

/* ---- slices --------------------------------------------------------------- */

bool has_work(const Request &r) { return r.state != FINISHED && !r.cancel && r.next_unit < r.n_units; }

/* Requests of `kind` (-1: both) with pieces left, earliest deadline first. */
std::vector<Request *> edf_order(lyap_server *s, int kind)
{
    std::vector<Request *> v;
    for (auto &kv : s->reqs)
        if ((kind < 0 || kv.second.rq.kind == kind) && has_work(kv.second)) v.push_back(&kv.second);
    std::stable_sort(v.begin(), v.end(), [](const Request *a, const Request *b) {
        return a->deadline < b->deadline;
    });
    return v;
}

/* Pixels inside the image among rays [begin, begin + count) of a view (its
 * 64-ray blocks are 8x8 pixels, clipped at the right and bottom edges). */
uint64_t piece_pixels(const LyapView &v, uint32_t begin, uint32_t count)
{
    const uint32_t w = v.width, h = v.height, bx = (w + 7) / 8;
    uint64_t px = 0;
    for (uint32_t blk = begin / 64; blk < (begin + count) / 64; ++blk) {
        const uint32_t x0 = (blk % bx) * 8, y0 = (blk / bx) * 8;
        px += (uint64_t)std::min(8u, w - x0) * std::min(8u, h - y0);
    }
    return px;
}

/* Takes the earliest-deadline pieces the slot can run, up to about slice_ms
 * of estimated device time: views' ray ranges (a multiple of chunk_rays, or
 * what is left), volumes' z slabs (at most kMaxSlabs per launch). */
int take_pieces(lyap_server *s, Slot &sl, std::vector<LyapView> &views, std::vector<uint32_t> &begin,
                std::vector<uint32_t> &count, std::vector<LyapVolumeBlock> &blocks)
{
    const double budget = s->cfg.slice_ms;
    const uint32_t chunk = s->cfg.chunk_rays;
    const int want = sl.kind == SLOT_MIXED ? -1 : (sl.kind == SLOT_RENDER ? LYAP_REQ_VIEW : LYAP_REQ_VOLUME);
    double est = 0;
    for (Request *r : edf_order(s, want)) {
        if (est >= budget) break;
        const bool view = r->rq.kind == LYAP_REQ_VIEW;
        if (!view && blocks.size() >= kMaxSlabs) continue;
        if (r->state == QUEUED) {
            int rc = admit(s, *r);
            if (rc == LYAP_ERR_BUDGET) continue;
            if (rc) return rc;
        }
        const uint32_t left = r->n_units - r->next_unit;
        const double rate = view ? s->render_ms_per_unit : s->volume_ms_per_unit;
        const double room = (budget - est) / (r->unit_cost * rate);
        uint32_t take = left;
        if (view) {
            if (room < (double)left) {
                take = (uint32_t)std::max<double>(chunk, std::ceil(room / chunk) * chunk);
                if (take > left) take = left;
            }
            views.push_back(r->rq.view);
            begin.push_back(r->next_unit);
            count.push_back(take);
            sl.pieces.push_back({r->ticket, r->next_unit, take, LYAP_REQ_VIEW, 0});
            sl.rays += piece_pixels(r->rq.view, r->next_unit, take);
            sl.est_view_ms += take * r->unit_cost * rate;
        } else {
            if (room < (double)left) take = std::max<uint32_t>(1u, (uint32_t)room);
            LyapVolumeBlock b;
            std::memset(&b, 0, sizeof(b));
            b.prm = r->rq.vol_prm;
            std::memcpy(b.seq, r->rq.vol_seq, sizeof(b.seq));
            for (int k = 0; k < 3; ++k) {
                b.lo[k] = r->rq.vol_lo[k];
                b.hi[k] = r->rq.vol_hi[k];
                b.n[k] = r->rq.vol_n[k];
            }
            b.z_begin = r->next_unit;
            b.z_end = r->next_unit + take;
            b.out = r->rq.vol_out + (size_t)b.n[0] * b.n[1] * r->next_unit;
            sl.pieces.push_back({r->ticket, r->next_unit, take, LYAP_REQ_VOLUME, (uint32_t)blocks.size()});
            blocks.push_back(b);
            sl.voxels += (uint64_t)b.n[0] * b.n[1] * take;
            sl.est_vol_ms += take * r->unit_cost * rate;
        }
        r->next_unit += take;
        r->outstanding++;
        est = sl.est_view_ms + sl.est_vol_ms;
    }
    return LYAP_OK;
}

int issue(lyap_server *s, Slot &sl)
{
    std::vector<LyapView> views;
    std::vector<uint32_t> begin, count;
    std::vector<LyapVolumeBlock> blocks;
    sl.pieces.clear();
    sl.est_view_ms = sl.est_vol_ms = 0;
    sl.rays = sl.voxels = 0;
    int rc = take_pieces(s, sl, views, begin, count, blocks);
    if (rc) return rc;
    if (sl.pieces.empty()) return LYAP_OK;
    LYAP_HIP(sl.ctx, hipEventRecord(sl.start, sl.ctx->stream));
    if (sl.kind == SLOT_VOLUME) {
        /* separate volume slot: the volume kernels, slab by slab; each slab's
         * re-evaluation count (the counter restarts per slab) */
        for (size_t i = 0; i < blocks.size(); ++i) {
            const LyapVolumeBlock &b = blocks[i];
            const Request &r = s->reqs.at(sl.pieces[i].ticket);
            rc = launch_volume_slab(sl.ctx, b.out, r.vol_e, b.lo, b.hi, b.n[0], b.n[1], b.n[2], b.z_begin,
                                    b.z_end, nullptr, 0, s->cfg.vol_mode);
            if (rc) return rc;
            LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_reeval + i * kItemCountStride, sl.ctx->counters, 4,
                                            hipMemcpyDeviceToHost, sl.ctx->stream));
        }
        *sl.host_evals = sl.voxels;
    } else {
        rc = render_slot_launch(sl.ctx, sl.rs, views.data(), begin.data(), count.data(),
                                (unsigned)views.size(), blocks.data(), (unsigned)blocks.size(),
                                s->cfg.vol_mode, s->cfg.mode, s->ctx->tuned);
        if (rc) return rc;
        LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_evals, render_slot_counters(sl.rs) + 2, 8,
                                        hipMemcpyDeviceToHost, sl.ctx->stream));
        if (!blocks.empty())
            LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_reeval,
                                            render_slot_vox_counts(sl.rs) + views.size() * kItemCountStride,
                                            blocks.size() * kItemCountStride * sizeof(uint32_t),
                                            hipMemcpyDeviceToHost, sl.ctx->stream));
    }
    LYAP_HIP(sl.ctx, hipEventRecord(sl.end, sl.ctx->stream));
    sl.busy = true;
    s->stats.launches++;
    return LYAP_OK;
}

/* Marks the pieces of a finished launch done. */
int retire(lyap_server *s, Slot &sl, bool *changed)
{
    if (!sl.busy) return LYAP_OK;
    hipError_t q = hipEventQuery(sl.end);
    if (q == hipErrorNotReady) return LYAP_OK;
    sl.busy = false;
    *changed = true;
    const int status = q == hipSuccess ? LYAP_OK : LYAP_ERR_DEVICE;
    if (status != LYAP_OK) s->ctx->last_error = hipGetErrorString(q);
    float ms = 0.0f;
    const double est = sl.est_view_ms + sl.est_vol_ms;
    if (status == LYAP_OK && hipEventElapsedTime(&ms, sl.start, sl.end) == hipSuccess && est > 0) {
        /* running measure of milliseconds per iteration unit, scaled by how
         * far this launch's estimate was off; launches of several slots
         * overlap, so it is the cost under sharing, which the next slice
         * sees */
        const double f = 0.75 + 0.25 * ((double)ms / est);
        if (sl.est_view_ms > 0) s->render_ms_per_unit *= f;
        if (sl.est_vol_ms > 0) s->volume_ms_per_unit *= f;
        s->stats.device_ms += ms;
    }
    s->stats.rays += sl.rays;
    s->stats.voxels += sl.voxels;
    if (status == LYAP_OK) s->stats.evals += (uint64_t)*sl.host_evals;
    for (size_t i = 0; i < sl.pieces.size(); ++i) {
        const Piece &p = sl.pieces[i];
        uint64_t reeval = 0;
        if (p.kind == LYAP_REQ_VOLUME && status == LYAP_OK) {
            reeval = sl.host_reeval[p.item * kItemCountStride];
            s->stats.evals += reeval;
        }
        auto it = s->reqs.find(p.ticket);
        if (it == s->reqs.end()) continue;
        Request &r = it->second;
        r.outstanding--;
        if (p.kind == LYAP_REQ_VOLUME)
            r.evals += (uint64_t)r.rq.vol_n[0] * r.rq.vol_n[1] * p.count + reeval;
        if (status != LYAP_OK) {
            r.cancel = true;
            if (r.outstanding == 0) finish(s, r, status);
            continue;
        }
        if (r.outstanding == 0 && (r.cancel || r.next_unit >= r.n_units))
            finish(s, r, r.cancel ? LYAP_ERR_CANCELLED : LYAP_OK);
    }
    sl.pieces.clear();
    return status;
}

int step(lyap_server *s, bool *changed)
{
    int first_error = LYAP_OK;
    for (auto &sl : s->slots) {
        int rc = retire(s, sl, changed);
        if (rc && !first_error) first_error = rc;
    }
    for (auto &sl : s->slots) {
        if (sl.busy) continue;
        int rc = issue(s, sl);
        if (rc) {
            /* the pieces taken by the failed launch end their requests */
            for (const Piece &p : sl.pieces) {
                auto it = s->reqs.find(p.ticket);
                if (it == s->reqs.end()) continue;
                it->second.outstanding--;
                it->second.cancel = true;
                if (it->second.outstanding == 0) finish(s, it->second, rc);
            }
            sl.pieces.clear();
            s->ctx->last_error = lyap_last_error(sl.ctx);
            if (!first_error) first_error = rc;
            *changed = true;
            continue;
        }
        if (sl.busy) *changed = true;
    }
    return first_error;
}

int validate(lyap_server *s, const LyapRequest &rq, Request &r)
{
    if (rq.kind == LYAP_REQ_VIEW) {
        const LyapView &v = rq.view;
        if (v.width == 0 || v.height == 0) return fail(s->ctx, LYAP_ERR_INVALID_ARGUMENT, "empty view");
        if ((uint64_t)v.width * v.height > (1ull << 31))
            return fail(s->ctx, LYAP_ERR_UNSUPPORTED, "view larger than 2^31 pixels");
        if (v.num_lights > LYAP_MAX_LIGHTS || (v.num_lights && !v.lights))
            return fail(s->ctx, LYAP_ERR_INVALID_ARGUMENT, "bad lights");
        EvalDesc e;
        if (!make_eval_desc(&e, v.prm, v.seq, 16))
            return fail(s->ctx, LYAP_ERR_BAD_SEQUENCE, "bad sequence");
        r.n_units = 64u * ((v.width + 7) / 8) * ((v.height + 7) / 8);
        r.unit_cost = (double)v.prm.settle + (double)v.prm.accum;
        /* a ray takes tens of samples; the measured rate absorbs the rest */
        r.unit_cost *= 16.0;
    } else if (rq.kind == LYAP_REQ_VOLUME) {
        const uint64_t n = (uint64_t)rq.vol_n[0] * rq.vol_n[1] * rq.vol_n[2];
        if (n == 0) return fail(s->ctx, LYAP_ERR_INVALID_ARGUMENT, "empty volume");
        if (n > 0xFFFFFFFFull) return fail(s->ctx, LYAP_ERR_UNSUPPORTED, "more than 2^32-1 voxels");
        if (!make_eval_desc(&r.vol_e, rq.vol_prm, rq.vol_seq, s->cfg.vol_mode.renorm))
            return fail(s->ctx, LYAP_ERR_BAD_SEQUENCE, "bad sequence");
        r.n_units = rq.vol_n[2];
        r.unit_cost = (double)rq.vol_n[0] * rq.vol_n[1] *
                      ((double)rq.vol_prm.settle + (double)rq.vol_prm.accum);
    } else {
        return fail(s->ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown request kind %d", rq.kind);
    }
    if (s->cfg.hbm_budget_bytes) {
        r.rq = rq;
        if (request_bytes(r) > s->cfg.hbm_budget_bytes)
            return fail(s->ctx, LYAP_ERR_BUDGET, "request needs more than the server budget");
    }
    return LYAP_OK;
}

} /* namespace */

extern "C" {

int lyap_server_create(lyap_ctx *ctx, const LyapServerConfig *cfg_in, lyap_server **out)
{
    if (!out) return LYAP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    LyapServerConfig cfg{};
    cfg.mode.refine = LYAP_REFINE_BISECT;
    cfg.mode.bisect_steps = 8;
    cfg.mode.normal_h = 1e-3f;
    cfg.mode.precision = LYAP_PREC_PRODUCT;
    cfg.vol_mode.precision = LYAP_PREC_PRODUCT;
    if (cfg_in) cfg = *cfg_in;
    if (cfg.render_slots == 0) cfg.render_slots = 2;
    if (cfg.volume_slots == 0) cfg.volume_slots = 1;
    if (cfg.flags & ~(uint32_t)LYAP_SERVER_SEPARATE_VOLUMES)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown server flags 0x%x", cfg.flags);
    if (cfg.render_slots > 4 || cfg.volume_slots > 4)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "at most 4 render and 4 volume slots");
    if (cfg.slice_ms <= 0) cfg.slice_ms = 12.0;
    if (cfg.chunk_rays == 0) cfg.chunk_rays = 4096;
    if (cfg.chunk_rays % 64u) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "chunk_rays must be a multiple of 64");
    if (cfg.mode.refine != LYAP_REFINE_WALK && cfg.mode.refine != LYAP_REFINE_BISECT)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown refine mode");
    if (cfg.mode.schedule != LYAP_SCHEDULE_BATCHED && cfg.mode.schedule != LYAP_SCHEDULE_DIRECT)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown schedule");
    if (cfg.mode.bisect_steps < 0 || cfg.mode.bisect_steps > 64 || cfg.mode.lanes_per_ray > 64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "bad render mode");
    if (cfg.mode.precision < LYAP_PREC_PLAIN || cfg.mode.precision > LYAP_PREC_FP64 ||
        cfg.vol_mode.precision < LYAP_PREC_PLAIN || cfg.vol_mode.precision > LYAP_PREC_FP64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown precision");

    lyap_server *s = new lyap_server();
    s->ctx = ctx;
    s->cfg = cfg;
    s->t0 = std::chrono::steady_clock::now();
    int rc = LYAP_OK;
    const unsigned n = cfg.render_slots + cfg.volume_slots;
    /* one launch carries both kinds only in one precision */
    const bool separate = (cfg.flags & LYAP_SERVER_SEPARATE_VOLUMES) ||
                          cfg.mode.precision != cfg.vol_mode.precision;
    s->slots.resize(n);
    for (unsigned i = 0; i < n && rc == LYAP_OK; ++i) {
        Slot &sl = s->slots[i];
        sl.kind = !separate ? SLOT_MIXED : (i < cfg.render_slots ? SLOT_RENDER : SLOT_VOLUME);
        rc = lyap_create(&sl.ctx, ctx->device, 0);
        if (rc) break;
        const size_t reeval_bytes = kMaxSlabs * kItemCountStride * sizeof(uint32_t);
        if (hipEventCreate(&sl.start) != hipSuccess || hipEventCreate(&sl.end) != hipSuccess ||
            hipHostMalloc((void **)&sl.host_evals, sizeof(unsigned long long), hipHostMallocDefault) !=
                hipSuccess ||
            hipHostMalloc((void **)&sl.host_reeval, reeval_bytes, hipHostMallocDefault) != hipSuccess) {
            rc = fail(ctx, LYAP_ERR_DEVICE, "server slot set-up failed");
            break;
        }
        *sl.host_evals = 0;
        std::memset(sl.host_reeval, 0, reeval_bytes);
        if (sl.kind != SLOT_VOLUME) {
            sl.rs = render_slot_create();
        } else if (cfg.vol_mode.precision == LYAP_PREC_PRODUCT &&
                   (cfg.vol_mode.fp64_margin > 0.0f || cfg.vol_mode.converge_tau > 0.0f)) {
            /* re-evaluation list for the largest slab (256 x 256 x 256) up front,
             * so a slab never reallocates while others run */
            rc = ensure_reeval_list(sl.ctx, 256u * 256u * 256u);
        }
    }
    if (rc) {
        lyap_server_destroy(s);
        return rc;
    }
    *out = s;
    return LYAP_OK;
}

void lyap_server_destroy(lyap_server *s)
{
    if (!s) return;
    for (auto &sl : s->slots) {
        if (sl.ctx) (void)hipStreamSynchronize(sl.ctx->stream);
    }
    for (auto &kv : s->reqs) release_outputs(s, kv.second);
    for (auto &kv : s->pool) ctx_free(s->ctx, kv.second);
    s->pool.clear();
    for (auto &sl : s->slots) {
        if (sl.rs) render_slot_destroy(sl.ctx, sl.rs);
        if (sl.start) (void)hipEventDestroy(sl.start);
        if (sl.end) (void)hipEventDestroy(sl.end);
        if (sl.host_evals) (void)hipHostFree(sl.host_evals);
        if (sl.host_reeval) (void)hipHostFree(sl.host_reeval);
        if (sl.ctx) lyap_destroy(sl.ctx);
    }
    delete s;
}

int lyap_server_submit(lyap_server *s, const LyapRequest *req, uint64_t *ticket)
{
    if (!s || !req || !ticket) return LYAP_ERR_INVALID_ARGUMENT;
    Request r;
    int rc = validate(s, *req, r);
    if (rc) return rc;
    r.rq = *req;
    r.ticket = s->next_ticket++;
    r.submit_ms = now_ms(s->t0);
    r.deadline = r.submit_ms + (req->deadline_ms > 0 ? req->deadline_ms : 1.0e9);
    *ticket = r.ticket;
    s->reqs.emplace(r.ticket, r);
    s->stats.submitted++;
    s->stats.queued++;
    return LYAP_OK;
}

int lyap_server_progress(lyap_server *s, double wait_ms)
{
    if (!s) return LYAP_ERR_INVALID_ARGUMENT;
    const double until = now_ms(s->t0) + (wait_ms > 0 ? wait_ms : 0);
    for (;;) {
        bool changed = false;
        int rc = step(s, &changed);
        if (rc) return rc;
        if (changed || now_ms(s->t0) >= until) return LYAP_OK;
        bool busy = false;
        for (auto &sl : s->slots) busy = busy || sl.busy;
        if (!busy) return LYAP_OK;   /* nothing to wait for */
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

int lyap_server_poll(lyap_server *s, uint64_t ticket, LyapResult *out)
{
    if (!s || !out) return LYAP_ERR_INVALID_ARGUMENT;
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LYAP_ERR_NOT_FOUND;
    const Request &r = it->second;
    std::memset(out, 0, sizeof(*out));
    out->status = r.state == FINISHED ? r.status : LYAP_ERR_PENDING;
    out->kind = (uint32_t)r.rq.kind;
    out->submit_ms = r.submit_ms;
    out->start_ms = r.start_ms;
    out->finish_ms = r.finish_ms;
    if (r.rq.kind == LYAP_REQ_VIEW) {
        out->rays = (uint64_t)r.rq.view.width * r.rq.view.height;
        out->rgba = r.rq.view.rgba;
        out->lambda = r.rq.view.lambda;
    } else {
        out->voxels = (uint64_t)r.rq.vol_n[0] * r.rq.vol_n[1] * r.rq.vol_n[2];
        out->evals = r.evals;
        out->lambda = r.rq.vol_out;
    }
    return LYAP_OK;
}

int lyap_server_release(lyap_server *s, uint64_t ticket)
{
    if (!s) return LYAP_ERR_INVALID_ARGUMENT;
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LYAP_ERR_NOT_FOUND;
    if (it->second.state != FINISHED) return LYAP_ERR_PENDING;
    release_outputs(s, it->second);
    s->reqs.erase(it);
    return LYAP_OK;
}

int lyap_server_cancel(lyap_server *s, uint64_t ticket)
{
    if (!s) return LYAP_ERR_INVALID_ARGUMENT;
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LYAP_ERR_NOT_FOUND;
    Request &r = it->second;
    if (r.state == FINISHED) return LYAP_OK;
    r.cancel = true;
    if (r.outstanding == 0) finish(s, r, LYAP_ERR_CANCELLED);
    return LYAP_OK;
}

double lyap_server_clock_ms(lyap_server *s) { return s ? now_ms(s->t0) : 0.0; }

int lyap_server_stats(lyap_server *s, LyapServerStats *out)
{
    if (!s || !out) return LYAP_ERR_INVALID_ARGUMENT;
    *out = s->stats;
    out->resident_bytes = s->resident;
    out->peak_resident_bytes = s->peak;
    return LYAP_OK;
}

int lyap_server_finished(lyap_server *s, uint64_t *tickets, unsigned cap, unsigned *n)
{
    if (!s || !n || (cap && !tickets)) return LYAP_ERR_INVALID_ARGUMENT;
    unsigned k = 0;
    while (k < cap && s->finished_head < s->finished.size())
        tickets[k++] = s->finished[s->finished_head++];
    if (s->finished_head == s->finished.size()) {
        s->finished.clear();
        s->finished_head = 0;
    }
    *n = k;
    return LYAP_OK;
}

int lyap_server_drain(lyap_server *s, double timeout_ms)
{
    if (!s) return LYAP_ERR_INVALID_ARGUMENT;
    const double until = now_ms(s->t0) + timeout_ms;
    for (;;) {
        if (s->stats.queued == 0 && s->stats.active == 0) return LYAP_OK;
        int rc = lyap_server_progress(s, 1.0);
        if (rc) return rc;
        if (timeout_ms > 0 && now_ms(s->t0) >= until) return LYAP_ERR_PENDING;
        bool busy = false;
        for (auto &sl : s->slots) busy = busy || sl.busy;
        if (!busy && s->stats.queued + s->stats.active > 0) {
            /* nothing runs and nothing could be issued: the budget holds the
             * queued requests back until a release */
            return LYAP_ERR_BUDGET;
        }
    }
}

} /* extern "C" */
