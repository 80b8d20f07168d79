/*
 * server.hip - continuous batching of view and volume requests on one device.
 *
 * The reference renders one view per launch from an interactive loop
 * (lyap_interactive.cu:711) and one volume per process (lyap_calculate.cu:72);
 * it has no request queue. This is the serving layer over the same kernels:
 *
 *  - a view is a range of 64-ray blocks (8x8 pixels), cut into pieces of a
 *    multiple of 64 rays; a volume is a range of z slices;
 *  - a free render slot takes the pending pieces with the earliest deadlines
 *    up to about slice_ms of estimated device time and runs them as one
 *    persistent launch (the batched path: lane groups, continuation levels);
 *    a free volume slot takes z slabs the same way;
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
};

struct Slot {
    lyap_ctx *ctx = nullptr;
    bool render = true;
    RenderSlot *rs = nullptr;
    hipEvent_t start = nullptr, end = nullptr;
    bool busy = false;
    std::vector<Piece> pieces;
    double est_units = 0;
    /* pinned: the render launch's evaluations, or each slab's re-evaluations */
    unsigned long long *host_count = nullptr;
    uint64_t work = 0;                          /* rays or voxels */
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

/* Requests of `kind` with pieces left, earliest deadline first. */
std::vector<Request *> edf_order(lyap_server *s, int kind)
{
    std::vector<Request *> v;
    for (auto &kv : s->reqs)
        if (kv.second.rq.kind == kind && has_work(kv.second)) v.push_back(&kv.second);
    std::stable_sort(v.begin(), v.end(), [](const Request *a, const Request *b) {
        return a->deadline < b->deadline;
    });
    return v;
}

int issue_render(lyap_server *s, Slot &sl)
{
    const double budget = s->cfg.slice_ms / s->render_ms_per_unit;
    const uint32_t chunk = s->cfg.chunk_rays;
    std::vector<LyapView> views;
    std::vector<uint32_t> begin, count;
    sl.pieces.clear();
    sl.est_units = 0;
    sl.work = 0;
    for (Request *r : edf_order(s, LYAP_REQ_VIEW)) {
        if (sl.est_units >= budget) break;
        if (r->state == QUEUED) {
            int rc = admit(s, *r);
            if (rc == LYAP_ERR_BUDGET) continue;
            if (rc) return rc;
        }
        const uint32_t left = r->n_units - r->next_unit;
        double room = (budget - sl.est_units) / r->unit_cost;
        uint32_t take = left;
        if (room < (double)left) {
            take = (uint32_t)std::max<double>(chunk, std::ceil(room / chunk) * chunk);
            if (take > left) take = left;
        }
        views.push_back(r->rq.view);
        begin.push_back(r->next_unit);
        count.push_back(take);
        sl.pieces.push_back({r->ticket, r->next_unit, take});
        r->next_unit += take;
        r->outstanding++;
        sl.est_units += take * r->unit_cost;
        uint64_t px = 0;   /* pixels inside the image among the piece's rays */
        {
            const uint32_t w = r->rq.view.width, h = r->rq.view.height, bx = (w + 7) / 8;
            for (uint32_t blk = begin.back() / 64; blk < (begin.back() + take) / 64; ++blk) {
                const uint32_t x0 = (blk % bx) * 8, y0 = (blk / bx) * 8;
                px += (uint64_t)std::min(8u, w - x0) * std::min(8u, h - y0);
            }
        }
        sl.work += px;
    }
    if (views.empty()) return LYAP_OK;
    LYAP_HIP(sl.ctx, hipEventRecord(sl.start, sl.ctx->stream));
    int rc = render_slot_launch(sl.ctx, sl.rs, views.data(), begin.data(), count.data(),
                                (unsigned)views.size(), s->cfg.mode, s->ctx->tuned);
    if (rc) return rc;
    LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_count, render_slot_counters(sl.rs) + 2, 8,
                                    hipMemcpyDeviceToHost, sl.ctx->stream));
    LYAP_HIP(sl.ctx, hipEventRecord(sl.end, sl.ctx->stream));
    sl.busy = true;
    s->stats.launches++;
    return LYAP_OK;
}

int issue_volume(lyap_server *s, Slot &sl)
{
    const double budget = s->cfg.slice_ms / s->volume_ms_per_unit;
    sl.pieces.clear();
    sl.est_units = 0;
    sl.work = 0;
    for (Request *r : edf_order(s, LYAP_REQ_VOLUME)) {
        if (sl.est_units >= budget || sl.pieces.size() >= kMaxSlabs) break;
        if (r->state == QUEUED) {
            int rc = admit(s, *r);
            if (rc == LYAP_ERR_BUDGET) continue;
            if (rc) return rc;
        }
        const uint32_t left = r->n_units - r->next_unit;
        const double room = (budget - sl.est_units) / r->unit_cost;
        const uint32_t take = room < (double)left ? std::max<uint32_t>(1u, (uint32_t)room) : left;
        const uint32_t nx = r->rq.vol_n[0], ny = r->rq.vol_n[1], nz = r->rq.vol_n[2];
        if (sl.pieces.empty()) LYAP_HIP(sl.ctx, hipEventRecord(sl.start, sl.ctx->stream));
        float *out = r->rq.vol_out + (size_t)nx * ny * r->next_unit;
        int rc = launch_volume_slab(sl.ctx, out, r->vol_e, r->rq.vol_lo, r->rq.vol_hi, nx, ny, nz,
                                    r->next_unit, r->next_unit + take, nullptr, 0, s->cfg.vol_mode);
        if (rc) return rc;
        /* the slab's re-evaluation count (the counter restarts per slab) */
        LYAP_HIP(sl.ctx, hipMemcpyAsync(sl.host_count + sl.pieces.size(), sl.ctx->counters, 4,
                                        hipMemcpyDeviceToHost, sl.ctx->stream));
        sl.pieces.push_back({r->ticket, r->next_unit, take});
        r->next_unit += take;
        r->outstanding++;
        sl.est_units += take * r->unit_cost;
        sl.work += (uint64_t)nx * ny * take;
    }
    if (sl.pieces.empty()) return LYAP_OK;
    LYAP_HIP(sl.ctx, hipEventRecord(sl.end, sl.ctx->stream));
    sl.busy = true;
    s->stats.launches++;
    return LYAP_OK;
}

/* Marks the pieces of a finished launch done; returns true when it was. */
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
    if (status == LYAP_OK && hipEventElapsedTime(&ms, sl.start, sl.end) == hipSuccess &&
        sl.est_units > 0) {
        /* running measure; launches of several slots overlap, so this is the
         * per-unit cost under sharing, which is what the next slice sees */
        double &m = sl.render ? s->render_ms_per_unit : s->volume_ms_per_unit;
        m = 0.75 * m + 0.25 * ((double)ms / sl.est_units);
        s->stats.device_ms += ms;
    }
    if (sl.render) {
        s->stats.rays += sl.work;
        if (status == LYAP_OK) s->stats.evals += (uint64_t)sl.host_count[0];
    } else {
        s->stats.voxels += sl.work;
    }
    for (size_t i = 0; i < sl.pieces.size(); ++i) {
        const Piece &p = sl.pieces[i];
        uint64_t vol_evals = 0;
        if (!sl.render && status == LYAP_OK) {
            /* the counter holds 32 bits; a slab has fewer than 2^32 voxels */
            vol_evals = (uint64_t)sl.host_count[i] & 0xFFFFFFFFull;
            s->stats.evals += vol_evals;
        }
        auto it = s->reqs.find(p.ticket);
        if (it == s->reqs.end()) continue;
        Request &r = it->second;
        r.outstanding--;
        if (!sl.render) {
            const uint64_t vox = (uint64_t)r.rq.vol_n[0] * r.rq.vol_n[1] * p.count;
            vol_evals += vox;
            r.evals += vol_evals;
            if (status == LYAP_OK) s->stats.evals += vox;
        }
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
        int rc = sl.render ? issue_render(s, sl) : issue_volume(s, sl);
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
    s->slots.resize(n);
    for (unsigned i = 0; i < n && rc == LYAP_OK; ++i) {
        Slot &sl = s->slots[i];
        sl.render = i < cfg.render_slots;
        rc = lyap_create(&sl.ctx, ctx->device, 0);
        if (rc) break;
        if (hipEventCreate(&sl.start) != hipSuccess || hipEventCreate(&sl.end) != hipSuccess ||
            hipHostMalloc((void **)&sl.host_count, kMaxSlabs * sizeof(unsigned long long),
                          hipHostMallocDefault) != hipSuccess) {
            rc = fail(ctx, LYAP_ERR_DEVICE, "server slot set-up failed");
            break;
        }
        std::memset(sl.host_count, 0, kMaxSlabs * sizeof(unsigned long long));
        if (sl.render) {
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
        if (sl.host_count) (void)hipHostFree(sl.host_count);
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
