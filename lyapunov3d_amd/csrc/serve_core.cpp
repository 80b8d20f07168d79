/*
 * serve_core.cpp - scheduling of the serving loop (host only, no HIP).
 *
 *  - a view is a range of 64-ray blocks (8x8 pixels), cut into pieces of a
 *    multiple of 64 rays; a volume is a range of z slices;
 *  - a free slot takes the pending pieces of the requests with the earliest
 *    deadlines, up to about slice_ms of estimated device time, and the
 *    backend runs them as one launch;
 *  - a request is admitted when its first piece is taken: its resident
 *    outputs come from a pool of device buffers against the budget, and
 *    released buffers go back to the pool, so serving does not free device
 *    memory (a free waits for the whole device);
 *  - costs are estimated in "iteration units" (rays or voxels times
 *    settle + accum) and converted to milliseconds with a running measure of
 *    the slots' launch durations.
 */
#include "serve_core.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "sequence.hpp"

namespace lyap {
namespace serve {

void SlotLaunch::clear()
{
    views.clear();
    begin.clear();
    count.clear();
    blocks.clear();
    pieces.clear();
    est_view_ms = est_vol_ms = 0;
    rays = voxels = 0;
}

int check_config(const LyapServerConfig *in, LyapServerConfig *out, std::string *why)
{
    LyapServerConfig cfg{};
    cfg.mode.refine = LYAP_REFINE_BISECT;
    cfg.mode.bisect_steps = 8;
    cfg.mode.normal_h = 1e-3f;
    cfg.mode.precision = LYAP_PREC_PRODUCT;
    cfg.vol_mode.precision = LYAP_PREC_PRODUCT;
    if (in) cfg = *in;
    if (cfg.render_slots == 0) cfg.render_slots = 2;
    if (cfg.volume_slots == 0) cfg.volume_slots = 1;
    if (cfg.slice_ms <= 0) cfg.slice_ms = 12.0;
    if (cfg.chunk_rays == 0) cfg.chunk_rays = 4096;
    const char *bad = nullptr;
    char buf[96];
    if (cfg.flags & ~(uint32_t)LYAP_SERVER_SEPARATE_VOLUMES) {
        std::snprintf(buf, sizeof(buf), "unknown server flags 0x%x", cfg.flags);
        bad = buf;
    } else if (cfg.render_slots > 4 || cfg.volume_slots > 4) {
        bad = "at most 4 render and 4 volume slots";
    } else if (cfg.chunk_rays % 64u) {
        bad = "chunk_rays must be a multiple of 64";
    } else if (cfg.mode.refine != LYAP_REFINE_WALK && cfg.mode.refine != LYAP_REFINE_BISECT) {
        bad = "unknown refine mode";
    } else if (cfg.mode.schedule != LYAP_SCHEDULE_BATCHED && cfg.mode.schedule != LYAP_SCHEDULE_DIRECT &&
               cfg.mode.schedule != LYAP_SCHEDULE_LEVELS) {
        bad = "unknown schedule";
    } else if (cfg.mode.bisect_steps < 0 || cfg.mode.bisect_steps > 64 || cfg.mode.lanes_per_ray > 64) {
        bad = "bad render mode";
    } else if (cfg.mode.precision < LYAP_PREC_PLAIN || cfg.mode.precision > LYAP_PREC_FP64 ||
               cfg.vol_mode.precision < LYAP_PREC_PLAIN || cfg.vol_mode.precision > LYAP_PREC_FP64) {
        bad = "unknown precision";
    }
    if (bad) {
        if (why) *why = bad;
        return LYAP_ERR_INVALID_ARGUMENT;
    }
    *out = cfg;
    return LYAP_OK;
}

std::vector<int> slot_kinds(const LyapServerConfig &cfg)
{
    /* one launch carries both kinds only in one precision */
    const bool separate = (cfg.flags & LYAP_SERVER_SEPARATE_VOLUMES) ||
                          cfg.mode.precision != cfg.vol_mode.precision;
    const unsigned n = cfg.render_slots + cfg.volume_slots;
    std::vector<int> k(n);
    for (unsigned i = 0; i < n; ++i)
        k[i] = !separate ? SLOT_MIXED : (i < cfg.render_slots ? SLOT_RENDER : SLOT_VOLUME);
    return k;
}

lyap_server *server_new(const LyapServerConfig &cfg, ServeBackend *dev)
{
    lyap_server *s = new lyap_server();
    s->dev = dev;
    s->cfg = cfg;
    const std::vector<int> kinds = slot_kinds(cfg);
    s->slots.resize(kinds.size());
    for (size_t i = 0; i < kinds.size(); ++i) s->slots[i].kind = kinds[i];
    return s;
}

namespace {

int err(lyap_server *s, int code, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return s->dev->error(code, buf);
}

double now(lyap_server *s) { return s->dev->now_ms(); }

/* ---- resident memory ------------------------------------------------------ */

/* Frees pooled buffers: the largest first while the pool and `need` more
 * bytes would pass the budget, or all of them. */
void trim_pool(lyap_server *s, size_t need, bool all)
{
    const size_t cap = s->cfg.hbm_budget_bytes;
    while (!s->pool.empty() && (all || (cap && s->resident + s->pooled + need > cap))) {
        auto big = std::prev(s->pool.end());
        s->dev->release(big->second);
        s->pooled -= big->first;
        s->pool.erase(big);
    }
}

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
        /* pooled buffers of other sizes count against the budget */
        trim_pool(s, bytes, false);
        int rc = s->dev->alloc(bytes, out);
        if (rc && !s->pool.empty()) {
            /* the device (or its context's cap) has no room: give it the pool */
            trim_pool(s, bytes, true);
            rc = s->dev->alloc(bytes, out);
        }
        if (rc) {
            *out = nullptr;
            return rc;
        }
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
 * they do not fit the budget now; the allocation's error when the device
 * has no room. */
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
        int rc = pool_take(s, need, &p);
        if (rc) return rc;
        r.own_volume = (float *)p;
        r.volume_bytes = need;
        r.rq.vol_out = r.own_volume;
    }
    r.state = ACTIVE;
    r.start_ms = now(s);
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
    r.finish_ms = now(s);
    s->finished.push_back(r.ticket);
    if (status == LYAP_OK) s->stats.completed++;
    else if (status == LYAP_ERR_CANCELLED) s->stats.cancelled++;
    else s->stats.failed++;
}

/* Ends a request whose last piece on the device is done, or that has no
 * piece there: with the first launch failure it had, else cancelled, else
 * done once every piece was issued. */
void settle(lyap_server *s, Request &r)
{
    if (r.state == FINISHED || r.outstanding != 0) return;
    if (r.error != LYAP_OK) finish(s, r, r.error);
    else if (r.cancel) finish(s, r, LYAP_ERR_CANCELLED);
    else if (r.next_unit >= r.n_units) finish(s, r, LYAP_OK);
}

/* ---- slices --------------------------------------------------------------- */

bool has_work(const Request &r) { return r.state != FINISHED && !r.cancel && r.next_unit < r.n_units; }

/* Requests of `kind` (-1: both) with pieces left, earliest deadline first
 * (ties in submission order). */
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
 * what is left), volumes' z slabs (at most kMaxSlabs per launch). A queued
 * request whose outputs do not fit yet holds back the admission of the
 * queued requests behind it. */
void take_pieces(lyap_server *s, Slot &sl)
{
    SlotLaunch &l = sl.launch;
    const double budget = s->cfg.slice_ms;
    const uint32_t chunk = s->cfg.chunk_rays;
    const int want = sl.kind == SLOT_MIXED ? -1 : (sl.kind == SLOT_RENDER ? LYAP_REQ_VIEW : LYAP_REQ_VOLUME);
    bool held = false;
    double est = 0;
    for (Request *r : edf_order(s, want)) {
        if (est >= budget) break;
        const bool view = r->rq.kind == LYAP_REQ_VIEW;
        if (!view && l.blocks.size() >= kMaxSlabs) continue;
        if (r->state == QUEUED) {
            if (held) continue;
            const int rc = admit(s, *r);
            if (rc != LYAP_OK) {
                /* no room now: it waits while memory may come back (the
                 * budget, or outputs other requests hold); a device that
                 * cannot hold it with nothing else resident ends it */
                if (rc == LYAP_ERR_BUDGET || s->resident > 0) held = true;
                else finish(s, *r, rc);
                continue;
            }
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
            l.pieces.push_back({r->ticket, r->next_unit, take, LYAP_REQ_VIEW, (uint32_t)l.views.size()});
            l.views.push_back(r->rq.view);
            l.begin.push_back(r->next_unit);
            l.count.push_back(take);
            l.rays += piece_pixels(r->rq.view, r->next_unit, take);
            l.est_view_ms += take * r->unit_cost * rate;
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
            l.pieces.push_back({r->ticket, r->next_unit, take, LYAP_REQ_VOLUME, (uint32_t)l.blocks.size()});
            l.blocks.push_back(b);
            l.voxels += (uint64_t)b.n[0] * b.n[1] * take;
            l.est_vol_ms += take * r->unit_cost * rate;
        }
        r->next_unit += take;
        r->outstanding++;
        est = l.est_view_ms + l.est_vol_ms;
    }
}

/* Ends the pieces of a launch that failed: each request ends with `status`
 * once its other pieces are done. */
void fail_pieces(lyap_server *s, const SlotLaunch &l, int status)
{
    for (const Piece &p : l.pieces) {
        auto it = s->reqs.find(p.ticket);
        if (it == s->reqs.end()) continue;
        Request &r = it->second;
        r.outstanding--;
        r.cancel = true;
        if (r.error == LYAP_OK) r.error = status;
        settle(s, r);
    }
}

int issue(lyap_server *s, unsigned i)
{
    Slot &sl = s->slots[i];
    sl.launch.clear();
    take_pieces(s, sl);
    if (sl.launch.pieces.empty()) return LYAP_OK;
    const int rc = s->dev->launch(i, sl.launch);
    if (rc) return rc;
    sl.busy = true;
    s->stats.launches++;
    return LYAP_OK;
}

/* Marks the pieces of a finished launch done. */
int retire(lyap_server *s, unsigned i, bool *changed)
{
    Slot &sl = s->slots[i];
    if (!sl.busy) return LYAP_OK;
    LaunchResult res;
    const int status = s->dev->query(i, &res);
    if (status == LYAP_ERR_PENDING) return LYAP_OK;
    sl.busy = false;
    *changed = true;
    const SlotLaunch &l = sl.launch;
    const double est = l.est_view_ms + l.est_vol_ms;
    if (status == LYAP_OK && res.ms >= 0 && est > 0) {
        /* running measure of milliseconds per iteration unit, scaled by how
         * far this launch's estimate was off; launches of several slots
         * overlap, so it is the cost under sharing, which the next slice
         * sees */
        const double f = 0.75 + 0.25 * (res.ms / est);
        if (l.est_view_ms > 0) s->render_ms_per_unit *= f;
        if (l.est_vol_ms > 0) s->volume_ms_per_unit *= f;
        s->stats.device_ms += res.ms;
    }
    s->stats.rays += l.rays;
    s->stats.voxels += l.voxels;
    if (status != LYAP_OK) {
        fail_pieces(s, l, status);
        sl.launch.clear();
        return status;
    }
    s->stats.evals += res.evals;
    for (const Piece &p : l.pieces) {
        uint64_t reeval = 0;
        if (p.kind == LYAP_REQ_VOLUME) {
            reeval = p.item < kMaxSlabs ? res.reeval[p.item] : 0;
            s->stats.evals += reeval;
        }
        auto it = s->reqs.find(p.ticket);
        if (it == s->reqs.end()) continue;
        Request &r = it->second;
        r.outstanding--;
        if (p.kind == LYAP_REQ_VOLUME)
            r.evals += (uint64_t)r.rq.vol_n[0] * r.rq.vol_n[1] * p.count + reeval;
        settle(s, r);
    }
    sl.launch.clear();
    return LYAP_OK;
}

int step(lyap_server *s, bool *changed)
{
    int first_error = LYAP_OK;
    for (unsigned i = 0; i < s->slots.size(); ++i) {
        const int rc = retire(s, i, changed);
        if (rc && !first_error) first_error = rc;
    }
    for (unsigned i = 0; i < s->slots.size(); ++i) {
        Slot &sl = s->slots[i];
        if (sl.busy) continue;
        const int rc = issue(s, i);
        if (rc) {
            /* the pieces taken by the failed launch end their requests */
            fail_pieces(s, sl.launch, rc);
            sl.launch.clear();
            if (!first_error) first_error = rc;
            *changed = true;
            continue;
        }
        if (sl.busy) *changed = true;
    }
    return first_error;
}

bool any_busy(const lyap_server *s)
{
    for (const Slot &sl : s->slots)
        if (sl.busy) return true;
    return false;
}

int validate(lyap_server *s, const LyapRequest &rq, Request &r)
{
    if (rq.kind == LYAP_REQ_VIEW) {
        const LyapView &v = rq.view;
        if (v.width == 0 || v.height == 0) return err(s, LYAP_ERR_INVALID_ARGUMENT, "empty view");
        if ((uint64_t)v.width * v.height > (1ull << 31))
            return err(s, LYAP_ERR_UNSUPPORTED, "view larger than 2^31 pixels");
        if (v.num_lights > LYAP_MAX_LIGHTS || (v.num_lights && !v.lights))
            return err(s, LYAP_ERR_INVALID_ARGUMENT, "bad lights");
        if (sequence_period(v.seq) == 0) return err(s, LYAP_ERR_BAD_SEQUENCE, "bad sequence");
        r.n_units = 64u * ((v.width + 7) / 8) * ((v.height + 7) / 8);
        r.unit_cost = (double)v.prm.settle + (double)v.prm.accum;
        /* a ray takes tens of samples; the measured rate absorbs the rest */
        r.unit_cost *= 16.0;
    } else if (rq.kind == LYAP_REQ_VOLUME) {
        const uint64_t n = (uint64_t)rq.vol_n[0] * rq.vol_n[1] * rq.vol_n[2];
        if (n == 0) return err(s, LYAP_ERR_INVALID_ARGUMENT, "empty volume");
        if (n > 0xFFFFFFFFull) return err(s, LYAP_ERR_UNSUPPORTED, "more than 2^32-1 voxels");
        if (sequence_period(rq.vol_seq) == 0) return err(s, LYAP_ERR_BAD_SEQUENCE, "bad sequence");
        r.n_units = rq.vol_n[2];
        r.unit_cost = (double)rq.vol_n[0] * rq.vol_n[1] *
                      ((double)rq.vol_prm.settle + (double)rq.vol_prm.accum);
    } else {
        return err(s, LYAP_ERR_INVALID_ARGUMENT, "unknown request kind %d", rq.kind);
    }
    r.rq = rq;
    if (s->cfg.hbm_budget_bytes && request_bytes(r) > s->cfg.hbm_budget_bytes)
        return err(s, LYAP_ERR_BUDGET, "request needs more than the server budget");
    return LYAP_OK;
}

} /* namespace */
} /* namespace serve */
} /* namespace lyap */

using namespace lyap::serve;

extern "C" {

void lyap_server_destroy(lyap_server *s)
{
    if (!s) return;
    for (auto &kv : s->reqs) release_outputs(s, kv.second);
    for (auto &kv : s->pool) s->dev->release(kv.second);
    s->pool.clear();
    delete s->dev;
    delete s;
}

int lyap_server_submit(lyap_server *s, const LyapRequest *req, uint64_t *ticket)
{
    if (!s || !req || !ticket) return LYAP_ERR_INVALID_ARGUMENT;
    if (s->cfg.max_queued && s->stats.queued >= s->cfg.max_queued)
        return err(s, LYAP_ERR_QUEUE_FULL, "%u requests already wait for admission", s->stats.queued);
    Request r;
    int rc = validate(s, *req, r);
    if (rc) return rc;
    r.ticket = s->next_ticket++;
    r.submit_ms = now(s);
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
    const double until = now(s) + (wait_ms > 0 ? wait_ms : 0);
    for (;;) {
        bool changed = false;
        int rc = step(s, &changed);
        if (rc) return rc;
        if (changed || now(s) >= until) return LYAP_OK;
        if (!any_busy(s)) return LYAP_OK;   /* nothing to wait for */
        s->dev->idle();
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
    settle(s, r);
    return LYAP_OK;
}

double lyap_server_clock_ms(lyap_server *s) { return s ? now(s) : 0.0; }

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
    const double until = now(s) + timeout_ms;
    for (;;) {
        if (s->stats.queued == 0 && s->stats.active == 0) return LYAP_OK;
        int rc = lyap_server_progress(s, 1.0);
        if (rc) return rc;
        if (timeout_ms > 0 && now(s) >= until) return LYAP_ERR_PENDING;
        if (!any_busy(s) && s->stats.queued + s->stats.active > 0) {
            /* nothing runs and nothing could be issued: the budget holds the
             * queued requests back until a release */
            return LYAP_ERR_BUDGET;
        }
    }
}

} /* extern "C" */
