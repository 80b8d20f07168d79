/*
 * fake_server.cpp - the serving scheduler (serve_core.cpp) over a host-memory
 * device, for CPU tests of admission, ordering, completion order,
 * cancellation and failures.
 *
 * Device memory is host memory. A launch "runs" when it is issued: it writes
 * every view pixel's lambda as the pixel's index and its RGBA as the
 * ticket's low bytes, every voxel as its index in the volume, and counts how
 * often each pixel or voxel of each ticket was written. It finishes
 *  - in automatic mode, on a virtual clock: after `ms_per_launch` (or, when
 *    that is 0, the scheduler's own estimate) times the next overrun factor;
 *    idle() moves the clock to the next finish;
 *  - in manual mode, when fake_complete() says so, with the status given;
 *    idle() moves the clock by `tick_ms`.
 * Failures are injected from the host: the next launch's issue error, a
 * launch's device error, a device memory size.
 */
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "fake_server.h"
#include "serve_core.hpp"

using namespace lyap::serve;

namespace {

struct FakeSlot {
    bool running = false;
    bool done = false;                 /* manual mode: completed by the caller */
    int status = LYAP_OK;
    double start = 0, finish = 0;
    uint64_t evals = 0;
    unsigned nblocks = 0;
};

typedef FakeLogEntry LogEntry;

size_t g_live_bytes = 0;               /* over every fake device */

class FakeBackend final : public ServeBackend {
public:
    bool manual = false;
    size_t device_bytes = 0;           /* 0: unlimited */
    double ms_per_launch = 0.0, tick_ms = 0.05;
    double t = 0.0;
    size_t in_use = 0;
    std::map<void *, size_t> live;
    std::vector<FakeSlot> slots;
    std::deque<int> fail_launch;       /* errors of the next launches' issue */
    std::deque<double> overrun;        /* duration factors of the next launches */
    std::vector<LogEntry> log;
    std::map<uint64_t, std::vector<uint8_t>> coverage;
    uint32_t launches = 0;
    std::string last;

    ~FakeBackend() override
    {
        for (auto &kv : live) {
            std::free(kv.first);
            g_live_bytes -= kv.second;
        }
    }

    double now_ms() override { return t; }

    int alloc(size_t bytes, void **out) override
    {
        *out = nullptr;
        if (device_bytes && in_use + bytes > device_bytes) return error(LYAP_ERR_DEVICE, "out of device memory");
        void *p = std::malloc(bytes);
        if (!p) return error(LYAP_ERR_DEVICE, "host malloc failed");
        std::memset(p, 0xA5, bytes);   /* never zero: outputs must be written */
        live[p] = bytes;
        in_use += bytes;
        g_live_bytes += bytes;
        *out = p;
        return LYAP_OK;
    }

    void release(void *p) override
    {
        auto it = live.find(p);
        if (it == live.end()) {
            last = "release of memory the device did not give";
            std::abort();
        }
        in_use -= it->second;
        g_live_bytes -= it->second;
        live.erase(it);
        std::free(p);
    }

    int launch(unsigned i, const SlotLaunch &l) override
    {
        if (i >= slots.size() || slots[i].running) {
            last = "launch on a busy or unknown slot";
            std::abort();
        }
        if (!fail_launch.empty()) {
            const int code = fail_launch.front();
            fail_launch.pop_front();
            if (code != LYAP_OK) return error(code, "injected launch failure");
        }
        FakeSlot &sl = slots[i];
        sl.running = true;
        sl.done = false;
        sl.status = LYAP_OK;
        sl.start = t;
        double factor = 1.0;
        if (!overrun.empty()) {
            factor = overrun.front();
            overrun.pop_front();
        }
        const double base = ms_per_launch > 0 ? ms_per_launch : (l.est_view_ms + l.est_vol_ms);
        sl.finish = t + base * factor;
        sl.evals = l.rays + l.voxels;
        sl.nblocks = (unsigned)l.blocks.size();
        for (const Piece &p : l.pieces) {
            log.push_back({p.ticket, i, launches, p.kind, p.begin, p.count, t});
            if (p.kind == LYAP_REQ_VIEW) run_view(p.ticket, l.views[p.item], p.begin, p.count);
            else run_block(p.ticket, l.blocks[p.item]);
        }
        launches++;
        return LYAP_OK;
    }

    int query(unsigned i, LaunchResult *res) override
    {
        FakeSlot &sl = slots[i];
        if (!sl.running) {
            last = "query of a slot with no launch";
            std::abort();
        }
        if (manual ? !sl.done : t < sl.finish) return LYAP_ERR_PENDING;
        sl.running = false;
        if (manual) sl.finish = t;
        if (sl.status != LYAP_OK) return error(sl.status, "injected device error");
        res->ms = sl.finish - sl.start;
        res->evals = sl.evals;
        for (unsigned b = 0; b < sl.nblocks && b < kMaxSlabs; ++b) res->reeval[b] = 0;
        return LYAP_OK;
    }

    void idle() override
    {
        if (manual) {
            t += tick_ms;
            return;
        }
        double next = -1.0;
        for (const FakeSlot &sl : slots)
            if (sl.running && (next < 0 || sl.finish < next)) next = sl.finish;
        t = next > t ? next : t + tick_ms;
    }

    int error(int code, const char *msg) override
    {
        last = msg;
        return code;
    }

private:
    std::vector<uint8_t> &cover(uint64_t ticket, size_t n)
    {
        std::vector<uint8_t> &c = coverage[ticket];
        if (c.size() < n) c.resize(n, 0);
        return c;
    }

    /* rays [begin, begin + count): 64-ray blocks of 8x8 pixels, row-major
     * blocks, row-major pixels within a block, clipped at the edges */
    void run_view(uint64_t ticket, const LyapView &v, uint32_t begin, uint32_t count)
    {
        const uint32_t w = v.width, h = v.height, bx = (w + 7) / 8;
        std::vector<uint8_t> &c = cover(ticket, (size_t)w * h);
        for (uint32_t r = begin; r < begin + count; ++r) {
            const uint32_t blk = r / 64, k = r % 64;
            const uint32_t x = (blk % bx) * 8 + k % 8, y = (blk / bx) * 8 + k / 8;
            if (x >= w || y >= h) continue;
            const size_t px = (size_t)y * w + x;
            if (v.lambda) v.lambda[px] = (float)px;
            if (v.rgba) {
                v.rgba[px].r = (uint8_t)(ticket & 0xff);
                v.rgba[px].g = (uint8_t)((ticket >> 8) & 0xff);
                v.rgba[px].b = 7;
                v.rgba[px].a = 255;
            }
            if (c[px] < 255) c[px]++;
        }
    }

    void run_block(uint64_t ticket, const LyapVolumeBlock &b)
    {
        const size_t plane = (size_t)b.n[0] * b.n[1];
        std::vector<uint8_t> &c = cover(ticket, plane * b.n[2]);
        for (size_t i = 0; i < plane * (b.z_end - b.z_begin); ++i) {
            const size_t v = (size_t)b.z_begin * plane + i;
            b.out[i] = (float)v;
            if (c[v] < 255) c[v]++;
        }
    }
};

FakeBackend *fake(lyap_server *s) { return static_cast<FakeBackend *>(s->dev); }

} /* namespace */

extern "C" {

int fake_server_create(const LyapServerConfig *cfg_in, const FakeOptions *opt, lyap_server **out)
{
    std::string why;
    if (!out) return LYAP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    LyapServerConfig cfg;
    int rc = check_config(cfg_in, &cfg, &why);
    if (rc) return rc;
    FakeBackend *dev = new FakeBackend();
    if (opt) {
        dev->manual = opt->manual != 0;
        dev->device_bytes = opt->device_bytes;
        dev->ms_per_launch = opt->ms_per_launch;
        if (opt->tick_ms > 0) dev->tick_ms = opt->tick_ms;
    }
    dev->slots.resize(slot_kinds(cfg).size());
    *out = server_new(cfg, dev);
    return LYAP_OK;
}

const char *fake_last_error(lyap_server *s) { return fake(s)->last.c_str(); }

int fake_slot_count(lyap_server *s) { return (int)s->slots.size(); }

int fake_slot_kind(lyap_server *s, unsigned slot) { return slot < s->slots.size() ? s->slots[slot].kind : -1; }

int fake_running(lyap_server *s, unsigned slot)
{
    FakeBackend *f = fake(s);
    return slot < f->slots.size() && f->slots[slot].running ? 1 : 0;
}

/* Manual mode: the running launch of `slot` finishes with `status`. */
int fake_complete(lyap_server *s, unsigned slot, int status)
{
    FakeBackend *f = fake(s);
    if (slot >= f->slots.size() || !f->slots[slot].running || f->slots[slot].done) return LYAP_ERR_NOT_FOUND;
    f->slots[slot].done = true;
    f->slots[slot].status = status;
    return LYAP_OK;
}

void fake_fail_next_launch(lyap_server *s, int code) { fake(s)->fail_launch.push_back(code); }

void fake_push_overrun(lyap_server *s, double factor) { fake(s)->overrun.push_back(factor); }

void fake_advance(lyap_server *s, double ms) { fake(s)->t += ms; }

/* The scheduler's measured milliseconds per iteration unit of views
 * (kind LYAP_REQ_VIEW) or volumes. */
double fake_rate(lyap_server *s, int kind)
{
    return kind == LYAP_REQ_VIEW ? s->render_ms_per_unit : s->volume_ms_per_unit;
}

size_t fake_pool_bytes(lyap_server *s) { return s->pooled; }

size_t fake_device_in_use(lyap_server *s) { return fake(s)->in_use; }

size_t fake_live_bytes_all(void) { return g_live_bytes; }

/* Pieces issued so far, in issue order: their number; up to `cap` copied. */
unsigned fake_log(lyap_server *s, LogEntry *out, unsigned cap)
{
    FakeBackend *f = fake(s);
    for (unsigned i = 0; i < cap && i < f->log.size(); ++i) out[i] = f->log[i];
    return (unsigned)f->log.size();
}

/* How often each pixel or voxel of `ticket` was written (0 when none was). */
size_t fake_coverage(lyap_server *s, uint64_t ticket, uint8_t *out, size_t cap)
{
    FakeBackend *f = fake(s);
    auto it = f->coverage.find(ticket);
    if (it == f->coverage.end()) return 0;
    const size_t n = it->second.size();
    if (out) std::memcpy(out, it->second.data(), n < cap ? n : cap);
    return n;
}

} /* extern "C" */
