/*
 * serve_core.hpp - the serving loop's scheduling, apart from the device.
 *
 * Admission against the HBM budget, earliest-deadline-first pieces, slices
 * sized by the measured launch rates, retirement, cancellation and failure
 * handling are host code with no HIP in them (serve_core.cpp). They drive
 * the device through ServeBackend: server.hip implements it with one HIP
 * stream and event pair per slot over the render and volume kernels; a
 * host-memory backend whose launches finish when its caller says so runs
 * the same scheduler without a GPU.
 *
 * The reference has no request queue: it renders one view per launch from
 * its interactive loop (lyap_interactive.cu:711) and one volume per process
 * (lyap_calculate.cu:72).
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "lyap.h"

namespace lyap {
namespace serve {

enum SlotKind { SLOT_MIXED = 0, SLOT_RENDER = 1, SLOT_VOLUME = 2 };

/* z slabs one launch may carry */
constexpr size_t kMaxSlabs = 16;

/* One request's share of a launch: rays [begin, begin + count) of a view
 * (whole 64-ray blocks of 8x8 pixels) or z slices of a volume. `item` is its
 * index among the launch's views or among its volume blocks. */
struct Piece {
    uint64_t ticket;
    uint32_t begin, count;
    int kind;                          /* LYAP_REQ_* */
    uint32_t item;
};

/* What one slot launch runs. */
struct SlotLaunch {
    std::vector<LyapView> views;
    std::vector<uint32_t> begin, count;         /* per view: ray range */
    std::vector<LyapVolumeBlock> blocks;        /* volume z slabs */
    std::vector<Piece> pieces;                  /* in the order they were taken */
    double est_view_ms = 0, est_vol_ms = 0;     /* estimated device time of each part */
    uint64_t rays = 0, voxels = 0;              /* pixels inside the images, voxels */
    void clear();
};

/* What a finished launch reports. */
struct LaunchResult {
    double ms = -1.0;                  /* device time, < 0 when unknown */
    uint64_t evals = 0;                /* lambda evaluations of the launch */
    uint32_t reeval[kMaxSlabs] = {};   /* FP64 re-evaluations of each volume block */
};

/* The device side of a server. Every call comes from the one thread that
 * drives the server. */
class ServeBackend {
public:
    virtual ~ServeBackend() = default;
    /* the server clock, milliseconds */
    virtual double now_ms() = 0;
    /* device memory for resident outputs; LYAP_ERR_BUDGET when a cap of the
     * device context refuses it, another error when the device has no room */
    virtual int alloc(size_t bytes, void **out) = 0;
    /* returns memory alloc gave (no launch still running uses it) */
    virtual void release(void *p) = 0;
    /* issues slot `slot`'s launch (the slot has none running) */
    virtual int launch(unsigned slot, const SlotLaunch &l) = 0;
    /* LYAP_ERR_PENDING while slot's launch runs; LYAP_OK once it finished,
     * with `res` filled; another error when it failed on the device */
    virtual int query(unsigned slot, LaunchResult *res) = 0;
    /* called between queries while launches run and nothing changed */
    virtual void idle() = 0;
    /* records the message of an error; returns `code` */
    virtual int error(int code, const char *msg) = 0;
};

enum ReqState { QUEUED = 0, ACTIVE = 1, FINISHED = 2 };

struct Request {
    uint64_t ticket = 0;
    LyapRequest rq{};
    int state = QUEUED;
    int status = LYAP_ERR_PENDING;
    bool cancel = false;
    int error = LYAP_OK;               /* first failure of one of its launches */
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
};

struct Slot {
    int kind = SLOT_MIXED;
    bool busy = false;
    SlotLaunch launch;
};

/* Fills the defaults of a server configuration (a null `in` gives all of
 * them) and checks it; LYAP_OK or LYAP_ERR_INVALID_ARGUMENT with `why`. */
int check_config(const LyapServerConfig *in, LyapServerConfig *out, std::string *why);

/* Kind of each slot of a checked configuration. */
std::vector<int> slot_kinds(const LyapServerConfig &cfg);

} /* namespace serve */
} /* namespace lyap */

struct lyap_server {
    lyap::serve::ServeBackend *dev = nullptr;   /* owned */
    LyapServerConfig cfg{};
    uint64_t next_ticket = 1;
    std::map<uint64_t, lyap::serve::Request> reqs;
    std::vector<lyap::serve::Slot> slots;
    /* device buffers by size, not in use */
    std::multimap<size_t, void *> pool;
    size_t resident = 0, pooled = 0, peak = 0;
    /* measured device milliseconds per iteration unit */
    double render_ms_per_unit = 1.0e-9, volume_ms_per_unit = 1.0e-10;
    LyapServerStats stats{};
    std::vector<uint64_t> finished;    /* not yet handed out by lyap_server_finished */
    size_t finished_head = 0;
};

namespace lyap {
namespace serve {

/* A server over `dev` (taken over) with a checked configuration. */
lyap_server *server_new(const LyapServerConfig &cfg, ServeBackend *dev);

} /* namespace serve */
} /* namespace lyap */
