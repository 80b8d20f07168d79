/*
 * fake_server.h - C interface of the host-memory serving device used by the
 * scheduler tests (fake_server.cpp). The server it creates is driven with
 * the lyap_server_* calls of lyap.h.
 */
#ifndef FAKE_SERVER_H
#define FAKE_SERVER_H

#include <stddef.h>
#include <stdint.h>

#include "lyap.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct FakeOptions {
    int32_t manual;                    /* launches finish on fake_complete */
    uint32_t reserved;
    size_t device_bytes;               /* device memory size (0: unlimited) */
    double ms_per_launch;              /* automatic mode: duration (0: the estimate) */
    double tick_ms;                    /* idle step of the virtual clock (0: 0.05) */
} FakeOptions;

/* One issued piece. */
typedef struct FakeLogEntry {
    uint64_t ticket;
    uint32_t slot, launch;             /* slot index, launch number */
    int32_t kind;                      /* LYAP_REQ_* */
    uint32_t begin, count;             /* rays or z slices */
    double t;                          /* virtual clock at issue */
} FakeLogEntry;

int fake_server_create(const LyapServerConfig *cfg, const FakeOptions *opt, lyap_server **out);
const char *fake_last_error(lyap_server *s);
int fake_slot_count(lyap_server *s);
int fake_slot_kind(lyap_server *s, unsigned slot);
int fake_running(lyap_server *s, unsigned slot);
int fake_complete(lyap_server *s, unsigned slot, int status);
void fake_fail_next_launch(lyap_server *s, int code);
void fake_push_overrun(lyap_server *s, double factor);
void fake_advance(lyap_server *s, double ms);
double fake_rate(lyap_server *s, int kind);
size_t fake_pool_bytes(lyap_server *s);
size_t fake_device_in_use(lyap_server *s);
size_t fake_live_bytes_all(void);
unsigned fake_log(lyap_server *s, FakeLogEntry *out, unsigned cap);
size_t fake_coverage(lyap_server *s, uint64_t ticket, uint8_t *out, size_t cap);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* FAKE_SERVER_H */
