/*
 * stubs.cpp - entries of the C ABI whose implementation is not built yet.
 * Each returns LYAP_ERR_UNSUPPORTED so callers fail loudly.
 */
#include "lyap.h"

extern "C" {
int lyap_server_create(lyap_ctx *, const LyapServerConfig *, lyap_server **out) { if (out) *out = 0; return LYAP_ERR_UNSUPPORTED; }
void lyap_server_destroy(lyap_server *) {}
int lyap_server_submit(lyap_server *, const LyapRequest *, uint64_t *) { return LYAP_ERR_UNSUPPORTED; }
int lyap_server_progress(lyap_server *, double) { return LYAP_ERR_UNSUPPORTED; }
int lyap_server_poll(lyap_server *, uint64_t, LyapResult *) { return LYAP_ERR_UNSUPPORTED; }
int lyap_server_release(lyap_server *, uint64_t) { return LYAP_ERR_UNSUPPORTED; }
int lyap_server_cancel(lyap_server *, uint64_t) { return LYAP_ERR_UNSUPPORTED; }
double lyap_server_clock_ms(lyap_server *) { return 0.0; }
}
