// server.cpp -- serving loop of libtiltrec_hip.so (trec_server_*).
// Placeholder until the request loop lands: every call reports UNSUPPORTED.
#include "trec_host.h"

extern "C" {
trec_status trec_server_create(int, size_t, int, trec_server **out)
{
    if (out) *out = nullptr;
    return TREC_ERR_UNSUPPORTED;
}
trec_status trec_server_submit(trec_server *, const trec_request *, uint64_t *) { return TREC_ERR_UNSUPPORTED; }
trec_status trec_server_poll(trec_server *, uint64_t, int *) { return TREC_ERR_UNSUPPORTED; }
trec_status trec_server_wait(trec_server *, uint64_t, trec_result *) { return TREC_ERR_UNSUPPORTED; }
void trec_server_destroy(trec_server *) {}
}
