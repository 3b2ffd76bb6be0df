// Serving entry points (continuous batching). Implemented in a later step;
// until then every call reports ELQ_ERR_INVALID_ARG.
#include "../../include/elq.h"

extern "C" {
int elq_server_create(int, const elq_server_opts*, elq_server** out) {
    if (out) *out = nullptr;
    return ELQ_ERR_INVALID_ARG;
}
void elq_server_destroy(elq_server*) {}
int elq_server_submit(elq_server*, const elq_grid*, float, int64_t*) { return ELQ_ERR_INVALID_ARG; }
int elq_server_step(elq_server*) { return 0; }
int elq_server_poll(elq_server*, elq_result*, int) { return 0; }
size_t elq_server_footprint(const elq_server*, int32_t, int32_t, int32_t) { return 0; }
}
