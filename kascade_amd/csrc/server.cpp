// placeholder, replaced by the continuous-batching server
#include "common.hpp"
extern "C" {
kc_status kc_server_create(void*, const kc_server_cfg*, void** out) { if (out) *out = nullptr; return KC_EINVAL; }
void kc_server_destroy(void*) {}
kc_status kc_submit(void*, const kc_job*, uint64_t*) { return KC_EINVAL; }
kc_status kc_poll(void*, uint64_t) { return KC_EINVAL; }
kc_status kc_wait(void*, uint64_t) { return KC_EINVAL; }
kc_status kc_server_stats_get(void*, kc_server_stats*) { return KC_EINVAL; }
kc_status kc_server_reset_stats(void*) { return KC_EINVAL; }
}
