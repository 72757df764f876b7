// server.cpp -- resident serving loop (continuous batching). Placeholder until the
// serving runtime lands: every entry point reports SHIPFEM_UNSUPPORTED.
#include "shipfem_host.hpp"

#include <chrono>

extern "C" {
shipfem_status shipfem_server_create(const shipfem_server_opts*, shipfem_server** out, char* err,
                                     size_t err_len) {
    if (out) *out = nullptr;
    shipfem::setError(err, err_len, "serving loop not built");
    return SHIPFEM_UNSUPPORTED;
}
shipfem_status shipfem_server_submit(shipfem_server*, const shipfem_case_desc*, double, uint64_t*) {
    return SHIPFEM_UNSUPPORTED;
}
shipfem_status shipfem_server_poll(shipfem_server*, shipfem_result*, uint32_t, uint32_t*) {
    return SHIPFEM_UNSUPPORTED;
}
shipfem_status shipfem_server_keep(shipfem_server*, uint64_t) { return SHIPFEM_UNSUPPORTED; }
shipfem_status shipfem_server_fetch(shipfem_server*, uint64_t, shipfem_state_view*) {
    return SHIPFEM_UNSUPPORTED;
}
shipfem_status shipfem_server_drain(shipfem_server*, double) { return SHIPFEM_UNSUPPORTED; }
shipfem_status shipfem_server_stats_get(shipfem_server*, shipfem_server_stats*) {
    return SHIPFEM_UNSUPPORTED;
}
double shipfem_now(void) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
void shipfem_server_destroy(shipfem_server*) {}
}
