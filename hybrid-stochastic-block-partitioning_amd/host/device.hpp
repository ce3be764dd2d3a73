// The per-graph device binding shared by the host layer's translation units.
#pragma once
#include <cstdint>

#include "sbpx.h"

namespace sbpx_host {

struct Device {
    sbpx_ctx* ctx = nullptr;
    long live_id = -1;  // id of the Blockmodel whose state the device holds
    long next_id = 0;
    uint32_t sweep_call = 0;  // call id of the next MCMC sweep (randomness stream)
    uint32_t merge_call = 0;  // call id of the next merge phase
    int N = 0, E = 0;
    ~Device() {
        if (ctx) sbpx_destroy(ctx);
    }
};

}  // namespace sbpx_host
