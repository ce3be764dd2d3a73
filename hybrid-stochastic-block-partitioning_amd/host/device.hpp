// The per-graph device binding shared by the host layer's translation units: a graph either owns
// a context or is one member of a batch of graphs held in a shared context (batch.cpp), whose
// coordinator runs the members' device steps together.
#pragma once
#include <cstdint>

#include "sbpx.h"

namespace sbpx_host {

struct Batch;

struct Device {
    sbpx_ctx* ctx = nullptr;  // this graph's own context (null for a member of a batch)
    Batch* batch = nullptr;   // the batch this graph belongs to, if any
    int slot = -1;            // its index in the batch
    long live_id = -1;  // id of the Blockmodel whose state the device holds
    long next_id = 0;
    uint32_t sweep_call = 0;  // call id of the next MCMC sweep (randomness stream)
    uint32_t merge_call = 0;  // call id of the next merge phase
    int N = 0, E = 0;
    long stored_edges = 0;  // edges held (E may be the count a loaded file reports)

    // The device steps of the driver; each throws std::runtime_error with the library's message.
    void build(int num_blocks, const int32_t* assignment);
    void download(int32_t* assignment);
    double log_posterior();
    void sweep(int algo, int batches, uint64_t seed, uint32_t call, sbpx_sweep_stats* stats);
    int merge(int num_to_merge, int mode, uint64_t seed, uint32_t call);  // returns the number of blocks left
    long trace();  // sum of M[b, b]
    sbpx_ctx* own_ctx(const char* what) const;  // this graph's context; throws for a batch member

    ~Device() {
        if (ctx) sbpx_destroy(ctx);
    }
};

}  // namespace sbpx_host
