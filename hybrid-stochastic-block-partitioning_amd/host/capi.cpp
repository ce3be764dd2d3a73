// Flat entry point over the C++ driver, for Python callers (bench.py, tests).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>

#include "device.hpp"
#include "sbp_host.hpp"

extern "C" {

typedef struct sbp_run_stats {
    int32_t num_blocks;
    int32_t outer_iterations;
    int64_t sweeps, attempts, evals, moves;
    double seconds;       // wall time of stochastic_block_partition, graph already resident
    double mcmc_seconds;  // wall time inside the finetune phases
    double eval_ms[4];    // device time of the vertex-move evaluation kernels: degree bins 0-2, sequential MH
    double eval_bytes[4]; // their algorithmic bytes (when counting is on)
    int64_t eval_launches[4];
    double mdl;           // final MDL (float, as Blockmodel caches it)
    int32_t stopped_at_limit;  // 1 when the run stopped at the time limit (sbp_set_time_limit)
} sbp_run_stats;

// Time limit in seconds of the following partitions on this thread (Args::max_seconds; 0: none).
static thread_local double g_time_limit = 0.0;
extern "C" void sbp_set_time_limit(double seconds) { g_time_limit = seconds > 0.0 ? seconds : 0.0; }

static void set_err(char* err, int errlen, const char* msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg, errlen - 1);
        err[errlen - 1] = 0;
    }
}

// A graph resident on one GPU, for repeated partitioning runs.
// num_edges < 0: the stored edges count in the MDL; otherwise that count (see Graph).
void* sbp_graph_new(int32_t N, const int32_t* out_ptr, const int32_t* out_idx, const int32_t* in_ptr,
                    const int32_t* in_idx, int device, int64_t num_edges, char* err, int errlen) {
    try {
        int64_t E = out_ptr[N];
        return new Graph(N, std::vector<int32_t>(out_ptr, out_ptr + N + 1), std::vector<int32_t>(out_idx, out_idx + E),
                         std::vector<int32_t>(in_ptr, in_ptr + N + 1), std::vector<int32_t>(in_idx, in_idx + E),
                         device, (long)num_edges);
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}

void sbp_graph_free(void* g) { delete (Graph*)g; }

// Graph files as the reference reads them (graph_io.cpp): load, sizes, arrays, free.
void* sbp_file_load(const char* base_path, int undirected, char* err, int errlen) {
    try {
        return new graph_io::LoadedGraph(graph_io::load(base_path, undirected != 0));
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}
void sbp_file_info(void* h, int32_t* num_vertices, int64_t* stored_edges, int64_t* num_edges, int32_t* undirected,
                   int32_t* truth_len) {
    const graph_io::LoadedGraph& g = *(graph_io::LoadedGraph*)h;
    if (num_vertices) *num_vertices = g.num_vertices;
    if (stored_edges) *stored_edges = (int64_t)g.out_idx.size();
    if (num_edges) *num_edges = g.num_edges;
    if (undirected) *undirected = g.undirected ? 1 : 0;
    if (truth_len) *truth_len = (int32_t)g.truth.size();
}
void sbp_file_arrays(void* h, int32_t* out_ptr, int32_t* out_idx, int32_t* in_ptr, int32_t* in_idx, int32_t* truth) {
    const graph_io::LoadedGraph& g = *(graph_io::LoadedGraph*)h;
    std::copy(g.out_ptr.begin(), g.out_ptr.end(), out_ptr);
    std::copy(g.out_idx.begin(), g.out_idx.end(), out_idx);
    std::copy(g.in_ptr.begin(), g.in_ptr.end(), in_ptr);
    std::copy(g.in_idx.begin(), g.in_idx.end(), in_idx);
    if (truth) std::copy(g.truth.begin(), g.truth.end(), truth);
}
void sbp_file_free(void* h) { delete (graph_io::LoadedGraph*)h; }

// The last run's outer iterations (sbp::intermediate_results): wall seconds since the start,
// MDL and number of blocks after each merge + MCMC phase. Returns how many there are; fills at
// most cap entries.
int sbp_last_trajectory(double* seconds, double* mdl, int32_t* num_blocks, int cap) {
    const std::vector<sbp::Intermediate> r = sbp::get_intermediates();
    for (int i = 0; i < (int)r.size() && i < cap; ++i) {
        if (seconds) seconds[i] = r[i].seconds;
        if (mdl) mdl[i] = r[i].mdl;
        if (num_blocks) num_blocks[i] = r[i].num_blocks;
    }
    return (int)r.size();
}

// The last run's intermediates on this thread, 11 doubles per outer iteration: iteration, mdl,
// normalized_mdl_v1, normalized_mdl_v2, modularity (-1), interblock_edges, block_size_variation,
// num_blocks, mcmc_iterations, mcmc_time, seconds. Returns how many there are; fills at most cap.
int sbp_last_intermediates(double* rows, int cap) {
    const std::vector<sbp::Intermediate> r = sbp::get_intermediates();
    for (int i = 0; i < (int)r.size() && i < cap; ++i) {
        const sbp::Intermediate& x = r[i];
        const double row[11] = {x.iteration, x.mdl, x.normalized_mdl_v1, x.normalized_mdl_v2, x.modularity,
                                x.interblock_edges, x.block_size_variation, (double)x.num_blocks,
                                (double)x.mcmc_iterations, x.mcmc_time, x.seconds};
        std::copy(row, row + 11, rows + 11 * i);
    }
    return (int)r.size();
}

// A device option of the graph's context (sbpx_set_option), e.g. 4 = proposal (0 plain, 1 learned).
int sbp_graph_option(void* gp, int option, int64_t value, char* err, int errlen) {
    sbpx_ctx* ctx = ((Graph*)gp)->device()->ctx;
    if (sbpx_set_option(ctx, option, value) != SBPX_OK) {
        set_err(err, errlen, sbpx_last_error(ctx));
        return 1;
    }
    return 0;
}

// Installs the proposal network's parameters (SBPX_LEARNED_PARAMS floats) on the graph's context.
int sbp_graph_learned(void* gp, const float* params, char* err, int errlen) {
    sbpx_ctx* ctx = ((Graph*)gp)->device()->ctx;
    if (sbpx_learned_init(ctx, 0) != SBPX_OK || sbpx_learned_params(ctx, nullptr, params) != SBPX_OK) {
        set_err(err, errlen, sbpx_last_error(ctx));
        return 1;
    }
    return 0;
}

int sbp_partition_checkpointed(void* gp, int algorithm, int32_t batches, int approximate, uint64_t seed,
                               int count_bytes, const char* checkpoint, int resume, int32_t max_iterations,
                               int32_t* assignment, sbp_run_stats* stats, char* err, int errlen);

// sbp::stochastic_block_partition on a resident graph. algorithm: 0 = metropolis_hastings,
// 1 = async_gibbs, 2 = hybrid_mcmc. Returns 0 on success, otherwise writes a message into err.
int sbp_partition(void* gp, int algorithm, int32_t batches, int approximate, uint64_t seed, int count_bytes,
                  int32_t* assignment, sbp_run_stats* stats, char* err, int errlen) {
    return sbp_partition_checkpointed(gp, algorithm, batches, approximate, seed, count_bytes, nullptr, 0, 0, assignment,
                                      stats, err, errlen);
}

// As sbp_partition, writing the driver state to `checkpoint` after every outer iteration (when
// not null), continuing from it when `resume` is set, and stopping after `max_iterations` outer
// iterations when that is positive.
int sbp_partition_checkpointed(void* gp, int algorithm, int32_t batches, int approximate, uint64_t seed,
                               int count_bytes, const char* checkpoint, int resume, int32_t max_iterations,
                               int32_t* assignment, sbp_run_stats* stats, char* err, int errlen) {
    try {
        Graph& graph = *(Graph*)gp;
        const int N = graph.num_vertices();
        sbpx_set_option(graph.device()->ctx, 1, count_bytes ? 1 : 0);
        Args a;
        a.algorithm = algorithm == 0 ? "metropolis_hastings" : (algorithm == 1 ? "async_gibbs" : "hybrid_mcmc");
        a.batches = batches;
        a.approximate = approximate != 0;
        a.seed = seed;
        if (checkpoint) a.checkpoint = checkpoint;
        a.resume = resume != 0;
        a.max_iterations = max_iterations;
        a.max_seconds = g_time_limit;
        auto t0 = std::chrono::steady_clock::now();
        Blockmodel bm = sbp::stochastic_block_partition(graph, a);
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<int> assign = bm.block_assignment();
        if (assignment) std::memcpy(assignment, assign.data(), sizeof(int32_t) * N);
        if (stats) {
            stats->num_blocks = bm.getNum_blocks();
            stats->outer_iterations = (int)sbp::get_intermediates().size() - 2;
            stats->sweeps = finetune::totals.sweeps;
            stats->attempts = finetune::totals.attempts;
            stats->evals = finetune::totals.evals;
            stats->moves = finetune::totals.moves;
            stats->seconds = secs;
            stats->mcmc_seconds = finetune::MCMC_time;
            for (int b = 0; b < 4; ++b) {
                stats->eval_ms[b] = finetune::totals.eval_ms[b];
                stats->eval_bytes[b] = finetune::totals.eval_bytes[b];
                stats->eval_launches[b] = finetune::totals.eval_launches[b];
            }
            stats->mdl = bm.getOverall_entropy();
            stats->stopped_at_limit = sbp::stopped_at_time_limit() ? 1 : 0;
        }
        return 0;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return 1;
    }
}

// sbp::partition_batch: G graphs given as one union (vertex_offsets[G + 1]), each with its seed
// and edge count (num_edges may be null: stored edges), async_gibbs with approximate merges and
// `batches` batches; eval_variant, count_bytes and graph_replay as options 3, 1 and 9 of
// sbpx_set_option. Writes graph-local assignments (N entries), blocks and MDL per graph, and
// totals[18] = {wall seconds, rounds, batched sweeps, vertex-move evaluations, eval_ms[4],
// eval_bytes[4], eval_launches[4] (per kernel class, as in sbpx_sweep_stats), sweep rounds replayed
// from the captured graph, graph captures}.
int sbp_partition_batch(int32_t G, const int32_t* vertex_offsets, const int32_t* out_ptr, const int32_t* out_idx,
                        const int32_t* in_ptr, const int32_t* in_idx, const int64_t* num_edges, int32_t batches,
                        const uint64_t* seeds, int device, int eval_variant, int count_bytes, int graph_replay,
                        int32_t* assignment,
                        int32_t* num_blocks,
                        double* mdl, int32_t* mcmc_iterations, double* totals, char* err, int errlen) {
    try {
        const int N = vertex_offsets[G];
        const int64_t E = out_ptr[N];
        std::vector<long> ne(G, -1);
        if (num_edges)
            for (int g = 0; g < G; ++g) ne[g] = (long)num_edges[g];
        std::vector<Args> args(G);
        for (int g = 0; g < G; ++g) {
            args[g].algorithm = "async_gibbs";
            args[g].batches = batches;
            args[g].approximate = true;
            args[g].seed = seeds[g];
        }
        sbp::BatchRun run = sbp::partition_batch(
            std::vector<int32_t>(vertex_offsets, vertex_offsets + G + 1), std::vector<int32_t>(out_ptr, out_ptr + N + 1),
            std::vector<int32_t>(out_idx, out_idx + E), std::vector<int32_t>(in_ptr, in_ptr + N + 1),
            std::vector<int32_t>(in_idx, in_idx + E), ne, args, device, eval_variant, count_bytes != 0,
            graph_replay != 0);
        for (int g = 0; g < G; ++g) {
            const sbp::BatchRun::Graph& r = run.graphs[g];
            if (assignment) std::copy(r.assignment.begin(), r.assignment.end(), assignment + vertex_offsets[g]);
            if (num_blocks) num_blocks[g] = r.num_blocks;
            if (mdl) mdl[g] = r.mdl;
            if (mcmc_iterations) mcmc_iterations[g] = r.mcmc_iterations;
        }
        if (totals) {
            totals[0] = run.seconds;
            totals[1] = (double)run.rounds;
            totals[2] = (double)run.sweeps;
            totals[3] = (double)run.totals.evals;
            for (int k = 0; k < 4; ++k) {
                totals[4 + k] = run.totals.eval_ms[k];
                totals[8 + k] = run.totals.eval_bytes[k];
                totals[12 + k] = (double)run.totals.eval_launches[k];
            }
            totals[16] = (double)run.replayed;
            totals[17] = (double)run.captures;
        }
        return 0;
    } catch (const std::exception& e) {
        set_err(err, errlen, e.what());
        return 1;
    }
}

int sbp_run(int32_t N, const int32_t* out_ptr, const int32_t* out_idx, const int32_t* in_ptr, const int32_t* in_idx,
            int algorithm, int32_t batches, int approximate, uint64_t seed, int device, int count_bytes,
            int32_t* assignment, sbp_run_stats* stats, char* err, int errlen) {
    void* g = sbp_graph_new(N, out_ptr, out_idx, in_ptr, in_idx, device, -1, err, errlen);
    if (!g) return 1;
    int rc = sbp_partition(g, algorithm, batches, approximate, seed, count_bytes, assignment, stats, err, errlen);
    sbp_graph_free(g);
    return rc;
}

// finetune::early_stop (finetune.cpp:144-162) of this library on a list of per-sweep MDL changes;
// bracket != 0: the golden-section bracket has formed. No device work.
int sbp_early_stop(int iteration, int bracket, double initial, const double* deltas, int n) {
    if (n <= 0 || !deltas) return -1;
    auto dev = std::make_shared<sbpx_host::Device>();
    BlockmodelTriplet t;
    if (bracket) t.get(2) = Blockmodel::detached(1, 0.0f, dev);
    std::vector<double> d(deltas, deltas + n);
    return finetune::early_stop(iteration, t, initial, d) ? 1 : 0;
}

// The driver's search over the number of blocks (sbp.cpp:61-112, blockmodel_triplet.cpp:3-100)
// as this library runs it, with the merge phase and the sweeps replaced by a table: the
// blockmodel starts at n blocks; each outer iteration ends at B - num_blocks_to_merge blocks
// with overall entropy (float)score[B] and goes to BlockmodelTriplet::get_next_blockmodel.
// Writes the block count of every outer iteration, then the one the search returns; returns
// how many were written (-1 when cap is too small). No device work.
int sbp_golden_trace(int n, const double* score, int* out, int cap) {
    if (n <= 0 || !score || !out) return -1;
    auto dev = std::make_shared<sbpx_host::Device>();
    Blockmodel bm = Blockmodel::detached(n, 0.0f, dev);
    BlockmodelTriplet t;
    int k = 0;
    while (!sbp::done_blockmodeling(bm, t)) {
        const int B = bm.getNum_blocks() - bm.getNum_blocks_to_merge();
        Blockmodel next = Blockmodel::detached(B, (float)score[B], dev);
        if (k >= cap) return -1;
        out[k++] = B;
        bm = t.get_next_blockmodel(next);
    }
    if (k >= cap) return -1;
    out[k++] = bm.getNum_blocks();
    return k;
}

}  // extern "C"
