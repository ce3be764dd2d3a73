// The reference's C++ API for the stochastic-block-partitioning path, on top of the sbpx C ABI.
// Names, argument meaning and control flow follow the reference (include/sbp.hpp,
// include/finetune.hpp, include/block_merge.hpp, include/entropy.hpp,
// include/blockmodel/blockmodel.hpp, include/blockmodel/blockmodel_triplet.hpp); the data live
// on one GPU. Errors raise std::runtime_error with the library's message.
#pragma once
#include <climits>
#include <cstdio>
#include <cstdint>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "sbpx.h"

/// Hot-path flags of include/args.hpp:13-130 plus the seed and device this build adds.
struct Args {
    std::string algorithm = "async_gibbs";  // async_gibbs | metropolis_hastings | hybrid_mcmc
    bool approximate = false;
    int asynciterations = INT_MAX;
    int batches = 10;
    // The reference defaults to false (include/args.hpp:98; the published runs pass --transpose).
    // The device always keeps both orientations, so the flag has no effect here; it defaults to
    // true to say so.
    bool transpose = true;
    int threads = 0;        // unused on the device path
    uint64_t seed = 0;
    int device = 0;
    // checkpointing of the outer loop (no counterpart in the reference): after every outer
    // iteration the driver state is written to `checkpoint` (when set); with `resume` a run
    // continues from that file; `max_iterations` > 0 stops after that many outer iterations.
    std::string checkpoint;
    bool resume = false;
    int max_iterations = 0;
    // time limit (no counterpart in the reference): > 0 stops the partition, after the sweep in
    // progress, once it has run this many seconds; sbp::stopped_at_time_limit() then reports it
    double max_seconds = 0.0;
};

namespace sbpx_host {
struct Device;  // one context per graph
}

namespace graph_io {
/// A graph read as the reference reads it (see graph_io.cpp).
struct LoadedGraph {
    int num_vertices = 0;
    long num_edges = 0;  // the count the reference's MDL uses
    bool undirected = false;
    std::vector<int32_t> out_ptr, out_idx, in_ptr, in_idx;
    std::vector<int> truth;  // planted blocks (0-based; -1 where the file names none)
};
/// Reads base_path + ".tsv" (else ".mtx") and base_path + "_truePartition.tsv" (graph.cpp:24-62).
LoadedGraph load(const std::string& base_path, bool undirected);
}  // namespace graph_io

/// typedefs.hpp:20, utils.hpp:21-28 and the EdgeWeights of the reference.
using NeighborList = std::vector<std::vector<int>>;
struct EdgeWeights {
    std::vector<int> indices;
    std::vector<int> values;
};
namespace utils {
typedef struct proposal_and_edge_counts_t {
    int proposal;
    int num_out_neighbor_edges;
    int num_in_neighbor_edges;
    int num_neighbor_edges;
} ProposalAndEdgeCounts;
}  // namespace utils

/// Delta (delta.hpp:22-149): the changes a move makes to the cells of rows and columns r
/// (current block) and s (proposed block), routed row first. A cell keeps its key when its change
/// returns to zero, as the reference's hash maps do; entries() lists them in ascending order.
class Delta {
public:
    Delta() = default;
    Delta(int current_block, int proposed_block, int buckets = 10)
        : current_(current_block), proposed_(proposed_block) { (void)buckets; }
    void add(int row, int col, int value) { int key; route(row, col, key)[key] += value; }
    void sub(int row, int col, int value) { int key; route(row, col, key)[key] -= value; }
    int get(int row, int col) const;
    std::vector<std::tuple<int, int, int>> entries() const;
    int self_edge_weight() const { return self_weight_; }
    void self_edge_weight(int weight) { self_weight_ = weight; }
    int current_block() const { return current_; }
    int proposed_block() const { return proposed_; }

private:
    std::map<int, int>& route(int row, int col, int& key);
    int current_ = -1, proposed_ = -1, self_weight_ = 0;
    std::map<int, int> current_row_, proposed_row_, current_col_, proposed_col_;
};

class Graph {
public:
    Graph() = default;
    /// num_edges < 0: the number of stored edges; otherwise the count the MDL uses (as a loaded
    /// MatrixMarket or undirected graph reports it).
    Graph(int num_vertices, std::vector<int32_t> out_ptr, std::vector<int32_t> out_idx, std::vector<int32_t> in_ptr,
          std::vector<int32_t> in_idx, int device = 0, long num_edges = -1);
    /// Graph::load (graph.cpp:24-62) from a path without extension.
    static Graph load(const std::string& base_path, bool undirected = false, int device = 0);
    /// A graph of a batch: its device steps run through the batch's shared context.
    Graph(int num_vertices, long num_edges, std::shared_ptr<sbpx_host::Device> dev)
        : num_vertices_(num_vertices), num_edges_((int)num_edges), dev_(std::move(dev)) {}
    int num_vertices() const { return num_vertices_; }
    int num_edges() const { return num_edges_; }
    const std::vector<int>& high_degree_vertices() const { return high_; }
    const std::vector<int>& low_degree_vertices() const { return low_; }
    std::shared_ptr<sbpx_host::Device> device() const { return dev_; }
    /// Graph::out_neighbors / in_neighbors (graph.hpp): adjacency lists in stored order, built on
    /// first use from the host copy of the CSR / CSC (not available for a graph of a batch).
    const NeighborList& out_neighbors() const;
    const NeighborList& in_neighbors() const;

private:
    int num_vertices_ = 0, num_edges_ = 0;
    std::vector<int> high_, low_;
    std::shared_ptr<sbpx_host::Device> dev_;
    std::shared_ptr<const std::vector<int32_t>> out_ptr_, out_idx_, in_ptr_, in_idx_;
    mutable std::shared_ptr<NeighborList> out_lists_, in_lists_;
};

class Blockmodel;
/// The block matrix as Blockmodel::blockmatrix() returns it (ISparseMatrix, csparse_matrix.hpp:
/// 62-139), read from the device holding the blockmodel.
class BlockMatrix {
public:
    explicit BlockMatrix(const Blockmodel& owner) : owner_(owner) {}
    int get(int row, int col) const;
    /// Nonzeros of a row / column as (index, value) in ascending index order.
    std::map<int, int> getrow_sparse(int row) const;
    std::map<int, int> getcol_sparse(int col) const;

private:
    const Blockmodel& owner_;
};

static const double BLOCK_REDUCTION_RATE = 0.5;

class Blockmodel {
public:
    Blockmodel() : empty(true) {}
    Blockmodel(int num_blocks, const Graph& graph, float block_reduction_rate);
    Blockmodel(int num_blocks, const Graph& graph, float block_reduction_rate, std::vector<int>& block_assignment);
    Blockmodel copy();
    std::vector<int> block_assignment() const;
    std::vector<int> degrees() const;
    std::vector<int> degrees_in() const;
    std::vector<int> degrees_out() const;
    int getNum_blocks() const { return num_blocks; }
    void setNum_blocks(int n) { num_blocks = n; }
    int& getNum_blocks_to_merge() { return num_blocks_to_merge; }
    void setNum_blocks_to_merge(int n) { num_blocks_to_merge = n; }
    float& getOverall_entropy() { return overall_entropy; }
    void setOverall_entropy(double e) { overall_entropy = (float)e; }
    float& getBlock_reduction_rate() { return block_reduction_rate; }
    double block_size_variation() const;
    double interblock_edges() const;
    /// block_assignment(vertex) (blockmodel.hpp:124).
    int block_assignment(int vertex) const;
    /// blockmatrix() (blockmodel.hpp:120): cell reads go to the device.
    std::shared_ptr<BlockMatrix> blockmatrix() const { return std::make_shared<BlockMatrix>(*this); }
    /// move_vertex(vertex, delta, proposal) (blockmodel.cpp:344-360) on the device: the vertex takes
    /// proposal.proposal, the matrix and dout/din follow, and the total degrees of the two blocks
    /// keep the reference's drift. `delta` must be the move's own (finetune::blockmodel_delta).
    void move_vertex(int vertex, const Delta& delta, utils::ProposalAndEdgeCounts& proposal);
    bool empty;

    // device binding
    void make_live() const;
    sbpx_ctx* ctx() const;  // the graph's own context (throws for a graph of a batch)
    sbpx_host::Device& device() const { return *dev_; }
    void mark_live() const;
    // checkpoint I/O of the fields and the assignment (the device state is rebuilt when used)
    void write(std::FILE* f) const;
    static Blockmodel read(std::FILE* f, const Graph& graph);
    /// A blockmodel that holds only its block count and overall entropy, with no device state:
    /// for running the search over the number of blocks alone (it supports copy() and the
    /// fields, not the device steps).
    static Blockmodel detached(int num_blocks, float overall_entropy, std::shared_ptr<sbpx_host::Device> dev);

private:
    std::shared_ptr<sbpx_host::Device> dev_;
    long id_ = -1;
    int num_blocks = 0;
    int num_blocks_to_merge = 0;
    float block_reduction_rate = 0.5f;
    float overall_entropy = std::numeric_limits<float>::max();
    mutable std::vector<int> snapshot_;  // assignment, when another blockmodel holds the device
};

class BlockmodelTriplet {
public:
    BlockmodelTriplet() : optimal_num_blocks_found(false) {}
    bool optimal_num_blocks_found;
    void update(Blockmodel& blockmodel);
    Blockmodel& get(int i) { return blockmodels[i]; }
    bool golden_ratio_not_reached();
    bool is_done();
    Blockmodel get_next_blockmodel(Blockmodel& old_blockmodel);

    void write(std::FILE* f) const;
    void read(std::FILE* f, const Graph& graph);

protected:
    Blockmodel blockmodels[3];
    int lower_difference();
    int upper_difference();
};

namespace entropy {
double mdl(const Blockmodel& blockmodel, int num_vertices, int num_edges);
/// delta_mdl(blockmodel, delta, proposal) (entropy.hpp:43, entropy.cpp:296-401), on the device.
double delta_mdl(const Blockmodel& blockmodel, const Delta& delta, const utils::ProposalAndEdgeCounts& proposal);
/// hastings_correction(vertex, graph, blockmodel, delta, current_block, proposal) (entropy.hpp:56,
/// entropy.cpp:495-557): the device evaluates the move of `vertex` to proposal.proposal; `delta`
/// must be that move's Delta, as in every call the reference makes.
double hastings_correction(int vertex, const Graph& graph, const Blockmodel& blockmodel, const Delta& delta,
                           int current_block, const utils::ProposalAndEdgeCounts& proposal);
/// block_merge_delta_mdl(current_block, proposal, blockmodel, delta) (entropy.hpp:27,
/// entropy.cpp:127-180): the device scores merging current_block into proposal.proposal; `delta`
/// must be that merge's (block_merge::blockmodel_delta).
double block_merge_delta_mdl(int current_block, utils::ProposalAndEdgeCounts proposal, const Blockmodel& blockmodel,
                             const Delta& delta);
double null_mdl_v1(long num_edges);
double null_mdl_v2(long num_vertices, long num_edges);
double normalize_mdl_v1(double mdl, long num_edges);
double normalize_mdl_v2(double mdl, long num_vertices, long num_edges);
}  // namespace entropy

namespace finetune {
extern thread_local int MCMC_iterations;
extern thread_local double MCMC_time;
static const int MAX_NUM_ITERATIONS = 100;
struct Totals {
    long long sweeps = 0, attempts = 0, evals = 0, moves = 0;
    double eval_ms[4] = {0, 0, 0, 0}, eval_bytes[4] = {0, 0, 0, 0};
    long long eval_launches[4] = {0, 0, 0, 0};
};
extern thread_local Totals totals;
Blockmodel& asynchronous_gibbs(Blockmodel& blockmodel, const Graph& graph, BlockmodelTriplet& blockmodels);
Blockmodel& hybrid_mcmc(Blockmodel& blockmodel, const Graph& graph, BlockmodelTriplet& blockmodels);
Blockmodel& metropolis_hastings(Blockmodel& blockmodel, const Graph& graph, BlockmodelTriplet& blockmodels);
bool early_stop(int iteration, BlockmodelTriplet& blockmodels, double initial_entropy,
                std::vector<double>& delta_entropies);
/// edge_weights (finetune.cpp:275-290): the neighbours of `vertex`, weight 1 each.
EdgeWeights edge_weights(const NeighborList& neighbors, int vertex, bool ignore_self = false);
/// blockmodel_delta (finetune.hpp:67-68, finetune.cpp:108-142): the Delta of moving `vertex` from
/// current_block to proposed_block, from the blocks its neighbours hold on the device.
Delta blockmodel_delta(int vertex, int current_block, int proposed_block, const EdgeWeights& out_edges,
                       const EdgeWeights& in_edges, const Blockmodel& blockmodel);
}  // namespace finetune

namespace block_merge {
static const int NUM_AGG_PROPOSALS_PER_BLOCK = 10;
/// blockmodel_delta (block_merge.hpp:37, block_merge.cpp:13-37): the Delta of merging
/// current_block into proposed_block, from row and column current_block on the device.
Delta blockmodel_delta(int current_block, int proposed_block, const Blockmodel& blockmodel);
Blockmodel& merge_blocks(Blockmodel& blockmodel, const Graph& graph, int num_edges);
}  // namespace block_merge

namespace sbp {
/// sbp::Intermediate (sbp.hpp:18-28) plus the number of blocks and the wall time.
struct Intermediate {
    float iteration;
    double mdl;
    double normalized_mdl_v1;
    double normalized_mdl_v2;
    double modularity;  // -1 (computed by the caller for the final row)
    double interblock_edges;
    double block_size_variation;
    int num_blocks;
    int mcmc_iterations;
    double mcmc_time;
    double seconds;  // wall time since stochastic_block_partition began
};
std::vector<Intermediate> get_intermediates();
Blockmodel stochastic_block_partition(Graph& graph, Args& args);

/// Many independent graphs partitioned side by side on one GPU (batch.cpp). The graphs are given
/// as one union: graph g is the vertex range [vertex_offsets[g], vertex_offsets[g + 1]) and no edge
/// joins two graphs. Each runs stochastic_block_partition with its own Args (async_gibbs with
/// approximate merges) and ends as it would alone; num_edges[g] < 0 means its stored edges.
/// eval_variant: the evaluation kernel variant (option 3 of sbpx_set_option; 0 = tuned);
/// count_bytes: count the evaluation kernels' algorithmic bytes (option 1); graph_replay: rounds
/// replayed from a captured HIP graph once the variant timing settled (option 9).
struct BatchRun {
    struct Graph {
        std::vector<int> assignment;  // graph-local block ids
        int num_blocks = 0;
        double mdl = 0.0;             // as the driver caches it (float)
        int mcmc_iterations = 0;
        std::vector<Intermediate> intermediates;
    };
    std::vector<Graph> graphs;
    double seconds = 0.0;  // wall time of the drivers (graphs resident)
    long rounds = 0;       // rounds of batched device steps
    long long sweeps = 0;  // batched sweeps
    sbpx_sweep_stats totals{};
    long long replayed = 0, captures = 0;  // sweep rounds replayed from a captured graph; captures
};
BatchRun partition_batch(const std::vector<int32_t>& vertex_offsets, const std::vector<int32_t>& out_ptr,
                         const std::vector<int32_t>& out_idx, const std::vector<int32_t>& in_ptr,
                         const std::vector<int32_t>& in_idx, const std::vector<long>& num_edges,
                         const std::vector<Args>& graph_args, int device = 0, int eval_variant = 0,
                         bool count_bytes = false, bool graph_replay = false);
bool done_blockmodeling(Blockmodel& blockmodel, BlockmodelTriplet& blockmodel_triplet, int min_num_blocks = 0);
/// Whether the last stochastic_block_partition on this thread stopped at Args::max_seconds.
bool stopped_at_time_limit();
}  // namespace sbp

// The driver's globals are per thread, so that the graphs of a batch run their drivers side by side.
extern thread_local Args args;
