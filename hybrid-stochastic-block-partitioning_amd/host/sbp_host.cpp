// The reference's driver, finetune and block-merge control flow on top of the sbpx C ABI.
// Control flow follows src/sbp.cpp:61-112, src/finetune.cpp:31-89, 342-444, 144-162,
// src/block_merge.cpp:278-311 and src/blockmodel/blockmodel_triplet.cpp:3-100; each sweep,
// merge phase and MDL runs on the device.
//
// Attribution: the golden-section search (BlockmodelTriplet), finetune::early_stop, the null
// models, Blockmodel::block_size_variation and the driver loop restate, step for step, logic of
// vtsynergy/Hybrid-Stochastic-Block-Partitioning (GNU LGPL 2.1), because the search over the
// number of blocks must take the same steps as the reference's; that logic remains under the
// reference's licence.
#include "sbp_host.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <numeric>
#include <stdexcept>

#include "device.hpp"

thread_local Args args;

// The time limit of the partition in progress (Args::max_seconds), measured from its start.
static thread_local std::chrono::steady_clock::time_point limit_start;
static thread_local bool limit_reached = false;
// With SBP_TRACE set: a line on stderr at most every 20 s from inside the sweeps of a phase.
static thread_local std::chrono::steady_clock::time_point last_trace;
static void trace_progress(int iteration, int num_blocks, double entropy) {
    static const bool on = std::getenv("SBP_TRACE") != nullptr;
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - last_trace).count() < 20.0) return;
    last_trace = now;
    std::fprintf(stderr, "sbp sweeps: B=%d sweep %d of this phase, mdl=%.3f, %.1f s into the partition\n", num_blocks,
                 iteration + 1, entropy, std::chrono::duration<double>(now - limit_start).count());
}

static bool past_time_limit() {
    if (args.max_seconds <= 0.0) return false;
    if (std::chrono::duration<double>(std::chrono::steady_clock::now() - limit_start).count() < args.max_seconds)
        return false;
    limit_reached = true;
    return true;
}

namespace sbpx_host {

static void check(sbpx_ctx* ctx, sbpx_status st, const char* what) {
    if (st != SBPX_OK) throw std::runtime_error(std::string(what) + ": " + sbpx_last_error(ctx));
}

// Requests of a batch member go to its coordinator (batch.cpp).
void batch_build(Batch& b, int slot, int num_blocks, const int32_t* assignment);
void batch_download(Batch& b, int slot, int32_t* assignment);
double batch_log_posterior(Batch& b, int slot);
void batch_sweep(Batch& b, int slot, int algo, int batches, uint64_t seed, uint32_t call, sbpx_sweep_stats* stats);
int batch_merge(Batch& b, int slot, int num_to_merge, int mode, uint64_t seed, uint32_t call);
long batch_trace(Batch& b, int slot);

sbpx_ctx* Device::own_ctx(const char* what) const {
    if (!ctx) throw std::runtime_error(std::string(what) + ": not available for a graph of a batch");
    return ctx;
}

void Device::build(int num_blocks, const int32_t* assignment) {
    if (batch) return batch_build(*batch, slot, num_blocks, assignment);
    check(ctx, sbpx_blockmodel_build(ctx, num_blocks, assignment), "sbpx_blockmodel_build");
}

void Device::download(int32_t* assignment) {
    if (batch) return batch_download(*batch, slot, assignment);
    check(ctx, sbpx_blockmodel_download(ctx, assignment, nullptr, nullptr, nullptr), "download");
}

double Device::log_posterior() {
    if (batch) return batch_log_posterior(*batch, slot);
    double lp = 0.0;
    check(ctx, sbpx_log_posterior(ctx, &lp), "sbpx_log_posterior");
    return lp;
}

void Device::sweep(int algo, int batches, uint64_t seed, uint32_t call, sbpx_sweep_stats* stats) {
    if (batch) return batch_sweep(*batch, slot, algo, batches, seed, call, stats);
    check(ctx, sbpx_sweep(ctx, algo, batches, seed, call, stats), "sbpx_sweep");
}

long Device::trace() {
    if (batch) return batch_trace(*batch, slot);
    int64_t t = 0;
    check(ctx, sbpx_blockmodel_trace(ctx, &t, nullptr), "sbpx_blockmodel_trace");
    return (long)t;
}

int Device::merge(int num_to_merge, int mode, uint64_t seed, uint32_t call) {
    if (batch) return batch_merge(*batch, slot, num_to_merge, mode, seed, call);
    check(ctx, sbpx_merge_blocks(ctx, num_to_merge, mode, seed, call), "sbpx_merge_blocks");
    int32_t B = 0;
    check(ctx, sbpx_blockmodel_info(ctx, &B, nullptr), "sbpx_blockmodel_info");
    return B;
}

}  // namespace sbpx_host

using sbpx_host::check;
using sbpx_host::Device;

// ---------------------------------------------------------------------------- Graph
Graph::Graph(int num_vertices, std::vector<int32_t> out_ptr, std::vector<int32_t> out_idx, std::vector<int32_t> in_ptr,
             std::vector<int32_t> in_idx, int device, long num_edges) {
    num_vertices_ = num_vertices;
    num_edges_ = num_edges < 0 ? (int)out_idx.size() : (int)num_edges;
    dev_ = std::make_shared<Device>();
    check(nullptr, sbpx_create(&dev_->ctx, device), "sbpx_create");
    dev_->N = num_vertices_;
    dev_->E = num_edges_;
    dev_->stored_edges = (long)out_idx.size();
    check(dev_->ctx,
          sbpx_graph_upload(dev_->ctx, num_vertices, (int64_t)out_idx.size(), out_ptr.data(), out_idx.data(),
                            in_ptr.data(), in_idx.data()),
          "sbpx_graph_upload");
    int nh = 0;
    check(dev_->ctx, sbpx_graph_split(dev_->ctx, nullptr, &nh, nullptr), "sbpx_graph_split");
    high_.resize(nh);
    low_.resize(num_vertices - nh);
    check(dev_->ctx, sbpx_graph_split(dev_->ctx, high_.data(), &nh, low_.data()), "sbpx_graph_split");
    out_ptr_ = std::make_shared<const std::vector<int32_t>>(std::move(out_ptr));
    out_idx_ = std::make_shared<const std::vector<int32_t>>(std::move(out_idx));
    in_ptr_ = std::make_shared<const std::vector<int32_t>>(std::move(in_ptr));
    in_idx_ = std::make_shared<const std::vector<int32_t>>(std::move(in_idx));
}

static std::shared_ptr<NeighborList> neighbor_lists(const std::vector<int32_t>& ptr, const std::vector<int32_t>& idx,
                                                    int n) {
    auto lists = std::make_shared<NeighborList>(n);
    for (int v = 0; v < n; ++v) (*lists)[v].assign(idx.begin() + ptr[v], idx.begin() + ptr[v + 1]);
    return lists;
}

const NeighborList& Graph::out_neighbors() const {
    if (!out_ptr_) throw std::runtime_error("Graph::out_neighbors: no host adjacency for a graph of a batch");
    if (!out_lists_) out_lists_ = neighbor_lists(*out_ptr_, *out_idx_, num_vertices_);
    return *out_lists_;
}

const NeighborList& Graph::in_neighbors() const {
    if (!in_ptr_) throw std::runtime_error("Graph::in_neighbors: no host adjacency for a graph of a batch");
    if (!in_lists_) in_lists_ = neighbor_lists(*in_ptr_, *in_idx_, num_vertices_);
    return *in_lists_;
}

// ---------------------------------------------------------------------------- Delta
std::map<int, int>& Delta::route(int row, int col, int& key) {
    if (row == current_) { key = col; return current_row_; }
    if (row == proposed_) { key = col; return proposed_row_; }
    if (col == current_) { key = row; return current_col_; }
    if (col == proposed_) { key = row; return proposed_col_; }
    throw std::logic_error("Neither the row nor column are current_block or proposed_block.");
}

int Delta::get(int row, int col) const {
    const std::map<int, int>* m;
    int key;
    if (row == current_) { m = &current_row_; key = col; }
    else if (row == proposed_) { m = &proposed_row_; key = col; }
    else if (col == current_) { m = &current_col_; key = row; }
    else if (col == proposed_) { m = &proposed_col_; key = row; }
    else throw std::logic_error("Neither the row nor column are current_block or proposed_block.");
    auto it = m->find(key);
    return it == m->end() ? 0 : it->second;
}

std::vector<std::tuple<int, int, int>> Delta::entries() const {
    std::vector<std::tuple<int, int, int>> out;
    for (const auto& e : current_row_) out.emplace_back(current_, e.first, e.second);
    for (const auto& e : proposed_row_) out.emplace_back(proposed_, e.first, e.second);
    for (const auto& e : current_col_) out.emplace_back(e.first, current_, e.second);
    for (const auto& e : proposed_col_) out.emplace_back(e.first, proposed_, e.second);
    return out;
}

// ---------------------------------------------------------------------------- BlockMatrix
int BlockMatrix::get(int row, int col) const {
    owner_.make_live();
    sbpx_ctx* c = owner_.ctx();
    int32_t r = row, k = col, v = 0;
    check(c, sbpx_blockmodel_cells(c, 1, &r, &k, &v), "blockmatrix()->get");
    return v;
}

static std::map<int, int> matrix_line(const Blockmodel& owner, int x, int by_column) {
    owner.make_live();
    sbpx_ctx* c = owner.ctx();
    int32_t n = 0;
    check(c, sbpx_blockmodel_line(c, x, by_column, 0, nullptr, nullptr, &n), "blockmatrix() line");
    std::vector<int32_t> idx(n), val(n);
    if (n > 0) check(c, sbpx_blockmodel_line(c, x, by_column, n, idx.data(), val.data(), &n), "blockmatrix() line");
    std::map<int, int> out;
    for (int i = 0; i < n; ++i) out[idx[i]] = val[i];
    return out;
}

std::map<int, int> BlockMatrix::getrow_sparse(int row) const { return matrix_line(owner_, row, 0); }
std::map<int, int> BlockMatrix::getcol_sparse(int col) const { return matrix_line(owner_, col, 1); }

// ---------------------------------------------------------------------------- Blockmodel
Blockmodel::Blockmodel(int num_blocks_, const Graph& graph, float rate) : empty(false) {
    std::vector<int> identity(graph.num_vertices());
    std::iota(identity.begin(), identity.end(), 0);
    *this = Blockmodel(num_blocks_, graph, rate, identity);
}

// Blockmodel(num_blocks, graph, rate, assignment) (blockmodel.hpp:63-73).
Blockmodel::Blockmodel(int num_blocks_, const Graph& graph, float rate, std::vector<int>& block_assignment)
    : empty(false) {
    dev_ = graph.device();
    id_ = dev_->next_id++;
    num_blocks = num_blocks_;
    block_reduction_rate = rate;
    num_blocks_to_merge = (int)(num_blocks * block_reduction_rate);
    dev_->build(num_blocks, block_assignment.data());
    dev_->live_id = id_;
}

sbpx_ctx* Blockmodel::ctx() const { return dev_->own_ctx("Blockmodel::ctx"); }
void Blockmodel::mark_live() const { dev_->live_id = id_; }

void Blockmodel::make_live() const {
    if (empty || !dev_) throw std::runtime_error("empty blockmodel");
    if (dev_->live_id == id_) return;
    if ((int)snapshot_.size() != dev_->N) throw std::runtime_error("blockmodel state was overwritten");
    dev_->build(num_blocks, snapshot_.data());
    dev_->live_id = id_;
}

// Blockmodel::copy (blockmodel.cpp:95-105): a deep copy with num_blocks_to_merge = 0.
Blockmodel Blockmodel::detached(int num_blocks, float overall_entropy, std::shared_ptr<sbpx_host::Device> dev) {
    Blockmodel b;
    b.empty = false;
    b.dev_ = std::move(dev);
    b.id_ = b.dev_->next_id++;
    b.num_blocks = num_blocks;
    b.num_blocks_to_merge = int(num_blocks * BLOCK_REDUCTION_RATE);
    b.overall_entropy = overall_entropy;
    return b;
}

Blockmodel Blockmodel::copy() {
    Blockmodel c;
    c.empty = empty;
    if (empty) return c;
    c.dev_ = dev_;
    c.id_ = dev_->next_id++;
    c.num_blocks = num_blocks;
    c.num_blocks_to_merge = 0;
    c.block_reduction_rate = block_reduction_rate;
    c.overall_entropy = overall_entropy;
    c.snapshot_ = block_assignment();
    return c;
}

std::vector<int> Blockmodel::block_assignment() const {
    if (dev_->live_id != id_) return snapshot_;
    std::vector<int> a(dev_->N);
    dev_->download(a.data());
    return a;
}
int Blockmodel::block_assignment(int vertex) const {
    if (vertex < 0 || vertex >= dev_->N) throw std::out_of_range("vertex out of range");
    return block_assignment()[vertex];
}

void Blockmodel::move_vertex(int vertex, const Delta& delta, utils::ProposalAndEdgeCounts& proposal) {
    make_live();
    sbpx_ctx* c = ctx();
    if (delta.current_block() != block_assignment(vertex) || delta.proposed_block() != proposal.proposal)
        throw std::invalid_argument("move_vertex: the Delta is not this vertex's move to the proposal");
    int32_t v = vertex, s = proposal.proposal;
    check(c, sbpx_apply_moves(c, 1, &v, &s), "move_vertex");
    snapshot_.clear();
    mark_live();
}

std::vector<int> Blockmodel::degrees() const {
    make_live();
    std::vector<int> d(num_blocks);
    sbpx_ctx* c = dev_->own_ctx("Blockmodel::degrees");
    check(c, sbpx_blockmodel_download(c, nullptr, nullptr, nullptr, d.data()), "download");
    return d;
}
std::vector<int> Blockmodel::degrees_in() const {
    make_live();
    std::vector<int> d(num_blocks);
    sbpx_ctx* c = dev_->own_ctx("Blockmodel::degrees_in");
    check(c, sbpx_blockmodel_download(c, nullptr, nullptr, d.data(), nullptr), "download");
    return d;
}
std::vector<int> Blockmodel::degrees_out() const {
    make_live();
    std::vector<int> d(num_blocks);
    sbpx_ctx* c = dev_->own_ctx("Blockmodel::degrees_out");
    check(c, sbpx_blockmodel_download(c, nullptr, d.data(), nullptr, nullptr), "download");
    return d;
}

// Blockmodel::block_size_variation (blockmodel.cpp:9-27): variance of the block sizes over
// (max - min)(max - mean) / 4, with the reference's initial min/max and float rounding.
double Blockmodel::block_size_variation() const {
    std::vector<int> block_sizes(num_blocks, 0);
    for (int block : block_assignment()) block_sizes[block]++;
    double total = 0;
    for (int x : block_sizes) total += x;
    double mean = total / double(num_blocks);
    double min = std::numeric_limits<double>::max(), max = std::numeric_limits<double>::min(), variance = 0;
    for (int block_size : block_sizes) {
        if (block_size < min) min = block_size;
        if (block_size > max) max = block_size;
        variance += double(block_size - mean) * double(block_size - mean);
    }
    variance /= double(num_blocks);
    double max_variance = (double(max - min) * double(max - mean)) / 4.0;
    return float(variance / max_variance);
}

// Blockmodel::interblock_edges (blockmodel.cpp:257-261): the share of edges between blocks.
double Blockmodel::interblock_edges() const {
    make_live();
    double num_edges = (double)dev_->stored_edges;
    double interblock = num_edges - double(dev_->trace());
    return interblock / num_edges;
}

// ---------------------------------------------------------------------------- triplet
bool BlockmodelTriplet::golden_ratio_not_reached() { return get(2).empty; }

bool BlockmodelTriplet::is_done() {
    if ((!get(0).empty && get(0).getNum_blocks() - get(2).getNum_blocks() == 2) ||
        (get(0).empty && get(1).getNum_blocks() - get(2).getNum_blocks() == 1)) {
        optimal_num_blocks_found = true;
    }
    return optimal_num_blocks_found;
}

int BlockmodelTriplet::lower_difference() { return get(1).getNum_blocks() - get(2).getNum_blocks(); }
int BlockmodelTriplet::upper_difference() { return get(0).getNum_blocks() - get(1).getNum_blocks(); }

void BlockmodelTriplet::update(Blockmodel& blockmodel) {
    int index;
    if (blockmodels[1].empty) {
        index = 1;
    } else {
        if (blockmodel.getOverall_entropy() <= blockmodels[1].getOverall_entropy()) {
            int old_index = (get(1).getNum_blocks() > blockmodel.getNum_blocks()) ? 0 : 2;
            blockmodels[old_index] = blockmodels[1].copy();
            index = 1;
        } else {
            index = (blockmodels[1].getNum_blocks() > blockmodel.getNum_blocks()) ? 2 : 0;
        }
    }
    blockmodels[index] = blockmodel.copy();
}

Blockmodel BlockmodelTriplet::get_next_blockmodel(Blockmodel& old_blockmodel) {
    old_blockmodel.setNum_blocks_to_merge(0);
    update(old_blockmodel);
    if (golden_ratio_not_reached()) {
        Blockmodel blockmodel = get(1).copy();
        blockmodel.setNum_blocks_to_merge(int(blockmodel.getNum_blocks() * BLOCK_REDUCTION_RATE));
        if (blockmodel.getNum_blocks_to_merge() == 0) optimal_num_blocks_found = true;
        return blockmodel;
    }
    if (is_done()) return get(1).copy();
    int index = 1;
    if (get(0).empty && get(1).getNum_blocks() > get(2).getNum_blocks()) index = 1;
    else if (upper_difference() >= lower_difference()) index = 0;
    else index = 1;
    int next_num_blocks_to_try = get(index + 1).getNum_blocks();
    next_num_blocks_to_try += int((get(index).getNum_blocks() - get(index + 1).getNum_blocks()) * 0.618);
    Blockmodel blockmodel = get(index).copy();
    blockmodel.setNum_blocks_to_merge(blockmodel.getNum_blocks() - next_num_blocks_to_try);
    return blockmodel;
}

// ---------------------------------------------------------------------------- checkpoints
namespace {
template <class T>
void put(std::FILE* f, const T& v) {
    if (std::fwrite(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("checkpoint write failed");
}
template <class T>
T take(std::FILE* f) {
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("checkpoint truncated");
    return v;
}
const uint64_t CHECKPOINT_MAGIC = 0x3150434B50425353ull;  // "SSBPKCP1"
}  // namespace

void Blockmodel::write(std::FILE* f) const {
    put<int32_t>(f, empty ? 1 : 0);
    if (empty) return;
    put<int32_t>(f, num_blocks);
    put<int32_t>(f, num_blocks_to_merge);
    put<float>(f, block_reduction_rate);
    put<float>(f, overall_entropy);
    const std::vector<int> a = block_assignment();
    put<int64_t>(f, (int64_t)a.size());
    if (!a.empty() && std::fwrite(a.data(), sizeof(int), a.size(), f) != a.size())
        throw std::runtime_error("checkpoint write failed");
}

Blockmodel Blockmodel::read(std::FILE* f, const Graph& graph) {
    Blockmodel b;
    if (take<int32_t>(f)) return b;
    b.empty = false;
    b.dev_ = graph.device();
    b.id_ = b.dev_->next_id++;
    b.num_blocks = take<int32_t>(f);
    b.num_blocks_to_merge = take<int32_t>(f);
    b.block_reduction_rate = take<float>(f);
    b.overall_entropy = take<float>(f);
    const int64_t n = take<int64_t>(f);
    if (n != graph.num_vertices()) throw std::runtime_error("checkpoint is for another graph");
    b.snapshot_.resize(n);
    if (n > 0 && std::fread(b.snapshot_.data(), sizeof(int), n, f) != (size_t)n)
        throw std::runtime_error("checkpoint truncated");
    return b;
}

void BlockmodelTriplet::write(std::FILE* f) const {
    put<int32_t>(f, optimal_num_blocks_found ? 1 : 0);
    for (const Blockmodel& b : blockmodels) b.write(f);
}

void BlockmodelTriplet::read(std::FILE* f, const Graph& graph) {
    optimal_num_blocks_found = take<int32_t>(f) != 0;
    for (Blockmodel& b : blockmodels) b = Blockmodel::read(f, graph);
}

// ---------------------------------------------------------------------------- entropy
namespace entropy {
// entropy::mdl (entropy.cpp:584-589): the model term for (num_vertices, num_edges, B) on the host,
// minus the device's log posterior.
double mdl(const Blockmodel& blockmodel, int num_vertices, int num_edges) {
    blockmodel.make_live();
    const double log_posterior = blockmodel.device().log_posterior();
    const double B = blockmodel.getNum_blocks();
    const double x = pow(B, 2) / num_edges;
    const double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (num_edges * h) + (num_vertices * log(B)) - log_posterior;
}
double null_mdl_v1(long num_edges) {
    double log_posterior_p = num_edges * log(1.0 / num_edges);
    double x = 1.0 / num_edges;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (num_edges * h) - log_posterior_p;
}
double null_mdl_v2(long num_vertices, long num_edges) {
    double log_posterior_p = num_edges * log(1.0 / num_edges);
    double x = pow(num_vertices, 2) / num_edges;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (num_edges * h) + (num_vertices * log(num_vertices)) - log_posterior_p;
}
double normalize_mdl_v1(double mdl, long num_edges) { return mdl / null_mdl_v1(num_edges); }
double normalize_mdl_v2(double mdl, long num_vertices, long num_edges) {
    return mdl / null_mdl_v2(num_vertices, num_edges);
}
}  // namespace entropy

// ---------------------------------------------------------------------------- finetune
namespace finetune {
thread_local int MCMC_iterations = 0;
thread_local double MCMC_time = 0.0;
thread_local Totals totals;

bool early_stop(int iteration, BlockmodelTriplet& blockmodels, double initial_entropy,
                std::vector<double>& delta_entropies) {
    size_t last_index = delta_entropies.size() - 1;
    if (delta_entropies[last_index] == 0.0) return true;
    if (iteration < 3) return false;
    double average = delta_entropies[last_index] + delta_entropies[last_index - 1] + delta_entropies[last_index - 2];
    average /= -3.0;
    double threshold = blockmodels.get(2).empty ? 5e-4 * initial_entropy : 1e-4 * initial_entropy;
    return average < threshold;
}

static void add_totals(const sbpx_sweep_stats& st) {
    totals.sweeps++;
    totals.attempts += st.attempts;
    totals.evals += st.evals;
    totals.moves += st.moves;
    for (int b = 0; b < 4; ++b) {
        totals.eval_ms[b] += st.eval_ms[b];
        totals.eval_bytes[b] += st.eval_bytes[b];
        totals.eval_launches[b] += st.eval_launches[b];
    }
}

// asynchronous_gibbs (finetune.cpp:31-89) and hybrid_mcmc (finetune.cpp:342-410): the entropy
// change is measured from the phase's start (last_iteration_entropy is never advanced).
static Blockmodel& batched_mcmc(int algo, Blockmodel& blockmodel, const Graph& graph, BlockmodelTriplet& blockmodels) {
    if (blockmodel.getNum_blocks() == 1) return blockmodel;
    blockmodel.make_live();
    Device& dev = *graph.device();
    std::vector<double> delta_entropies;
    blockmodel.setOverall_entropy(entropy::mdl(blockmodel, graph.num_vertices(), graph.num_edges()));
    double initial_entropy = blockmodel.getOverall_entropy();
    double last_iteration_entropy = initial_entropy;
    for (int iteration = 0; iteration < MAX_NUM_ITERATIONS; ++iteration) {
        sbpx_sweep_stats st{};
        dev.sweep(algo, args.batches, args.seed, dev.sweep_call++, &st);
        add_totals(st);
        double entropy = entropy::mdl(blockmodel, graph.num_vertices(), graph.num_edges());
        delta_entropies.push_back(entropy - last_iteration_entropy);
        MCMC_iterations++;
        trace_progress(iteration, blockmodel.getNum_blocks(), entropy);
        if (early_stop(iteration, blockmodels, initial_entropy, delta_entropies)) break;
        if (past_time_limit()) break;
    }
    blockmodel.setOverall_entropy(entropy::mdl(blockmodel, graph.num_vertices(), graph.num_edges()));
    return blockmodel;
}

Blockmodel& asynchronous_gibbs(Blockmodel& blockmodel, const Graph& graph, BlockmodelTriplet& blockmodels) {
    return batched_mcmc(SBPX_ALGO_ASYNC_GIBBS, blockmodel, graph, blockmodels);
}

Blockmodel& hybrid_mcmc(Blockmodel& blockmodel, const Graph& graph, BlockmodelTriplet& blockmodels) {
    return batched_mcmc(SBPX_ALGO_HYBRID, blockmodel, graph, blockmodels);
}

// metropolis_hastings (finetune.cpp:412-444): per-iteration sums of accepted delta MDL, measured
// against the float entropy cached at the start.
Blockmodel& metropolis_hastings(Blockmodel& blockmodel, const Graph& graph, BlockmodelTriplet& blockmodels) {
    if (blockmodel.getNum_blocks() == 1) return blockmodel;
    blockmodel.make_live();
    Device& dev = *graph.device();
    std::vector<double> delta_entropies;
    blockmodel.setOverall_entropy(entropy::mdl(blockmodel, graph.num_vertices(), graph.num_edges()));
    for (int iteration = 0; iteration < MAX_NUM_ITERATIONS; ++iteration) {
        sbpx_sweep_stats st{};
        dev.sweep(SBPX_ALGO_METROPOLIS_HASTINGS, args.batches, args.seed, dev.sweep_call++, &st);
        add_totals(st);
        delta_entropies.push_back(st.dS_accepted);
        MCMC_iterations++;
        if (early_stop(iteration, blockmodels, blockmodel.getOverall_entropy(), delta_entropies)) break;
        if (past_time_limit()) break;
    }
    blockmodel.setOverall_entropy(entropy::mdl(blockmodel, graph.num_vertices(), graph.num_edges()));
    return blockmodel;
}

EdgeWeights edge_weights(const NeighborList& neighbors, int vertex, bool ignore_self) {
    EdgeWeights w;
    for (int neighbor : neighbors[vertex]) {
        if (ignore_self && neighbor == vertex) continue;
        w.indices.push_back(neighbor);
        w.values.push_back(1);
    }
    return w;
}

// The out loop adds each edge to the proposed block's row and takes it from the current block's
// row; the in loop does the same for the columns; a self edge goes to (proposed, proposed).
Delta blockmodel_delta(int vertex, int current_block, int proposed_block, const EdgeWeights& out_edges,
                       const EdgeWeights& in_edges, const Blockmodel& blockmodel) {
    const std::vector<int> assignment = blockmodel.block_assignment();
    Delta delta(current_block, proposed_block);
    for (size_t i = 0; i < out_edges.indices.size(); ++i) {
        const int w = out_edges.indices[i], b = assignment.at(w), weight = out_edges.values[i];
        if (w == vertex) {
            delta.add(proposed_block, proposed_block, weight);
            delta.self_edge_weight(1);
        } else {
            delta.add(proposed_block, b, weight);
        }
        delta.sub(current_block, b, weight);
    }
    for (size_t i = 0; i < in_edges.indices.size(); ++i) {
        const int w = in_edges.indices[i], b = assignment.at(w), weight = in_edges.values[i];
        if (w == vertex) {
            delta.add(proposed_block, proposed_block, weight);
            delta.self_edge_weight(1);
        } else {
            delta.add(b, proposed_block, weight);
        }
        delta.sub(b, current_block, weight);
    }
    return delta;
}
}  // namespace finetune

namespace entropy {
double delta_mdl(const Blockmodel& blockmodel, const Delta& delta, const utils::ProposalAndEdgeCounts& proposal) {
    blockmodel.make_live();
    sbpx_ctx* c = blockmodel.ctx();
    std::vector<int32_t> rows, cols, vals;
    for (const auto& e : delta.entries()) {
        rows.push_back(std::get<0>(e));
        cols.push_back(std::get<1>(e));
        vals.push_back(std::get<2>(e));
    }
    double dS = 0.0;
    sbpx_host::check(c,
                     sbpx_delta_mdl(c, delta.current_block(), delta.proposed_block(), (int32_t)rows.size(), rows.data(),
                                    cols.data(), vals.data(), proposal.num_out_neighbor_edges,
                                    proposal.num_in_neighbor_edges, delta.self_edge_weight(), &dS),
                     "entropy::delta_mdl");
    return dS;
}

double hastings_correction(int vertex, const Graph& graph, const Blockmodel& blockmodel, const Delta& delta,
                           int current_block, const utils::ProposalAndEdgeCounts& proposal) {
    (void)graph;
    blockmodel.make_live();
    sbpx_ctx* c = blockmodel.ctx();
    if (current_block != blockmodel.block_assignment(vertex) || delta.proposed_block() != proposal.proposal)
        throw std::invalid_argument("hastings_correction: the Delta is not this vertex's move to the proposal");
    int32_t v = vertex, s = proposal.proposal;
    double u = 0.5, dS = 0.0, H = 1.0;
    uint8_t acc = 0;
    sbpx_host::check(c, sbpx_eval_moves(c, 1, &v, &s, &u, &dS, &H, &acc), "entropy::hastings_correction");
    return H;
}

double block_merge_delta_mdl(int current_block, utils::ProposalAndEdgeCounts proposal, const Blockmodel& blockmodel,
                             const Delta& delta) {
    blockmodel.make_live();
    sbpx_ctx* c = blockmodel.ctx();
    if (delta.current_block() != current_block || delta.proposed_block() != proposal.proposal)
        throw std::invalid_argument("block_merge_delta_mdl: the Delta is not this merge's");
    int32_t r = current_block, s = proposal.proposal;
    double dS = 0.0;
    sbpx_host::check(c, sbpx_merge_eval(c, 1, &r, &s, &dS), "entropy::block_merge_delta_mdl");
    return dS;
}
}  // namespace entropy

// ---------------------------------------------------------------------------- block merge
namespace block_merge {
Delta blockmodel_delta(int current_block, int proposed_block, const Blockmodel& blockmodel) {
    Delta delta(current_block, proposed_block);
    const std::shared_ptr<BlockMatrix> m = blockmodel.blockmatrix();
    for (const auto& e : m->getrow_sparse(current_block)) {
        const int col = e.first, value = e.second;
        if (col == current_block || col == proposed_block) delta.add(proposed_block, proposed_block, value);
        else delta.add(proposed_block, col, value);
        delta.sub(current_block, col, value);
    }
    for (const auto& e : m->getcol_sparse(current_block)) {
        const int row = e.first, value = e.second;
        if (row == current_block) continue;
        if (row == proposed_block) delta.add(proposed_block, proposed_block, value);
        else delta.add(row, proposed_block, value);
        delta.sub(row, current_block, value);
    }
    return delta;
}

Blockmodel& merge_blocks(Blockmodel& blockmodel, const Graph& graph, int num_edges) {
    (void)num_edges;
    blockmodel.make_live();
    Device& dev = *graph.device();
    int mode = args.approximate ? SBPX_MERGE_APPROXIMATE : SBPX_MERGE_ADVANCED;
    const int B = dev.merge(blockmodel.getNum_blocks_to_merge(), mode, args.seed, dev.merge_call++);
    blockmodel.setNum_blocks(B);
    blockmodel.mark_live();
    return blockmodel;
}
}  // namespace block_merge

// ---------------------------------------------------------------------------- driver
namespace sbp {
static thread_local std::vector<Intermediate> intermediate_results;
static thread_local std::chrono::steady_clock::time_point partition_start;
std::vector<Intermediate> get_intermediates() { return intermediate_results; }

// sbp::add_intermediate (sbp.cpp:33-59). Modularity stays -1 here: the reference computes it only
// for the final row, with an O(N^2) loop over the graph, which the caller reproduces in closed
// form from the final assignment.
static void add_intermediate(float iteration, const Graph& graph, const Blockmodel& blockmodel, double mdl) {
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - partition_start).count();
    Intermediate r{};
    r.iteration = iteration;
    r.mdl = mdl;
    r.normalized_mdl_v1 = entropy::normalize_mdl_v1(mdl, graph.num_edges());
    r.normalized_mdl_v2 = entropy::normalize_mdl_v2(mdl, graph.num_vertices(), graph.num_edges());
    r.modularity = -1;
    r.interblock_edges = blockmodel.interblock_edges();
    r.block_size_variation = blockmodel.block_size_variation();
    r.num_blocks = blockmodel.getNum_blocks();
    r.mcmc_iterations = finetune::MCMC_iterations;
    r.mcmc_time = finetune::MCMC_time;
    r.seconds = t;
    intermediate_results.push_back(r);
}

bool done_blockmodeling(Blockmodel& blockmodel, BlockmodelTriplet& blockmodel_triplet, int min_num_blocks) {
    if (min_num_blocks > 0) {
        if ((blockmodel.getNum_blocks() <= min_num_blocks) || !blockmodel_triplet.get(2).empty) return true;
    }
    return blockmodel_triplet.optimal_num_blocks_found;
}

// sbp::stochastic_block_partition (sbp.cpp:61-98).
bool stopped_at_time_limit() { return limit_reached; }

Blockmodel stochastic_block_partition(Graph& graph, Args& a) {
    args = a;
    partition_start = std::chrono::steady_clock::now();
    limit_start = partition_start;
    limit_reached = false;
    intermediate_results.clear();
    finetune::MCMC_iterations = 0;
    finetune::MCMC_time = 0.0;
    finetune::totals = finetune::Totals();
    Device& dev = *graph.device();
    dev.sweep_call = 0;
    dev.merge_call = 0;
    Blockmodel blockmodel;
    BlockmodelTriplet blockmodel_triplet = BlockmodelTriplet();
    float iteration = 0;
    std::FILE* cf = (args.resume && !args.checkpoint.empty()) ? std::fopen(args.checkpoint.c_str(), "rb") : nullptr;
    if (cf) {  // continue a checkpointed run where it stopped
        try {
            if (take<uint64_t>(cf) != CHECKPOINT_MAGIC) throw std::runtime_error("not a checkpoint file");
            if (take<int64_t>(cf) != graph.num_vertices() || take<int64_t>(cf) != graph.num_edges())
                throw std::runtime_error("checkpoint is for another graph");
            iteration = take<float>(cf);
            dev.sweep_call = take<uint32_t>(cf);
            dev.merge_call = take<uint32_t>(cf);
            finetune::MCMC_iterations = take<int32_t>(cf);
            finetune::MCMC_time = take<double>(cf);
            blockmodel_triplet.read(cf, graph);
            blockmodel = Blockmodel::read(cf, graph);
        } catch (...) {
            std::fclose(cf);
            throw;
        }
        std::fclose(cf);
    } else {
        blockmodel = Blockmodel(graph.num_vertices(), graph, float(BLOCK_REDUCTION_RATE));
        double initial_mdl = entropy::mdl(blockmodel, graph.num_vertices(), graph.num_edges());
        add_intermediate(0, graph, blockmodel, initial_mdl);
    }
    int iterations_this_call = 0;
    while (!done_blockmodeling(blockmodel, blockmodel_triplet)) {
        if (args.max_iterations > 0 && iterations_this_call++ >= args.max_iterations) break;
        if (past_time_limit()) break;
        blockmodel = block_merge::merge_blocks(blockmodel, graph, graph.num_edges());
        auto start = std::chrono::steady_clock::now();
        if (args.algorithm == "async_gibbs" && iteration < float(args.asynciterations))
            blockmodel = finetune::asynchronous_gibbs(blockmodel, graph, blockmodel_triplet);
        else if (args.algorithm == "hybrid_mcmc")
            blockmodel = finetune::hybrid_mcmc(blockmodel, graph, blockmodel_triplet);
        else
            blockmodel = finetune::metropolis_hastings(blockmodel, graph, blockmodel_triplet);
        finetune::MCMC_time +=
            std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        add_intermediate(++iteration, graph, blockmodel, blockmodel.getOverall_entropy());
        if (std::getenv("SBP_TRACE"))
            std::fprintf(stderr, "sbp iter %.0f: B=%d mdl=%.3f mcmc_iterations=%d\n", iteration,
                         blockmodel.getNum_blocks(), (double)blockmodel.getOverall_entropy(), finetune::MCMC_iterations);
        if (limit_reached) break;  // keep the state the sweeps reached
        blockmodel = blockmodel_triplet.get_next_blockmodel(blockmodel);
        if (!args.checkpoint.empty()) {  // written to a temporary file, then renamed over the old one
            const std::string tmp = args.checkpoint + ".tmp";
            std::FILE* f = std::fopen(tmp.c_str(), "wb");
            if (!f) throw std::runtime_error("cannot write " + tmp);
            try {
                put<uint64_t>(f, CHECKPOINT_MAGIC);
                put<int64_t>(f, graph.num_vertices());
                put<int64_t>(f, graph.num_edges());
                put<float>(f, iteration);
                put<uint32_t>(f, dev.sweep_call);
                put<uint32_t>(f, dev.merge_call);
                put<int32_t>(f, finetune::MCMC_iterations);
                put<double>(f, finetune::MCMC_time);
                blockmodel_triplet.write(f);
                blockmodel.write(f);
            } catch (...) {
                std::fclose(f);
                throw;
            }
            if (std::fclose(f) != 0 || std::rename(tmp.c_str(), args.checkpoint.c_str()) != 0)
                throw std::runtime_error("cannot write " + args.checkpoint);
        }
    }
    add_intermediate(-1, graph, blockmodel, blockmodel.getOverall_entropy());
    return blockmodel;
}
}  // namespace sbp
