// ORACLE — test infrastructure only.
//
// A from-scratch CPU restatement of the stochastic-block-partitioning hot path of
// vtsynergy/Hybrid-Stochastic-Block-Partitioning (the reference). It exists to check the
// MI355X library and to be timed as the CPU baseline. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it. The product library never links it.
//
// What is restated (reference file:line):
//   Blockmodel build          src/blockmodel/blockmodel.cpp:187-255
//   log posterior / MDL       src/blockmodel/blockmodel.cpp:267-285, src/entropy.cpp:584-589
//   null models               src/entropy.cpp:559-582
//   vertex-move Delta         src/finetune.cpp:108-142, include/blockmodel/sparse/delta.hpp:22-149
//   vertex-move delta MDL     src/entropy.cpp:296-401 (incl. the net-zero double count)
//   Hastings correction       src/entropy.cpp:495-557
//   accept                    src/finetune.cpp:19-29
//   move_vertex               src/blockmodel/blockmodel.cpp:344-360 (incl. total-degree drift)
//   proposal                  src/common.cpp:131-192 (randomness injected, see below)
//   async Gibbs / hybrid / MH src/finetune.cpp:31-89, 342-444, 488-545
//   early stop                src/finetune.cpp:144-162
//   merge Delta / delta MDL   src/block_merge.cpp:13-37, src/entropy.cpp:127-180
//   merge_blocks              src/block_merge.cpp:278-368
//   merge selection           src/blockmodel/blockmodel.cpp:29-78, src/block_merge.cpp:123-217
//   triplet / driver          src/blockmodel/blockmodel_triplet.cpp:3-100, src/sbp.cpp:61-112
//   learned mixture Hastings  no reference counterpart: the plain terms of src/entropy.cpp:495-557
//                             mixed with a softmax over given scores
//   degree split              src/graph.cpp:15-22, 215-228
//
// Randomness: the reference draws from unseeded std engines (include/common.hpp:25-26), so its
// runs cannot be replayed. This restatement draws every random number from Philox4x32-10
// (Salmon et al., SC'11) keyed by (seed) and indexed by (vertex or block, call id, tag), with
// the same distributions the reference samples: a uniform neighbour slot, then a block drawn
// in proportion to row t + column t of the block matrix, walked in ascending index order.
// The MI355X library uses the identical counter layout, so whole sweeps can be compared.
#pragma once
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------- Philox4x32-10
struct Draw { uint32_t x[4]; };
Draw philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3);
inline uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline double unit53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
enum : uint32_t { TAG_ASYNC = 0x4D430000u, TAG_MH = 0x4D480000u, TAG_MERGE = 0x4D520000u };

// ---------------------------------------------------------------- graph
struct Graph {
    int N = 0;
    long E = 0;                               // number of (deduplicated) edges incl. self-loops
    std::vector<std::vector<int>> out, in;    // adjacency in CSR/CSC order as given
    std::vector<char> self;                   // self-loop flag
    std::vector<int> high, low;               // degree split (graph.cpp:215-228)
};
Graph make_graph(int N, const int32_t* out_ptr, const int32_t* out_idx, const int32_t* in_ptr,
                 const int32_t* in_idx);

// ---------------------------------------------------------------- blockmodel
using Row = std::unordered_map<int, int>;
struct Blockmodel {
    bool empty = true;
    int B = 0;
    int num_to_merge = 0;
    float overall_entropy = 0.0f;             // float, as blockmodel.hpp:182
    std::vector<int> assign;
    std::vector<Row> rows, cols;              // M and its transpose (DictTransposeMatrix)
    std::vector<int> dout, din, deg;
    int get(int r, int c) const;
};
void build(Blockmodel& bm, const Graph& g);   // initialize_edge_counts
Blockmodel make_blockmodel(int B, const Graph& g, const std::vector<int>& assign, float rate = 0.5f);
double log_posterior(const Blockmodel& bm);
double mdl(const Blockmodel& bm, long N, long E);
double null_mdl_v1(long E);
double null_mdl_v2(long N, long E);

// ---------------------------------------------------------------- Delta
struct Delta {
    int r = -1, s = -1, self_w = 0;
    Row cur_row, prop_row, cur_col, prop_col;
    Delta(int r_, int s_) : r(r_), s(s_) {}
    Row& route(int row, int col, int& key);
    void add(int row, int col, int v);
    void sub(int row, int col, int v) { add(row, col, -v); }
    int get(int row, int col) const;
    struct Entry { int row, col, d; };
    std::vector<Entry> entries() const;
};
struct Proposal { int s, k_out, k_in, k; };

Delta vertex_delta(const Graph& g, const Blockmodel& bm, int v, int r, int s);
double delta_mdl(const Blockmodel& bm, const Delta& d, const Proposal& p);
double hastings(const Graph& g, const Blockmodel& bm, int v, const Delta& d, int r, const Proposal& p);
void hastings_terms(const Graph& g, const Blockmodel& bm, int v, const Delta& d, int r, const Proposal& p,
                    double* fwd, double* bwd);
// learned mixture proposal (this project's own; see cpu_ref.cpp)
void learned_features(const Graph& g, const Blockmodel& bm, int v, int cand[32], int cnt[32], float x[64]);
void plain_probabilities(const Graph& g, const Blockmodel& bm, int v, int s, double* fwd, double* bwd);
double mixture_hastings(const Graph& g, const Blockmodel& bm, int v, int s, const float z[32], const int cand[32],
                        double w, double* qf, double* qb);
bool accept(double dS, double H, double u);
void move_vertex(Blockmodel& bm, int v, const Delta& d, const Proposal& p);
Proposal vertex_counts(const Graph& g, int v, int s);

// sampling (canonical order, injected uniforms)
Proposal propose_vertex(const Graph& g, const Blockmodel& bm, int v, const Draw& dr);

struct MoveEval { int v, s; bool proposed, accepted; double dS, H; };
MoveEval eval_move(const Graph& g, const Blockmodel& bm, int v, int s, double u);

// ---------------------------------------------------------------- finetune
struct SweepStats { long attempts = 0, evals = 0, moves = 0; double dS_sum = 0.0; };
struct Triplet;
SweepStats async_sweep(const Graph& g, Blockmodel& bm, int batches, uint64_t seed, uint32_t call,
                       bool hybrid_low_only);
SweepStats mh_pass(const Graph& g, Blockmodel& bm, const std::vector<int>& order, uint64_t seed,
                   uint32_t call);
bool early_stop(int iteration, bool bracket, double initial, const std::vector<double>& deltas);
struct FinetuneResult { int sweeps = 0; SweepStats stats; };
FinetuneResult finetune(const Graph& g, Blockmodel& bm, int algo, int batches, bool bracket,
                        uint64_t seed, uint32_t& call);

// ---------------------------------------------------------------- merge
Delta merge_delta(const Blockmodel& bm, int r, int s);
double merge_delta_mdl(const Blockmodel& bm, int r, const Proposal& p, const Delta& d);
struct MergeEval { int s; double dS; };
MergeEval propose_merge(const Blockmodel& bm, int r, int i, std::unordered_map<int, bool>& past,
                        uint64_t seed, uint32_t call);
void merge_proposals(const Blockmodel& bm, uint64_t seed, uint32_t call, std::vector<double>& dS,
                     std::vector<int>& best);
void carry_out_best_merges(Blockmodel& bm, const std::vector<double>& dS, const std::vector<int>& best);
void carry_out_best_merges_advanced(Blockmodel& bm, const std::vector<double>& dS,
                                    const std::vector<int>& best);
void merge_blocks(const Graph& g, Blockmodel& bm, bool approximate, uint64_t seed, uint32_t call);

// ---------------------------------------------------------------- triplet + driver
struct Triplet {
    bool optimal = false;
    Blockmodel bms[3];
    void update(const Blockmodel& bm);
    bool golden_ratio_not_reached() const { return bms[2].empty; }
    bool is_done();
    Blockmodel next(Blockmodel& old);
};
struct RunResult {
    Blockmodel bm;
    int outer_iterations = 0;
    long sweeps = 0, attempts = 0, evals = 0, moves = 0;
    double seconds = 0.0, mcmc_seconds = 0.0, merge_seconds = 0.0;
    // per outer iteration: blocks after the merge phase and the float MDL after the sweeps
    std::vector<int> trace_B;
    std::vector<float> trace_mdl;
};
RunResult stochastic_block_partition(const Graph& g, int algo, int batches, bool approximate,
                                     uint64_t seed);

}  // namespace oracle
