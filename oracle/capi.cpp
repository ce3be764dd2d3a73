// ORACLE — test infrastructure only. Flat C entry points over cpu_ref.cpp for the Python tests
// and the CPU baseline leg of bench.py (loaded with ctypes).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

#include "cpu_ref.hpp"

using namespace oracle;

static thread_local std::string g_err;

#define GUARD_BEGIN try {
#define GUARD_END                 \
    }                             \
    catch (const std::exception& e) { \
        g_err = e.what();         \
        return -1;                \
    }                             \
    return 0;

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

void oracle_philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t* out) {
    Draw d = philox(seed, c0, c1, c2, c3);
    std::memcpy(out, d.x, 16);
}

float oracle_logf(float x) { return logf(x); }

// Count mismatches between a caller-supplied table of logf results and host logf for the
// 32-bit patterns [lo, hi).
long oracle_logf_check(uint32_t lo, uint32_t hi, const float* got) {
    long bad = 0;
    #pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long u = lo; u < (long)hi; ++u) {
        float x;
        uint32_t b = (uint32_t)u;
        std::memcpy(&x, &b, 4);
        float h = logf(x);
        uint32_t hb, gb;
        std::memcpy(&hb, &h, 4);
        std::memcpy(&gb, &got[u - lo], 4);
        if (hb != gb) bad++;
    }
    return bad;
}

void* oracle_graph_new(int N, const int32_t* out_ptr, const int32_t* out_idx, const int32_t* in_ptr,
                       const int32_t* in_idx) {
    try {
        return new Graph(make_graph(N, out_ptr, out_idx, in_ptr, in_idx));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void oracle_graph_free(void* g) { delete (Graph*)g; }
long oracle_graph_E(void* g) { return ((Graph*)g)->E; }
int oracle_graph_nhigh(void* g) { return (int)((Graph*)g)->high.size(); }
void oracle_graph_split(void* gp, int32_t* high, int32_t* low) {
    Graph* g = (Graph*)gp;
    for (size_t i = 0; i < g->high.size(); ++i) high[i] = g->high[i];
    for (size_t i = 0; i < g->low.size(); ++i) low[i] = g->low[i];
}

void* oracle_bm_new(void* gp, int B, const int32_t* assign) {
    try {
        Graph* g = (Graph*)gp;
        std::vector<int> a(assign, assign + g->N);
        return new Blockmodel(make_blockmodel(B, *g, a, 0.5f));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void* oracle_bm_clone(void* bm) { return new Blockmodel(*(Blockmodel*)bm); }
void oracle_bm_free(void* bm) { delete (Blockmodel*)bm; }
int oracle_bm_B(void* bm) { return ((Blockmodel*)bm)->B; }
int oracle_bm_num_to_merge(void* bm) { return ((Blockmodel*)bm)->num_to_merge; }
void oracle_bm_set_num_to_merge(void* bm, int n) { ((Blockmodel*)bm)->num_to_merge = n; }

void oracle_bm_state(void* bmp, int32_t* assign, int32_t* dout, int32_t* din, int32_t* deg) {
    Blockmodel* bm = (Blockmodel*)bmp;
    if (assign) std::memcpy(assign, bm->assign.data(), bm->assign.size() * 4);
    if (dout) std::memcpy(dout, bm->dout.data(), bm->B * 4);
    if (din) std::memcpy(din, bm->din.data(), bm->B * 4);
    if (deg) std::memcpy(deg, bm->deg.data(), bm->B * 4);
}

long oracle_bm_nnz(void* bmp) {
    Blockmodel* bm = (Blockmodel*)bmp;
    long n = 0;
    for (const auto& r : bm->rows) n += (long)r.size();
    return n;
}

// Dense B x B download (small B only).
void oracle_bm_dense(void* bmp, int32_t* M) {
    Blockmodel* bm = (Blockmodel*)bmp;
    const long B = bm->B;
    std::memset(M, 0, sizeof(int32_t) * B * B);
    for (long r = 0; r < B; ++r)
        for (const auto& e : bm->rows[r]) M[r * B + e.first] = e.second;
}

// Row-major sorted coordinate list of the nonzeros; also checks the transpose agrees.
int oracle_bm_coo(void* bmp, int32_t* rows, int32_t* cols, int32_t* vals) {
    GUARD_BEGIN
    Blockmodel* bm = (Blockmodel*)bmp;
    long k = 0;
    for (int r = 0; r < bm->B; ++r) {
        std::vector<std::pair<int, int>> row(bm->rows[r].begin(), bm->rows[r].end());
        std::sort(row.begin(), row.end());
        for (const auto& e : row) {
            auto it = bm->cols[e.first].find(r);
            if (it == bm->cols[e.first].end() || it->second != e.second)
                throw std::logic_error("transpose disagrees with matrix");
            rows[k] = r;
            cols[k] = e.first;
            vals[k] = e.second;
            ++k;
        }
    }
    GUARD_END
}

int oracle_mdl(void* bm, long N, long E, double* out) {
    GUARD_BEGIN
    *out = mdl(*(Blockmodel*)bm, N, E);
    GUARD_END
}
int oracle_log_posterior(void* bm, double* out) {
    GUARD_BEGIN
    *out = log_posterior(*(Blockmodel*)bm);
    GUARD_END
}
double oracle_null_mdl_v1(long E) { return null_mdl_v1(E); }
double oracle_null_mdl_v2(long N, long E) { return null_mdl_v2(N, E); }

// Vertex-move Delta for v -> s on the current state. Writes up to cap entries; returns count.
int oracle_vertex_delta(void* gp, void* bmp, int v, int s, int cap, int32_t* rows, int32_t* cols,
                        int32_t* ds, int32_t* self_w) {
    try {
        Graph* g = (Graph*)gp;
        Blockmodel* bm = (Blockmodel*)bmp;
        Delta d = vertex_delta(*g, *bm, v, bm->assign[v], s);
        auto en = d.entries();
        int n = 0;
        for (const auto& e : en) {
            if (n < cap) {
                rows[n] = e.row;
                cols[n] = e.col;
                ds[n] = e.d;
            }
            ++n;
        }
        if (self_w) *self_w = d.self_w;
        return n;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int oracle_merge_delta(void* bmp, int r, int s, int cap, int32_t* rows, int32_t* cols, int32_t* ds) {
    try {
        Blockmodel* bm = (Blockmodel*)bmp;
        Delta d = merge_delta(*bm, r, s);
        auto en = d.entries();
        int n = 0;
        for (const auto& e : en) {
            if (n < cap) {
                rows[n] = e.row;
                cols[n] = e.col;
                ds[n] = e.d;
            }
            ++n;
        }
        return n;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int oracle_merge_delta_mdl(void* bmp, int r, int s, double* out) {
    GUARD_BEGIN
    Blockmodel* bm = (Blockmodel*)bmp;
    int k_out = 0, k_in = 0;
    for (const auto& e : bm->rows[r]) k_out += e.second;
    for (const auto& e : bm->cols[r]) k_in += e.second;
    Delta d = merge_delta(*bm, r, s);
    *out = merge_delta_mdl(*bm, r, Proposal{s, k_out, k_in, k_out + k_in}, d);
    GUARD_END
}

// One merge proposal of block r (draw i of merge phase `call`), before the repeat check:
// the target block the sampler of propose_merge_sparse draws.
int oracle_propose_merge_draw(void* bmp, int r, int i, uint64_t seed, uint32_t call, int32_t* s) {
    GUARD_BEGIN
    std::unordered_map<int, bool> past;
    *s = propose_merge(*(Blockmodel*)bmp, r, i, past, seed, call).s;
    GUARD_END
}

// Learned proposal: candidates, counts and features of vertex v on the current state.
int oracle_learned_features(void* gp, void* bmp, int v, int32_t* cand, int32_t* cnt, float* x) {
    GUARD_BEGIN
    int c[32], k[32];
    learned_features(*(Graph*)gp, *(Blockmodel*)bmp, v, c, k, x);
    std::memcpy(cand, c, sizeof(c));
    std::memcpy(cnt, k, sizeof(k));
    GUARD_END
}

// Hastings ratio of the learned mixture (weight w on the plain proposal) for v -> s with the given
// scores z[32] of the candidates cand[32]; qf / qb receive the net's forward / reverse probability.
int oracle_mixture_hastings(void* gp, void* bmp, int v, int s, const float* z, const int32_t* cand, double w,
                            double* H, double* qf, double* qb) {
    GUARD_BEGIN
    int c[32];
    float zz[32];
    for (int j = 0; j < 32; ++j) {
        c[j] = cand[j];
        zz[j] = z[j];
    }
    *H = mixture_hastings(*(Graph*)gp, *(Blockmodel*)bmp, v, s, zz, c, w, qf, qb);
    GUARD_END
}

// The two sums of the plain Hastings correction for v -> s.
// P(s | current state) and P(r | state after the move) of the plain sampler for v -> s.
int oracle_plain_probabilities(void* gp, void* bmp, int v, int s, double* fwd, double* bwd) {
    GUARD_BEGIN
    plain_probabilities(*(Graph*)gp, *(Blockmodel*)bmp, v, s, fwd, bwd);
    GUARD_END
}

int oracle_hastings_terms(void* gp, void* bmp, int v, int s, double* fwd, double* bwd) {
    GUARD_BEGIN
    Graph* g = (Graph*)gp;
    Blockmodel* bm = (Blockmodel*)bmp;
    const int r = bm->assign[v];
    Delta d = vertex_delta(*g, *bm, v, r, s);
    hastings_terms(*g, *bm, v, d, r, vertex_counts(*g, v, s), fwd, bwd);
    GUARD_END
}

// Evaluate fixed (v, s, u) proposals against the current state without applying them.
int oracle_eval_moves(void* gp, void* bmp, int n, const int32_t* v, const int32_t* s, const double* u,
                      double* dS, double* H, uint8_t* acc) {
    GUARD_BEGIN
    Graph* g = (Graph*)gp;
    Blockmodel* bm = (Blockmodel*)bmp;
    for (int i = 0; i < n; ++i) {
        MoveEval ev = eval_move(*g, *bm, v[i], s[i], u[i]);
        dS[i] = ev.dS;
        H[i] = ev.H;
        acc[i] = ev.accepted ? 1 : 0;
    }
    GUARD_END
}

// Apply moves one after another with Blockmodel::move_vertex semantics (Delta on the current
// assignment), in the given order.
int oracle_apply_moves(void* gp, void* bmp, int n, const int32_t* v, const int32_t* s) {
    GUARD_BEGIN
    Graph* g = (Graph*)gp;
    Blockmodel* bm = (Blockmodel*)bmp;
    for (int i = 0; i < n; ++i) {
        int r = bm->assign[v[i]];
        if (r == s[i]) continue;
        Delta d = vertex_delta(*g, *bm, v[i], r, s[i]);
        move_vertex(*bm, v[i], d, vertex_counts(*g, v[i], s[i]));
    }
    GUARD_END
}

// The proposal each vertex would draw on the current state (MCMC sampler), tag 0 = async,
// 1 = Metropolis-Hastings.
int oracle_propose(void* gp, void* bmp, int n, const int32_t* v, uint64_t seed, uint32_t call,
                   int tag, int32_t* s, double* u) {
    GUARD_BEGIN
    Graph* g = (Graph*)gp;
    Blockmodel* bm = (Blockmodel*)bmp;
    for (int i = 0; i < n; ++i) {
        Draw dr = philox(seed, (uint32_t)v[i], call, 0u, tag ? TAG_MH : TAG_ASYNC);
        s[i] = propose_vertex(*g, *bm, v[i], dr).s;
        u[i] = unit53(dr.x[2], dr.x[3]);
    }
    GUARD_END
}

struct OracleStats {
    long attempts, evals, moves;
    double dS_sum;
};

int oracle_async_sweep(void* gp, void* bmp, int batches, uint64_t seed, uint32_t call, int hybrid_low,
                       OracleStats* st) {
    GUARD_BEGIN
    SweepStats s = async_sweep(*(Graph*)gp, *(Blockmodel*)bmp, batches, seed, call, hybrid_low != 0);
    if (st) *st = OracleStats{s.attempts, s.evals, s.moves, s.dS_sum};
    GUARD_END
}

// Metropolis-Hastings over `order` (n vertices), or over the graph's high-degree list when
// order is NULL.
int oracle_mh_pass(void* gp, void* bmp, int n, const int32_t* order, uint64_t seed, uint32_t call,
                   OracleStats* st) {
    GUARD_BEGIN
    Graph* g = (Graph*)gp;
    std::vector<int> ord = order ? std::vector<int>(order, order + n) : g->high;
    SweepStats s = mh_pass(*g, *(Blockmodel*)bmp, ord, seed, call);
    if (st) *st = OracleStats{s.attempts, s.evals, s.moves, s.dS_sum};
    GUARD_END
}

int oracle_finetune(void* gp, void* bmp, int algo, int batches, int bracket, uint64_t seed,
                    uint32_t* call, int* sweeps, OracleStats* st) {
    GUARD_BEGIN
    FinetuneResult f = finetune(*(Graph*)gp, *(Blockmodel*)bmp, algo, batches, bracket != 0, seed, *call);
    if (sweeps) *sweeps = f.sweeps;
    if (st) *st = OracleStats{f.stats.attempts, f.stats.evals, f.stats.moves, 0.0};
    GUARD_END
}

float oracle_bm_overall(void* bm) { return ((Blockmodel*)bm)->overall_entropy; }

int oracle_merge_proposals(void* bmp, uint64_t seed, uint32_t call, double* dS, int32_t* best) {
    GUARD_BEGIN
    std::vector<double> d;
    std::vector<int> b;
    merge_proposals(*(Blockmodel*)bmp, seed, call, d, b);
    std::memcpy(dS, d.data(), d.size() * 8);
    std::memcpy(best, b.data(), b.size() * 4);
    GUARD_END
}

int oracle_carry_out(void* gp, void* bmp, const double* dS, const int32_t* best, int advanced) {
    GUARD_BEGIN
    Blockmodel* bm = (Blockmodel*)bmp;
    std::vector<double> d(dS, dS + bm->B);
    std::vector<int> b(best, best + bm->B);
    if (advanced) carry_out_best_merges_advanced(*bm, d, b);
    else carry_out_best_merges(*bm, d, b);
    build(*bm, *(Graph*)gp);
    GUARD_END
}

int oracle_merge_blocks(void* gp, void* bmp, int approximate, uint64_t seed, uint32_t call) {
    GUARD_BEGIN
    merge_blocks(*(Graph*)gp, *(Blockmodel*)bmp, approximate != 0, seed, call);
    GUARD_END
}

struct OracleRun {
    int B, outer_iterations;
    long sweeps, attempts, evals, moves;
    double seconds, mcmc_seconds, merge_seconds, mdl;
};

// trace_B / trace_mdl (may be NULL) receive up to trace_cap entries of the per-outer-iteration
// trace (blocks after the merge phase, float MDL after the sweeps).
int oracle_sbp_traced(void* gp, int algo, int batches, int approximate, uint64_t seed, int32_t* assign,
                      OracleRun* out, int32_t* trace_B, float* trace_mdl, int trace_cap) {
    GUARD_BEGIN
    Graph* g = (Graph*)gp;
    RunResult r = stochastic_block_partition(*g, algo, batches, approximate != 0, seed);
    if (assign) std::memcpy(assign, r.bm.assign.data(), g->N * 4);
    for (int i = 0; i < (int)r.trace_B.size() && i < trace_cap; ++i) {
        if (trace_B) trace_B[i] = r.trace_B[i];
        if (trace_mdl) trace_mdl[i] = r.trace_mdl[i];
    }
    out->B = r.bm.B;
    out->outer_iterations = r.outer_iterations;
    out->sweeps = r.sweeps;
    out->attempts = r.attempts;
    out->evals = r.evals;
    out->moves = r.moves;
    out->seconds = r.seconds;
    out->mcmc_seconds = r.mcmc_seconds;
    out->merge_seconds = r.merge_seconds;
    out->mdl = r.bm.overall_entropy;
    GUARD_END
}

int oracle_sbp(void* gp, int algo, int batches, int approximate, uint64_t seed, int32_t* assign,
               OracleRun* out) {
    return oracle_sbp_traced(gp, algo, batches, approximate, seed, assign, out, nullptr, nullptr, 0);
}

// entropy::mdl (entropy.cpp:584-589) of a block matrix given as nonzero cells (any order) with its
// block degrees, B blocks, for a graph of N vertices and E edges: the log posterior summed in
// cell order with the reference's FP32 logf rule (blockmodel.cpp:267-285).
int oracle_mdl_cells(long nnz, const int32_t* rows, const int32_t* cols, const int32_t* vals,
                     const int32_t* dout, const int32_t* din, int B, long N, long E, double* out) {
    GUARD_BEGIN
    double lp = 0.0;
    #pragma omp parallel for reduction(+ : lp) schedule(static)
    for (long i = 0; i < nnz; ++i) {
        double value = vals[i];
        double d_out = dout[rows[i]], d_in = din[cols[i]];
        lp += value * (double)logf((float)(value / (d_out * d_in)));
    }
    double x = pow(B, 2) / E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    *out = (E * h) + (N * log(B)) - lp;
    GUARD_END
}

// finetune::early_stop (finetune.cpp:144-162) on a list of per-sweep MDL changes.
int oracle_early_stop(int iteration, int bracket, double initial, const double* deltas, int n) {
    if (n <= 0) return -1;
    std::vector<double> d(deltas, deltas + n);
    return early_stop(iteration, bracket != 0, initial, d) ? 1 : 0;
}

// The golden-section search of the driver (sbp.cpp:61-112, blockmodel_triplet.cpp:3-100) with
// the merge phase and the sweeps replaced by a table: the blockmodel starts at n blocks with
// num_to_merge = int(n * 0.5); each outer iteration ends at B - num_to_merge blocks with overall
// entropy (float)score[B] and hands it to Triplet::next. Writes the block count of every outer
// iteration and then the block count the search returns; returns how many were written (-1 when
// cap is too small).
int oracle_golden_trace(int n, const double* score, int* out, int cap) {
    Blockmodel bm;
    bm.empty = false;
    bm.B = n;
    bm.num_to_merge = int(n * 0.5f);
    Triplet t;
    int k = 0;
    while (!t.optimal) {
        bm.B -= bm.num_to_merge;
        bm.num_to_merge = 0;
        bm.overall_entropy = (float)score[bm.B];
        if (k >= cap) return -1;
        out[k++] = bm.B;
        bm = t.next(bm);
    }
    if (k >= cap) return -1;
    out[k++] = bm.B;
    return k;
}

}  // extern "C"
