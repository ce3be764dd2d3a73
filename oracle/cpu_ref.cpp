// ORACLE — test infrastructure only (see cpu_ref.hpp for scope and citations).
#include "cpu_ref.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <limits>
#include <map>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <tuple>

namespace oracle {

// ------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3").
Draw philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Draw{{c0, c1, c2, c3}};
}

// ------------------------------------------------------------------------------------------
// Graph (graph.hpp:15-113). Degree = |out| + |in| - self (graph.cpp:15-22); the top 15% by
// degree, ordered by std::sort with a '>' comparator, are the high-degree vertices
// (graph.cpp:215-228).
Graph make_graph(int N, const int32_t* out_ptr, const int32_t* out_idx, const int32_t* in_ptr,
                 const int32_t* in_idx) {
    Graph g;
    g.N = N;
    g.out.resize(N);
    g.in.resize(N);
    g.self.assign(N, 0);
    for (int v = 0; v < N; ++v) {
        for (int j = out_ptr[v]; j < out_ptr[v + 1]; ++j) {
            g.out[v].push_back(out_idx[j]);
            if (out_idx[j] == v) g.self[v] = 1;
        }
        for (int j = in_ptr[v]; j < in_ptr[v + 1]; ++j) g.in[v].push_back(in_idx[j]);
    }
    g.E = out_ptr[N];
    std::vector<int> degrees(N);
    for (int v = 0; v < N; ++v) degrees[v] = int(g.out[v].size() + g.in[v].size() - g.self[v]);
    std::vector<int> idx(N);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.data(), idx.data() + idx.size(),
              [&degrees](size_t i1, size_t i2) { return degrees[i1] > degrees[i2]; });
    for (int index = 0; index < N; ++index) {
        if (index < 0.15 * N) g.high.push_back(idx[index]);
        else g.low.push_back(idx[index]);
    }
    return g;
}

// ------------------------------------------------------------------------------------------
// Blockmodel (blockmodel.hpp:45-184; DictTransposeMatrix: rows + transposed rows).
int Blockmodel::get(int r, int c) const {
    auto it = rows[r].find(c);
    return it == rows[r].end() ? 0 : it->second;
}

// initialize_edge_counts (blockmodel.cpp:187-255): M[b(u), b(w)] += 1 per out-edge, the
// transpose from in-edges, dout/din per block, deg = out edges + in edges from other blocks.
void build(Blockmodel& bm, const Graph& g) {
    const int B = bm.B;
    bm.rows.assign(B, Row());
    bm.cols.assign(B, Row());
    bm.dout.assign(B, 0);
    bm.din.assign(B, 0);
    bm.deg.assign(B, 0);
    for (int v = 0; v < g.N; ++v) {
        int b = bm.assign[v];
        if (b < 0 || b >= B) throw std::out_of_range("block id out of range");
        for (int w : g.out[v]) {
            int nb = bm.assign[w];
            bm.rows[b][nb] += 1;
            bm.dout[b] += 1;
            bm.deg[b] += 1;
        }
        for (int w : g.in[v]) {
            int nb = bm.assign[w];
            bm.cols[b][nb] += 1;
            bm.din[b] += 1;
            if (b != nb) bm.deg[b] += 1;
        }
    }
}

Blockmodel make_blockmodel(int B, const Graph& g, const std::vector<int>& assign, float rate) {
    Blockmodel bm;
    bm.empty = false;
    bm.B = B;
    bm.overall_entropy = std::numeric_limits<float>::max();
    bm.num_to_merge = (int)(B * rate);
    bm.assign = assign;
    build(bm, g);
    return bm;
}

// log_posterior_probability (blockmodel.cpp:267-285) with utils::nat_log = logf
// (utils.hpp:145-151): sum over nonzeros of M * (double)logf((float)(M / (dout * din))).
double log_posterior(const Blockmodel& bm) {
    double sum = 0.0;
    for (int r = 0; r < bm.B; ++r) {
        for (const auto& e : bm.rows[r]) {
            double value = e.second;
            double dout = bm.dout[r], din = bm.din[e.first];
            if (din == 0.0 || dout == 0.0) throw std::runtime_error("zero block degree with a nonzero cell");
            double ratio = value / (dout * din);
            double lg = logf((float)ratio);
            sum += value * lg;
        }
    }
    return sum;
}

// entropy::mdl (entropy.cpp:584-589).
double mdl(const Blockmodel& bm, long N, long E) {
    double log_posterior_p = log_posterior(bm);
    double x = pow(bm.B, 2) / E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (E * h) + (N * log(bm.B)) - log_posterior_p;
}

// entropy.cpp:567-582.
double null_mdl_v1(long E) {
    double log_posterior_p = E * log(1.0 / E);
    double x = 1.0 / E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (E * h) - log_posterior_p;
}
double null_mdl_v2(long N, long E) {
    double log_posterior_p = E * log(1.0 / E);
    double x = pow(N, 2) / E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (E * h) + (N * log(N)) - log_posterior_p;
}

// ------------------------------------------------------------------------------------------
// Delta (delta.hpp:22-149): four change maps, routed row-first; keys are kept even when the
// net change returns to zero.
Row& Delta::route(int row, int col, int& key) {
    if (row == r) { key = col; return cur_row; }
    if (row == s) { key = col; return prop_row; }
    if (col == r) { key = row; return cur_col; }
    if (col == s) { key = row; return prop_col; }
    throw std::logic_error("Delta cell in neither row/col r nor s");
}
void Delta::add(int row, int col, int v) {
    int key;
    Row& m = route(row, col, key);
    m[key] += v;
}
int Delta::get(int row, int col) const {
    const Row* m;
    int key;
    if (row == r) { m = &cur_row; key = col; }
    else if (row == s) { m = &prop_row; key = col; }
    else if (col == r) { m = &cur_col; key = row; }
    else if (col == s) { m = &prop_col; key = row; }
    else throw std::logic_error("Delta cell in neither row/col r nor s");
    auto it = m->find(key);
    return it == m->end() ? 0 : it->second;
}
std::vector<Delta::Entry> Delta::entries() const {
    std::vector<Entry> out;
    for (const auto& e : cur_row) out.push_back({r, e.first, e.second});
    for (const auto& e : prop_row) out.push_back({s, e.first, e.second});
    for (const auto& e : cur_col) out.push_back({e.first, r, e.second});
    for (const auto& e : prop_col) out.push_back({e.first, s, e.second});
    return out;
}

// k_out counts the out list including a self-loop; k_in the in list without it
// (finetune.cpp:535-536 with edge_weights :275-290).
Proposal vertex_counts(const Graph& g, int v, int s) {
    int k_out = (int)g.out[v].size();
    int k_in = (int)g.in[v].size() - g.self[v];
    return Proposal{s, k_out, k_in, k_out + k_in};
}

// finetune::blockmodel_delta (finetune.cpp:108-142), in-list without the self-loop.
Delta vertex_delta(const Graph& g, const Blockmodel& bm, int v, int r, int s) {
    Delta d(r, s);
    for (int w : g.out[v]) {
        int b = bm.assign[w];
        if (w == v) {
            d.add(s, s, 1);
            d.self_w = 1;
        } else {
            d.add(s, b, 1);
        }
        d.sub(r, b, 1);
    }
    for (int w : g.in[v]) {
        if (w == v) continue;
        int b = bm.assign[w];
        d.add(b, s, 1);
        d.sub(b, r, 1);
    }
    return d;
}

// common::cell_entropy (common.hpp:35-44): FP32 throughout.
static inline float cell(float value, float degree_in, float degree_out) {
    if (value == 0.0) return 0.0;
    return value * logf(value / (degree_in * degree_out));
}

// entropy::delta_mdl(bm, Delta, proposal) (entropy.cpp:296-401). Keys with a net-zero change
// are counted in the entries loop and again in the row/column loops, which skip only cells
// whose delta reads non-zero.
double delta_mdl(const Blockmodel& bm, const Delta& d, const Proposal& p) {
    const int r = d.r, s = d.s;
    auto deg_in = [&](int index) -> float {
        int value = bm.din[index];
        if (index == r) value -= (p.k_in + d.self_w);
        else if (index == s) value += (p.k_in + d.self_w);
        return float(value);
    };
    auto deg_out = [&](int index) -> float {
        int value = bm.dout[index];
        if (index == r) value -= p.k_out;
        else if (index == s) value += p.k_out;
        return float(value);
    };
    double dS = 0.0;
    for (const auto& e : d.entries()) {
        int m = bm.get(e.row, e.col);
        dS += cell((float)m, (float)bm.din[e.col], (float)bm.dout[e.row]);
        dS -= cell(float(m + e.d), deg_in(e.col), deg_out(e.row));
    }
    for (const auto& e : bm.rows[r]) {
        if (d.get(r, e.first) != 0) continue;
        dS += cell((float)e.second, (float)bm.din[e.first], (float)bm.dout[r]);
        dS -= cell((float)e.second, deg_in(e.first), deg_out(r));
    }
    for (const auto& e : bm.rows[s]) {
        if (d.get(s, e.first) != 0) continue;
        dS += cell((float)e.second, (float)bm.din[e.first], (float)bm.dout[s]);
        dS -= cell((float)e.second, deg_in(e.first), deg_out(s));
    }
    for (const auto& e : bm.cols[r]) {
        int row = e.first;
        if (d.get(row, r) != 0 || row == r || row == s) continue;
        dS += cell((float)e.second, (float)bm.din[r], (float)bm.dout[row]);
        dS -= cell((float)e.second, deg_in(r), deg_out(row));
    }
    for (const auto& e : bm.cols[s]) {
        int row = e.first;
        if (d.get(row, s) != 0 || row == r || row == s) continue;
        dS += cell((float)e.second, (float)bm.din[s], (float)bm.dout[row]);
        dS -= cell((float)e.second, deg_in(s), deg_out(row));
    }
    if (std::isnan(dS) || std::isinf(dS)) throw std::invalid_argument("nan/inf in delta MDL");
    return dS;
}

// entropy::hastings_correction(vertex, ...) (entropy.cpp:495-557): dense row/column of s,
// neighbour-block counts from out-neighbours (with a self-loop) and in-neighbours (without).
double hastings(const Graph& g, const Blockmodel& bm, int v, const Delta& d, int r, const Proposal& p) {
    if (p.k == 0) return 1.0;
    double p_fwd, p_bwd;
    hastings_terms(g, bm, v, d, r, p, &p_fwd, &p_bwd);
    return p_bwd / p_fwd;
}

// The two sums of hastings_correction: p_fwd over the current matrix, p_bwd over the matrix
// after the move.
void hastings_terms(const Graph& g, const Blockmodel& bm, int v, const Delta& d, int r, const Proposal& p,
                    double* fwd, double* bwd) {
    Row block_counts;
    for (int w : g.out[v]) block_counts[bm.assign[w]] += 1;
    for (int w : g.in[v]) {
        if (w == v) continue;
        block_counts[bm.assign[w]] += 1;
    }
    std::vector<int> prow(bm.B, 0), pcol(bm.B, 0);
    for (const auto& e : bm.rows[p.s]) prow[e.first] = e.second;
    for (const auto& e : bm.cols[p.s]) pcol[e.first] = e.second;
    const int B = bm.B;
    double p_fwd = 0.0, p_bwd = 0.0;
    for (const auto& e : block_counts) {
        int b = e.first;
        double count = e.second;
        double proposal_weight = prow[b] + pcol[b] + 1.0;
        double block_degree = bm.deg[b] + B;
        double block_weight = bm.get(r, b) + d.get(r, b) + bm.get(b, r) + d.get(b, r) + 1.0;
        int new_degree = bm.deg[b];
        if (b == r) {
            int self_edges = bm.get(r, r) + d.get(r, r);
            new_degree = (bm.dout[r] - p.k_out) + (bm.din[r] - p.k_in) - self_edges;
        } else if (b == p.s) {
            int self_edges = bm.get(p.s, p.s) + d.get(p.s, p.s);
            new_degree = (bm.dout[p.s] + p.k_out) + (bm.din[p.s] + p.k_in) - self_edges;
        }
        double proposal_degree = new_degree + B;
        p_fwd += count * proposal_weight / block_degree;
        p_bwd += count * block_weight / proposal_degree;
    }
    *fwd = p_fwd;
    *bwd = p_bwd;
}

// ------------------------------------------------------------------------------------------
// The learned mixture proposal (no counterpart in the reference). Candidates: the 32 blocks v's
// neighbours fall in most often (self-loop excluded; ties: smaller block id). Features: for
// candidate i < 16 its share of v's edges and log1p of its in/out degree without v over
// log1p(E); then log1p(k_out), log1p(k_in) over log(4097), the share of v's edges into
// candidates 16-31, distinct neighbour blocks / k, log(B) / log(N), 1, zeros. The scores z come
// from the caller (the network runs elsewhere).
void learned_features(const Graph& g, const Blockmodel& bm, int v, int cand[32], int cnt[32], float x[64]) {
    std::map<int, int> counts;
    for (int w : g.out[v])
        if (w != v) counts[bm.assign[w]] += 1;
    for (int w : g.in[v])
        if (w != v) counts[bm.assign[w]] += 1;
    std::vector<std::pair<int, int>> order(counts.begin(), counts.end());  // (block, count)
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.second > b.second; });
    for (int i = 0; i < 32; ++i) {
        cand[i] = i < (int)order.size() ? order[i].first : -1;
        cnt[i] = i < (int)order.size() ? order[i].second : 0;
    }
    const int own = bm.assign[v];
    const int k_out = (int)g.out[v].size(), self = g.self[v];
    const int k_in = (int)g.in[v].size() - self;
    const float k = (float)(k_out + k_in);
    const float invk = k > 0 ? 1.0f / k : 0.0f;
    const float inv_logE = 1.0f / log1pf((float)g.E);
    for (int i = 0; i < 64; ++i) x[i] = 0.0f;
    for (int i = 0; i < 16; ++i) {
        const int b = cand[i];
        if (b < 0) continue;
        x[i] = (float)cnt[i] * invk;
        x[16 + i] = log1pf((float)(bm.din[b] - (b == own ? k_in + self : 0))) * inv_logE;
        x[32 + i] = log1pf((float)(bm.dout[b] - (b == own ? k_out : 0))) * inv_logE;
    }
    const float inv_logD = 1.0f / logf(4097.0f);
    x[48] = log1pf((float)k_out) * inv_logD;
    x[49] = log1pf((float)k_in) * inv_logD;
    int rest = 0;
    for (int i = 16; i < 32; ++i) rest += cand[i] >= 0 ? cnt[i] : 0;
    x[50] = (float)rest * invk;
    x[51] = (float)counts.size() * invk;
    x[52] = g.N > 1 ? (float)(log((double)bm.B) / log((double)g.N)) : 0.0f;
    x[53] = 1.0f;
}

// Softmax of z over the candidates other than `own`, at `pick` (-1: no candidate qualifies).
static double net_prob(const float z[32], const int cand[32], int own, int pick) {
    double mx = -1e300;
    int any = 0;
    for (int j = 0; j < 32; ++j)
        if (cand[j] >= 0 && cand[j] != own) {
            mx = std::max(mx, (double)z[j]);
            any = 1;
        }
    if (!any) return -1.0;
    double den = 0.0, num = 0.0;
    for (int j = 0; j < 32; ++j)
        if (cand[j] >= 0 && cand[j] != own) {
            const double e = exp((double)z[j] - mx);
            den += e;
            if (cand[j] == pick) num = e;
        }
    return num / den;
}

// The plain sampler's law (common.cpp:131-192 as this implementation draws it: the random-block
// branch never taken, Q3): a uniform neighbour slot of v (out-neighbours with the self-loop,
// in-neighbours without it), t its block, then s in proportion to row t + column t. fwd =
// P(s | current state), bwd = P(r | the state after v moves to s), the self-loop's slot then in s.
void plain_probabilities(const Graph& g, const Blockmodel& bm, int v, int s, double* fwd, double* bwd) {
    const int r = bm.assign[v];
    const Delta d = vertex_delta(g, bm, v, r, s);
    const int k_out = (int)g.out[v].size(), k_in_all = (int)g.in[v].size();
    const int self = g.self[v];
    const int k = k_out + k_in_all - self;
    *fwd = 0.0;
    *bwd = 0.0;
    if (k == 0) return;
    auto slots = [&](auto&& f) {
        for (int w : g.out[v]) f(w);
        for (int w : g.in[v])
            if (w != v) f(w);
    };
    auto dout_after = [&](int t) { return bm.dout[t] - (t == r ? k_out : 0) + (t == s ? k_out : 0); };
    auto din_after = [&](int t) { return bm.din[t] - (t == r ? k_in_all : 0) + (t == s ? k_in_all : 0); };
    double f = 0.0, b = 0.0;
    slots([&](int w) {
        const int t = bm.assign[w];  // the self-loop's slot: r now
        const double tf = (double)bm.dout[t] + bm.din[t];
        if (tf > 0) f += (bm.get(t, s) + bm.get(s, t)) / tf;
        const int t2 = w == v ? s : t;  // after the move
        const double tb = (double)dout_after(t2) + din_after(t2);
        if (tb > 0) b += (bm.get(t2, r) + d.get(t2, r) + bm.get(r, t2) + d.get(r, t2)) / tb;
    });
    *fwd = f / k;
    *bwd = b / k;
}

// Hastings ratio of the mixture w q_plain + (1 - w) q_net for the move v: r -> s, with q_plain the
// plain sampler's law (plain_probabilities) and q_net the softmax of the scores over the
// candidates other than the vertex's block in each direction; without learned mass (no candidate
// outside the block) the plain part alone.
double mixture_hastings(const Graph& g, const Blockmodel& bm, int v, int s, const float z[32], const int cand[32],
                        double w, double* qf, double* qb) {
    const int r = bm.assign[v];
    Proposal p = vertex_counts(g, v, s);
    *qf = net_prob(z, cand, r, s);
    *qb = net_prob(z, cand, s, r);
    if (p.k == 0) return 1.0;
    double plain_f, plain_b;
    plain_probabilities(g, bm, v, s, &plain_f, &plain_b);
    const double fwd = w * plain_f + (1.0 - w) * (*qf >= 0.0 ? *qf : plain_f);
    const double bwd = w * plain_b + (1.0 - w) * (*qb >= 0.0 ? *qb : plain_b);
    return bwd / fwd;
}

// finetune::accept (finetune.cpp:19-29), with the uniform injected.
bool accept(double dS, double H, double u) {
    double a = exp(-3.0 * dS) * H;
    a = (a >= 1.0) ? 1.0 : a;
    return u <= a;
}

// Blockmodel::move_vertex(v, Delta, proposal) (blockmodel.cpp:344-360) with
// DictTransposeMatrix::update_edge_counts(Delta) (dict_transpose_matrix.cpp:398-414). The
// total degrees of r and s are recomputed from the updated diagonal plus the Delta diagonal.
static void apply_delta(Blockmodel& bm, const Delta& d) {
    for (const auto& e : d.entries()) {
        int& m = bm.rows[e.row][e.col];
        m += e.d;
        bm.cols[e.col][e.row] += e.d;
        if (m == 0) {
            bm.rows[e.row].erase(e.col);
            bm.cols[e.col].erase(e.row);
        }
    }
}
void move_vertex(Blockmodel& bm, int v, const Delta& d, const Proposal& p) {
    bm.assign[v] = p.s;
    apply_delta(bm, d);
    const int r = d.r;
    int r_self = bm.get(r, r) + d.get(r, r);
    int s_self = bm.get(p.s, p.s) + d.get(p.s, p.s);
    bm.dout[r] -= p.k_out;
    bm.dout[p.s] += p.k_out;
    bm.din[r] -= (p.k_in + d.self_w);
    bm.din[p.s] += (p.k_in + d.self_w);
    bm.deg[r] = bm.dout[r] + bm.din[r] - r_self;
    bm.deg[p.s] = bm.dout[p.s] + bm.din[p.s] - s_self;
}

// ------------------------------------------------------------------------------------------
// Proposal (common.cpp:131-192). The random-block branch at common.cpp:154 compares an int in
// [0, RAND_MAX] with a value in (0, 1] and so never fires; it is left out. A vertex with no
// neighbours proposes a uniform block other than its own.
static int random_block(int r, int B, uint32_t x) {
    int proposed = (int)mulhi32(x, (uint32_t)(B - 1));
    if (proposed >= r) proposed++;
    return proposed;
}

struct Cell { int idx, val; };
static std::vector<Cell> sorted_cells(const Row& m) {
    std::vector<Cell> out;
    out.reserve(m.size());
    for (const auto& e : m) out.push_back({e.first, e.second});
    std::sort(out.begin(), out.end(), [](const Cell& a, const Cell& b) { return a.idx < b.idx; });
    return out;
}

// Draw a block in proportion to row t + column t (ascending index, row first), optionally
// leaving out index `skip` in both.
static int walk_row_col(const Blockmodel& bm, int t, uint32_t x, int skip, long* total_out) {
    std::vector<Cell> row = sorted_cells(bm.rows[t]);
    std::vector<Cell> col = sorted_cells(bm.cols[t]);
    long total = 0;
    for (const Cell& c : row) if (c.idx != skip) total += c.val;
    for (const Cell& c : col) if (c.idx != skip) total += c.val;
    *total_out = total;
    if (total == 0) return -1;
    uint32_t target = mulhi32(x, (uint32_t)total);
    long cum = 0;
    for (const Cell& c : row) {
        if (c.idx == skip) continue;
        cum += c.val;
        if ((long)target < cum) return c.idx;
    }
    for (const Cell& c : col) {
        if (c.idx == skip) continue;
        cum += c.val;
        if ((long)target < cum) return c.idx;
    }
    throw std::logic_error("proposal walk ran past the end");
}

Proposal propose_vertex(const Graph& g, const Blockmodel& bm, int v, const Draw& dr) {
    const int r = bm.assign[v];
    Proposal p = vertex_counts(g, v, r);
    if (p.k == 0) {
        p.s = random_block(r, bm.B, dr.x[0]);
        return p;
    }
    int slot = (int)mulhi32(dr.x[0], (uint32_t)p.k);
    int w;
    if (slot < p.k_out) {
        w = g.out[v][slot];
    } else {
        int j = slot - p.k_out;
        w = -1;
        for (int x : g.in[v]) {
            if (x == v) continue;
            if (j == 0) { w = x; break; }
            --j;
        }
    }
    int t = bm.assign[w];
    long total;
    int s = walk_row_col(bm, t, dr.x[1], -1, &total);
    if (total != (long)bm.dout[t] + bm.din[t]) throw std::logic_error("row+col sum != dout+din");
    p.s = s;
    return p;
}

// eval_vertex_move_v2 (finetune.cpp:306-318) for a given proposal s and uniform u.
MoveEval eval_move(const Graph& g, const Blockmodel& bm, int v, int s, double u) {
    const int r = bm.assign[v];
    MoveEval ev{v, s, false, false, 0.0, 1.0};
    if (s == r) return ev;
    ev.proposed = true;
    Proposal p = vertex_counts(g, v, s);
    Delta d = vertex_delta(g, bm, v, r, s);
    ev.H = hastings(g, bm, v, d, r, p);
    ev.dS = delta_mdl(bm, d, p);
    ev.accepted = accept(ev.dS, ev.H, u);
    return ev;
}

// ------------------------------------------------------------------------------------------
// Asynchronous Gibbs batch (finetune.cpp:41-72, hybrid :364-393): proposals on a stale
// snapshot in parallel, then the accepted moves applied in ascending vertex id with the Delta
// recomputed on the current assignment.
SweepStats async_sweep(const Graph& g, Blockmodel& bm, int batches, uint64_t seed, uint32_t call,
                       bool hybrid_low_only) {
    SweepStats st;
    std::vector<int> all;
    const std::vector<int>* list = &g.low;
    if (!hybrid_low_only) {
        all.resize(g.N);
        std::iota(all.begin(), all.end(), 0);
        list = &all;
    }
    const int n = (int)list->size();
    if (n == 0) return st;
    double num_batches = batches;
    int batch_size = int(ceil(n / num_batches));
    std::vector<MoveEval> moves(g.N);
    for (int batch = 0; batch < n / batch_size; ++batch) {
        int start = batch * batch_size;
        int end = std::min(n, (batch + 1) * batch_size);
        for (auto& m : moves) m.accepted = false;
        long attempts = 0, evals = 0;
        #pragma omp parallel for schedule(dynamic, 64) reduction(+ : attempts, evals)
        for (int index = start; index < end; ++index) {
            int v = (*list)[index];
            Draw dr = philox(seed, (uint32_t)v, call, 0u, TAG_ASYNC);
            Proposal p = propose_vertex(g, bm, v, dr);
            attempts++;
            MoveEval ev{v, p.s, false, false, 0.0, 1.0};
            if (p.s != bm.assign[v]) {
                evals++;
                ev = eval_move(g, bm, v, p.s, unit53(dr.x[2], dr.x[3]));
            }
            moves[v] = ev;
        }
        st.attempts += attempts;
        st.evals += evals;
        for (int v = 0; v < g.N; ++v) {
            const MoveEval& m = moves[v];
            if (!m.accepted) continue;
            int r = bm.assign[v];
            Delta d = vertex_delta(g, bm, v, r, m.s);
            Proposal p = vertex_counts(g, v, m.s);
            move_vertex(bm, v, d, p);
            st.moves++;
            st.dS_sum += m.dS;
        }
    }
    return st;
}

// Sequential Metropolis-Hastings with immediate moves (finetune.cpp:446-461, 488-501).
SweepStats mh_pass(const Graph& g, Blockmodel& bm, const std::vector<int>& order, uint64_t seed,
                   uint32_t call) {
    SweepStats st;
    for (int v : order) {
        Draw dr = philox(seed, (uint32_t)v, call, 0u, TAG_MH);
        Proposal p = propose_vertex(g, bm, v, dr);
        st.attempts++;
        const int r = bm.assign[v];
        if (p.s == r) continue;
        st.evals++;
        Delta d = vertex_delta(g, bm, v, r, p.s);
        double H = hastings(g, bm, v, d, r, p);
        double dS = delta_mdl(bm, d, p);
        if (accept(dS, H, unit53(dr.x[2], dr.x[3]))) {
            move_vertex(bm, v, d, p);
            st.moves++;
            st.dS_sum += dS;
        }
    }
    return st;
}

// finetune::early_stop (finetune.cpp:144-162).
bool early_stop(int iteration, bool bracket, double initial, const std::vector<double>& deltas) {
    size_t last = deltas.size() - 1;
    if (deltas[last] == 0.0) return true;
    if (iteration < 3) return false;
    double average = deltas[last] + deltas[last - 1] + deltas[last - 2];
    average /= -3.0;
    double threshold = bracket ? 1e-4 * initial : 5e-4 * initial;
    return average < threshold;
}

// asynchronous_gibbs (finetune.cpp:31-89), hybrid_mcmc (:342-410), metropolis_hastings
// (:412-444). algo: 0 = Metropolis-Hastings, 1 = asynchronous Gibbs, 2 = hybrid.
FinetuneResult finetune(const Graph& g, Blockmodel& bm, int algo, int batches, bool bracket,
                        uint64_t seed, uint32_t& call) {
    FinetuneResult res;
    if (bm.B == 1) return res;
    std::vector<double> deltas;
    const long N = g.N, E = g.E;
    if (algo == 0) {
        bm.overall_entropy = (float)mdl(bm, N, E);
        std::vector<int> order(g.N);
        std::iota(order.begin(), order.end(), 0);
        for (int it = 0; it < 100; ++it) {
            SweepStats st = mh_pass(g, bm, order, seed, call++);
            res.sweeps++;
            res.stats.attempts += st.attempts;
            res.stats.evals += st.evals;
            res.stats.moves += st.moves;
            deltas.push_back(st.dS_sum);
            if (early_stop(it, bracket, (double)bm.overall_entropy, deltas)) break;
        }
    } else {
        bm.overall_entropy = (float)mdl(bm, N, E);
        double initial = bm.overall_entropy;
        double last = initial;
        for (int it = 0; it < 100; ++it) {
            SweepStats a, b;
            if (algo == 2) {
                a = mh_pass(g, bm, g.high, seed, call);
                b = async_sweep(g, bm, batches, seed, call, true);
            } else {
                b = async_sweep(g, bm, batches, seed, call, false);
            }
            call++;
            res.sweeps++;
            res.stats.attempts += a.attempts + b.attempts;
            res.stats.evals += a.evals + b.evals;
            res.stats.moves += a.moves + b.moves;
            double e = mdl(bm, N, E);
            deltas.push_back(e - last);
            if (early_stop(it, bracket, initial, deltas)) break;
        }
    }
    bm.overall_entropy = (float)mdl(bm, N, E);
    return res;
}

// ------------------------------------------------------------------------------------------
// block_merge::blockmodel_delta (block_merge.cpp:13-37).
Delta merge_delta(const Blockmodel& bm, int r, int s) {
    Delta d(r, s);
    for (const auto& e : bm.rows[r]) {
        int col = e.first, value = e.second;
        if (col == r || col == s) d.add(s, s, value);
        else d.add(s, col, value);
        d.sub(r, col, value);
    }
    for (const auto& e : bm.cols[r]) {
        int row = e.first;
        if (row == r) continue;
        int value = e.second;
        if (row == s) d.add(s, s, value);
        else d.add(row, s, value);
        d.sub(row, r, value);
    }
    return d;
}

// entropy::block_merge_delta_mdl(r, proposal, bm, Delta) (entropy.cpp:127-180).
double merge_delta_mdl(const Blockmodel& bm, int r, const Proposal& p, const Delta& d) {
    const int s = d.s;
    auto deg_in = [&](int index) -> float {
        int value = bm.din[index];
        if (index == r) value -= p.k_in;
        else if (index == s) value += p.k_in;
        return float(value);
    };
    auto deg_out = [&](int index) -> float {
        int value = bm.dout[index];
        if (index == r) value -= p.k_out;
        else if (index == s) value += p.k_out;
        return float(value);
    };
    double dS = 0.0;
    for (const auto& e : d.entries()) {
        float change = (float)e.d;
        float value = (float)bm.get(e.row, e.col);
        dS += cell(value, (float)bm.din[e.col], (float)bm.dout[e.row]);
        if (e.row == r || e.col == r) continue;
        dS -= cell(value + change, deg_in(e.col), deg_out(e.row));
    }
    for (const auto& e : bm.rows[s]) {
        int col = e.first;
        float value = (float)e.second;
        if (d.get(s, col) != 0) continue;
        dS += cell(value, (float)bm.din[col], (float)bm.dout[s]);
        dS -= cell(value, deg_in(col), deg_out(s));
    }
    for (const auto& e : bm.cols[s]) {
        int row = e.first;
        float value = (float)e.second;
        if (d.get(row, s) != 0 || row == r || row == s) continue;
        dS += cell(value, (float)bm.din[s], (float)bm.dout[row]);
        dS -= cell(value, deg_in(s), deg_out(row));
    }
    return dS;
}

// propose_merge_sparse (block_merge.cpp:335-368) with propose_new_block(block_merge=true)
// (common.cpp:131-182): t in proportion to row r + column r, s in proportion to row t +
// column t without index r, a uniform other block when either total is zero; a repeated
// proposal scores DBL_MAX.
MergeEval propose_merge(const Blockmodel& bm, int r, int i, std::unordered_map<int, bool>& past,
                        uint64_t seed, uint32_t call) {
    Draw dr = philox(seed, (uint32_t)r, call, (uint32_t)i, TAG_MERGE);
    std::vector<Cell> row = sorted_cells(bm.rows[r]);
    std::vector<Cell> col = sorted_cells(bm.cols[r]);
    long k_out = 0, k_in = 0;
    for (const Cell& c : row) k_out += c.val;
    for (const Cell& c : col) k_in += c.val;
    long k = k_out + k_in;
    int s;
    if (k == 0) {
        s = random_block(r, bm.B, dr.x[2]);
    } else {
        uint32_t target = mulhi32(dr.x[0], (uint32_t)k);
        long cum = 0;
        int t = -1;
        for (const Cell& c : row) {
            cum += c.val;
            if ((long)target < cum) { t = c.idx; break; }
        }
        if (t < 0) {
            for (const Cell& c : col) {
                cum += c.val;
                if ((long)target < cum) { t = c.idx; break; }
            }
        }
        long total;
        s = walk_row_col(bm, t, dr.x[1], r, &total);
        if (total == 0) s = random_block(r, bm.B, dr.x[2]);
    }
    if (past[s]) return MergeEval{s, std::numeric_limits<double>::max()};
    Delta d = merge_delta(bm, r, s);
    Proposal p{s, (int)k_out, (int)k_in, (int)k};
    double dS = merge_delta_mdl(bm, r, p, d);
    past[s] = true;
    return MergeEval{s, dS};
}

// merge_blocks proposal loop (block_merge.cpp:278-301): best per block by strict '<'.
void merge_proposals(const Blockmodel& bm, uint64_t seed, uint32_t call, std::vector<double>& dS,
                     std::vector<int>& best) {
    const int B = bm.B;
    dS.assign(B, std::numeric_limits<double>::max());
    best.assign(B, -1);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int r = 0; r < B; ++r) {
        std::unordered_map<int, bool> past;
        for (int i = 0; i < 10; ++i) {
            MergeEval ev = propose_merge(bm, r, i, past, seed, call);
            if (ev.dS < dS[r]) {
                best[r] = ev.s;
                dS[r] = ev.dS;
            }
        }
    }
}

// Blockmodel::build_mapping (blockmodel.cpp:29-41).
static std::vector<int> build_mapping(const std::vector<int>& values) {
    std::map<int, bool> unique_map;
    for (int v : values) unique_map[v] = true;
    std::vector<int> mapping(values.size(), -1);
    int counter = 0;
    for (const auto& e : unique_map) mapping[e.first] = counter++;
    return mapping;
}

static void relabel(Blockmodel& bm) {
    std::vector<int> mapping = build_mapping(bm.assign);
    for (int& b : bm.assign) b = mapping[b];
    bm.B -= bm.num_to_merge;
}

// Blockmodel::carry_out_best_merges (blockmodel.cpp:50-78), utils::sort_indices
// (utils.hpp:94-101): std::sort ascending by delta MDL, then greedy merges through block_map.
void carry_out_best_merges(Blockmodel& bm, const std::vector<double>& dS, const std::vector<int>& best) {
    std::vector<int> order(dS.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.data(), order.data() + order.size(),
              [&dS](size_t i1, size_t i2) { return dS[i1] < dS[i2]; });
    std::vector<int> block_map(bm.B);
    std::iota(block_map.begin(), block_map.end(), 0);
    int num_merged = 0;
    size_t counter = 0;
    while (num_merged < bm.num_to_merge) {
        if (counter >= order.size()) throw std::runtime_error("ran out of merge candidates");
        int from = order[counter];
        if (best[from] < 0) throw std::runtime_error("block without a merge proposal");
        int to = block_map[best[from]];
        counter++;
        if (to != from) {
            for (size_t i = 0; i < block_map.size(); ++i)
                if (block_map[i] == from) block_map[i] = to;
            for (int& b : bm.assign)
                if (b == from) b = to;
            num_merged++;
        }
    }
    relabel(bm);
}

// carry_out_best_merges_advanced (block_merge.cpp:123-217): min-heap on delta MDL with lazy
// re-evaluation against the current matrix.
void carry_out_best_merges_advanced(Blockmodel& bm, const std::vector<double>& dS,
                                    const std::vector<int>& best) {
    typedef std::tuple<int, int, double> merge_t;
    auto cmp = [](merge_t left, merge_t right) { return std::get<2>(left) > std::get<2>(right); };
    std::priority_queue<merge_t, std::vector<merge_t>, decltype(cmp)> queue(cmp);
    for (int i = 0; i < (int)dS.size(); ++i) queue.push(std::make_tuple(i, best[i], dS[i]));
    int num_merged = 0;
    std::vector<int> block_map(bm.B);
    std::iota(block_map.begin(), block_map.end(), 0);
    while (num_merged < bm.num_to_merge && !queue.empty()) {
        merge_t merge = queue.top();
        queue.pop();
        int from = std::get<0>(merge);
        if (std::get<1>(merge) < 0) throw std::runtime_error("block without a merge proposal");
        int to = block_map[std::get<1>(merge)];
        if (from == to) continue;
        int k_out = 0, k_in = 0;
        for (const auto& e : bm.rows[from]) k_out += e.second;
        for (const auto& e : bm.cols[from]) k_in += e.second;
        Proposal p{to, k_out, k_in, k_out + k_in};
        Delta d = merge_delta(bm, from, to);
        int to_self = bm.get(to, to) + d.get(to, to);
        double actual = merge_delta_mdl(bm, from, p, d);
        if (!queue.empty() && actual > std::get<2>(queue.top())) {
            std::get<2>(merge) = actual;
            queue.push(merge);
            continue;
        }
        for (size_t i = 0; i < block_map.size(); ++i)
            if (block_map[i] == from) block_map[i] = to;
        for (int& b : bm.assign)
            if (b == from) b = to;
        apply_delta(bm, d);
        bm.dout[from] = 0;
        bm.dout[to] += p.k_out;
        bm.din[from] = 0;
        bm.din[to] += p.k_in;
        bm.deg[from] = 0;
        bm.deg[to] = bm.dout[to] + bm.din[to] - to_self;
        num_merged++;
    }
    relabel(bm);
}

// block_merge::merge_blocks (block_merge.cpp:278-311).
void merge_blocks(const Graph& g, Blockmodel& bm, bool approximate, uint64_t seed, uint32_t call) {
    std::vector<double> dS;
    std::vector<int> best;
    merge_proposals(bm, seed, call, dS, best);
    if (approximate) carry_out_best_merges(bm, dS, best);
    else carry_out_best_merges_advanced(bm, dS, best);
    build(bm, g);
}

// ------------------------------------------------------------------------------------------
// BlockmodelTriplet (blockmodel_triplet.cpp:3-100).
static Blockmodel copy_bm(const Blockmodel& bm) {
    Blockmodel c = bm;
    c.num_to_merge = 0;
    return c;
}

void Triplet::update(const Blockmodel& bm) {
    int index;
    if (bms[1].empty) {
        index = 1;
    } else if (bm.overall_entropy <= bms[1].overall_entropy) {
        int old_index = (bms[1].B > bm.B) ? 0 : 2;
        bms[old_index] = copy_bm(bms[1]);
        index = 1;
    } else {
        index = (bms[1].B > bm.B) ? 2 : 0;
    }
    bms[index] = copy_bm(bm);
}

bool Triplet::is_done() {
    if ((!bms[0].empty && bms[0].B - bms[2].B == 2) || (bms[0].empty && bms[1].B - bms[2].B == 1))
        optimal = true;
    return optimal;
}

Blockmodel Triplet::next(Blockmodel& old) {
    old.num_to_merge = 0;
    update(old);
    if (golden_ratio_not_reached()) {
        Blockmodel bm = copy_bm(bms[1]);
        bm.num_to_merge = int(bm.B * 0.5);
        if (bm.num_to_merge == 0) optimal = true;
        return bm;
    }
    if (is_done()) return copy_bm(bms[1]);
    int index;
    if (bms[0].empty && bms[1].B > bms[2].B) index = 1;
    else if ((bms[0].B - bms[1].B) >= (bms[1].B - bms[2].B)) index = 0;
    else index = 1;
    int next_num = bms[index + 1].B;
    next_num += int((bms[index].B - bms[index + 1].B) * 0.618);
    Blockmodel bm = copy_bm(bms[index]);
    bm.num_to_merge = bm.B - next_num;
    return bm;
}

// sbp::stochastic_block_partition (sbp.cpp:61-98) with done_blockmodeling (:100-112).
RunResult stochastic_block_partition(const Graph& g, int algo, int batches, bool approximate,
                                     uint64_t seed) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    RunResult res;
    std::vector<int> identity(g.N);
    std::iota(identity.begin(), identity.end(), 0);
    Blockmodel bm = make_blockmodel(g.N, g, identity, 0.5f);
    Triplet triplet;
    uint32_t sweep_call = 0, merge_call = 0;
    while (!triplet.optimal) {
        auto m0 = clk::now();
        merge_blocks(g, bm, approximate, seed, merge_call++);
        auto m1 = clk::now();
        FinetuneResult f = finetune(g, bm, algo, batches, !triplet.bms[2].empty, seed, sweep_call);
        auto m2 = clk::now();
        res.merge_seconds += std::chrono::duration<double>(m1 - m0).count();
        res.mcmc_seconds += std::chrono::duration<double>(m2 - m1).count();
        res.sweeps += f.sweeps;
        res.attempts += f.stats.attempts;
        res.evals += f.stats.evals;
        res.moves += f.stats.moves;
        res.outer_iterations++;
        res.trace_B.push_back(bm.B);
        res.trace_mdl.push_back(bm.overall_entropy);
        if (getenv("ORACLE_TRACE"))
            fprintf(stderr, "oracle iter %d: B=%d mdl=%.3f sweeps=%d moves=%ld\n", res.outer_iterations, bm.B,
                    (double)bm.overall_entropy, f.sweeps, f.stats.moves);
        bm = triplet.next(bm);
    }
    res.bm = bm;
    res.seconds = std::chrono::duration<double>(clk::now() - t0).count();
    return res;
}

}  // namespace oracle
