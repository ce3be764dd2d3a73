// Seeded directed degree-corrected SBM generator (host only).
//   1. block proportions pi ~ Dirichlet(alpha 1_K); sizes by largest remainder, at least 1;
//   2. out-degree k_v ~ truncated discrete power law P(k) ~ k^-gamma on [kmin, kmax]
//      (inverse CDF);
//   3. in-propensity theta_v from an independent draw of the same law;
//   4. each out-stub picks its target block (own block with probability p_in, otherwise a
//      uniform other block) and a destination in proportion to theta inside that block;
//   5. self-loops and duplicate edges are dropped.
// Vertex ids are a random permutation of the block layout. Out lists are sorted by destination
// and in lists by source. The reference's generator (graph-tool) is not part of its tree.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/sbpx.h"

namespace {

struct Rng {  // xoshiro256** seeded by splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // Marsaglia-Tsang gamma(alpha, 1)
    double gamma(double alpha) {
        if (alpha < 1.0) {
            double u = uniform();
            return gamma(alpha + 1.0) * std::pow(u < 1e-300 ? 1e-300 : u, 1.0 / alpha);
        }
        double d = alpha - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x = normal(), v = 1.0 + c * x;
            if (v <= 0) continue;
            v = v * v * v;
            double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
        }
    }
};

struct PowerLaw {
    int kmin, kmax;
    std::vector<double> cdf;
    PowerLaw(int lo, int hi, double gamma) : kmin(lo), kmax(hi) {
        double acc = 0.0;
        for (int k = lo; k <= hi; ++k) {
            acc += std::pow((double)k, -gamma);
            cdf.push_back(acc);
        }
        for (double& x : cdf) x /= acc;
    }
    int draw(Rng& rng) const {
        double u = rng.uniform();
        size_t i = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        if (i >= cdf.size()) i = cdf.size() - 1;
        return kmin + (int)i;
    }
};

}  // namespace

extern "C" sbpx_status sbpx_generate_dcsbm(const sbpx_dcsbm_params* p, int64_t* num_edges, int32_t* out_ptr,
                                           int32_t* out_idx, int32_t* in_ptr, int32_t* in_idx, int32_t* truth) {
    if (!p || !num_edges || p->num_vertices < 2 || p->num_blocks < 1 || p->num_blocks > p->num_vertices ||
        p->deg_min < 1 || p->deg_max < p->deg_min || p->p_in < 0 || p->p_in > 1 || p->dirichlet <= 0)
        return SBPX_ERR_ARG;
    const int N = p->num_vertices, K = p->num_blocks;
    Rng rng(p->seed);
    // 1. block sizes
    std::vector<double> pi(K);
    double tot = 0.0;
    for (int k = 0; k < K; ++k) tot += (pi[k] = rng.gamma(p->dirichlet));
    for (double& x : pi) x /= tot;
    std::vector<int> size(K, 1);
    int left = N - K;
    std::vector<double> rem(K);
    int assigned = 0;
    for (int k = 0; k < K; ++k) {
        double want = pi[k] * left;
        int base = (int)std::floor(want);
        size[k] += base;
        assigned += base;
        rem[k] = want - base;
    }
    std::vector<int> order(K);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return rem[a] > rem[b] || (rem[a] == rem[b] && a < b); });
    for (int i = 0; assigned < left; ++i, ++assigned) size[order[i % K]]++;
    // vertex ids: random permutation of the block layout
    std::vector<int> block_of(N);
    {
        int v = 0;
        for (int k = 0; k < K; ++k)
            for (int i = 0; i < size[k]; ++i) block_of[v++] = k;
        for (int i = N - 1; i > 0; --i) std::swap(block_of[i], block_of[rng.below(i + 1)]);
    }
    std::vector<std::vector<int>> members(K);
    for (int v = 0; v < N; ++v) members[block_of[v]].push_back(v);
    // 2-3. degrees and propensities
    PowerLaw law(p->deg_min, p->deg_max, p->exponent);
    std::vector<int> kout(N);
    std::vector<double> theta(N);
    for (int v = 0; v < N; ++v) kout[v] = law.draw(rng);
    for (int v = 0; v < N; ++v) theta[v] = law.draw(rng);
    std::vector<std::vector<double>> cum(K);
    for (int k = 0; k < K; ++k) {
        double acc = 0.0;
        for (int v : members[k]) cum[k].push_back(acc += theta[v]);
    }
    // 4. stubs
    std::vector<uint64_t> edges;
    edges.reserve((size_t)N * 24);
    for (int v = 0; v < N; ++v) {
        const int b = block_of[v];
        for (int e = 0; e < kout[v]; ++e) {
            int tb = b;
            if (K > 1 && rng.uniform() >= p->p_in) {
                tb = (int)rng.below(K - 1);
                if (tb >= b) tb++;
            }
            const std::vector<double>& cb = cum[tb];
            double x = rng.uniform() * cb.back();
            size_t i = std::upper_bound(cb.begin(), cb.end(), x) - cb.begin();
            if (i >= cb.size()) i = cb.size() - 1;
            int w = members[tb][i];
            if (w == v) continue;
            edges.push_back(((uint64_t)v << 32) | (uint32_t)w);
        }
    }
    // 5. dedup
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    const int64_t E = (int64_t)edges.size();
    *num_edges = E;
    if (truth)
        for (int v = 0; v < N; ++v) truth[v] = block_of[v];
    if (!out_ptr) return SBPX_OK;
    std::vector<int> indeg(N, 0);
    for (int v = 0; v <= N; ++v) out_ptr[v] = 0;
    for (uint64_t e : edges) {
        out_ptr[(e >> 32) + 1]++;
        indeg[(uint32_t)e]++;
    }
    for (int v = 0; v < N; ++v) out_ptr[v + 1] += out_ptr[v];
    for (int64_t i = 0; i < E; ++i) out_idx[i] = (int32_t)(uint32_t)edges[i];
    in_ptr[0] = 0;
    for (int v = 0; v < N; ++v) in_ptr[v + 1] = in_ptr[v] + indeg[v];
    std::vector<int> fill(in_ptr, in_ptr + N);
    for (uint64_t e : edges) in_idx[fill[(uint32_t)e]++] = (int32_t)(e >> 32);  // sources ascending
    return SBPX_OK;
}
