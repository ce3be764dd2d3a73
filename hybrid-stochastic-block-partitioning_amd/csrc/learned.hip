// Learned proposal. For a vertex v in block r with degree <= the context's learned degree bound,
// the candidates are the 32 blocks v's neighbours fall in most often (self-loop excluded; ties:
// smaller block id), and the proposal is the mixture
//   q(s | v) = w * q_plain(s | v) + (1 - w) * q_net(s | v),   w = LEARN_W_PLAIN,
// where q_plain is the reference's proposal (common.cpp:131-182) and q_net a softmax of the
// network's scores over the candidates other than r. The features do not change when v itself
// moves: the candidates are v's neighbour blocks, and every block degree is taken without v's own
// contribution (din[b] - (k_in + self), dout[b] - k_out for b = v's block). So one set of scores
// serves both directions of the Hastings ratio: forward q_net(s) is the softmax over the
// candidates other than r, reverse q_net(r) the softmax over the candidates other than s. q_plain
// enters through the reference's closed form (entropy.cpp:495-557, divided by the vertex
// degree). The chain therefore keeps the reference's target: every move is still judged by its
// delta MDL. A vertex whose neighbours all sit in its own block has no learned mass and uses the
// plain proposal alone.
//
// Features (64, BF16): for candidate i < 16: count_i / k, log1p(din'[b_i]) / log1p(E),
// log1p(dout'[b_i]) / log1p(E) (degrees without v); then log1p(k_out) / log(4097),
// log1p(k_in) / log(4097), the share of v's edges into candidates 16-31, the number of distinct
// neighbour blocks / k, log(B) / log(N), 1, and ten zeros. Scores: output i < 32 scores
// candidate i.
#include <cmath>

#include "eval_core.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace sbpx {

constexpr int FEAT_WAVES = 4;
// neighbour-block table per wavefront: 256 slots for degree <= LEARN_MAXDEG, 1024 for degree
// <= LEARN_MH_MAXDEG (at most half full either way)
constexpr int FEAT_CAP = 256;
constexpr int FEAT_CAP_MH = 1024;

__device__ __forceinline__ __bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    const unsigned short h = (unsigned short)(u >> 16);
    return __builtin_bit_cast(__bf16, h);
}

struct FeatParams {
    GraphView G;
    BlockState S;
    const int* list;  // vertices
    int n;
    float inv_logE;   // 1 / log1p(E)
    float log_ratio;  // log(B) / log(N)
    __bf16* x;        // [n][64]
    int* cand;        // [n][32]
    int* cnt;         // [n][32]
    int skip_deg;     // vertices with more list entries (out + in) get an empty row
};

template <int CAP>
__global__ void __launch_bounds__(FEAT_WAVES * WAVE) features_kernel(FeatParams P) {
    __shared__ int tkey[FEAT_WAVES][CAP];
    __shared__ int tcnt[FEAT_WAVES][CAP];
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    const int i = blockIdx.x * FEAT_WAVES + wave;
    if (i >= P.n) return;
    int* key = tkey[wave];
    int* cntt = tcnt[wave];
    for (int j = lane; j < CAP; j += WAVE) {
        key[j] = -1;
        cntt[j] = 0;
    }
    fence_wave();
    const int v = P.list[i];
    const int r = P.S.assign[v];
    const int ob = P.G.out_ptr[v], dO = P.G.out_ptr[v + 1] - ob;
    const int ib = P.G.in_ptr[v], dI = P.G.in_ptr[v + 1] - ib;
    const int self = P.G.self[v];
    if (dO + dI > P.skip_deg) {  // no learned part for this vertex: its row is never read
        if (lane < MLP_CAND) {
            P.cand[(long)i * MLP_CAND + lane] = -1;
            P.cnt[(long)i * MLP_CAND + lane] = 0;
        }
        P.x[(long)i * MLP_IN + lane] = f2bf(0.0f);
        return;
    }
    for (int j = lane; j < dO + dI; j += WAVE) {
        const int w = j < dO ? P.G.out_idx[ob + j] : P.G.in_idx[ib + (j - dO)];
        if (w == v) continue;
        const int slot = tab_insert(key, CAP, P.S.assign[w]);
        if (slot >= 0) atomicAdd(&cntt[slot], 1);
    }
    fence_wave();
    // top 32 by (count desc, block asc): each lane holds the scores of its CAP / 64 slots
    // (count << 32 | ~block, unique per block) and the wavefront takes the largest 32 times
    int* cand = P.cand + (long)i * MLP_CAND;
    int* cnt = P.cnt + (long)i * MLP_CAND;
    constexpr int PER = CAP / WAVE;
    long long sc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int kk = key[lane + q * WAVE], c = cntt[lane + q * WAVE];
        sc[q] = (kk >= 0 && c > 0) ? (((long long)c << 32) | (long long)(0x7FFFFFFF - kk)) : -1;
    }
    for (int t = 0; t < MLP_CAND; ++t) {
        long long best = -1;
#pragma unroll
        for (int q = 0; q < PER; ++q) best = sc[q] > best ? sc[q] : best;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const long long o = __shfl_xor(best, off, WAVE);
            best = o > best ? o : best;
        }
        if (best < 0) {  // no block left: the remaining candidates are empty
            if (lane == 0)
                for (int u = t; u < MLP_CAND; ++u) {
                    cand[u] = -1;
                    cnt[u] = 0;
                }
            break;
        }
        int b = -1, c = 0;
        if (best >= 0) {
            b = 0x7FFFFFFF - (int)(best & 0xFFFFFFFF);
            c = (int)(best >> 32);
#pragma unroll
            for (int q = 0; q < PER; ++q)
                if (sc[q] == best) sc[q] = -1;  // taken
        }
        if (lane == 0) {
            cand[t] = b;
            cnt[t] = c;
        }
    }
    fence_wave();  // lane 0's candidate rows before every lane reads them
    const int k_out = dO, k_in = dI - self;
    int nuniq = 0;
    for (int q = 0; q < CAP / WAVE; ++q) nuniq += (key[lane + q * WAVE] >= 0) ? 1 : 0;
    nuniq = (int)wave_sum_ll(nuniq);
    P.x[(long)i * MLP_IN + lane] =
        f2bf(learned_feature(lane, cand, cnt, k_out, k_in, self, r, nuniq, P.S.din, P.S.dout, P.inv_logE, P.log_ratio));
}

struct MixParams {
    GraphView G;
    BlockState S;
    const int* list;
    int n;
    uint64_t seed;
    uint32_t call;
    float w_plain;
    const float* z;       // [n][32] scores
    const int* cand;
    const int* s_plain;   // [N] plain proposal (vertex-indexed)
    const int* s_given;   // [n] the target to evaluate (instead of a draw), or null
    int* ps;              // [N] mixture proposal
    float* qf;            // [N] learned forward probability (< 0: no learned mass)
    float* qb;            // [N] learned reverse probability
};

// 32 lanes per vertex, one per candidate slot (two vertices per wavefront): draws the mixture's
// proposal (the branch, then on the learned branch a candidate from the softmax over the
// candidates other than r), or takes the given target, and the net's forward and reverse
// probabilities of the move from the same scores. The maxima are order-free; the softmax sums and
// the draw's running sum are folded in slot order by every lane of the half, so each value has
// the bits of cand_prob and learned_pick (which the sequential pass uses).
__global__ void __launch_bounds__(256) mixture_kernel(MixParams P) {
    const int lane = lane_id();
    const int half = lane & 32;  // first lane of this vertex's half-wavefront
    const int j = lane & 31;     // candidate slot
    const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 5);
    const bool on = i < P.n;
    const int v = on ? P.list[i] : 0;
    const int r = on ? P.S.assign[v] : -1;
    const int c = on ? P.cand[(long)i * MLP_CAND + j] : -1;
    const float zj = on ? P.z[(long)i * MLP_OUT + j] : 0.0f;
    auto half_max = [&](float x) {
        for (int off = 16; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off, WAVE));
        return x;
    };
    auto half_any = [&](bool b) { return ((__ballot(b) >> half) & 0xFFFFFFFFull) != 0ull; };
    // forward: the candidates other than r
    const bool qf_ok = c >= 0 && c != r;
    const bool any = half_any(qf_ok);
    const float mx_f = half_max(qf_ok ? zj : -1e30f);
    const float e_f = qf_ok ? expf(zj - mx_f) : 0.0f;
    float den_f = 0.0f;
    for (int k = 0; k < MLP_CAND; ++k) den_f += __shfl(e_f, half + k, WAVE);  // slot order; + 0 is exact
    int s = on ? (P.s_given ? P.s_given[i] : P.s_plain[v]) : -1;
    if (any && P.s_given == nullptr && on) {
        const Draw4 dr = philox4x32_10(P.seed, (uint32_t)v, P.call, 0u, TAG_LEARN);
        if (unit53(dr.x0, dr.x1) >= (double)P.w_plain) {
            const float target = (float)unit53(dr.x2, dr.x3) * den_f;
            float run = 0.0f;
            int pick = -1;
            bool done = false;
            for (int k = 0; k < MLP_CAND; ++k) {
                const float ek = __shfl(e_f, half + k, WAVE);
                const int ck = __shfl(c, half + k, WAVE);
                if (!done && ck >= 0 && ck != r) {
                    run += ek;
                    pick = ck;
                    done = run > target;
                }
            }
            s = pick;
        }
    }
    // the forward probability of s, and the reverse one of r over the candidates other than s
    const float num_f = (qf_ok && c == s) ? e_f : 0.0f;
    const bool qb_ok = c >= 0 && c != s;
    const bool any_b = half_any(qb_ok);
    const float mx_b = half_max(qb_ok ? zj : -1e30f);
    const float e_b = qb_ok ? expf(zj - mx_b) : 0.0f;
    float den_b = 0.0f, nf = 0.0f, nb = 0.0f;
    for (int k = 0; k < MLP_CAND; ++k) {
        den_b += __shfl(e_b, half + k, WAVE);
        nf = fmaxf(nf, __shfl(num_f, half + k, WAVE));
        const int ck = __shfl(c, half + k, WAVE);
        const float ebk = __shfl(e_b, half + k, WAVE);
        if (ck == r && ck >= 0 && ck != s) nb = ebk;
    }
    if (on && j == 0) {
        P.ps[v] = s;
        P.qf[v] = any ? nf / den_f : -1.0f;
        P.qb[v] = s == r ? -1.0f : (any_b ? nb / den_b : -1.0f);
    }
}

// Appends the accepted moves whose target is a candidate to the training set.
// Whether the move of list entry i is a training row: accepted, with its target among the
// candidates (mask: the candidates other than the own block; label: the target's index).
__device__ __forceinline__ bool log_row(const int* __restrict__ list, int i, const int* __restrict__ assign,
                                        const uint8_t* __restrict__ stat, const int* __restrict__ out_s,
                                        const int* __restrict__ cand, uint32_t* mask_out, int* label_out) {
    const int v = list[i];
    if (stat[v] != 2) return false;
    const int s = out_s[v], r = assign[v];
    const int* c = cand + (long)i * MLP_CAND;
    int label = -1;
    uint32_t mask = 0;
    for (int j = 0; j < MLP_CAND; ++j) {
        if (c[j] >= 0 && c[j] != r) mask |= 1u << j;
        if (c[j] == s) label = j;
    }
    *mask_out = mask;
    *label_out = label;
    return label >= 0 && (mask >> label & 1u);
}

__global__ void log_flags_kernel(const int* __restrict__ list, int n, const int* __restrict__ assign,
                                 const uint8_t* __restrict__ stat, const int* __restrict__ out_s,
                                 const int* __restrict__ cand, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t mask;
    int label;
    flag[i] = log_row(list, i, assign, stat, out_s, cand, &mask, &label) ? 1 : 0;
}

// Rows go to count + (their rank in the list), so the log holds the same rows in the same order
// on every run; rows past the capacity are dropped.
__global__ void log_moves_kernel(const int* __restrict__ list, int n, const int* __restrict__ assign,
                                 const uint8_t* __restrict__ stat, const int* __restrict__ out_s,
                                 const __bf16* __restrict__ x, const int* __restrict__ cand,
                                 const int* __restrict__ rank, __bf16* ds_x, uint32_t* ds_mask, int* ds_y,
                                 const int* __restrict__ count, int cap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t mask;
    int label;
    if (!log_row(list, i, assign, stat, out_s, cand, &mask, &label)) return;
    const int slot = count[0] + rank[i];
    if (slot >= cap) return;
    for (int j = 0; j < MLP_IN; ++j) ds_x[(long)slot * MLP_IN + j] = x[(long)i * MLP_IN + j];
    ds_mask[slot] = mask;
    ds_y[slot] = label;
}

__global__ void log_advance_kernel(int* count, const int* __restrict__ rank, int n) { count[0] += rank[n]; }

// ------------------------------------------------------------------------------------------
static FeatParams feat_params(Ctx& c, const int* list, int n) {
    FeatParams P{};
    P.G = GraphView{c.g.out_ptr, c.g.out_idx, c.g.in_ptr, c.g.in_idx, c.g.self_flag};
    P.S = BlockState{c.bm.assign, c.bm.dout, c.bm.din, c.bm.deg, c.bm.B};
    P.list = list;
    P.n = n;
    P.inv_logE = 1.0f / log1pf((float)c.g.E);
    P.log_ratio = c.g.N > 1 ? (float)(log((double)c.bm.B) / log((double)c.g.N)) : 0.0f;
    P.x = c.mlp.x;
    P.cand = c.mlp.cand;
    P.cnt = c.mlp.cnt;
    return P;
}

void learned_features(Ctx& c, const int* list, int n, int max_deg, int skip_deg) {
    if (n <= 0) return;
    FeatParams P = feat_params(c, list, n);
    P.skip_deg = skip_deg;
    // vertices above the table's degree bound get features from a partial table; their scores
    // are not used (the mixture gives them no learned part)
    if (max_deg > LEARN_MAXDEG)
        features_kernel<FEAT_CAP_MH><<<(n + FEAT_WAVES - 1) / FEAT_WAVES, FEAT_WAVES * WAVE, 0, c.stream>>>(P);
    else
        features_kernel<FEAT_CAP><<<(n + FEAT_WAVES - 1) / FEAT_WAVES, FEAT_WAVES * WAVE, 0, c.stream>>>(P);
    SBPX_HIP(hipGetLastError());
}

void learned_log(Ctx& c, const int* list, int n) {
    MlpState& m = c.mlp;
    if (n <= 0 || !m.logging) return;
    if ((size_t)n + 1 > m.log_slots) {
        if (m.log_flag) SBPX_HIP(hipFree(m.log_flag));
        if (m.log_rank) SBPX_HIP(hipFree(m.log_rank));
        m.log_flag = m.log_rank = nullptr;
        m.log_slots = 0;
        SBPX_HIP(hipMalloc((void**)&m.log_flag, sizeof(int) * ((size_t)n + 1)));
        SBPX_HIP(hipMalloc((void**)&m.log_rank, sizeof(int) * ((size_t)n + 1)));
        m.log_slots = (size_t)n + 1;
    }
    const int nb = (n + 255) / 256;
    log_flags_kernel<<<nb, 256, 0, c.stream>>>(list, n, c.bm.assign, c.ev_acc, c.ev_s, m.cand, m.log_flag);
    scan_exclusive(m.log_flag, m.log_rank, n, c.scan_tmp, c.stream);
    log_moves_kernel<<<nb, 256, 0, c.stream>>>(list, n, c.bm.assign, c.ev_acc, c.ev_s, m.x, m.cand, m.log_rank,
                                               m.ds_x, m.ds_mask, m.ds_y, m.ds_count, m.ds_cap);
    log_advance_kernel<<<1, 1, 0, c.stream>>>(m.ds_count, m.log_rank, n);
    SBPX_HIP(hipGetLastError());
}

void learned_mixture(Ctx& c, const int* list, int n, uint64_t seed, uint32_t call, const int* s_given) {
    if (n <= 0) return;
    MlpState& m = c.mlp;
    mlp_forward(c, m.x, n, m.sh1, m.sh2, m.sz);
    MixParams P{};
    P.G = GraphView{c.g.out_ptr, c.g.out_idx, c.g.in_ptr, c.g.in_idx, c.g.self_flag};
    P.S = BlockState{c.bm.assign, c.bm.dout, c.bm.din, c.bm.deg, c.bm.B};
    P.list = list;
    P.n = n;
    P.seed = seed;
    P.call = call;
    P.w_plain = LEARN_W_PLAIN;
    P.z = m.sz;
    P.cand = m.cand;
    P.s_plain = c.ev_s;
    P.s_given = s_given;
    P.ps = m.ps;
    P.qf = m.qf;
    P.qb = m.qb;
    mixture_kernel<<<(n + 7) / 8, 256, 0, c.stream>>>(P);  // 8 vertices per workgroup
    SBPX_HIP(hipGetLastError());
}

}  // namespace sbpx
