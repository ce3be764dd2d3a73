// Learned proposal for the batched (asynchronous Gibbs) evaluation. For a vertex v in block r
// with degree <= LEARN_MAXDEG the network scores the 16 blocks its neighbours fall in most often
// (ties: smaller block id); the proposal is the mixture
//   q(s | v) = w * q_plain(s | v) + (1 - w) * q_net(s | v),   w = LEARN_W_PLAIN,
// where q_plain is the reference's proposal (common.cpp:131-182) and q_net a softmax over the
// candidates other than r. The Hastings ratio uses the mixture in both directions: the forward
// q_net(s) from the scores before the move, the reverse q_net'(r) from the scores of the
// features after the move (the candidates do not change when v moves; the degrees of r and s,
// and the own block, do). q_plain enters through the reference's closed form (entropy.cpp:
// 495-557, divided by the vertex degree). The chain therefore keeps the reference's target:
// every move is still judged by its delta MDL. A vertex whose neighbours all sit in its own
// block has no learned mass and uses the plain proposal alone.
//
// Features (64, BF16): for candidate i < 16: count_i / k, log1p(din[b_i]) / log1p(E),
// log1p(dout[b_i]) / log1p(E); then log1p(k_out) / log(4097), log1p(k_in) / log(4097),
// log1p(din[own]) / log1p(E), log1p(dout[own]) / log1p(E), count_own / k (0 unless the own block
// is a candidate), [own block is a candidate], log(B) / log(N), 1, and eight zeros.
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
    int* cand;        // [n][16]
    int* cnt;         // [n][16]
    int skip_deg;     // vertices with more list entries (out + in) get an empty row
};

// One lane writes feature j (< 64) of a vertex whose own block is `own`; din_of / dout_of give
// the block degrees to use (before or after the move).
template <class DinOf, class DoutOf>
__device__ __forceinline__ float feature(int j, const int* cand, const int* cnt, int k_out, int k_in, int own,
                                         DinOf din_of, DoutOf dout_of, float inv_logE, float log_ratio) {
    const int k = k_out + k_in;
    const float invk = k > 0 ? 1.0f / (float)k : 0.0f;
    if (j < 16) return cand[j] >= 0 ? (float)cnt[j] * invk : 0.0f;
    if (j < 32) return cand[j - 16] >= 0 ? log1pf((float)din_of(cand[j - 16])) * inv_logE : 0.0f;
    if (j < 48) return cand[j - 32] >= 0 ? log1pf((float)dout_of(cand[j - 32])) * inv_logE : 0.0f;
    const float inv_logD = 1.0f / logf(4097.0f);
    int own_i = -1;
    for (int i = 0; i < 16; ++i)
        if (cand[i] == own) own_i = i;
    switch (j - 48) {
        case 0: return log1pf((float)k_out) * inv_logD;
        case 1: return log1pf((float)k_in) * inv_logD;
        case 2: return log1pf((float)din_of(own)) * inv_logE;
        case 3: return log1pf((float)dout_of(own)) * inv_logE;
        case 4: return own_i >= 0 ? (float)cnt[own_i] * invk : 0.0f;
        case 5: return own_i >= 0 ? 1.0f : 0.0f;
        case 6: return log_ratio;
        case 7: return 1.0f;
        default: return 0.0f;
    }
}

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
    // top 16 by (count desc, block asc): each lane holds the scores of its CAP / 64 slots
    // (count << 32 | ~block, unique per block) and the wavefront takes the largest 16 times
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
    auto din_of = [&](int b) { return P.S.din[b]; };
    auto dout_of = [&](int b) { return P.S.dout[b]; };
    P.x[(long)i * MLP_IN + lane] =
        f2bf(feature(lane, cand, cnt, k_out, k_in, r, din_of, dout_of, P.inv_logE, P.log_ratio));
}

struct MixParams {
    GraphView G;
    BlockState S;
    const int* list;
    int n;
    uint64_t seed;
    uint32_t call;
    float w_plain;
    const float* z;       // [n][32] scores before the move
    const int* cand;
    const int* cnt;
    const int* s_plain;   // [N] plain proposal (vertex-indexed)
    int* ps;              // [N] mixture proposal
    float* qf;            // [N]
    __bf16* x2;           // [n][64] features after the move
    float inv_logE;
    float log_ratio;
};

// One wavefront per vertex: draws the mixture's proposal, its learned forward probability, and
// the features of the moved vertex.
__global__ void __launch_bounds__(FEAT_WAVES * WAVE) mixture_kernel(MixParams P) {
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    const int i = blockIdx.x * FEAT_WAVES + wave;
    if (i >= P.n) return;
    const int v = P.list[i];
    const int r = P.S.assign[v];
    const int* cand = P.cand + (long)i * MLP_CAND;
    const int* cnt = P.cnt + (long)i * MLP_CAND;
    const float* z = P.z + (long)i * MLP_OUT;
    const Draw4 dr = philox4x32_10(P.seed, (uint32_t)v, P.call, 0u, TAG_LEARN);
    int s = P.s_plain[v];
    // softmax weights of the candidates other than r, formed once (the values and summation
    // order of cand_prob): the categorical draw and the forward probability both use them
    float qf = -1.0f;  // no learned mass
    float mx = -1e30f;
    int any = 0;
#pragma unroll
    for (int j = 0; j < MLP_CAND; ++j)
        if (cand[j] >= 0 && cand[j] != r) {
            mx = fmaxf(mx, z[j]);
            any = 1;
        }
    if (any) {
        float e[MLP_CAND];
        float den = 0.0f;
#pragma unroll
        for (int j = 0; j < MLP_CAND; ++j) {
            e[j] = 0.0f;
            if (cand[j] >= 0 && cand[j] != r) {
                e[j] = expf(z[j] - mx);
                den += e[j];
            }
        }
        if (unit53(dr.x0, dr.x1) >= (double)P.w_plain) {
            // categorical draw over the softmax, candidates in slot order
            const float target = (float)unit53(dr.x2, dr.x3) * den;
            float run = 0.0f;
            int pick = -1;
#pragma unroll
            for (int j = 0; j < MLP_CAND; ++j)
                if (cand[j] >= 0 && cand[j] != r) {
                    run += e[j];
                    pick = cand[j];
                    if (run > target) break;
                }
            s = pick;
        }
        float num = 0.0f;
#pragma unroll
        for (int j = 0; j < MLP_CAND; ++j)
            if (cand[j] >= 0 && cand[j] != r && cand[j] == s) num = e[j];
        qf = num / den;
    }
    if (lane == 0) {
        P.ps[v] = s;
        P.qf[v] = qf;
    }
    const int ob = P.G.out_ptr[v], dO = P.G.out_ptr[v + 1] - ob;
    const int dI = P.G.in_ptr[v + 1] - P.G.in_ptr[v];
    const int self = P.G.self[v];
    const int k_out = dO, k_in = dI - self;
    // block degrees after v moves r -> s (Blockmodel::move_vertex, blockmodel.cpp:344-360)
    auto din_of = [&](int b) {
        int d = P.S.din[b];
        if (s != r) d += b == s ? (k_in + self) : (b == r ? -(k_in + self) : 0);
        return d;
    };
    auto dout_of = [&](int b) {
        int d = P.S.dout[b];
        if (s != r) d += b == s ? k_out : (b == r ? -k_out : 0);
        return d;
    };
    P.x2[(long)i * MLP_IN + lane] =
        f2bf(feature(lane, cand, cnt, k_out, k_in, s, din_of, dout_of, P.inv_logE, P.log_ratio));
}

// The learned part for a sequential pass, scores fixed for the pass: per queue vertex the
// branch and, on the learned branch, the candidate drawn and the accept uniform of the plain
// branch's stream. ps[v] = -2: no learned part (degree above maxdeg); -1: the plain branch,
// drawn during the pass; else the target.
__global__ void mh_mixture_kernel(const int* __restrict__ list, int n, GraphView G, const int* __restrict__ assign,
                                  const float* __restrict__ z, const int* __restrict__ cand, uint64_t seed,
                                  uint32_t call, uint32_t tag, float w_plain, int maxdeg, int* __restrict__ ps,
                                  double* __restrict__ pu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = list[i];
    const int deg = (G.out_ptr[v + 1] - G.out_ptr[v]) + (G.in_ptr[v + 1] - G.in_ptr[v]);
    if (deg > maxdeg) {
        ps[v] = -2;
        return;
    }
    const float* zi = z + (long)i * MLP_OUT;
    const int* ci = cand + (long)i * MLP_CAND;
    const int r = assign[v];
    const Draw4 dl = philox4x32_10(seed, (uint32_t)v, call, 0u, TAG_LEARN);
    if (cand_prob(zi, ci, r, -1) < 0.0f || unit53(dl.x0, dl.x1) < (double)w_plain) {
        ps[v] = -1;
        return;
    }
    float mx = -1e30f;
    for (int j = 0; j < MLP_CAND; ++j)
        if (ci[j] >= 0 && ci[j] != r) mx = fmaxf(mx, zi[j]);
    float den = 0.0f;
    for (int j = 0; j < MLP_CAND; ++j)
        if (ci[j] >= 0 && ci[j] != r) den += expf(zi[j] - mx);
    const float target = (float)unit53(dl.x2, dl.x3) * den;
    float run = 0.0f;
    int pick = -1;
    for (int j = 0; j < MLP_CAND; ++j)
        if (ci[j] >= 0 && ci[j] != r) {
            run += expf(zi[j] - mx);
            pick = ci[j];
            if (run > target) break;
        }
    const Draw4 dr = philox4x32_10(seed, (uint32_t)v, call, 0u, tag);
    ps[v] = pick;
    pu[v] = unit53(dr.x2, dr.x3);
}

void learned_mh_draws(Ctx& c, const int* list, int n, int max_deg, uint64_t seed, uint32_t call, uint32_t tag) {
    if (n <= 0) return;
    MlpState& m = c.mlp;
    const int maxdeg = std::min(c.learn_maxdeg, LEARN_MH_MAXDEG);
    learned_features(c, list, n, std::min(max_deg, maxdeg), maxdeg);
    mlp_forward(c, m.x, n, m.sh1, m.sh2, m.sz);
    const GraphView G{c.g.out_ptr, c.g.out_idx, c.g.in_ptr, c.g.in_idx, c.g.self_flag};
    mh_mixture_kernel<<<(n + 255) / 256, 256, 0, c.stream>>>(list, n, G, c.bm.assign, m.sz, m.cand, seed, call, tag,
                                                            LEARN_W_PLAIN, maxdeg, m.ps, m.pu);
    SBPX_HIP(hipGetLastError());
}

// Learned reverse probability q_net'(r | v in s) from the scores after the move.
__global__ void reverse_kernel(const int* __restrict__ list, int n, const int* __restrict__ assign,
                               const float* __restrict__ z2, const int* __restrict__ cand, const int* __restrict__ ps,
                               float* __restrict__ qb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = list[i];
    qb[v] = cand_prob(z2 + (long)i * MLP_OUT, cand + (long)i * MLP_CAND, ps[v], assign[v]);
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

void learned_mixture(Ctx& c, const int* list, int n, uint64_t seed, uint32_t call) {
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
    P.cnt = m.cnt;
    P.s_plain = c.ev_s;
    P.ps = m.ps;
    P.qf = m.qf;
    P.x2 = m.x2;
    P.inv_logE = 1.0f / log1pf((float)c.g.E);
    P.log_ratio = c.g.N > 1 ? (float)(log((double)c.bm.B) / log((double)c.g.N)) : 0.0f;
    mixture_kernel<<<(n + FEAT_WAVES - 1) / FEAT_WAVES, FEAT_WAVES * WAVE, 0, c.stream>>>(P);
    SBPX_HIP(hipGetLastError());
    mlp_forward(c, m.x2, n, m.sh1, m.sh2, m.sz2);
    reverse_kernel<<<(n + 255) / 256, 256, 0, c.stream>>>(list, n, c.bm.assign, m.sz2, m.cand, m.ps, m.qb);
    SBPX_HIP(hipGetLastError());
}

}  // namespace sbpx
