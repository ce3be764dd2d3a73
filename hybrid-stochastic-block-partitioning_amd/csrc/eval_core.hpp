// Vertex-move evaluation by one wavefront, shared by the batched (asynchronous Gibbs) kernels
// and the sequential Metropolis-Hastings kernel. The block matrix is read through a view:
//   CsrView  — the immutable CSR/CSC snapshot, rows sorted, with inclusive prefix sums;
//   HashView — a mutable per-row open-addressing table (CSR and CSC orientations), used while
//              moves are applied one at a time; empty slots have key -1 and cells that dropped
//              to zero keep their slot with value 0.
//
// Steps (one wavefront, vertex v in block r):
//   1. neighbour blocks into a table: per block b, o_b out-neighbours and i_b in-neighbours in b
//      (self-loop kept apart);
//   2. proposal (common.cpp:131-192): t = block of a uniform neighbour slot, then s in proportion
//      to row t + column t taken in ascending index order (row first);
//   3. one pass over rows r, s and columns r, s. Every cell the move changes is a function of
//      (o_b, i_b, self), so the Delta of finetune::blockmodel_delta (finetune.cpp:108-142) needs
//      no hash of its own: a scanned cell is a Delta key (handled in step 4) or keeps its value
//      and sees new degrees (added here). A key whose net change is zero is counted in both
//      places, as entropy::delta_mdl does (entropy.cpp:317-401). The pass also records
//      M[s,b], M[b,s], M[r,b], M[b,r] per neighbour block for the Hastings correction;
//   4. Delta-key terms and Hastings sums (entropy.cpp:495-557), wave reduction, accept rule
//      (finetune.cpp:19-29).
// FP32 cell entropies, FP64 accumulation, as the reference.
#pragma once
#include <climits>

#include "device_common.hpp"
#include "mlp.hpp"

namespace sbpx {

struct CsrView {
    const int* __restrict__ row_ptr;
    const int* __restrict__ row_col;
    const int* __restrict__ row_val;
    const int* __restrict__ row_cum;
    const int* __restrict__ col_ptr;
    const int* __restrict__ col_row;
    const int* __restrict__ col_val;
    const int* __restrict__ col_cum;
    // Current entropy of every nonzero, cell_entropy(M, din[col], dout[row]), computed when the
    // snapshot is built (CSR and CSC order): the "before" term of every cell an evaluation reads.
    const float* __restrict__ row_h = nullptr;
    const float* __restrict__ col_h = nullptr;
    static constexpr bool kSlots = false;  // every entry in range is a nonzero
    static constexpr bool kHasH = true;
    __device__ float rh(int j) const { return row_h[j]; }
    __device__ float ch(int j) const { return col_h[j]; }
    __device__ int rbeg(int x) const { return row_ptr[x]; }
    __device__ int rend(int x) const { return row_ptr[x + 1]; }
    __device__ int rkey(int j) const { return row_col[j]; }
    __device__ int rval(int j) const { return row_val[j]; }
    __device__ int cbeg(int x) const { return col_ptr[x]; }
    __device__ int cend(int x) const { return col_ptr[x + 1]; }
    __device__ int ckey(int j) const { return col_row[j]; }
    __device__ int cval(int j) const { return col_val[j]; }
    __device__ int nnz_row(int x) const { return row_ptr[x + 1] - row_ptr[x]; }
    __device__ int nnz_col(int x) const { return col_ptr[x + 1] - col_ptr[x]; }
    // M[x, c] and M[c, x] (one lane)
    __device__ int rlookup(int x, int c) const { return lookup_sorted(row_col, row_val, row_ptr[x], row_ptr[x + 1], c); }
    __device__ int clookup(int x, int c) const { return lookup_sorted(col_row, col_val, col_ptr[x], col_ptr[x + 1], c); }
    // Block at unit `target` of row t (target < dout[t]) or of column t (target - dout[t]), found
    // by a group of W lanes.
    template <int W = WAVE>
    __device__ int sample(int t, int target, int dout_t) const {
        if (target < dout_t) return row_col[group_upper_bound<W>(row_cum, row_ptr[t], row_ptr[t + 1], target)];
        return col_row[group_upper_bound<W>(col_cum, col_ptr[t], col_ptr[t + 1], target - dout_t)];
    }
};

// Smallest index c (in [0, B)) with sum of weights over keys <= c greater than target, for the
// unordered (key, value) slots [lo, hi): a bisection over the index range, each step one
// wave reduction. Slots beyond 4 per lane are re-read from memory.
__device__ __forceinline__ int wave_weighted_select(const int* keys, const int* vals,
                                                    int lo, int hi, int target, int B) {
    const int lane = lane_id();
    int k[4], w[4];
    const bool in_regs = (hi - lo) <= 4 * WAVE;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int j = lo + q * WAVE + lane;
        bool ok = in_regs && j < hi;
        int kk = ok ? keys[j] : -1;
        int vv = ok ? vals[j] : 0;
        k[q] = kk < 0 ? INT_MAX : kk;
        w[q] = kk < 0 ? 0 : vv;
    }
    int a = 0, b = B - 1;
    while (a < b) {
        int mid = (a + b) >> 1;
        long long part = 0;
        if (in_regs) {
#pragma unroll
            for (int q = 0; q < 4; ++q) part += (k[q] <= mid) ? w[q] : 0;
        } else {
            for (int j = lo + lane; j < hi; j += WAVE) {
                int kk = keys[j];
                if (kk >= 0 && kk <= mid) part += vals[j];
            }
        }
        long long below = wave_sum_ll(part);
        if (below > target) b = mid;
        else a = mid + 1;
    }
    return a;
}

// The same selection by a radix descent over 8-bit digits of the block index, most significant
// first: each pass sums, per digit, the weights of the keys that share the digits chosen so far,
// in an LDS histogram of 256 counters owned by this wavefront, and picks the digit where the
// running sum passes the target. ceil(bits(B - 1) / 8) passes over the slots instead of
// log2(B) for the bisection.
__device__ __forceinline__ int wave_weighted_select_radix(const int* keys, const int* vals, int lo, int hi,
                                                          int target, int B, int* hist) {
    const int lane = lane_id();
    int nbits = 1;
    while (nbits < 31 && (1 << nbits) < B) ++nbits;
    int shift = ((nbits + 7) / 8 - 1) * 8;  // lowest bit of the top digit
    int prefix = 0;                         // digits chosen so far
    long long below = 0;                    // weight of the keys below the chosen prefix
    for (;;) {
        for (int d = lane; d < 256; d += WAVE) hist[d] = 0;
        fence_wave();
        const int up = shift + 8;
        for (int j = lo + lane; j < hi; j += WAVE) {
            const int k = keys[j];
            if (k < 0) continue;
            const int w = vals[j];
            if (w == 0) continue;
            const int hi_part = up >= 31 ? 0 : (k >> up);
            if (hi_part == prefix) atomicAdd(&hist[(k >> shift) & 255], w);
        }
        fence_wave();
        long long h[4];
        long long local = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            h[q] = hist[lane * 4 + q];
            local += h[q];
        }
        long long incl = local;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            long long y = __shfl_up(incl, off, WAVE);
            if (lane >= off) incl += y;
        }
        long long run = below + (incl - local);
        int digit = INT_MAX;
        long long before = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (digit == INT_MAX && run + h[q] > target) {
                digit = lane * 4 + q;
                before = run;
            }
            run += h[q];
        }
        const int d = wave_min(digit);
        if (d == INT_MAX) return B - 1;  // target beyond the total weight: not reached for valid input
        below = __shfl(before, d / 4, WAVE);
        prefix = (prefix << 8) | d;
        if (shift == 0) return prefix;
        shift -= 8;
        fence_wave();
    }
}

// No __restrict__ here: the Metropolis-Hastings kernel writes these arrays between evaluations.
// Row x occupies slots [base[x], base[x] + cap[x]) of the key/value pools (cap a power of two).
struct HashView {
    const int* rbase;
    const int* rcap;
    const int* rkeys;
    const int* rvals;
    const int* cbase;
    const int* ccap;
    const int* ckeys;
    const int* cvals;
    int B;
    int* hist = nullptr;  // 256 LDS counters of this wavefront for the radix selection, or null
    static constexpr bool kSlots = true;  // skip key < 0 and value 0
    static constexpr bool kHasH = false;  // cells change as moves commit: entropies are computed
    __device__ float rh(int) const { return 0.0f; }
    __device__ float ch(int) const { return 0.0f; }
    __device__ int rbeg(int x) const { return rbase[x]; }
    __device__ int rend(int x) const { return rbase[x] + rcap[x]; }
    __device__ int rkey(int j) const { return rkeys[j]; }
    __device__ int rval(int j) const { return rvals[j]; }
    __device__ int cbeg(int x) const { return cbase[x]; }
    __device__ int cend(int x) const { return cbase[x] + ccap[x]; }
    __device__ int ckey(int j) const { return ckeys[j]; }
    __device__ int cval(int j) const { return cvals[j]; }
    __device__ int nnz_row(int x) const { return rcap[x]; }
    __device__ int nnz_col(int x) const { return ccap[x]; }
    static __device__ int probe(const int* base, const int* capa, const int* keys, const int* vals, int x, int c) {
        const int b0 = base[x], cap = capa[x];
        int slot = (int)(hash_block(c) & (uint32_t)(cap - 1));
        for (int p = 0; p < cap; ++p) {
            int k = keys[b0 + slot];
            if (k == c) return vals[b0 + slot];
            if (k == -1) return 0;
            slot = (slot + 1) & (cap - 1);
        }
        return 0;
    }
    __device__ int rlookup(int x, int c) const { return probe(rbase, rcap, rkeys, rvals, x, c); }
    __device__ int clookup(int x, int c) const { return probe(cbase, ccap, ckeys, cvals, x, c); }
    template <int W = WAVE>
    __device__ int sample(int t, int target, int dout_t) const {
        static_assert(W == WAVE, "hash rows are sampled by whole wavefronts");
        const bool row = target < dout_t;
        const int* keys = row ? rkeys : ckeys;
        const int* vals = row ? rvals : cvals;
        const int lo = row ? rbase[t] : cbase[t];
        const int hi = lo + (row ? rcap[t] : ccap[t]);
        const int tg = row ? target : target - dout_t;
        if (hist != nullptr && hi - lo > 4 * WAVE) return wave_weighted_select_radix(keys, vals, lo, hi, tg, B, hist);
        return wave_weighted_select(keys, vals, lo, hi, tg, B);
    }
};

struct GraphView {
    const int* __restrict__ out_ptr;
    const int* __restrict__ out_idx;
    const int* __restrict__ in_ptr;
    const int* __restrict__ in_idx;
    const uint8_t* __restrict__ self;  // 1 where the vertex has a self-loop
};

// No __restrict__: mutable during a Metropolis-Hastings pass.
struct BlockState {
    const int* assign;
    const int* dout;
    const int* din;
    const int* deg;
    int B;
    // Several graphs at once (null for one graph): graph g owns the vertices [gvo[g], gvo[g + 1])
    // and the block ids from gbo[g], of which its first gB[g] are in use; its randomness is keyed
    // by its own seed and call and by vertex ids local to it.
    const int* vgraph = nullptr;  // graph of each vertex
    const int* gvo = nullptr;
    const int* gbo = nullptr;
    const int* gB = nullptr;
    const uint64_t* gseed = nullptr;
    const uint32_t* gcall = nullptr;
};

// The graph-local view of vertex v: block count, first block id, local vertex id, seed, call.
struct LocalKey {
    int B, base, id;
    uint64_t seed;
    uint32_t call;
};
__device__ __forceinline__ LocalKey local_key(const BlockState& S, int v, uint64_t seed, uint32_t call) {
    if (S.vgraph == nullptr) return LocalKey{S.B, 0, v, seed, call};
    const int g = S.vgraph[v];
    return LocalKey{S.gB[g], S.gbo[g], v - S.gvo[g], S.gseed[g], S.gcall[g]};
}

// Wave-uniform result of one evaluation; the neighbour table stays valid for a commit.
struct MoveResult {
    int v, r, s, t;
    int k_out, k_in, self;
    int status;  // 0 no move proposed, 1 evaluated and rejected, 2 accepted, -1 table overflow
    double dS, H, u;
    int m_ss, m_rr, m_sr, m_rs;
    int d_ss, d_rr, d_sr, d_rs;
    int nuniq;
    long long nnz_rs;  // nonzeros read in rows r, s and columns r, s
    int tslots;        // slots of the row or column of t the proposal's selection scans (hash rows)
};

// Insert (or find) block b; returns its slot, or -1 when the table is full.
__device__ __forceinline__ int tab_insert(int* key, int cap, int b) {
    int slot = (int)(hash_block(b) & (uint32_t)(cap - 1));
    for (int probe = 0; probe < cap; ++probe) {
        int old = atomicCAS(&key[slot], -1, b);
        if (old == -1 || old == b) return slot;
        slot = (slot + 1) & (cap - 1);
    }
    return -1;
}
__device__ __forceinline__ int tab_find(const int* key, int cap, int b) {
    int slot = (int)(hash_block(b) & (uint32_t)(cap - 1));
    for (int probe = 0; probe < cap; ++probe) {
        int k = key[slot];
        if (k == b) return slot;
        if (k == -1) return -1;
        slot = (slot + 1) & (cap - 1);
    }
    return -1;
}

__device__ __forceinline__ float fi(int x) { return (float)x; }

// Feature j (< 64) of the learned proposal for a vertex in block `own` with candidates cand[32]
// (blocks, -1 past the last) and their edge counts cnt[32], nuniq distinct neighbour blocks. Block
// degrees are taken without the vertex's own contribution, so the features, and the scores, are
// the same before and after the vertex moves.
__device__ inline float learned_feature(int j, const int* cand, const int* cnt, int k_out, int k_in, int self, int own,
                                        int nuniq, const int* din, const int* dout, float inv_logE,
                                        float log_ratio) {
    const int k = k_out + k_in;
    const float invk = k > 0 ? 1.0f / (float)k : 0.0f;
    if (j < 16) return cand[j] >= 0 ? (float)cnt[j] * invk : 0.0f;
    if (j < 32) {
        const int b = cand[j - 16];
        return b >= 0 ? log1pf((float)(din[b] - (b == own ? k_in + self : 0))) * inv_logE : 0.0f;
    }
    if (j < 48) {
        const int b = cand[j - 32];
        return b >= 0 ? log1pf((float)(dout[b] - (b == own ? k_out : 0))) * inv_logE : 0.0f;
    }
    const float inv_logD = 1.0f / logf(4097.0f);
    switch (j - 48) {
        case 0: return log1pf((float)k_out) * inv_logD;
        case 1: return log1pf((float)k_in) * inv_logD;
        case 2: {
            int rest = 0;
            for (int i = 16; i < MLP_CAND; ++i) rest += cand[i] >= 0 ? cnt[i] : 0;
            return (float)rest * invk;
        }
        case 3: return (float)nuniq * invk;
        case 4: return log_ratio;
        case 5: return 1.0f;
        default: return 0.0f;
    }
}

// Softmax probability of candidate `pick` among the candidates other than `own` (-1 when no
// candidate qualifies; 0 when pick is not one of them).
__device__ inline float cand_prob(const float* z, const int* cand, int own, int pick) {
    float mx = -1e30f;
    int any = 0;
#pragma unroll 1
    for (int j = 0; j < MLP_CAND; ++j)
        if (cand[j] >= 0 && cand[j] != own) {
            mx = fmaxf(mx, z[j]);
            any = 1;
        }
    if (!any) return -1.0f;
    float den = 0.0f, num = 0.0f;
#pragma unroll 1
    for (int j = 0; j < MLP_CAND; ++j)
        if (cand[j] >= 0 && cand[j] != own) {
            const float e = expf(z[j] - mx);
            den += e;
            if (cand[j] == pick) num = e;
        }
    return num / den;
}

// A draw from the softmax of the scores over the candidates other than `own`, walked in slot
// order with the weights cand_prob forms (u01 uniform in [0, 1)); -1 when no candidate qualifies.
__device__ inline int learned_pick(const float* z, const int* cand, int own, double u01) {
    float mx = -1e30f;
#pragma unroll 1
    for (int j = 0; j < MLP_CAND; ++j)
        if (cand[j] >= 0 && cand[j] != own) mx = fmaxf(mx, z[j]);
    float den = 0.0f;
#pragma unroll 1
    for (int j = 0; j < MLP_CAND; ++j)
        if (cand[j] >= 0 && cand[j] != own) den += expf(z[j] - mx);
    const float target = (float)u01 * den;
    float run = 0.0f;
    int pick = -1;
#pragma unroll 1
    for (int j = 0; j < MLP_CAND; ++j)
        if (cand[j] >= 0 && cand[j] != own) {
            run += expf(z[j] - mx);
            pick = cand[j];
            if (run > target) break;
        }
    return pick;
}

// Table layout: 7 arrays of cap ints: key, o, i, M[s,b], M[b,s], M[r,b], M[b,r].
// given_s >= 0 fixes the proposal (with uniform given_u); propose_only stops after step 2.
//
// Latency is the bound (each step waits on a gather), so the steps are arranged to keep the
// chain of dependent loads short: the proposal's random draws come first and the lane that
// gathers the drawn neighbour slot records its block during step 1; the four scans of step 3
// run as one loop over the concatenated ranges, two slots per lane per trip, with the keys,
// then the degrees, loaded for both slots before either is used.
// U: slots per lane per trip of the row scans and the Hastings pass (loads in flight per lane).
// lw >= 0: the proposal was the learned mixture w * plain + (1 - w) * net with w = lw; lqf / lqb
// are the net's forward and reverse probabilities (negative: no learned mass in that state).
// With lz / lcand (the vertex's candidate scores and blocks) they are computed here once the
// proposal is known instead: the net's part is then the softmax of those fixed scores over the
// candidates other than the vertex's current block, in both directions.
// W: lanes per vertex (a group of W consecutive lanes; 64 = the whole wavefront).
template <class View, int U = 2, int W = WAVE>
__device__ __forceinline__ MoveResult eval_move(const GraphView& G, const BlockState& S, const View& M, int v,
                                                int* tab, int cap, uint64_t seed, uint32_t call, uint32_t tag,
                                                int given_s, double given_u, bool propose_only, double lw = -1.0,
                                                double lqf = 0.0, double lqb = 0.0, const float* lz = nullptr,
                                                const int* lcand = nullptr) {
    const int lane = group_lane<W>();
    MoveResult R{};
    R.v = v;
    const int r = S.assign[v];
    R.r = r;
    R.t = -1;
    const LocalKey lk = local_key(S, v, seed, call);
    const int B = lk.B;  // blocks of this vertex's graph (Hastings smoothing, random block)
    int* tkey = tab;
    int* to = tab + cap;
    int* ti = tab + 2 * cap;
    int* tmsb = tab + 3 * cap;
    int* tmbs = tab + 4 * cap;
    int* tmrb = tab + 5 * cap;
    int* tmbr = tab + 6 * cap;
    for (int j = lane; j < cap; j += W) {
        tkey[j] = -1;
        to[j] = 0;
        ti[j] = 0;
        tmsb[j] = 0;
        tmbs[j] = 0;
        tmrb[j] = 0;
        tmbr[j] = 0;
    }
    const int ob = G.out_ptr[v], dO = G.out_ptr[v + 1] - ob;
    const int ib = G.in_ptr[v], dI = G.in_ptr[v + 1] - ib;
    const int self = G.self[v];
    const int k_out = dO;
    const int k_in = dI - self;
    const int k = k_out + k_in;
    R.k_out = k_out;
    R.k_in = k_in;
    R.self = self;

    // proposal draws (common.cpp:131-182): neighbour slot ps among the out list (self included)
    // and the in list (self excluded)
    Draw4 dr{0u, 0u, 0u, 0u};
    double u = given_u;
    int pick0 = -1, pick1 = -1;  // combined-list positions whose block may be t
    if (given_s < 0) {
        dr = philox4x32_10(lk.seed, (uint32_t)lk.id, lk.call, 0u, tag);
        u = unit53(dr.x2, dr.x3);
        if (k > 0) {
            const int ps = (int)mulhi32(dr.x0, (uint32_t)k);
            if (ps < k_out) {
                pick0 = ps;
            } else {
                pick0 = dO + (ps - k_out);  // in-list position when the self-loop comes later
                pick1 = pick0 + 1;          // ... and when it comes at or before it
            }
        }
    }
    fence_wave();

    // 1. neighbour blocks over the concatenated out and in lists
    int self_pos = INT_MAX, full = 0, cand0 = -1, cand1 = -1;
    for (int j = lane; j < dO + dI; j += W) {
        const bool is_out = j < dO;
        const int w = is_out ? G.out_idx[ob + j] : G.in_idx[ib + (j - dO)];
        const int bw = S.assign[w];
        if (j == pick0) cand0 = bw;
        if (j == pick1) cand1 = bw;
        if (w == v) {
            if (!is_out) self_pos = j - dO;
        } else {
            int slot = tab_insert(tkey, cap, bw);
            if (slot < 0) full = 1;
            else atomicAdd(is_out ? &to[slot] : &ti[slot], 1);
        }
    }
    self_pos = group_min<W>(self_pos);
    cand0 = group_max<W>(cand0);
    cand1 = group_max<W>(cand1);
    if (self && lane == 0) {
        if (tab_insert(tkey, cap, r) < 0) full = 1;
    }
    full = group_or<W>(full);
    fence_wave();
    if (full) {
        R.status = -1;
        return R;
    }

    // 2. proposal
    int s;
    if (given_s >= 0) {
        s = given_s;
    } else if (k == 0) {
        s = lk.base + random_block(r - lk.base, B, dr.x0);
    } else {
        int t = cand0;
        if (pick1 >= 0 && pick0 - dO >= self_pos) t = cand1;  // in list: skip over the self-loop
        R.t = t;
        const int trow = S.dout[t];
        const int total = trow + S.din[t];
        const int target = (int)mulhi32(dr.x1, (uint32_t)total);
        s = M.template sample<W>(t, target, trow);
        if constexpr (View::kSlots) R.tslots = target < trow ? M.nnz_row(t) : M.nnz_col(t);
    }
    R.s = s;
    R.u = u;
    if (lw >= 0.0 && lz != nullptr) {
        lqf = cand_prob(lz, lcand, r, s);
        lqb = cand_prob(lz, lcand, s, r);
    }
    R.H = 1.0;
    R.dS = 0.0;
    if (propose_only || s == r) {
        R.status = 0;
        return R;
    }

    // 3. one pass over rows s, r and columns s, r
    const int slot_r = tab_find(tkey, cap, r), slot_s = tab_find(tkey, cap, s);
    const int o_r = slot_r >= 0 ? to[slot_r] : 0, i_r = slot_r >= 0 ? ti[slot_r] : 0;
    const int o_s = slot_s >= 0 ? to[slot_s] : 0, i_s = slot_s >= 0 ? ti[slot_s] : 0;
    const int dout_r = S.dout[r], dout_s = S.dout[s], din_r = S.din[r], din_s = S.din[s];
    const int ndout_r = dout_r - k_out, ndout_s = dout_s + k_out;
    const int ndin_r = din_r - (k_in + self), ndin_s = din_s + (k_in + self);
    const bool t_ss = o_s > 0 || i_s > 0 || self, t_rr = o_r > 0 || i_r > 0 || self;
    const bool t_sr = o_r > 0 || i_s > 0, t_rs = o_s > 0 || i_r > 0;
    const int d_ss = o_s + i_s + self, d_rr = -(o_r + i_r + self);
    const int d_sr = o_r - i_s, d_rs = i_r - o_s;
    int m_ss = 0, m_rr = 0, m_sr = 0, m_rs = 0;
    double acc = 0.0;

    const int b0 = M.rbeg(s), L0 = M.rend(s) - b0;  // row s
    const int b1 = M.rbeg(r), L1 = M.rend(r) - b1;  // row r
    const int b2 = M.cbeg(s), L2 = M.cend(s) - b2;  // column s
    const int b3 = M.cbeg(r), L3 = M.cend(r) - b3;  // column r
    const int T1 = L0 + L1, T2 = T1 + L2, T = T2 + L3;
    long long live = 0;
    for (int base = 0; base < T; base += U * W) {
        int key[U], val[U], seg[U];
        float h[U];  // the cell's current entropy, when the view keeps it
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int i = base + q * W + lane;
            seg[q] = i < L0 ? 0 : i < T1 ? 1 : i < T2 ? 2 : i < T ? 3 : -1;
            key[q] = -1;
            val[q] = 0;
            h[q] = 0.0f;
            if (seg[q] == 0) {
                key[q] = M.rkey(b0 + i);
                val[q] = M.rval(b0 + i);
                if constexpr (View::kHasH) h[q] = M.rh(b0 + i);
            } else if (seg[q] == 1) {
                key[q] = M.rkey(b1 + (i - L0));
                val[q] = M.rval(b1 + (i - L0));
                if constexpr (View::kHasH) h[q] = M.rh(b1 + (i - L0));
            } else if (seg[q] == 2) {
                key[q] = M.ckey(b2 + (i - T1));
                val[q] = M.cval(b2 + (i - T1));
                if constexpr (View::kHasH) h[q] = M.ch(b2 + (i - T1));
            } else if (seg[q] == 3) {
                key[q] = M.ckey(b3 + (i - T2));
                val[q] = M.cval(b3 + (i - T2));
                if constexpr (View::kHasH) h[q] = M.ch(b3 + (i - T2));
            }
        }
        int dsel[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            dsel[q] = -1;
            const int c = key[q], m = val[q];
            if (seg[q] < 0 || (View::kSlots && (c < 0 || m == 0))) continue;
            ++live;
            bool unchanged = true;
            if (seg[q] == 0) {  // row s: cell (s, c)
                if (c == s) {
                    m_ss = m;
                    unchanged = !t_ss;
                } else if (c == r) {
                    m_sr = m;
                    unchanged = !t_sr || d_sr == 0;
                } else {
                    const int sl = tab_find(tkey, cap, c);
                    if (sl >= 0) {
                        tmsb[sl] = m;
                        unchanged = to[sl] == 0;
                        if (View::kHasH && !unchanged) acc += (double)h[q];  // a Delta key's "before" term
                    }
                }
            } else if (seg[q] == 1) {  // row r: cell (r, c)
                if (c == r) {
                    m_rr = m;
                    unchanged = !t_rr;
                } else if (c == s) {
                    m_rs = m;
                    unchanged = !t_rs || d_rs == 0;
                } else {
                    const int sl = tab_find(tkey, cap, c);
                    if (sl >= 0) {
                        tmrb[sl] = m;
                        unchanged = to[sl] == 0;
                        if (View::kHasH && !unchanged) acc += (double)h[q];
                    }
                }
            } else {  // column s or r: cell (c, s) or (c, r), rows r and s left to the rows
                if (c == r || c == s) continue;
                const int sl = tab_find(tkey, cap, c);
                if (sl >= 0) {
                    (seg[q] == 2 ? tmbs : tmbr)[sl] = m;
                    unchanged = ti[sl] == 0;
                    if (View::kHasH && !unchanged) acc += (double)h[q];
                }
            }
            if (unchanged) dsel[q] = c;
        }
        int dv[U];
#pragma unroll
        for (int q = 0; q < U; ++q) dv[q] = dsel[q] >= 0 ? (seg[q] < 2 ? S.din[dsel[q]] : S.dout[dsel[q]]) : 0;
#pragma unroll
        for (int q = 0; q < U; ++q) {
            if (dsel[q] < 0) continue;
            const int c = dsel[q];
            const float m = fi(val[q]);
            const int d = dv[q];
            if (seg[q] < 2) {  // unchanged cell of row x in {s, r}: din[c] may change, dout[x] does
                const int xo = seg[q] == 0 ? dout_s : dout_r, nxo = seg[q] == 0 ? ndout_s : ndout_r;
                const int nd = c == r ? ndin_r : (c == s ? ndin_s : d);
                acc += View::kHasH ? (double)h[q] : (double)cell_entropy(m, fi(d), fi(xo));
                acc -= (double)cell_entropy(m, fi(nd), fi(nxo));
            } else {  // unchanged cell of column x: din[x] changes
                const int xi = seg[q] == 2 ? din_s : din_r, nxi = seg[q] == 2 ? ndin_s : ndin_r;
                acc += View::kHasH ? (double)h[q] : (double)cell_entropy(m, fi(xi), fi(d));
                acc -= (double)cell_entropy(m, fi(nxi), fi(d));
            }
        }
    }
    m_ss = group_max<W>(m_ss);
    m_rr = group_max<W>(m_rr);
    m_sr = group_max<W>(m_sr);
    m_rs = group_max<W>(m_rs);
    fence_wave();

    // 4. Delta keys and Hastings sums. With the learned mixture also pf0 and pb0: the plain
    // sampler's own probabilities of s now and of r after the move, times k (the reference's
    // closed form smooths its weights by 1 and B for a random-block branch the sampler never
    // takes, so it cannot be mixed with the network's probabilities)
    const bool mix = lw >= 0.0;
    double pf = 0.0, pb = 0.0, pf0 = 0.0, pb0 = 0.0;
    int nuniq = 0;
    for (int j0 = 0; j0 < cap; j0 += U * W) {
        int bk[U], degb[U], dinb[U], doutb[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int j = j0 + q * W + lane;
            bk[q] = j < cap ? tkey[j] : -1;
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int b = bk[q];
            const bool other = b >= 0 && b != r && b != s;
            degb[q] = b >= 0 ? S.deg[b] : 0;
            dinb[q] = other ? S.din[b] : 0;
            doutb[q] = other ? S.dout[b] : 0;
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int b = bk[q];
            if (b < 0) continue;
            const int j = j0 + q * W + lane;
            nuniq++;
            const int ob_ = to[j], ib_ = ti[j];
            int Msb, Mbs, Mrb, Mbr, Drb, Dbr, nd;
            if (b == r) {
                Msb = m_sr; Mbs = m_rs; Mrb = m_rr; Mbr = m_rr; Drb = d_rr; Dbr = d_rr;
                nd = (dout_r - k_out) + (din_r - k_in) - (m_rr + d_rr);
            } else if (b == s) {
                Msb = m_ss; Mbs = m_ss; Mrb = m_rs; Mbr = m_sr; Drb = d_rs; Dbr = d_sr;
                nd = (dout_s + k_out) + (din_s + k_in) - (m_ss + d_ss);
            } else {
                Msb = tmsb[j]; Mbs = tmbs[j]; Mrb = tmrb[j]; Mbr = tmbr[j]; Drb = -ob_; Dbr = -ib_;
                nd = degb[q];
                const int db = dinb[q], ub = doutb[q];
                // "before" terms of the Delta keys: added by the scan when the view keeps them
                if (ob_ > 0) {
                    if constexpr (!View::kHasH) acc += (double)cell_entropy(fi(Msb), fi(db), fi(dout_s));
                    acc -= (double)cell_entropy(fi(Msb + ob_), fi(db), fi(ndout_s));
                    if constexpr (!View::kHasH) acc += (double)cell_entropy(fi(Mrb), fi(db), fi(dout_r));
                    acc -= (double)cell_entropy(fi(Mrb - ob_), fi(db), fi(ndout_r));
                }
                if (ib_ > 0) {
                    if constexpr (!View::kHasH) acc += (double)cell_entropy(fi(Mbs), fi(din_s), fi(ub));
                    acc -= (double)cell_entropy(fi(Mbs + ib_), fi(ndin_s), fi(ub));
                    if constexpr (!View::kHasH) acc += (double)cell_entropy(fi(Mbr), fi(din_r), fi(ub));
                    acc -= (double)cell_entropy(fi(Mbr - ib_), fi(ndin_r), fi(ub));
                }
            }
            if (mix && ob_ + ib_ > 0) {  // v's neighbour slots in b (the self-loop's slot: below)
                const int tf = b == r ? dout_r + din_r : (b == s ? dout_s + din_s : doutb[q] + dinb[q]);
                const int tb = b == r ? ndout_r + ndin_r : (b == s ? ndout_s + ndin_s : doutb[q] + dinb[q]);
                if (tf > 0) pf0 += (double)(ob_ + ib_) * (double)(Msb + Mbs) / (double)tf;
                if (tb > 0) pb0 += (double)(ob_ + ib_) * (double)(Mrb + Drb + Mbr + Dbr) / (double)tb;
            }
            const int nb = ob_ + ib_ + (b == r ? self : 0);
            if (nb == 0) continue;
            const double count = (double)nb;
            const double pw = (double)(Msb + Mbs) + 1.0;
            const double bd = (double)(degb[q] + B);
            const double bw = (double)(Mrb + Drb + Mbr + Dbr) + 1.0;
            const double pd = (double)(nd + B);
            pf += count * pw / bd;
            pb += count * bw / pd;
        }
    }
    if (lane == 0) {
        if (t_ss) {
            acc += (double)cell_entropy(fi(m_ss), fi(din_s), fi(dout_s));
            acc -= (double)cell_entropy(fi(m_ss + d_ss), fi(ndin_s), fi(ndout_s));
        }
        if (t_rr) {
            acc += (double)cell_entropy(fi(m_rr), fi(din_r), fi(dout_r));
            acc -= (double)cell_entropy(fi(m_rr + d_rr), fi(ndin_r), fi(ndout_r));
        }
        if (t_sr) {
            acc += (double)cell_entropy(fi(m_sr), fi(din_r), fi(dout_s));
            acc -= (double)cell_entropy(fi(m_sr + d_sr), fi(ndin_r), fi(ndout_s));
        }
        if (t_rs) {
            acc += (double)cell_entropy(fi(m_rs), fi(din_s), fi(dout_r));
            acc -= (double)cell_entropy(fi(m_rs + d_rs), fi(ndin_s), fi(ndout_r));
        }
    }
    if (mix && self && lane == 0) {  // the self-loop's slot: block r now, block s after the move
        const int tf = dout_r + din_r, tb = ndout_s + ndin_s;
        if (tf > 0) pf0 += (double)(m_sr + m_rs) / (double)tf;
        if (tb > 0) pb0 += (double)((m_rs + d_rs) + (m_sr + d_sr)) / (double)tb;
    }
    const double dS = group_sum<W>(acc);
    pf = group_sum<W>(pf);
    pb = group_sum<W>(pb);
    R.nuniq = (int)group_sum_ll<W>(nuniq);
    double H = (k == 0) ? 1.0 : pb / pf;
    if (mix && k > 0) {
        // mixture proposal w q_plain + (1 - w) q_net, both the samplers' own probabilities
        pf0 = group_sum<W>(pf0);
        pb0 = group_sum<W>(pb0);
        const double plain_f = pf0 / (double)k, plain_b = pb0 / (double)k;
        const double fwd = lw * plain_f + (1.0 - lw) * (lqf >= 0.0 ? lqf : plain_f);
        const double bwd = lw * plain_b + (1.0 - lw) * (lqb >= 0.0 ? lqb : plain_b);
        H = bwd / fwd;
    }
    double a = exp(-3.0 * dS) * H;
    a = (a >= 1.0) ? 1.0 : a;
    R.dS = dS;
    R.H = H;
    R.status = (u <= a) ? 2 : 1;
    R.m_ss = m_ss;
    R.m_rr = m_rr;
    R.m_sr = m_sr;
    R.m_rs = m_rs;
    R.d_ss = d_ss;
    R.d_rr = d_rr;
    R.d_sr = d_sr;
    R.d_rs = d_rs;
    R.nnz_rs = View::kSlots ? group_sum_ll<W>(live) : (long long)T;
    return R;
}

}  // namespace sbpx
