// Sequential Metropolis-Hastings (finetune.cpp:446-461, 488-501, used by metropolis_hastings
// :412-444 and by hybrid_mcmc for the high-degree vertices :358-363): each accepted move changes
// the block matrix before the next vertex is proposed.
//
// The block matrix is held, for the duration of the pass, as a per-row open-addressing table in
// both orientations (HashView), built from the CSR snapshot. A window of queue entries is
// evaluated speculatively against the current state, one wavefront (or a group of them) per
// entry, and walked in queue order: a rejected entry is final, an accepted one is committed, and
// a later entry is kept only if no commit before it changed anything its evaluation read. The
// first entry that is not kept ends the window; the next window starts at it. The outcome is
// that of evaluating one vertex at a time, because proposals and uniforms are keyed by vertex,
// not by evaluation order. mh_kernel and mh_group_kernel re-check the later entries after every
// commit (one workgroup, 16 entries per window); mh_wide_kernel checks every entry against the
// accepted entries before it at once, over K workgroups (8 K entries per window, K = 1 too).
// Before a commit, every row and column it may extend is given room: a row that is too full moves
// to a larger segment taken from the free end of the pool (the other wavefronts wait at the
// barrier meanwhile). Only when the pool itself runs out does the kernel stop; the host then
// rebuilds the tables with a larger pool and the pass resumes at the same vertex.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <cstdint>

#include "eval_core.hpp"
#include "internal.hpp"
#include "eval_group.hpp"
#include "merge_core.hpp"

namespace sbpx {

constexpr int MH_WAVES = 16;
constexpr int MH_CHUNK = 1024;  // queue entries per launch
#ifndef MH_UNROLL
#define MH_UNROLL 2  // slots per lane per trip of the row scans (4 spills at 128 VGPRs)
#endif

// ------------------------------------------------------------------------------------------
// Table build: cap(x) = next power of two >= 1.5 nnz(x) + 32, filled from the CSR. Scans walk
// every slot of a row, so the tables are kept dense (a load factor of 1/3 to 2/3 after a build;
// a row is moved to a larger segment before a commit would take it past 3/4).
__device__ __forceinline__ int table_cap(int keys) {
    int want = (3 * keys) / 2 + 32;
    int cap = 32;
    while (cap < want) cap <<= 1;
    return cap;
}

__global__ void hash_caps_kernel(const int* __restrict__ ptr, int B, int* __restrict__ caps, int* max_cap) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= B) return;
    const int cap = table_cap(ptr[x + 1] - ptr[x]);
    caps[x] = cap;
    atomicMax(max_cap, cap);
}

// One wavefront per row: clear the row's slots, then insert its CSR entries.
__global__ void hash_fill_kernel(const int* __restrict__ ptr, const int* __restrict__ idx, const int* __restrict__ val,
                                 int B, const int* __restrict__ base, int* keys, int* vals, int* occ) {
    const int x = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (x >= B) return;
    const int lane = lane_id();
    const int b0 = base[x], cap = base[x + 1] - b0;
    for (int j = lane; j < cap; j += WAVE) {
        keys[b0 + j] = -1;
        vals[b0 + j] = 0;
    }
    fence_wave();
    for (int j = ptr[x] + lane; j < ptr[x + 1]; j += WAVE) {
        int key = idx[j];
        int slot = (int)(hash_block(key) & (uint32_t)(cap - 1));
        for (;;) {
            int old = atomicCAS(&keys[b0 + slot], -1, key);
            if (old == -1) break;
            slot = (slot + 1) & (cap - 1);
        }
        vals[b0 + slot] = val[j];
    }
    if (lane == 0) occ[x] = ptr[x + 1] - ptr[x];
}

// ------------------------------------------------------------------------------------------
struct WideSlot;

// The proposal network as the sequential pass reads it (see learned_mlp_group).
struct LearnedNet {
    const __bf16* w1;   // [256][64], output-major (the batched products' layout)
    const __bf16* w2;   // [256][256]
    const __bf16* w3;   // [32][256]
    const float* b1;
    const float* b2;
    const float* b3;
    float inv_logE;     // 1 / log1p(E)
    float log_ratio;    // log(B) / log(N)
    int maxdeg;         // vertices of larger degree (out + in) keep the plain proposal
};

struct MhParams {
    GraphView G;
    // mutable state, read and written through the same (non-restrict) pointers
    int* assign;
    int* dout;
    int* din;
    int* deg;
    int* drift;
    int B;
    int* rbase;
    int* rcap;
    int* rkeys;
    int* rvals;
    int* rocc;
    int* cbase;
    int* ccap;
    int* ckeys;
    int* cvals;
    int* cocc;
    int pool;  // slots in each key/value array
    const int* __restrict__ list;
    int n;
    int start;
    uint64_t seed;
    uint32_t call;
    uint32_t tag;
    int* tables;
    int tcap;
    int* status;  // [0] next queue position, [1] 1 if the pool ran out, [2] / [3] row / column pool
                  // top, [4] commits so far (the touch serial)
    long long* counters;  // attempts, evals, moves
    long long* bytes;     // algorithmic bytes of the evaluations that stand, or null
    int* touch;           // [N] serial of the last commit that moved a neighbour of the vertex
    double* scan;         // [3] slots scanned in rows/columns r and s, and evaluations, that stand;
                          // windows evaluated
    int stop_at;          // queue position where this launch ends
    double* dS_sum;
    int* err;
    // the pass over several workgroups (mh_wide_kernel)
    WideSlot* wide;                 // [8 K] one record per window entry
    int* wst;                       // [8 K] the entries' evaluation status, dense
    unsigned long long* wstamp;     // [N] (window serial << 32 | entry) of the accepted entries
    int* ctrl;                      // [0] barrier arrivals, [1] first entry that does not hold,
                                    // [2] next queue position, [3] stop reason
    // workgroups per window entry (mh_wide_kernel<8, *, true>), their histograms, partial sums and
    // arrival counters
    int team;
    int* team_hist;
    double* team_part;
    int* team_arrivals;
    // the learned mixture (LEARNED kernels only): the network and the degree bound
    LearnedNet net;
    // optional phase clock of mh_wide_kernel (workgroup 0, 100 MHz ticks): evaluation and first
    // barrier, checks and second barrier, commits, third barrier; then the number of commits
    double* phase;
};

__device__ __forceinline__ int room(const int* capa, const int* occ, int x) { return (capa[x] * 3) / 4 - occ[x]; }

// Adds d to cell (x, key) of one orientation, claiming a slot when the key is new.
__device__ __forceinline__ void upsert(const int* base, const int* capa, int* keys, int* vals, int* occ, int x, int key,
                                       int d) {
    const int b0 = base[x], cap = capa[x];
    int slot = (int)(hash_block(key) & (uint32_t)(cap - 1));
    for (int probe = 0; probe < cap; ++probe) {
        int k = keys[b0 + slot];
        if (k == key) {
            atomicAdd(&vals[b0 + slot], d);
            return;
        }
        if (k == -1) {
            int old = atomicCAS(&keys[b0 + slot], -1, key);
            if (old == -1) {
                atomicAdd(&occ[x], 1);
                atomicAdd(&vals[b0 + slot], d);
                return;
            }
            if (old == key) {
                atomicAdd(&vals[b0 + slot], d);
                return;
            }
        }
        slot = (slot + 1) & (cap - 1);
    }
}

__device__ __forceinline__ void cell_add(const MhParams& P, int row, int col, int d) {
    upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, row, col, d);
    upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, col, row, d);
}

// Moves row x of one orientation to a new segment from the pool's free end, sized for its live
// keys plus `extra`, dropping keys whose count fell to zero (lookups read 0 for them either
// way). The whole wavefront takes part. Returns false, leaving the row where it was, when the
// pool has no segment of that size left.
__device__ bool relocate(int* base, int* capa, int* keys, int* vals, int* occ, int* top, int pool, int x, int extra) {
    const int lane = lane_id();
    const int b0 = base[x], ocap = capa[x];
    const int ncap = table_cap(occ[x] + extra);
    int nb = 0;
    if (lane == 0) nb = atomicAdd(top, ncap);
    nb = __shfl(nb, 0, WAVE);
    if (nb > pool - ncap) return false;
    for (int j = lane; j < ncap; j += WAVE) {
        keys[nb + j] = -1;
        vals[nb + j] = 0;
    }
    fence_wave();
    int live = 0;
    for (int j = lane; j < ocap; j += WAVE) {
        const int k = keys[b0 + j], m = vals[b0 + j];
        if (k < 0 || m == 0) continue;
        int slot = (int)(hash_block(k) & (uint32_t)(ncap - 1));
        for (;;) {
            int old = atomicCAS(&keys[nb + slot], -1, k);
            if (old == -1) break;
            slot = (slot + 1) & (ncap - 1);
        }
        vals[nb + slot] = m;
        ++live;
    }
    live = (int)wave_sum_ll(live);
    fence_wave();
    if (lane == 0) {
        base[x] = nb;
        capa[x] = ncap;
        occ[x] = live;
    }
    fence_wave();
    return true;
}

// Gives every row and column a commit may extend the room it needs: row s and column s one cell
// per neighbour block plus the four cells among {r, s}; rows/columns r four; every other
// neighbour block one cell in its row and one in its column (two kept, for slack). Returns false
// when the pool ran out.
// check_blocks = false skips the neighbour blocks (blocks_need_room found none short of room).
__device__ bool make_room(const MhParams& P, const MoveResult& R, const int* tab, int cap, bool check_blocks) {
    const int lane = lane_id();
    const int r = R.r, s = R.s;
    const int big = R.nuniq + 4;
    int* rtop = P.status + 2;
    int* ctop = P.status + 3;
    if (room(P.rcap, P.rocc, s) < big && !relocate(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, rtop, P.pool, s, big))
        return false;
    if (room(P.ccap, P.cocc, s) < big && !relocate(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, ctop, P.pool, s, big))
        return false;
    if (room(P.rcap, P.rocc, r) < 4 && !relocate(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, rtop, P.pool, r, 4))
        return false;
    if (room(P.ccap, P.cocc, r) < 4 && !relocate(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, ctop, P.pool, r, 4))
        return false;
    if (!check_blocks) return true;
    for (int j0 = 0; j0 < cap; j0 += WAVE) {
        const int j = j0 + lane;
        const int b = j < cap ? tab[j] : -1;
        const bool mine = b >= 0 && b != r && b != s;
        unsigned long long mr = __ballot(mine && room(P.rcap, P.rocc, b) < 2);
        unsigned long long mc = __ballot(mine && room(P.ccap, P.cocc, b) < 2);
        while (mr) {
            const int l = first_lane(mr);
            mr &= mr - 1;
            const int x = __shfl(b, l, WAVE);
            if (!relocate(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, rtop, P.pool, x, 2)) return false;
        }
        while (mc) {
            const int l = first_lane(mc);
            mc &= mc - 1;
            const int x = __shfl(b, l, WAVE);
            if (!relocate(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, ctop, P.pool, x, 2)) return false;
        }
    }
    return true;
}

// Whether any neighbour block b (other than r and s) of a move's table lacks the room make_room
// keeps in its row or column. The whole workgroup looks, one slot per lane, so that the committing
// wavefront walks the table (relocating, in slot order) only when one does.
__device__ bool blocks_need_room(const MhParams& P, const int* tab, int cap, int r, int s, int* s_need) {
    if (threadIdx.x == 0) *s_need = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < cap; j += blockDim.x) {
        const int b = tab[j];
        if (b >= 0 && b != r && b != s && (room(P.rcap, P.rocc, b) < 2 || room(P.ccap, P.cocc, b) < 2)) *s_need = 1;
    }
    __syncthreads();
    return *s_need != 0;
}

// The fields of an accepted move its commit needs, shared through LDS with every wavefront.
struct CommitInfo {
    int v, r, s, k_out, k_in, self, d_ss, d_rr, d_sr, d_rs, m_rr, m_ss;
};
__device__ __forceinline__ CommitInfo commit_info(const MoveResult& R) {
    return CommitInfo{R.v, R.r, R.s, R.k_out, R.k_in, R.self, R.d_ss, R.d_rr, R.d_sr, R.d_rs, R.m_rr, R.m_ss};
}

// Applies an accepted move with every wavefront of the workgroup: block matrix cells from
// (o_b, i_b, self), block degrees with the update rule of Blockmodel::move_vertex
// (blockmodel.cpp:344-360), the assignment, and the neighbours' touch stamps. Each upsert is a
// chain of dependent global reads and atomics, so the updates are spread over lanes rather than
// run one after another. Row s and column s are the only tables that can take several new keys
// in one commit: wavefront 0 fills row s and wavefront 1 column s, taking the neighbour blocks in
// slot order with one lane per block and then the cells among {r, s}, so the keys land in the
// same slots on every run and later scans see the same order. Every other update is to a key that
// exists or is the only new key of its table, and is spread over the remaining wavefronts. The
// committing wavefront's table `tab` is in LDS or global memory; make_room has run.
constexpr int COMMIT_LIST_MAX = 2048;  // neighbour-block table slots a commit compacts in LDS

// LDS a commit works in: the occupied slots of the committing table (other than r and s) in slot
// order, and per-wavefront counts for the compaction.
struct CommitLds {
    int list[COMMIT_LIST_MAX];
    int wcount[MH_WAVES];
};

// NWV: wavefronts of the workgroup (at least 4).
template <int NWV = MH_WAVES>
__device__ void commit_shared(const MhParams& P, const CommitInfo& C, const int* tab, int cap, int serial,
                              CommitLds& L) {
    static_assert(NWV >= 4 && NWV <= MH_WAVES, "commit_shared needs 4 to 16 wavefronts");
    constexpr int NT = NWV * WAVE;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const int r = C.r, s = C.s;
    const int* to = tab + cap;
    const int* ti = tab + 2 * cap;
    {  // stamp the neighbours, whose neighbour-block tables this move changes
        const int v = C.v;
        const int ob = P.G.out_ptr[v], oe = P.G.out_ptr[v + 1];
        const int ib = P.G.in_ptr[v], ie = P.G.in_ptr[v + 1];
        for (int j = ob + tid; j < oe; j += NT) P.touch[P.G.out_idx[j]] = serial;
        for (int j = ib + tid; j < ie; j += NT) P.touch[P.G.in_idx[j]] = serial;
    }
    // The neighbour blocks b (other than r and s), in slot order: compacted into LDS when the
    // table fits, so that the loops below run over the blocks rather than over every slot.
    const bool compact = cap <= COMMIT_LIST_MAX;
    int n = cap;
    if (compact) {
        int base = 0;
        for (int j0 = 0; j0 < cap; j0 += NT) {
            const int j = j0 + tid;
            const int b = j < cap ? tab[j] : -1;
            const bool keep = b >= 0 && b != r && b != s;
            const unsigned long long m = __ballot(keep);
            if (lane == 0) L.wcount[wave] = __popcll(m);
            __syncthreads();
            int off = base, total = base;
            for (int w = 0; w < NWV; ++w) {
                if (w < wave) off += L.wcount[w];
                total += L.wcount[w];
            }
            if (keep) L.list[off + __popcll(m & ((1ull << lane) - 1))] = j;
            __syncthreads();
            base = total;
        }
        n = base;
    }
    // slot of the k-th neighbour block (or -1 for a slot to skip when not compacted)
    auto slot_of = [&](int k) -> int {
        if (compact) return L.list[k];
        const int b = tab[k];
        return (b < 0 || b == r || b == s) ? -1 : k;
    };
    if (wave == 0) {  // row s: (s, b) += o_b, then (s, s) and (s, r)
        for (int k = lane; k < n; k += WAVE) {
            const int j = slot_of(k);
            if (j < 0) continue;
            const int o = to[j];
            if (o) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, s, tab[j], o);
        }
        if (lane == 0) {
            if (C.d_ss) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, s, s, C.d_ss);
            if (C.d_sr) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, s, r, C.d_sr);
        }
    } else if (wave == 1) {  // column s: (b, s) += i_b, then (s, s) and (r, s)
        for (int k = lane; k < n; k += WAVE) {
            const int j = slot_of(k);
            if (j < 0) continue;
            const int i = ti[j];
            if (i) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, s, tab[j], i);
        }
        if (lane == 0) {
            if (C.d_ss) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, s, s, C.d_ss);
            if (C.d_rs) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, s, r, C.d_rs);
        }
    } else if (wave == 2) {  // row r and column r at the diagonal and at s; degrees; assignment
        if (lane == 0) {
            if (C.d_rr) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, r, r, C.d_rr);
            if (C.d_rs) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, r, s, C.d_rs);
        } else if (lane == 1) {
            if (C.d_rr) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, r, r, C.d_rr);
            if (C.d_sr) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, r, s, C.d_sr);
        } else if (lane == 2) {
            P.assign[C.v] = s;
            const int nout_r = P.dout[r] - C.k_out, nout_s = P.dout[s] + C.k_out;
            const int nin_r = P.din[r] - (C.k_in + C.self), nin_s = P.din[s] + (C.k_in + C.self);
            P.dout[r] = nout_r;
            P.dout[s] = nout_s;
            P.din[r] = nin_r;
            P.din[s] = nin_s;
            // self edges read after the update, plus the Delta diagonal (blockmodel.cpp:348-359)
            P.deg[r] = nout_r + nin_r - ((C.m_rr + C.d_rr) + C.d_rr);
            P.deg[s] = nout_s + nin_s - ((C.m_ss + C.d_ss) + C.d_ss);
            P.drift[r] = C.d_rr;
            P.drift[s] = C.d_ss;
        }
    } else {  // the other six updates per neighbour block b
        constexpr int NW = (NWV - 3) * WAVE;
        for (int it = tid - 3 * WAVE; it < 6 * n; it += NW) {
            const int k = it / n, q = it - k * n;
            const int j = slot_of(q);
            if (j < 0) continue;
            const int b = tab[j];
            const int o = to[j], i = ti[j];
            switch (k) {
                case 0: if (o) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, b, s, o); break;   // (s, b)
                case 1: if (o) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, r, b, -o); break;  // (r, b)
                case 2: if (o) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, b, r, -o); break;  // (r, b)
                case 3: if (i) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, b, s, i); break;   // (b, s)
                case 4: if (i) upsert(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, b, r, -i); break;  // (b, r)
                default: if (i) upsert(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, r, b, -i); break; // (b, r)
            }
        }
    }
}

constexpr int MH_LDS_CAP = 256;  // neighbour-block table slots per wavefront held in LDS

// Table slots for vertex v: a power of two at least twice its neighbour count plus two.
__device__ __forceinline__ int vertex_table_cap(const GraphView& G, int v) {
    const int d = (G.out_ptr[v + 1] - G.out_ptr[v]) + (G.in_ptr[v + 1] - G.in_ptr[v]);
    int cap = 16;
    while (cap < 2 * (d + 2)) cap <<= 1;
    return cap;
}

// ------------------------------------------------------------------------------------------
// The learned mixture in the sequential pass. Every step's proposal must be a function of the
// state the step starts from, so each window scores its entries' candidates on the current state:
// each entry's first wavefront gathers the neighbour blocks, takes the top 32 and forms the 64
// features (learned_features_wave); then the whole workgroup runs the 64 -> 256 -> 256 -> 32
// network on its entries' rows with the matrix-core products of the batched forward pass
// (learned_mlp_group), so that the weights are read once per workgroup and window and the scores
// are the bits the batched path gives the same features. The features do not change when the
// vertex itself moves, so these scores give both the forward and the reverse learned
// probability. Everything they read (the neighbours' blocks, the degrees of the vertex's
// neighbour blocks and of its own block) is covered by the window checks, as the evaluation's own
// reads are.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
constexpr int MLP_ROWS = 16;  // feature rows a workgroup scores at once (one matrix-core tile)

__device__ __forceinline__ unsigned short bf16_bits(float f) {  // nearest BF16, ties to even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Candidates and features of vertex v (block r) in one wavefront. tab (cap slots, 2 * cap ints) is
// scratch for the neighbour-block counts; cand[32] and cnt[32] receive the candidates and their
// edge counts (LDS) and xrow[64] the features in BF16 (LDS). Returns whether some candidate is
// not r.
__device__ bool learned_features_wave(const GraphView& G, const BlockState& S, int v, int r, int* tab, int cap,
                                      const LearnedNet& N, int* cand, int* cnt, unsigned short* xrow) {
    const int lane = lane_id();
    int* key = tab;
    int* count = tab + cap;
    for (int j = lane; j < cap; j += WAVE) {
        key[j] = -1;
        count[j] = 0;
    }
    fence_wave();
    const int ob = G.out_ptr[v], dO = G.out_ptr[v + 1] - ob;
    const int ib = G.in_ptr[v], dI = G.in_ptr[v + 1] - ib;
    const int self = G.self[v];
    for (int j = lane; j < dO + dI; j += WAVE) {
        const int w = j < dO ? G.out_idx[ob + j] : G.in_idx[ib + (j - dO)];
        if (w == v) continue;
        const int slot = tab_insert(key, cap, S.assign[w]);
        if (slot >= 0) atomicAdd(&count[slot], 1);
    }
    fence_wave();
    int nuniq = 0;
    for (int j = lane; j < cap; j += WAVE) nuniq += key[j] >= 0 ? 1 : 0;
    nuniq = (int)wave_sum_ll(nuniq);
    // top 32 by (count desc, block asc); a taken slot's count is negated
    for (int t = 0; t < MLP_CAND; ++t) {
        long long best = -1;
        int bslot = -1;
        for (int j = lane; j < cap; j += WAVE) {
            const int kk = key[j], c = count[j];
            const long long sc = (kk >= 0 && c > 0) ? (((long long)c << 32) | (long long)(0x7FFFFFFF - kk)) : -1;
            if (sc > best) {
                best = sc;
                bslot = j;
            }
        }
        long long wb = best;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const long long o = __shfl_xor(wb, off, WAVE);
            wb = o > wb ? o : wb;
        }
        if (wb < 0) {  // no block left: the remaining candidates are empty
            if (lane == 0)
                for (int u = t; u < MLP_CAND; ++u) {
                    cand[u] = -1;
                    cnt[u] = 0;
                }
            fence_wave();
            break;
        }
        int b = -1, c = 0;
        if (wb >= 0) {
            b = 0x7FFFFFFF - (int)(wb & 0xFFFFFFFF);
            c = (int)(wb >> 32);
            if (best == wb) count[bslot] = -c;  // one slot per block: one lane owns it
        }
        if (lane == 0) {
            cand[t] = b;
            cnt[t] = c;
        }
        fence_wave();
    }
    const int k_out = dO, k_in = dI - self;
    static_assert(MLP_IN == WAVE, "one feature per lane");
    xrow[lane] = bf16_bits(
        learned_feature(lane, cand, cnt, k_out, k_in, self, r, nuniq, S.din, S.dout, N.inv_logE, N.log_ratio));
    fence_wave();
    int any = 0;
    for (int j = 0; j < MLP_CAND; ++j) any |= cand[j] >= 0 && cand[j] != r;
    return any != 0;
}

// One 16 x 16 tile of A W^T (columns n0 .. n0 + 15): A is MLP_ROWS rows of K BF16 in LDS, of which
// the first nrows are used (the others read as zero), W is output-major in global memory. The
// fragments, the k steps and the matrix-core instruction are those of the batched forward product
// (gemm_bf16_kcontig_kernel), so each row's sums are the same bits.
template <int K>
__device__ __forceinline__ f32x4_t mlp_tile(const unsigned short* A, int nrows, const __bf16* W, int n0) {
    const int lane = lane_id();
    const int row = lane & 15, kq = 8 * (lane >> 4);
    const bf16x8_t zero8 = {};
    const __bf16* wp = W + (long)(n0 + row) * K + kq;  // B fragment: column n0 + (lane & 15)
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bf16x8_t a = row < nrows ? *(const bf16x8_t*)(A + row * K + k0 + kq) : zero8;
        const bf16x8_t b = *(const bf16x8_t*)(wp + k0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// The network on feature rows [0, nrows) of x (MLP_ROWS x 64 BF16 in LDS): h1 and h2 (MLP_ROWS x 256
// BF16 in LDS; h2 may share x's space) receive the hidden layers, z[row][32] the scores of the rows
// below nrows. Every wavefront of the workgroup (NW of them) calls it; it ends with the scores
// visible to all of them.
template <int NW>
__device__ void learned_mlp_group(const LearnedNet& N, const unsigned short* x, unsigned short* h1, unsigned short* h2,
                                  float (*z)[MLP_OUT], int nrows) {
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    const int col = lane & 15, r4 = 4 * (lane >> 4);
    for (int n0 = 16 * wave; n0 < MLP_H; n0 += 16 * NW) {  // relu(x W1^T + b1)
        const f32x4_t acc = mlp_tile<MLP_IN>(x, nrows, N.w1, n0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float t = acc[i] + N.b1[n0 + col];
            h1[(r4 + i) * MLP_H + n0 + col] = bf16_bits(t > 0.0f ? t : 0.0f);
        }
    }
    __syncthreads();
    for (int n0 = 16 * wave; n0 < MLP_H; n0 += 16 * NW) {  // relu(h1 W2^T + b2)
        const f32x4_t acc = mlp_tile<MLP_H>(h1, nrows, N.w2, n0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float t = acc[i] + N.b2[n0 + col];
            h2[(r4 + i) * MLP_H + n0 + col] = bf16_bits(t > 0.0f ? t : 0.0f);
        }
    }
    __syncthreads();
    for (int n0 = 16 * wave; n0 < MLP_OUT; n0 += 16 * NW) {  // h2 W3^T + b3
        const f32x4_t acc = mlp_tile<MLP_H>(h2, nrows, N.w3, n0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (r4 + i < nrows) z[r4 + i][n0 + col] = acc[i] + N.b3[n0 + col];
    }
    __syncthreads();
}

// The mixture's choice for vertex v (block r) with scores in z / cand: on the learned branch the
// target and the accept uniform of the plain stream; otherwise gs = -1 (the plain draw follows).
__device__ __forceinline__ void learned_branch(uint64_t seed, uint32_t call, uint32_t tag, int v, int r,
                                               const float* z, const int* cand, bool has, int* gs, double* gu) {
    *gs = -1;
    *gu = 0.0;
    if (!has) return;
    const Draw4 dl = philox4x32_10(seed, (uint32_t)v, call, 0u, TAG_LEARN);
    if (unit53(dl.x0, dl.x1) < (double)LEARN_W_PLAIN) return;
    *gs = learned_pick(z, cand, r, unit53(dl.x2, dl.x3));
    const Draw4 dr = philox4x32_10(seed, (uint32_t)v, call, 0u, tag);
    *gu = unit53(dr.x2, dr.x3);
}

// The sequential pass's scorer on its own, one wavefront per listed vertex and PROBE_WAVES vertices
// per workgroup scored together (tables: 2 * tcap ints of scratch per vertex), for comparing its
// candidates and scores with the batched path's.
constexpr int PROBE_WAVES = 8;
__global__ void __launch_bounds__(PROBE_WAVES * WAVE) scorer_probe_kernel(GraphView G, BlockState S,
                                                                          const int* __restrict__ list, int n,
                                                                          int* tables, int tcap, LearnedNet N,
                                                                          int* __restrict__ cand_out,
                                                                          float* __restrict__ z_out) {
    __shared__ int lc[PROBE_WAVES][MLP_CAND], lcnt[PROBE_WAVES][MLP_CAND];
    __shared__ float lz[PROBE_WAVES][MLP_OUT];
    __shared__ __align__(16) unsigned short mlp[2][MLP_ROWS * MLP_H];
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    const int i = blockIdx.x * PROBE_WAVES + wave;
    unsigned short* x = mlp[1];
    if (i < n) {
        const int v = list[i];
        learned_features_wave(G, S, v, S.assign[v], tables + (size_t)i * 2 * tcap, vertex_table_cap(G, v), N,
                              lc[wave], lcnt[wave], x + wave * MLP_IN);
    }
    __syncthreads();
    learned_mlp_group<PROBE_WAVES>(N, x, mlp[0], mlp[1], lz, min(PROBE_WAVES, n - (int)blockIdx.x * PROBE_WAVES));
    if (i < n && lane < MLP_CAND) {
        cand_out[(long)i * MLP_CAND + lane] = lc[wave][lane];
        z_out[(long)i * MLP_OUT + lane] = lz[wave][lane];
    }
}

// Algorithmic bytes of an evaluation that stands: the formula of the batched kernels, with rows
// and columns of the hash tables counted by the keys they hold.
__device__ __forceinline__ long long eval_bytes(const MhParams& P, const MoveResult& R) {
    long long b = 16 + 4 + 8LL * (R.k_out + R.k_in);
    if (R.t >= 0) b += 8LL * (P.rocc[R.t] + P.cocc[R.t]);
    if (R.status >= 1) b += 8 * R.nnz_rs + 12LL * (R.nuniq + 2) + 8LL * R.nuniq + 16;
    return b;
}

// Whether wavefront w's evaluation (result R, own table tab/cap) still holds after the commit
// that moved a vertex from block cr to block cs with neighbour-block table ctab/ccap and stamped
// its neighbours with `serial`. The commit changed the assignment of one vertex, the cells of
// rows and columns cr and cs (which include every cell (b, cr), (b, cs), (cr, b), (cs, b) for
// its neighbour blocks b) and the degrees of cr and cs. The evaluation read the blocks of its
// neighbours, rows and columns r, s and t, and the degrees of t, of its neighbour blocks and of
// the blocks keyed in rows and columns r and s. It holds when none of these was touched; rows
// r, s and t were then neither changed nor moved, so a new evaluation would read the same
// values in the same order.
// Check q (1-13) of that test, the neighbour check (0) aside: whether the evaluation of (r, s, t)
// with neighbour-block table tab/cap read something the commit cr -> cs (table ctab/ccap) changes.
__device__ __forceinline__ bool reads_changed(const HashView& M, int q, int r, int s, int t, const int* tab, int cap,
                                              const int* ctab, int ccap, int cr, int cs) {
    if (q <= 3) {  // r, s or t is one of the blocks whose cells changed
        const int y = q == 1 ? r : (q == 2 ? s : t);
        return y >= 0 && (y == cr || y == cs || tab_find(ctab, ccap, y) >= 0);
    }
    if (q <= 11) {  // cr or cs keyed in rows / columns r and s
        const int k = q - 4;
        const int x = (k & 1) ? cs : cr;
        const int row = (k & 2) ? s : r;
        return ((k & 4) ? M.clookup(row, x) : M.rlookup(row, x)) != 0;
    }
    return tab_find(tab, cap, q == 12 ? cr : cs) >= 0;  // cr or cs a neighbour block
}

__device__ bool still_holds(const MhParams& P, const HashView& M, const MoveResult& R, const int* tab, int cap,
                            const int* ctab, int ccap, int cr, int cs, int serial) {
    const int lane = lane_id();
    int bad = 0;
    if (lane == 0) bad = P.touch[R.v] == serial;  // a neighbour moved
    else if (lane <= 13) bad = reads_changed(M, lane, R.r, R.s, R.t, tab, cap, ctab, ccap, cr, cs);
    return wave_or(bad) == 0;
}

__global__ void __launch_bounds__(MH_WAVES * WAVE) mh_kernel(MhParams P) {
    __shared__ int st[MH_WAVES];    // status of each wavefront's evaluation, 3 = idle
    __shared__ int holds[MH_WAVES]; // 1 while that evaluation still holds
    __shared__ double sdS[MH_WAVES];
    __shared__ int s_flag, s_cr, s_cs, s_ccap, s_need;
    __shared__ CommitInfo s_commit;
    __shared__ CommitLds s_clds;
    __shared__ int* s_ctab;
    __shared__ int lds_tab[MH_WAVES][7 * MH_LDS_CAP];
    __shared__ int lds_hist[MH_WAVES][256];
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    int* const gtab = P.tables + (size_t)wave * 7 * P.tcap;
    const GraphView G = P.G;
    const BlockState S{P.assign, P.dout, P.din, P.deg, P.B};
    const HashView M{P.rbase, P.rcap, P.rkeys, P.rvals, P.cbase, P.ccap, P.ckeys, P.cvals, P.B, lds_hist[wave]};
    long long att = 0, evs = 0, mvs = 0, bytes_local = 0;
    double dsum = 0.0, scan_local = 0.0, scan_count = 0.0;
    int pos = P.start;
    int stopped = 0;
    int serial = P.status[4];
    double windows = 0.0;
    while (pos < P.stop_at && stopped == 0) {
        // a window of up to 16 queue entries, each evaluated against the current state
        windows += 1.0;
        const int idx = pos + wave;
        MoveResult R{};
        const bool active = idx < P.n;
        int cap = 0;
        int* tab = gtab;
        if (active) {
            const int v = P.list[idx];
            cap = vertex_table_cap(G, v);
            if (cap <= MH_LDS_CAP) tab = lds_tab[wave];
            R = eval_move<HashView, MH_UNROLL>(G, S, M, v, tab, cap, P.seed, P.call, P.tag, -1, 0.0, false);
        }
        if (lane == 0) {
            st[wave] = active ? R.status : 3;
            holds[wave] = 1;
            sdS[wave] = active ? R.dS : 0.0;
        }
        __syncthreads();
        const int W = min(MH_WAVES, P.n - pos);
        // walk the window in queue order: rejected entries are final, an accepted one is committed
        // and the later entries are checked against the commit
        int cur = 0;
        for (;;) {
            int j = -1;
            for (int w = cur; w < W; ++w) {
                if (!holds[w] || st[w] == 2 || st[w] < 0) {
                    j = w;
                    break;
                }
            }
            const int end = j >= 0 ? j : W;
            if (lane == 0 && wave >= cur && wave < end && st[wave] != 3) {
                if (P.bytes != nullptr) bytes_local += eval_bytes(P, R);
                if (R.status >= 1) {
                    scan_local += (double)R.nnz_rs;
                    scan_count += 1.0;
                }
            }
            if (threadIdx.x == 0) {
                for (int w = cur; w < end; ++w) {
                    att += st[w] != 3;
                    evs += st[w] == 1;
                }
            }
            if (j < 0) {
                pos += W;
                break;
            }
            if (!holds[j]) {  // evaluate again from here
                pos += j;
                break;
            }
            if (st[j] < 0) {
                if (threadIdx.x == 0) atomicOr(P.err, DEV_ERR_TABLE);
                stopped = 2;
                pos += j;
                break;
            }
            if (wave == j && lane == 0) {
                s_cr = R.r;
                s_cs = R.s;
                s_ctab = tab;
                s_ccap = cap;
                s_commit = commit_info(R);
            }
            __syncthreads();
            {
                const bool need = blocks_need_room(P, s_ctab, s_ccap, s_cr, s_cs, &s_need);
                if (wave == j) {
                    const bool ok = make_room(P, R, tab, cap, need);
                    if (lane == 0) s_flag = ok ? 1 : 2;
                }
            }
            __syncthreads();
            if (s_flag == 2) {  // the pool is full: the host rebuilds it and resumes here
                stopped = 1;
                pos += j;
                break;
            }
            commit_shared(P, s_commit, s_ctab, s_ccap, serial + 1, s_clds);
            __syncthreads();
            ++serial;
            if (wave == j && lane == 0) {
                if (P.bytes != nullptr) bytes_local += eval_bytes(P, R);
                scan_local += (double)R.nnz_rs;
                scan_count += 1.0;
            }
            if (threadIdx.x == 0) {
                att++;
                evs++;
                mvs++;
                dsum += sdS[j];
            }
            if (wave > j && wave < W && holds[wave] && st[wave] != 3) {
                const bool h = still_holds(P, M, R, tab, cap, s_ctab, s_ccap, s_cr, s_cs, serial);
                if (lane == 0 && !h) holds[wave] = 0;
            }
            __syncthreads();
            cur = j + 1;
        }
        __syncthreads();
    }
    if (P.bytes != nullptr && lane == 0 && bytes_local > 0)
        atomicAdd((unsigned long long*)P.bytes, (unsigned long long)bytes_local);
    if (lane == 0 && scan_count > 0.0) {
        atomicAdd(&P.scan[0], scan_local);
        atomicAdd(&P.scan[1], scan_count);
    }
    if (threadIdx.x == 0) atomicAdd(&P.scan[2], windows);
    if (threadIdx.x == 0) {
        P.status[0] = pos;
        P.status[1] = stopped;
        P.status[4] = serial;
        atomicAdd((unsigned long long*)&P.counters[0], (unsigned long long)att);
        atomicAdd((unsigned long long*)&P.counters[1], (unsigned long long)evs);
        atomicAdd((unsigned long long*)&P.counters[2], (unsigned long long)mvs);
        P.dS_sum[0] += dsum;
    }
}

// LEARNED: the learned mixture, its entries scored together on the current state in each window
// (learned_features_wave, learned_mlp_group).
template <int G, bool LEARNED>
__global__ void __launch_bounds__(MH_WAVES * WAVE) mh_group_kernel(MhParams P) {
    constexpr int NG = MH_WAVES / G;
    constexpr int LCAP = MH_WAVES * MH_LDS_CAP / NG;  // LDS table slots per group
    __shared__ int st[NG];
    __shared__ int holds[NG];
    __shared__ double sdS[NG];
    __shared__ int s_flag, s_cr, s_cs, s_ccap, s_need;
    __shared__ CommitInfo s_commit;
    __shared__ CommitLds s_clds;
    __shared__ int* s_ctab;
    __shared__ int lds_tab[NG][7 * LCAP];
    __shared__ int lds_hist[NG][256];
    __shared__ double redd[MH_WAVES][RED_D];
    __shared__ int redi[MH_WAVES][RED_I];
    // the learned mixture: per group the candidates, counts, scores and the branch drawn
    __shared__ int s_lc[LEARNED ? NG : 1][MLP_CAND], s_lcnt[LEARNED ? NG : 1][MLP_CAND];
    __shared__ float s_lz[LEARNED ? NG : 1][MLP_OUT];
    __shared__ int s_lon[NG], s_lgs[NG], s_lhas[NG];
    __shared__ double s_lgu[NG];
    __shared__ __align__(16) unsigned short s_mlp[LEARNED ? 2 : 1][LEARNED ? MLP_ROWS * MLP_H : 1];
    static_assert(NG <= MLP_ROWS, "one feature row per group");
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    const int g = wave / G, wig = wave % G;
    const bool leader = wig == 0;  // the group's first wavefront commits and checks
    int* const gtab = P.tables + (size_t)g * 7 * P.tcap;
    const GraphView Gr = P.G;
    const BlockState S{P.assign, P.dout, P.din, P.deg, P.B};
    const HashView M{P.rbase, P.rcap, P.rkeys, P.rvals, P.cbase, P.ccap, P.ckeys, P.cvals, P.B, nullptr};
    long long att = 0, evs = 0, mvs = 0, bytes_local = 0;
    double dsum = 0.0, scan_local = 0.0, scan_count = 0.0;
    int pos = P.start;
    int stopped = 0;
    int serial = P.status[4];
    double windows = 0.0;
    while (pos < P.stop_at && stopped == 0) {
        windows += 1.0;
        const int idx = pos + g;
        const bool active = idx < P.n;
        const int v = active ? P.list[idx] : -1;
        const int cap = active ? vertex_table_cap(Gr, v) : 16;
        int* tab = cap <= LCAP ? lds_tab[g] : gtab;
        MoveResult R;
        if constexpr (LEARNED) {
            // each group's first wavefront forms its vertex's features on the current state, the
            // workgroup scores them together, then the first wavefront draws the branch
            if (leader) {
                int on = 0, has = 0;
                if (active && (Gr.out_ptr[v + 1] - Gr.out_ptr[v]) + (Gr.in_ptr[v + 1] - Gr.in_ptr[v]) <= P.net.maxdeg) {
                    on = 1;
                    has = learned_features_wave(Gr, S, v, S.assign[v], tab, cap, P.net, s_lc[g], s_lcnt[g],
                                                s_mlp[1] + g * MLP_IN);
                }
                if (lane == 0) {
                    s_lon[g] = on;
                    s_lhas[g] = has;
                }
            }
            __syncthreads();
            learned_mlp_group<MH_WAVES>(P.net, s_mlp[1], s_mlp[0], s_mlp[1], s_lz, NG);
            if (leader) {
                int gs = -1;
                double gu = 0.0;
                if (s_lon[g]) learned_branch(P.seed, P.call, P.tag, v, S.assign[v], s_lz[g], s_lc[g], s_lhas[g] != 0, &gs, &gu);
                if (lane == 0) {
                    s_lgs[g] = gs;
                    s_lgu[g] = gu;
                }
            }
            __syncthreads();
            const bool on = s_lon[g] != 0;
            R = eval_group<G, HashView, true>(Gr, S, M, v, tab, cap, P.seed, P.call, P.tag, on ? s_lgs[g] : -1,
                                              on ? s_lgu[g] : 0.0, false, lds_hist[g], redd, redi,
                                              on ? (double)LEARN_W_PLAIN : -1.0, on ? s_lz[g] : nullptr,
                                              on ? s_lc[g] : nullptr);
        } else {
            R = eval_group<G, HashView>(Gr, S, M, v, tab, cap, P.seed, P.call, P.tag, -1, 0.0, false, lds_hist[g], redd,
                                        redi);
        }
        if (leader && lane == 0) {
            st[g] = active ? R.status : 3;
            holds[g] = 1;
            sdS[g] = active ? R.dS : 0.0;
        }
        __syncthreads();
        const int W = min(NG, P.n - pos);
        int cur = 0;
        for (;;) {
            int j = -1;
            for (int w = cur; w < W; ++w) {
                if (!holds[w] || st[w] == 2 || st[w] < 0) {
                    j = w;
                    break;
                }
            }
            const int end = j >= 0 ? j : W;
            if (leader && lane == 0 && g >= cur && g < end && st[g] != 3) {
                if (P.bytes != nullptr) bytes_local += eval_bytes(P, R);
                if (R.status >= 1) {
                    scan_local += (double)R.nnz_rs;
                    scan_count += 1.0;
                }
            }
            if (threadIdx.x == 0) {
                for (int w = cur; w < end; ++w) {
                    att += st[w] != 3;
                    evs += st[w] == 1;
                }
            }
            if (j < 0) {
                pos += W;
                break;
            }
            if (!holds[j]) {
                pos += j;
                break;
            }
            if (st[j] < 0) {
                if (threadIdx.x == 0) atomicOr(P.err, DEV_ERR_TABLE);
                stopped = 2;
                pos += j;
                break;
            }
            if (g == j && leader && lane == 0) {
                s_cr = R.r;
                s_cs = R.s;
                s_ctab = tab;
                s_ccap = cap;
                s_commit = commit_info(R);
            }
            __syncthreads();
            {
                const bool need = blocks_need_room(P, s_ctab, s_ccap, s_cr, s_cs, &s_need);
                if (g == j && leader) {
                    const bool ok = make_room(P, R, tab, cap, need);
                    if (lane == 0) s_flag = ok ? 1 : 2;
                }
            }
            __syncthreads();
            if (s_flag == 2) {
                stopped = 1;
                pos += j;
                break;
            }
            commit_shared(P, s_commit, s_ctab, s_ccap, serial + 1, s_clds);
            __syncthreads();
            ++serial;
            if (g == j && leader && lane == 0) {
                if (P.bytes != nullptr) bytes_local += eval_bytes(P, R);
                scan_local += (double)R.nnz_rs;
                scan_count += 1.0;
            }
            if (threadIdx.x == 0) {
                att++;
                evs++;
                mvs++;
                dsum += sdS[j];
            }
            if (leader && g > j && g < W && holds[g] && st[g] != 3) {
                const bool h = still_holds(P, M, R, tab, cap, s_ctab, s_ccap, s_cr, s_cs, serial);
                if (lane == 0 && !h) holds[g] = 0;
            }
            __syncthreads();
            cur = j + 1;
        }
        __syncthreads();
    }
    if (P.bytes != nullptr && lane == 0 && bytes_local > 0)
        atomicAdd((unsigned long long*)P.bytes, (unsigned long long)bytes_local);
    if (lane == 0 && scan_count > 0.0) {
        atomicAdd(&P.scan[0], scan_local);
        atomicAdd(&P.scan[1], scan_count);
    }
    if (threadIdx.x == 0) atomicAdd(&P.scan[2], windows);
    if (threadIdx.x == 0) {
        P.status[0] = pos;
        P.status[1] = stopped;
        P.status[4] = serial;
        atomicAdd((unsigned long long*)&P.counters[0], (unsigned long long)att);
        atomicAdd((unsigned long long*)&P.counters[1], (unsigned long long)evs);
        atomicAdd((unsigned long long*)&P.counters[2], (unsigned long long)mvs);
        P.dS_sum[0] += dsum;
    }
}

// ------------------------------------------------------------------------------------------
// The same pass over several compute units. K workgroups of 8 wavefronts evaluate a window of
// 8 K queue entries against the current state, one entry per wavefront, and the window ends as
// the one-workgroup walk ends it: an entry's evaluation stands unless an accepted entry before it
// commits a change to something it read. Every accepted entry before the first entry that fails
// this check holds and is committed, so that entry ends the window whatever the later entries'
// outcomes. Each wavefront checks its entry against every accepted entry before it (the checks of
// still_holds, on the state before the window: a row a commit extends is one the check already
// names), the first failing entry is an atomic minimum, and workgroup 0 commits the accepted
// entries before it in queue order with commit_shared, exactly as the one-workgroup kernels do.
// Three grid barriers per window: a counter in device memory with device-scope release/acquire
// fences (the L2 of each XCD is written back and invalidated) and a bounded wait; a workgroup
// whose wait runs out reports DEV_ERR_BARRIER and leaves, and so does every other one at its
// next barrier. With one workgroup they are workgroup barriers.
// Wavefronts per workgroup: 8, so that each has 256 registers (the evaluation, the window
// bookkeeping and the commit stay out of scratch memory) and two workgroups share a compute unit.
constexpr int MH_WIDE_WAVES = 8;
constexpr int MH_WIDE_MAX = 512;  // workgroups (co-resident: at most two per compute unit)
constexpr int MH_TEAM_SCAN = 8192;  // mean slots scanned per evaluation from which entries take teams
constexpr int MH_WIDE_SLOTS = MH_WIDE_MAX * MH_WIDE_WAVES;
// Accepted entries a window tracks; a window with more ends before the next one (its commits
// and checks see only these).
constexpr int MH_ACC_MAX = 512;
constexpr int BARRIER_SPINS = 1 << 22;

struct WideSlot {
    int st;  // the evaluation's status, 3 = idle
    int r, s, nuniq, cap;
    int bad;  // its evaluation read something an accepted entry before it changes
    CommitInfo ci;
    double dS;
    long long bytes, nnz_rs;
    int tslots;  // slots of the proposal's row or column of t
    int ticks;   // with the phase clock: the evaluation's duration (100 MHz ticks)
    int ticks0;  // of which the learned scoring before it
};

// The n-th barrier of the launch (n >= 1) completes when ctrl[0] reaches n * gridDim.x. rel: this
// workgroup's writes since its last barrier are read by other workgroups after this one (a
// device-scope release: the XCD's L2 written back); acq: it reads what other workgroups wrote
// before this barrier (a device-scope acquire: the L2 invalidated). A workgroup skips what it does
// not need, since every invalidation makes the next window fetch its rows again.
__device__ bool grid_barrier(int* ctrl, int n, int* s_ok, bool rel, bool acq) {
    if (gridDim.x == 1) {  // one workgroup: a workgroup barrier is enough
        __syncthreads();
        return true;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (rel) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        atomicAdd(&ctrl[0], 1);
        const int target = n * (int)gridDim.x;
        int ok = 1;
        for (int spin = 0; __hip_atomic_load(&ctrl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; ++spin) {
            if (spin >= BARRIER_SPINS) {
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        if (acq) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_ok = ok;
    }
    __syncthreads();
    return *s_ok != 0;
}

// The team of workgroups that evaluates one window entry (eval_group's Team). Its exchanges are
// barriers among the team's members only: an arrival counter per entry (zeroed before each launch),
// the n-th exchange waiting for n * size arrivals, with the same release/acquire fences and bounded
// wait as grid_barrier. After a barrier that timed out the others are skipped (the window loop
// then stops).
struct WideTeam {
    static constexpr bool on = true;
    int rank, size;
    int* hist;
    double* part;
    int* arrivals;  // this entry's counter
    int* nsync;     // exchanges so far in this launch (the same in every member)
    int* s_ok;
    __device__ bool sync() const {
        if (*s_ok == 0) return false;
        const int target = ++*nsync * size;
        __syncthreads();
        if (threadIdx.x == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            atomicAdd(arrivals, 1);
            int ok = 1;
            for (int spin = 0; __hip_atomic_load(arrivals, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;
                 ++spin) {
                if (spin >= BARRIER_SPINS) {
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            *s_ok = ok;
        }
        __syncthreads();
        return *s_ok != 0;
    }
};

// Resets the window kernel's control words and the first `teams` team arrival counters.
__global__ void wide_init_kernel(int* ctrl, int* team_arrivals, int teams) {
    if (threadIdx.x == 0) {
        ctrl[0] = 0;
        ctrl[1] = INT32_MAX;
        ctrl[2] = 0;
        ctrl[3] = 0;
    }
    for (int i = threadIdx.x; i < teams; i += blockDim.x) team_arrivals[i] = 0;
}

// GW: wavefronts per entry (1, 2, 4 or 8; a workgroup holds 8 / GW entries, evaluated by
// eval_move for GW = 1 and by eval_group otherwise; each group's first wavefront checks its entry).
// LEARNED: the proposals are the learned mixture, each window's entries scored together on the
// current state (learned_features_wave, learned_mlp_group). A separate instantiation keeps the
// plain kernel free of the scorer.
// TEAM (GW = 8): P.team consecutive workgroups evaluate each entry together (WideTeam), for rows
// of tens of thousands of slots; the team's first workgroup records and checks the entry.
template <int GW, bool LEARNED, bool TEAM = false>
__global__ void __launch_bounds__(MH_WIDE_WAVES * WAVE) mh_wide_kernel(MhParams P) {
    static_assert(GW == 1 || GW == 2 || GW == 4 || GW == 8, "1, 2, 4 or 8 wavefronts per entry");
    static_assert(!TEAM || GW == 8, "a team evaluates one entry per workgroup");
    constexpr int NG = MH_WIDE_WAVES / GW;  // entries per workgroup
    constexpr int LCAP = GW * MH_LDS_CAP;   // LDS table slots per entry
    __shared__ int s_ok, s_flag, s_n, s_end, s_stop, s_need;
    __shared__ CommitInfo s_commit;
    // the wavefronts' neighbour-block tables (evaluations and checks) share their LDS with the
    // commits' lists (workgroup 0, after the checks)
    __shared__ union {
        int tab[MH_WIDE_WAVES][7 * MH_LDS_CAP];
        CommitLds clds;
    } s_u;
    CommitLds& s_clds = s_u.clds;
    // the evaluations' histograms and the window's accepted list share their LDS with the learned
    // network's feature and hidden rows, which are done with before the evaluations start
    __shared__ __align__(16) union {
        struct {
            int hist[MH_WIDE_WAVES][256];
            int acc[MH_ACC_MAX], acc_r[MH_ACC_MAX], acc_s[MH_ACC_MAX], acc_cap[MH_ACC_MAX];
            double dS[MH_ACC_MAX];
        } w;
        unsigned short mlp[2][MLP_ROWS * MLP_H];
    } s_w;
    auto& lds_hist = s_w.w.hist;
    auto& s_acc = s_w.w.acc;
    auto& s_acc_r = s_w.w.acc_r;
    auto& s_acc_s = s_w.w.acc_s;
    auto& s_acc_cap = s_w.w.acc_cap;
    auto& s_dS = s_w.w.dS;
    __shared__ int s_wcnt[MH_WIDE_WAVES], s_cut;
    __shared__ long long s_red[MH_WIDE_WAVES][4];
    __shared__ double s_redd[MH_WIDE_WAVES][3];
    // the learned mixture's candidates, their counts and scores, per entry; with GW > 1 also the
    // branch the group's first wavefront drew
    __shared__ int s_lc[LEARNED ? NG : 1][MLP_CAND], s_lcnt[LEARNED ? NG : 1][MLP_CAND];
    __shared__ float s_lz[LEARNED ? NG : 1][MLP_OUT];
    __shared__ int s_lon[LEARNED ? NG : 1], s_lgs[LEARNED ? NG : 1], s_lhas[LEARNED ? NG : 1];
    __shared__ double s_lgu[LEARNED ? NG : 1];
    // eval_group's per-wavefront partial sums
    __shared__ double redd[GW > 1 ? MH_WIDE_WAVES : 1][RED_D];
    __shared__ int redi[GW > 1 ? MH_WIDE_WAVES : 1][RED_I];
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    const int g = wave / GW, wig = wave % GW;
    const int TS = TEAM ? P.team : 1;
    const int rank = TEAM ? (int)blockIdx.x % TS : 0;  // place in the entry's team
    const bool first = wig == 0;               // the group's first wavefront (features, learned branch)
    const bool leader = first && rank == 0;    // the entry's first wavefront records and checks it
    const int slot = TEAM ? (int)blockIdx.x / TS : (int)blockIdx.x * NG + g;
    const int WS = TEAM ? (int)gridDim.x / TS : (int)gridDim.x * NG;  // entries per window
    // the entry's table (read by the checks and the commit) is its first workgroup's; the other
    // members keep theirs after the entries' tables
    int* const gtab = P.tables + (size_t)(rank == 0 ? slot : WS + slot * (TS - 1) + rank - 1) * 7 * P.tcap;
    const GraphView G = P.G;
    const BlockState S{P.assign, P.dout, P.din, P.deg, P.B};
    const HashView M{P.rbase, P.rcap, P.rkeys, P.rvals, P.cbase, P.ccap, P.ckeys, P.cvals, P.B, lds_hist[g]};
    const bool lead = blockIdx.x == 0;
    long long att = 0, evs = 0, mvs = 0, bytes_sum = 0;
    double dsum = 0.0, scan_sum = 0.0, scan_count = 0.0, scan_prop = 0.0, scan_wmax = 0.0, windows = 0.0;
    int radix_passes = 1;  // passes of the proposal's radix selection (8 bits each)
    while (radix_passes < 4 && (1LL << (8 * radix_passes)) < (long long)P.B) ++radix_passes;
    int pos = P.start;
    int stopped = 0;
    int serial = P.status[4];
    unsigned wserial = (unsigned)P.status[9];
    int nbar = 0;
    [[maybe_unused]] int nsync = 0;  // team exchanges (TEAM)
    if (threadIdx.x == 0) s_ok = 1;
    __syncthreads();
    const bool clocked = P.phase != nullptr && blockIdx.x == 0 && threadIdx.x == 0;
    // phases (4), commits, then per window the slowest evaluation: its ticks, delta-MDL slots,
    // proposal row slots and vertex degree, and the windows counted
    double ph[11] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    __shared__ long long s_tmax[MH_WIDE_WAVES];
    unsigned long long tclk = clocked ? wall_clock64() : 0ull;
    auto tick = [&](int k) {
        if (clocked) {
            const unsigned long long t = wall_clock64();
            ph[k] += (double)(t - tclk);
            tclk = t;
        }
    };
    while (pos < P.stop_at && stopped == 0) {
        ++wserial;
        windows += 1.0;
        const int Wn = min(WS, P.n - pos);
        // 1. evaluate this group's entry
        const bool active = slot < Wn;
        MoveResult R{};
        int cap = 16;
        int* tab = gtab;
        const int v = active ? P.list[pos + slot] : -1;
        const unsigned long long te0 = P.phase != nullptr ? wall_clock64() : 0ull;
        if (active) {
            cap = vertex_table_cap(G, v);
            if (cap <= LCAP) tab = s_u.tab[g * GW];
        }
        if constexpr (LEARNED) {
            // each entry's first wavefront forms its features on the current state, the workgroup
            // scores its entries together, then the first wavefront draws the branch
            if (first) {
                int on = 0, has = 0;
                if (active && (G.out_ptr[v + 1] - G.out_ptr[v]) + (G.in_ptr[v + 1] - G.in_ptr[v]) <= P.net.maxdeg) {
                    on = 1;
                    has = learned_features_wave(G, S, v, S.assign[v], tab, cap, P.net, s_lc[g], s_lcnt[g],
                                                s_w.mlp[1] + g * MLP_IN);
                }
                if (lane == 0) {
                    s_lon[g] = on;
                    s_lhas[g] = has;
                }
            }
            __syncthreads();
            learned_mlp_group<MH_WIDE_WAVES>(P.net, s_w.mlp[1], s_w.mlp[0], s_w.mlp[1], s_lz, NG);
            if (first) {
                int gs = -1;
                double gu = 0.0;
                if (s_lon[g]) learned_branch(P.seed, P.call, P.tag, v, S.assign[v], s_lz[g], s_lc[g], s_lhas[g] != 0, &gs, &gu);
                if (lane == 0) {
                    s_lgs[g] = gs;
                    s_lgu[g] = gu;
                }
            }
            __syncthreads();
            const bool on = s_lon[g] != 0;
            if (P.phase != nullptr && leader && lane == 0) P.wide[slot].ticks0 = (int)(wall_clock64() - te0);
            if constexpr (GW == 1) {
                if (active)
                    R = eval_move<HashView, MH_UNROLL>(G, S, M, v, tab, cap, P.seed, P.call, P.tag, on ? s_lgs[g] : -1,
                                                       on ? s_lgu[g] : 0.0, false, on ? (double)LEARN_W_PLAIN : -1.0,
                                                       0.0, 0.0, on ? s_lz[g] : nullptr, on ? s_lc[g] : nullptr);
            } else if constexpr (TEAM) {
                const WideTeam tm{rank, TS, P.team_hist + (size_t)slot * 2 * TS * 256, P.team_part + (size_t)slot * 2 * TS,
                                  P.team_arrivals + slot, &nsync, &s_ok};
                R = eval_group<GW, HashView, true, 2, WideTeam>(
                    G, S, M, v, tab, cap, P.seed, P.call, P.tag, on ? s_lgs[g] : -1, on ? s_lgu[g] : 0.0, false,
                    lds_hist[g], redd, redi, on ? (double)LEARN_W_PLAIN : -1.0, on ? s_lz[g] : nullptr,
                    on ? s_lc[g] : nullptr, 0.0, 0.0, tm);
            } else {  // every wavefront of the workgroup takes part (eval_group's workgroup barriers)
                R = eval_group<GW, HashView, true, 2, NoTeam, (NG > 1 ? NG : 0)>(
                    G, S, M, v, tab, cap, P.seed, P.call, P.tag, on ? s_lgs[g] : -1, on ? s_lgu[g] : 0.0, false,
                    lds_hist[g], redd, redi, on ? (double)LEARN_W_PLAIN : -1.0, on ? s_lz[g] : nullptr,
                    on ? s_lc[g] : nullptr);
            }
        } else if constexpr (GW == 1) {
            if (active) R = eval_move<HashView, MH_UNROLL>(G, S, M, v, tab, cap, P.seed, P.call, P.tag, -1, 0.0, false);
        } else if constexpr (TEAM) {
            const WideTeam tm{rank, TS, P.team_hist + (size_t)slot * 2 * TS * 256, P.team_part + (size_t)slot * 2 * TS,
                              P.team_arrivals + slot, &nsync, &s_ok};
            R = eval_group<GW, HashView, false, 2, WideTeam>(G, S, M, v, tab, cap, P.seed, P.call, P.tag, -1, 0.0, false,
                                                             lds_hist[g], redd, redi, -1.0, nullptr, nullptr, 0.0, 0.0,
                                                             tm);
        } else {  // every wavefront of the workgroup takes part (eval_group's workgroup barriers); the
                  // workgroup's entries scan their rows together
            R = eval_group<GW, HashView, false, 2, NoTeam, (NG > 1 ? NG : 0)>(
                G, S, M, v, tab, cap, P.seed, P.call, P.tag, -1, 0.0, false, lds_hist[g], redd, redi);
        }
        if constexpr (TEAM) {
            if (!s_ok) break;  // a team barrier timed out
        }
        if (active && R.status == 2 && tab != gtab) {  // the commit and the other checks read it from memory
            for (int j = wig * WAVE + lane; j < 7 * cap; j += GW * WAVE) gtab[j] = tab[j];
        }
        if (leader && lane == 0) {
            WideSlot& o = P.wide[slot];
            P.wst[slot] = active ? R.status : 3;
            o.st = active ? R.status : 3;
            o.r = R.r;
            o.s = R.s;
            o.nuniq = R.nuniq;
            o.cap = cap;
            o.bad = 0;
            o.ci = commit_info(R);
            o.dS = R.dS;
            o.bytes = (active && P.bytes != nullptr) ? eval_bytes(P, R) : 0;
            o.nnz_rs = R.nnz_rs;
            o.tslots = R.tslots;
            o.ticks = P.phase != nullptr ? (int)(wall_clock64() - te0) : 0;
            if (active && R.status == 2) P.wstamp[R.v] = ((unsigned long long)wserial << 32) | (unsigned)slot;
        }
        // every entry's record and table, read by every workgroup's list and checks
        if (!grid_barrier(P.ctrl, ++nbar, &s_ok, true, true)) break;
        tick(0);
        // 2. the accepted entries of the window, in queue order: each wavefront counts a slice of
        // the dense status array, then fills its part of the list (the first MH_ACC_MAX; the
        // window ends before the next accepted entry)
        {
            const int per = ((Wn + MH_WIDE_WAVES * WAVE - 1) / (MH_WIDE_WAVES * WAVE)) * WAVE;
            const int lo = wave * per, hi = min(Wn, lo + per);
            int cnt = 0;
            for (int j = lo + lane; j < hi; j += WAVE) cnt += P.wst[j] == 2 ? 1 : 0;
            cnt = (int)wave_sum_ll(cnt);
            if (lane == 0) s_wcnt[wave] = cnt;
            if (threadIdx.x == 0) s_cut = INT32_MAX;
            __syncthreads();
            int base = 0, total = 0;
            for (int w = 0; w < MH_WIDE_WAVES; ++w) {
                base += w < wave ? s_wcnt[w] : 0;
                total += s_wcnt[w];
            }
            for (int b0 = lo; b0 < hi && base <= MH_ACC_MAX; b0 += WAVE) {
                const int j = b0 + lane;
                const bool acc = j < hi && P.wst[j] == 2;
                const unsigned long long m = __ballot(acc);
                if (acc) {
                    const int k = base + __popcll(m & ((1ull << lane) - 1));
                    if (k < MH_ACC_MAX) {
                        s_acc[k] = j;
                        s_acc_r[k] = P.wide[j].r;
                        s_acc_s[k] = P.wide[j].s;
                        s_acc_cap[k] = P.wide[j].cap;
                        s_dS[k] = P.wide[j].dS;
                    } else if (k == MH_ACC_MAX) {
                        s_cut = j;
                    }
                }
                base += __popcll(m);
            }
            if (threadIdx.x == 0) s_n = min(total, MH_ACC_MAX);
        }
        __syncthreads();
        if (threadIdx.x == 0 && s_cut < Wn) atomicMin(&P.ctrl[1], s_cut);
        // 3. whether this entry's evaluation holds after the accepted entries before it
        if (active && leader) {
            int nb = 0, hi = s_n;  // accepted entries before this one
            while (nb < hi) {
                const int mid = (nb + hi) >> 1;
                if (s_acc[mid] < slot) nb = mid + 1;
                else hi = mid;
            }
            int bad = 0;
            if (nb > 0) {
                const int v = R.v;
                for (int j = G.out_ptr[v] + lane; j < G.out_ptr[v + 1] && !bad; j += WAVE) {
                    const unsigned long long w = P.wstamp[G.out_idx[j]];
                    bad = (unsigned)(w >> 32) == wserial && (int)(unsigned)w < slot;
                }
                for (int j = G.in_ptr[v] + lane; j < G.in_ptr[v + 1] && !bad; j += WAVE) {
                    const unsigned long long w = P.wstamp[G.in_idx[j]];
                    bad = (unsigned)(w >> 32) == wserial && (int)(unsigned)w < slot;
                }
                bad = wave_or(bad);
                // four accepted entries per trip, checks 1-13 on lanes 1-13 of each quarter
                for (int k0 = 0; k0 < nb && !bad; k0 += 4) {
                    const int k = k0 + lane / 16, q = lane & 15;
                    int b = 0;
                    if (k < nb && q >= 1 && q <= 13) {
                        const int* ctab = P.tables + (size_t)s_acc[k] * 7 * P.tcap;
                        b = reads_changed(M, q, R.r, R.s, R.t, tab, cap, ctab, s_acc_cap[k], s_acc_r[k], s_acc_s[k]);
                    }
                    bad = wave_or(b);
                }
            }
            if (lane == 0 && (bad || R.status < 0)) {
                P.wide[slot].bad = bad;
                atomicMin(&P.ctrl[1], slot);
            }
        }
        // the checks' outcomes, read by workgroup 0 only
        if (!grid_barrier(P.ctrl, ++nbar, &s_ok, !lead, lead)) break;
        tick(1);
        // 4. workgroup 0 commits the accepted entries before the first that does not hold
        if (lead) {
            if (threadIdx.x == 0) {
                int end = min(__hip_atomic_load(&P.ctrl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), Wn);
                int stop = 0;
                if (end < Wn && P.wide[end].st < 0 && P.wide[end].bad == 0) {
                    atomicOr(P.err, DEV_ERR_TABLE);
                    stop = 2;
                }
                s_end = end;
                s_stop = stop;
            }
            __syncthreads();
            for (int k = 0; k < s_n && s_acc[k] < s_end; ++k) {
                const int j = s_acc[k];
                int* const ctab = P.tables + (size_t)j * 7 * P.tcap;
                const bool need = blocks_need_room(P, ctab, s_acc_cap[k], s_acc_r[k], s_acc_s[k], &s_need);
                if (wave == 0) {
                    MoveResult Rj{};
                    Rj.r = s_acc_r[k];
                    Rj.s = s_acc_s[k];
                    Rj.nuniq = P.wide[j].nuniq;
                    const bool ok = make_room(P, Rj, ctab, s_acc_cap[k], need);
                    if (lane == 0) {
                        s_flag = ok ? 1 : 2;
                        s_commit = P.wide[j].ci;
                    }
                }
                __syncthreads();
                if (s_flag == 2) {  // the pool is full: the host rebuilds it and resumes here
                    if (threadIdx.x == 0) {
                        s_end = j;
                        s_stop = 1;
                    }
                    __syncthreads();
                    break;
                }
                commit_shared<MH_WIDE_WAVES>(P, s_commit, ctab, s_acc_cap[k], serial + 1, s_clds);
                __syncthreads();
                ++serial;
                if (threadIdx.x == 0) {
                    mvs++;
                    dsum += s_dS[k];
                }
            }
            // counts over the entries that stand
            const int end = s_end;
            long long a = 0, e = 0, by = 0;
            double sc = 0.0, scn = 0.0, sp = 0.0;
            long long wm = 0;  // the most row slots an evaluation of the window visited
            for (int i = threadIdx.x; i < end; i += MH_WIDE_WAVES * WAVE) {
                const WideSlot& o = P.wide[i];
                if (o.st == 3) continue;
                a += 1;
                e += o.st >= 1;
                by += o.bytes;
                if (o.st >= 1) {
                    sc += (double)o.nnz_rs;
                    sp += (double)o.tslots;
                    scn += 1.0;
                    const long long visits = o.nnz_rs + (long long)radix_passes * o.tslots;
                    wm = visits > wm ? visits : wm;
                }
            }
            a = wave_sum_ll(a);
            e = wave_sum_ll(e);
            by = wave_sum_ll(by);
            for (int off = WAVE / 2; off > 0; off >>= 1) {
                const long long y = __shfl_xor(wm, off, WAVE);
                wm = y > wm ? y : wm;
            }
            sc = wave_sum(sc);
            sp = wave_sum(sp);
            scn = wave_sum(scn);
            if (lane == 0) {
                s_red[wave][0] = a;
                s_red[wave][1] = e;
                s_red[wave][2] = by;
                s_red[wave][3] = wm;
                s_redd[wave][0] = sc;
                s_redd[wave][1] = scn;
                s_redd[wave][2] = sp;
            }
            __syncthreads();
            if (P.phase != nullptr) {  // the window's slowest evaluation (every entry evaluated)
                long long best = -1;
                for (int i = threadIdx.x; i < Wn; i += MH_WIDE_WAVES * WAVE) {
                    const long long key = ((long long)P.wide[i].ticks << 20) | i;
                    best = key > best ? key : best;
                }
                for (int off = WAVE / 2; off > 0; off >>= 1) {
                    const long long y = __shfl_xor(best, off, WAVE);
                    best = y > best ? y : best;
                }
                if (lane == 0) s_tmax[wave] = best;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                if (P.phase != nullptr) {
                    long long best = -1;
                    for (int w = 0; w < MH_WIDE_WAVES; ++w) best = s_tmax[w] > best ? s_tmax[w] : best;
                    if (best >= 0) {
                        const int i = (int)(best & ((1 << 20) - 1));
                        const int u = P.list[pos + i];
                        ph[5] += (double)(best >> 20);
                        ph[6] += (double)P.wide[i].nnz_rs;
                        ph[7] += (double)P.wide[i].tslots;
                        ph[8] += (double)((G.out_ptr[u + 1] - G.out_ptr[u]) + (G.in_ptr[u + 1] - G.in_ptr[u]));
                        ph[9] += 1.0;
                        ph[10] += LEARNED ? (double)P.wide[i].ticks0 : 0.0;
                    }
                }
                for (int w = 0; w < MH_WIDE_WAVES; ++w) {
                    att += s_red[w][0];
                    evs += s_red[w][1];
                    bytes_sum += s_red[w][2];
                    scan_sum += s_redd[w][0];
                    scan_count += s_redd[w][1];
                    scan_prop += s_redd[w][2];
                    wm = s_red[w][3] > wm ? s_red[w][3] : wm;
                }
                scan_wmax += (double)wm;
                P.ctrl[1] = INT32_MAX;
                P.ctrl[2] = pos + end;
                P.ctrl[3] = s_stop;
            }
        }
        tick(2);
        // workgroup 0's commits, read by every workgroup's next evaluations
        if (!grid_barrier(P.ctrl, ++nbar, &s_ok, lead, !lead)) break;
        tick(3);
        pos = __hip_atomic_load(&P.ctrl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        stopped = __hip_atomic_load(&P.ctrl[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!s_ok) {
        if (threadIdx.x == 0) atomicOr(P.err, DEV_ERR_BARRIER);
        return;
    }
    if (lead && threadIdx.x == 0) {
        if (P.bytes != nullptr && bytes_sum > 0) atomicAdd((unsigned long long*)P.bytes, (unsigned long long)bytes_sum);
        atomicAdd(&P.scan[0], scan_sum);
        atomicAdd(&P.scan[1], scan_count);
        atomicAdd(&P.scan[3], scan_prop);
        atomicAdd(&P.scan[4], scan_wmax);
        atomicAdd(&P.scan[2], windows);
        P.status[0] = pos;
        P.status[1] = stopped;
        P.status[4] = serial;
        P.status[9] = (int)wserial;
        if (clocked) {
            for (int k = 0; k < 4; ++k) P.phase[k] += ph[k];
            P.phase[4] += (double)mvs;
            for (int k = 5; k < 11; ++k) P.phase[k] += ph[k];
        }
        atomicAdd((unsigned long long*)&P.counters[0], (unsigned long long)att);
        atomicAdd((unsigned long long*)&P.counters[1], (unsigned long long)evs);
        atomicAdd((unsigned long long*)&P.counters[2], (unsigned long long)mvs);
        P.dS_sum[0] += dsum;
    }
}

// ------------------------------------------------------------------------------------------
// Marks the queue vertices the pass moved (their block differs from the snapshot's) as a batch of
// moves for the incremental block-matrix update.
__global__ void mark_pass_moves_kernel(const int* __restrict__ list, int n, const int* __restrict__ assign0,
                                       const int* __restrict__ assign, int epoch, uint8_t* __restrict__ stat,
                                       int* __restrict__ new_block, int* __restrict__ moved_epoch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = list[i];
    const int a = assign[v];
    if (a != assign0[v]) {
        stat[v] = 2;
        new_block[v] = a;
        moved_epoch[v] = epoch;
    } else {
        stat[v] = 0;
    }
}

// Brings the CSR snapshot up to date with the pass's moves so far (dout/din exact, deg keeps the
// drift the commits left): from the moved vertices' edges when they are few, else by a full
// rebuild; then takes a new snapshot of the assignment.
static void update_after_pass(Ctx& c, const int* d_list, int n) {
    MhState& h = c.mh;
    const int N = c.g.N;
    c.epoch++;
    mark_pass_moves_kernel<<<(n + 255) / 256, 256, 0, c.stream>>>(d_list, n, h.assign0, c.bm.assign, c.epoch,
                                                                 c.ev_acc, c.new_block, c.moved_epoch);
    SBPX_HIP(hipGetLastError());
    const int total = prepare_update(c, d_list, n, h.assign0);
    if (total >= 0) finish_update(c, d_list, n, total);
    else rebuild_keep_drift(c);
    SBPX_HIP(hipMemcpyAsync(h.assign0, c.bm.assign, sizeof(int) * N, hipMemcpyDeviceToDevice, c.stream));
}

// ------------------------------------------------------------------------------------------
static void build_tables(Ctx& c) {
    const int B = c.bm.B;
    hipStream_t st = c.stream;
    MhState& h = c.mh;
    const int tb = 256;
    // capacities and bases, rows then columns
    SBPX_HIP(hipMemsetAsync(h.status + 5, 0, sizeof(int), st));
    hash_caps_kernel<<<(B + tb - 1) / tb, tb, 0, st>>>(c.bm.row_ptr, B, h.rcap, h.status + 5);
    scan_exclusive(h.rcap, h.rbase, B, c.scan_tmp, st);
    hash_caps_kernel<<<(B + tb - 1) / tb, tb, 0, st>>>(c.bm.col_ptr, B, h.ccap, h.status + 5);
    scan_exclusive(h.ccap, h.cbase, B, c.scan_tmp, st);
    int tot[2] = {0, 0};
    SBPX_HIP(hipMemcpyAsync(&tot[0], h.rbase + B, sizeof(int), hipMemcpyDeviceToHost, st));
    SBPX_HIP(hipMemcpyAsync(&tot[1], h.cbase + B, sizeof(int), hipMemcpyDeviceToHost, st));
    SBPX_HIP(hipMemcpyAsync(&h.max_cap, h.status + 5, sizeof(int), hipMemcpyDeviceToHost, st));
    SBPX_HIP(hipStreamSynchronize(st));
    const size_t used = (size_t)std::max(tot[0], tot[1]);
    if (c.mh_headroom > 0) h.headroom = std::max(h.headroom, (size_t)c.mh_headroom);
    else h.headroom = std::max(h.headroom, std::max(used, (size_t)1 << 20));
    // pool tops stay below 2^31 even after a failed claim
    const size_t need = std::min(used + h.headroom, (size_t)(INT32_MAX / 2));
    if (need < used + std::min(h.headroom, (size_t)64))
        throw StateError("block matrix too large for the Metropolis-Hastings tables");
    if (need > h.slots) {
        for (int** p : {&h.rkeys, &h.rvals, &h.ckeys, &h.cvals}) {
            if (*p) SBPX_HIP(hipFree(*p));
            SBPX_HIP(hipMalloc((void**)p, sizeof(int) * need));
        }
        h.slots = need;
    }
    // the pool the kernels may fill: the whole allocation, or exactly the headroom asked for
    h.pool_limit = c.mh_headroom > 0 ? need : h.slots;
    const int wpb = tb / WAVE;
    hash_fill_kernel<<<(B + wpb - 1) / wpb, tb, 0, st>>>(c.bm.row_ptr, c.bm.row_col, c.bm.row_val, B, h.rbase,
                                                          h.rkeys, h.rvals, h.rocc);
    hash_fill_kernel<<<(B + wpb - 1) / wpb, tb, 0, st>>>(c.bm.col_ptr, c.bm.col_row, c.bm.col_val, B, h.cbase,
                                                          h.ckeys, h.cvals, h.cocc);
    SBPX_HIP(hipGetLastError());
    // free ends of the pools
    SBPX_HIP(hipMemcpyAsync(h.status + 2, h.rbase + B, sizeof(int), hipMemcpyDeviceToDevice, st));
    SBPX_HIP(hipMemcpyAsync(h.status + 3, h.cbase + B, sizeof(int), hipMemcpyDeviceToDevice, st));
}

// ------------------------------------------------------------------------------------------
// Heap-ordered merge selection (carry_out_best_merges_advanced, block_merge.cpp:123-217): a
// min-heap of (delta MDL, block, proposed block); the top merge is re-scored on the current
// matrix and goes back into the heap when it got worse than the next one, otherwise it is
// applied at once. One wavefront runs the whole loop: lane 0 keeps the heap and the union-find
// of merged blocks, the wavefront scores merges and folds row and column r into s in the hash
// tables. The heap steps are those of libstdc++'s std::push_heap / std::pop_heap (used by the
// reference through std::priority_queue), so merges with equal delta MDL leave the heap in the
// same order.
struct MergeEntry {
    double dS;
    int r;
    int s;
};

// priority_queue's comparator: a greater delta MDL sorts first, so the top is the minimum
__device__ __forceinline__ bool heap_before(const MergeEntry& a, const MergeEntry& b) { return a.dS > b.dS; }

__device__ void heap_sift_up(MergeEntry* h, int hole, int top, MergeEntry value) {
    int parent = (hole - 1) / 2;
    while (hole > top && heap_before(h[parent], value)) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = value;
}

// Removes the top of a heap of len entries (len >= 1); the heap then holds len - 1.
__device__ void heap_pop(MergeEntry* h, int len) {
    if (len <= 1) return;
    const int n = len - 1;  // the last entry is re-inserted from the hole left at the top
    MergeEntry value = h[n];
    h[n] = h[0];
    int hole = 0, child = 0;
    while (child < (n - 1) / 2) {
        child = 2 * (child + 1);
        if (heap_before(h[child], h[child - 1])) child--;
        h[hole] = h[child];
        hole = child;
    }
    if ((n & 1) == 0 && child == (n - 2) / 2) {
        child = 2 * (child + 1);
        h[hole] = h[child - 1];
        hole = child - 1;
    }
    heap_sift_up(h, hole, 0, value);
}

__device__ int uf_find(int* parent, int x) {
    int root = x;
    while (parent[root] != root) root = parent[root];
    while (parent[x] != root) {
        const int nx = parent[x];
        parent[x] = root;
        x = nx;
    }
    return root;
}

// Room for folding row and column r into s: row and column s may gain every key of row and
// column r; each block keyed in row r gains key s in its column, each block keyed in column r
// gains key s in its row.
__device__ bool merge_room(const MhParams& P, int r, int s) {
    const int lane = lane_id();
    int* rtop = P.status + 2;
    int* ctop = P.status + 3;
    const int big = P.rocc[r] + P.cocc[r] + 4;
    if (room(P.rcap, P.rocc, s) < big && !relocate(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, rtop, P.pool, s, big))
        return false;
    if (room(P.ccap, P.cocc, s) < big && !relocate(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, ctop, P.pool, s, big))
        return false;
    for (int pass = 0; pass < 2; ++pass) {
        const int b0 = pass == 0 ? P.rbase[r] : P.cbase[r];
        const int n = pass == 0 ? P.rcap[r] : P.ccap[r];
        const int* keys = pass == 0 ? P.rkeys : P.ckeys;
        const int* vals = pass == 0 ? P.rvals : P.cvals;
        for (int j0 = 0; j0 < n; j0 += WAVE) {
            const int j = j0 + lane;
            int x = -1;
            if (j < n) {
                const int kk = keys[b0 + j];
                if (kk >= 0 && vals[b0 + j] != 0 && kk != r && kk != s) x = kk;
            }
            // a key of row r needs room in its column; a key of column r in its row
            unsigned long long need = pass == 0 ? __ballot(x >= 0 && room(P.ccap, P.cocc, x) < 2)
                                                : __ballot(x >= 0 && room(P.rcap, P.rocc, x) < 2);
            while (need) {
                const int l = first_lane(need);
                need &= need - 1;
                const int y = __shfl(x, l, WAVE);
                const bool ok = pass == 0 ? relocate(P.cbase, P.ccap, P.ckeys, P.cvals, P.cocc, ctop, P.pool, y, 2)
                                          : relocate(P.rbase, P.rcap, P.rkeys, P.rvals, P.rocc, rtop, P.pool, y, 2);
                if (!ok) return false;
            }
        }
    }
    return true;
}

// Folds row and column r into s (block_merge::blockmodel_delta, block_merge.cpp:13-37, applied as
// DictTransposeMatrix::update_edge_counts): cell (r, c) moves to (s, c), with (r, r) and (r, s)
// landing on (s, s); then cell (i, r) moves to (i, s). Degrees as block_merge.cpp:199-205.
__device__ void merge_apply(const MhParams& P, int r, int s, int k_out, int k_in) {
    const int lane = lane_id();
    {
        const int b0 = P.rbase[r], n = P.rcap[r];
        for (int j = lane; j < n; j += WAVE) {
            const int c = P.rkeys[b0 + j], m = P.rvals[b0 + j];
            if (c < 0 || m == 0) continue;
            cell_add(P, s, c == r ? s : c, m);
            cell_add(P, r, c, -m);
        }
    }
    fence_wave();
    {
        const int b0 = P.cbase[r], n = P.ccap[r];
        for (int j = lane; j < n; j += WAVE) {
            const int i = P.ckeys[b0 + j], m = P.cvals[b0 + j];
            if (i < 0 || m == 0) continue;
            cell_add(P, i, s, m);
            cell_add(P, i, r, -m);
        }
    }
    fence_wave();
    if (lane == 0) {
        P.dout[r] = 0;
        P.dout[s] += k_out;
        P.din[r] = 0;
        P.din[s] += k_in;
    }
    fence_wave();
}

// status: [0] merges done, [1] 1 when the table pool ran out before a merge (its fields in
// [6] = r, [7] = proposed block, and the heap top already removed), [2]/[3] pool tops,
// [8] entries left in the heap; err gets DEV_ERR_NO_PROPOSAL for a block without a proposal.
__global__ void __launch_bounds__(WAVE) merge_heap_kernel(MhParams P, MergeEntry* heap, int len, int* parent,
                                                         int num_merged, int num_to_merge, int pending_r,
                                                         int pending_s, double* merged_dS) {
    const int lane = lane_id();
    const HashView M{P.rbase, P.rcap, P.rkeys, P.rvals, P.cbase, P.ccap, P.ckeys, P.cvals, P.B, nullptr};
    int stopped = 0;
    double dsum = 0.0;
    // a merge that was accepted when the pool ran out is applied first, without a new score
    if (pending_r >= 0) {
        const int r = pending_r;
        int s = 0;
        if (lane == 0) s = uf_find(parent, pending_s);
        s = __shfl(s, 0, WAVE);
        const int k_out = P.dout[r], k_in = P.din[r];
        if (merge_room(P, r, s)) {
            merge_apply(P, r, s, k_out, k_in);
            if (lane == 0) parent[r] = s;
            fence_wave();
            num_merged++;
        } else {
            stopped = 1;
        }
    }
    while (!stopped && num_merged < num_to_merge && len > 0) {
        MergeEntry top{0.0, 0, 0};
        if (lane == 0) {
            top = heap[0];
            heap_pop(heap, len);
        }
        len--;
        top.dS = __shfl(top.dS, 0, WAVE);
        top.r = __shfl(top.r, 0, WAVE);
        top.s = __shfl(top.s, 0, WAVE);
        fence_wave();
        const int r = top.r;
        if (top.s < 0) {
            if (lane == 0) atomicOr(P.err, DEV_ERR_NO_PROPOSAL);
            break;
        }
        int s = 0;
        if (lane == 0) s = uf_find(parent, top.s);
        s = __shfl(s, 0, WAVE);
        if (r == s) continue;
        const int k_out = P.dout[r], k_in = P.din[r];
        const double actual = merge_delta_mdl_wave<HashView>(M, P.dout, P.din, r, s, k_out, k_in);
        if (len > 0) {
            double next = 0.0;
            if (lane == 0) next = heap[0].dS;
            next = __shfl(next, 0, WAVE);
            if (actual > next) {  // worse than anticipated: back into the heap with its new score
                if (lane == 0) {
                    heap[len] = MergeEntry{actual, top.r, top.s};
                    heap_sift_up(heap, len, 0, MergeEntry{actual, top.r, top.s});
                }
                len++;
                fence_wave();
                continue;
            }
        }
        if (!merge_room(P, r, s)) {
            if (lane == 0) {
                P.status[6] = r;
                P.status[7] = top.s;
            }
            stopped = 1;
            break;
        }
        merge_apply(P, r, s, k_out, k_in);
        if (lane == 0) parent[r] = s;
        fence_wave();
        dsum += actual;
        num_merged++;
    }
    if (lane == 0) {
        P.status[0] = num_merged;
        P.status[1] = stopped;
        P.status[8] = len;
        merged_dS[0] += dsum;
    }
}

void merge_heap_select(Ctx& c, int num_to_merge, const double* dS, const int32_t* best, std::vector<int>& parent) {
    const int B = c.bm.B;
    MhState& h = c.mh;
    // the initial heap, built by successive pushes as std::priority_queue does
    std::vector<MergeEntry> heap;
    heap.reserve(B);
    auto before = [](const MergeEntry& a, const MergeEntry& b) { return a.dS > b.dS; };
    for (int i = 0; i < B; ++i) {
        heap.push_back(MergeEntry{dS[i], i, best[i]});
        std::push_heap(heap.begin(), heap.end(), before);
    }
    MergeEntry* d_heap = nullptr;
    int* d_parent = nullptr;
    SBPX_HIP(hipMalloc((void**)&d_heap, sizeof(MergeEntry) * std::max(B, 1)));
    SBPX_HIP(hipMalloc((void**)&d_parent, sizeof(int) * std::max(B, 1)));
    std::vector<int> ident(B);
    std::iota(ident.begin(), ident.end(), 0);
    try {
        SBPX_HIP(hipMemcpyAsync(d_heap, heap.data(), sizeof(MergeEntry) * B, hipMemcpyHostToDevice, c.stream));
        SBPX_HIP(hipMemcpyAsync(d_parent, ident.data(), sizeof(int) * B, hipMemcpyHostToDevice, c.stream));
        int len = B, merged = 0, pending_r = -1, pending_s = -1, rounds = 0;
        double* merged_dS = c.red + 2056;
        SBPX_HIP(hipMemsetAsync(merged_dS, 0, sizeof(double), c.stream));
        for (;;) {
            if (rounds++ > 24) throw StateError("merge selection keeps running out of table room");
            build_tables(c);
            MhParams P{};
            P.assign = c.bm.assign;
            P.dout = c.bm.dout;
            P.din = c.bm.din;
            P.deg = c.bm.deg;
            P.drift = c.bm.drift;
            P.B = B;
            P.rbase = h.rbase;
            P.rcap = h.rcap;
            P.rkeys = h.rkeys;
            P.rvals = h.rvals;
            P.rocc = h.rocc;
            P.cbase = h.cbase;
            P.ccap = h.ccap;
            P.ckeys = h.ckeys;
            P.cvals = h.cvals;
            P.cocc = h.cocc;
            P.pool = (int)h.pool_limit;
            P.status = h.status;
            P.err = c.dev_err;
            merge_heap_kernel<<<1, WAVE, 0, c.stream>>>(P, d_heap, len, d_parent, merged, num_to_merge, pending_r,
                                                       pending_s, merged_dS);
            SBPX_HIP(hipGetLastError());
            int stat[9];
            SBPX_HIP(hipMemcpyAsync(stat, h.status, sizeof(stat), hipMemcpyDeviceToHost, c.stream));
            SBPX_HIP(hipStreamSynchronize(c.stream));
            check_device_error(c);
            merged = stat[0];
            len = stat[8];
            if (stat[1] == 0) break;
            // The pool ran out: bring the matrix snapshot up to date with the merges so far
            // (assignment through the union-find; merged-away blocks become empty), rebuild the
            // tables with more headroom and resume with the accepted merge.
            pending_r = stat[6];
            pending_s = stat[7];
            parent.resize(B);
            SBPX_HIP(hipMemcpyAsync(parent.data(), d_parent, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
            SBPX_HIP(hipStreamSynchronize(c.stream));
            std::vector<int> map(B);
            for (int x = 0; x < B; ++x) {
                int y = x;
                while (parent[y] != y) y = parent[y];
                map[x] = y;
            }
            SBPX_HIP(hipMemcpyAsync(c.m_map, map.data(), sizeof(int) * B, hipMemcpyHostToDevice, c.stream));
            launch_relabel(c);
            build_blockmodel(c);
            h.headroom *= 2;
        }
        parent.resize(B);
        SBPX_HIP(hipMemcpyAsync(parent.data(), d_parent, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipStreamSynchronize(c.stream));
    } catch (...) {
        (void)hipFree(d_heap);
        (void)hipFree(d_parent);
        throw;
    }
    SBPX_HIP(hipFree(d_heap));
    SBPX_HIP(hipFree(d_parent));
}

// The window-check kernel for GW wavefronts per entry, plain or learned.
template <int GW>
static const void* wide_kernel_gw(bool learned) {
    return learned ? (const void*)mh_wide_kernel<GW, true> : (const void*)mh_wide_kernel<GW, false>;
}
static const void* wide_kernel(int gw, bool learned, bool team = false) {
    if (team) return learned ? (const void*)mh_wide_kernel<8, true, true> : (const void*)mh_wide_kernel<8, false, true>;
    switch (gw) {
        case 1: return wide_kernel_gw<1>(learned);
        case 2: return wide_kernel_gw<2>(learned);
        case 4: return wide_kernel_gw<4>(learned);
        default: return wide_kernel_gw<8>(learned);
    }
}

// One chunk of the sequential pass on one workgroup with G wavefronts per vertex.
template <int G>
static void launch_group(const MhParams& P, bool learned, hipStream_t st) {
    if (learned)
        mh_group_kernel<G, true><<<1, MH_WAVES * WAVE, 0, st>>>(P);
    else
        mh_group_kernel<G, false><<<1, MH_WAVES * WAVE, 0, st>>>(P);
}

static LearnedNet learned_net(Ctx& c) {
    return LearnedNet{c.mlp.w16[0], c.mlp.w16[1], c.mlp.w16[2], c.mlp.b[0], c.mlp.b[1], c.mlp.b[2],
                      1.0f / log1pf((float)c.g.E),
                      c.g.N > 1 ? (float)(log((double)c.bm.B) / log((double)c.g.N)) : 0.0f,
                      std::min(c.learn_maxdeg, LEARN_MH_MAXDEG)};
}

void learned_scores_probe(Ctx& c, const int* d_list, int n, int max_deg, int* d_cand, float* d_z) {
    if (n <= 0) return;
    int tcap = 16;  // max_deg counts a self-loop once; vertex_table_cap counts it in both lists
    while (tcap < 2 * (max_deg + 3)) tcap <<= 1;
    int* tables = nullptr;
    SBPX_HIP(hipMalloc((void**)&tables, sizeof(int) * 2 * (size_t)tcap * n));
    const LearnedNet net = learned_net(c);
    const GraphView G{c.g.out_ptr, c.g.out_idx, c.g.in_ptr, c.g.in_idx, c.g.self_flag};
    const BlockState S{c.bm.assign, c.bm.dout, c.bm.din, c.bm.deg, c.bm.B};
    scorer_probe_kernel<<<(n + PROBE_WAVES - 1) / PROBE_WAVES, PROBE_WAVES * WAVE, 0, c.stream>>>(
        G, S, d_list, n, tables, tcap, net, d_cand, d_z);
    const hipError_t e = hipGetLastError();
    SBPX_HIP(hipStreamSynchronize(c.stream));
    SBPX_HIP(hipFree(tables));
    SBPX_HIP(e);
}

void mh_sequential(Ctx& c, const int* d_list, int n, int max_deg, uint64_t seed, uint32_t call, uint32_t tag,
                   long long* bytes, bool timed) {
    if (n <= 0) return;
    MhState& h = c.mh;
    const int B = c.bm.B;
    // max_deg counts a self-loop once; vertex_table_cap counts it in both lists
    int tcap = 16;
    while (tcap < 2 * (max_deg + 3)) tcap <<= 1;
    const bool wide = c.mh_wide != 1;  // the multi-workgroup pass may run
    if (wide && !h.wide) {
        SBPX_HIP(hipMalloc(&h.wide, sizeof(WideSlot) * MH_WIDE_SLOTS));
        SBPX_HIP(hipMalloc((void**)&h.wst, sizeof(int) * MH_WIDE_SLOTS));
        // the most workgroups of the window kernels that fit on the device at once (a
        // cooperative launch needs them all resident)
        int per_cu = 1 << 20, per_cu_l = 1 << 20, cus = 1;
        for (int gw = 1; gw <= 16; gw *= 2) {  // 16: the team kernels
            int p = 1, pl = 1;
            SBPX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, wide_kernel(std::min(gw, 8), false, gw == 16),
                                                                  MH_WIDE_WAVES * WAVE, 0));
            SBPX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pl, wide_kernel(std::min(gw, 8), true, gw == 16),
                                                                  MH_WIDE_WAVES * WAVE, 0));
            per_cu = std::min(per_cu, std::max(p, 1));
            per_cu_l = std::min(per_cu_l, std::max(pl, 1));
        }
        SBPX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
        h.k_max = 1;
        while (2 * h.k_max <= std::min(MH_WIDE_MAX, per_cu * cus)) h.k_max *= 2;
        h.k_max_l = 1;
        while (2 * h.k_max_l <= std::min(MH_WIDE_MAX, per_cu_l * cus)) h.k_max_l *= 2;
        SBPX_HIP(hipMalloc((void**)&h.ctrl, sizeof(int) * 4));
        SBPX_HIP(hipMalloc((void**)&h.team_hist, sizeof(int) * 512 * (size_t)std::max(h.k_max, h.k_max_l)));
        SBPX_HIP(hipMalloc((void**)&h.team_part, sizeof(double) * 2 * (size_t)std::max(h.k_max, h.k_max_l)));
        SBPX_HIP(hipMalloc((void**)&h.team_arrivals, sizeof(int) * (size_t)std::max(h.k_max, h.k_max_l)));
        SBPX_HIP(hipMalloc((void**)&h.wstamp, sizeof(unsigned long long) * std::max(c.g.N, 1)));
        SBPX_HIP(hipMemsetAsync(h.wstamp, 0, sizeof(unsigned long long) * std::max(c.g.N, 1), c.stream));
        SBPX_HIP(hipMemsetAsync(h.status + 9, 0, sizeof(int), c.stream));
    }
    // one table per wavefront of the widest window (or of one workgroup)
    const size_t tslots =
        (size_t)(wide ? std::max(std::max(h.k_max, h.k_max_l) * MH_WIDE_WAVES, MH_WAVES) : MH_WAVES) * 7 * tcap;
    if (tslots > h.table_slots) {
        if (h.tables) SBPX_HIP(hipFree(h.tables));
        SBPX_HIP(hipMalloc((void**)&h.tables, sizeof(int) * tslots));
        h.table_slots = tslots;
    }
    // the learned mixture: scored per evaluation on the current state, from input-major BF16 weights
    const bool learned = c.proposal >= 2;
    LearnedNet net{};
    if (learned) {
        if (!c.mlp.ready) throw StateError("learned proposal selected but no network initialised");
        if (c.mh_wide == 1) throw StateError("the learned sequential pass runs in the window-check kernel (option 6 != 1)");
        net = learned_net(c);
    }
    if (!h.assign0) SBPX_HIP(hipMalloc((void**)&h.assign0, sizeof(int) * std::max(c.g.N, 1)));
    SBPX_HIP(hipMemcpyAsync(h.assign0, c.bm.assign, sizeof(int) * c.g.N, hipMemcpyDeviceToDevice, c.stream));
    int pos = 0;
    int rounds = 0;
    // [5] delta-MDL row/column slots scanned, evaluations, windows, proposal row slots, and the sum
    // over windows of the most row slots one of its evaluations visited, in the current chunk
    double* scan = c.red + 2064;
    // SBPX_TRACE_MH=1: one line per pass on stderr (windows, entries per window, workgroups and
    // wavefronts per vertex chosen, kernel time)
    static const bool trace = std::getenv("SBPX_TRACE_MH") != nullptr;
    double tr_windows = 0.0, tr_scan = 0.0, tr_prop = 0.0, tr_evals = 0.0, tr_k = 0.0, tr_ms = 0.0;
    int tr_chunks = 0, tr_g1 = 0, tr_team = 0;
    hipEvent_t tr_e0 = nullptr, tr_e1 = nullptr;
    if (trace) {
        SBPX_HIP(hipEventCreate(&tr_e0));
        SBPX_HIP(hipEventCreate(&tr_e1));
        SBPX_HIP(hipMemsetAsync(c.red + 3000, 0, 11 * sizeof(double), c.stream));
    }
    while (pos < n) {
        // each restart doubles the pool headroom, so the number of restarts is logarithmic
        if (rounds++ > 24) throw StateError("Metropolis-Hastings pass keeps running out of table room");
        build_tables(c);
        MhParams P{};
        P.G = GraphView{c.g.out_ptr, c.g.out_idx, c.g.in_ptr, c.g.in_idx, c.g.self_flag};
        P.assign = c.bm.assign;
        P.dout = c.bm.dout;
        P.din = c.bm.din;
        P.deg = c.bm.deg;
        P.drift = c.bm.drift;
        P.B = B;
        P.rbase = h.rbase;
        P.rcap = h.rcap;
        P.rkeys = h.rkeys;
        P.rvals = h.rvals;
        P.rocc = h.rocc;
        P.cbase = h.cbase;
        P.ccap = h.ccap;
        P.ckeys = h.ckeys;
        P.cvals = h.cvals;
        P.cocc = h.cocc;
        P.pool = (int)h.pool_limit;
        P.list = d_list;
        P.n = n;
        P.seed = seed;
        P.call = call;
        P.tag = tag;
        P.tables = h.tables;
        P.tcap = tcap;
        P.status = h.status;
        P.counters = c.counters;
        P.bytes = bytes;
        P.touch = h.touch;
        P.scan = scan;
        P.dS_sum = c.red + 2048;
        P.err = c.dev_err;
        P.wide = (WideSlot*)h.wide;
        P.wst = h.wst;
        P.wstamp = h.wstamp;
        P.ctrl = h.ctrl;
        P.team = 1;
        P.team_hist = h.team_hist;
        P.team_part = h.team_part;
        P.team_arrivals = h.team_arrivals;
        P.net = net;
        P.phase = trace ? c.red + 3000 : nullptr;
        bool out_of_room = false;
        // The queue runs in chunks; before each, the wavefronts per entry (G) and the workgroups
        // (K) are chosen from the previous chunk: G by the mean number of row/column slots its
        // evaluations scanned (long scans gain from more lanes per vertex, short ones from more
        // vertices in flight), K so that a window holds about 1.5 times the queue entries the
        // previous chunk's windows advanced before a commit invalidated the rest. Windows of about
        // one entry with long scans take one workgroup of 16 wavefronts. The first chunk at a new
        // number of blocks goes by the longest row. With option 6 = 1 the one-workgroup kernels
        // run instead (16 / G entries per window).
        // Scans of tens of thousands of slots (rows of a block that holds a large share of the
        // graph, as a merge can leave behind) take teams: TS workgroups of 8 wavefronts share each
        // entry, about 4096 slots per member of the largest evaluation a window held in the
        // previous chunk (on average over its windows), up to 32 (option 10 fixes TS).
        while (pos < n) {
            int G = c.mh_group;
            const int k_cap = learned ? h.k_max_l : h.k_max;
            // slots an evaluation visits in the window kernel: the delta-MDL rows once, the
            // proposal's row once per pass of its radix selection
            int nbits = 1;
            while (nbits < 31 && (1 << nbits) < B) ++nbits;
            const double visits = h.mean_scan + ((nbits + 7) / 8) * h.mean_prop;
            const bool fresh = h.last_B != B || h.mean_scan < 0.0;
            if (fresh) h.y_est = std::min(h.y_est, 16.0);
            if (G == 0) {
                if (fresh) {
                    G = h.max_cap >= 16384 ? 16 : (h.max_cap >= 2048 ? 4 : (wide ? 2 : 1));
                } else if (wide) {
                    G = h.mean_scan >= 2400.0 ? 8 : (h.mean_scan >= 1200.0 ? 4 : 2);
                    if (G == 8 && h.y_est < 1.5) G = 16;
                } else {
                    G = h.mean_scan >= 4000.0 ? 16 : (h.mean_scan >= 1500.0 ? 4 : 1);
                    G = std::max(G, h.g_yield);
                }
            }
            int TS = 1;
            if (wide && c.mh_team >= 2) {
                TS = c.mh_team;
            } else if (wide && c.mh_team == 0 && c.mh_group == 0) {
                // a window waits for its evaluation with the most row slots: its typical size in
                // the previous chunk sets the team (from the window kernel's counts; else the mean)
                const double sc = fresh ? (double)h.max_cap : (h.mean_wmax > 0.0 ? h.mean_wmax : visits);
                if (sc >= MH_TEAM_SCAN) {
                    TS = 2;
                    while (TS < 32 && TS * (MH_TEAM_SCAN / 2) < sc) TS *= 2;
                }
                // a team shrinks the window to k_cap / TS entries: only as far as the entries the
                // windows advance still fit
                const double want = std::max(1.0, 1.5 * h.y_est);
                int ts_cap = 1;
                while (2.0 * ts_cap * want <= (double)k_cap) ts_cap *= 2;
                TS = std::min(TS, ts_cap);
                if (TS < 2) TS = 1;
            }
            TS = std::min(TS, k_cap);
            if (TS > 1) G = 8;
            const bool wk = wide && G <= 8;  // the window-check kernel
            int K = 1;
            if (wk && TS > 1) {
                if (c.mh_wide) {
                    K = std::min(c.mh_wide, k_cap);
                    TS = std::min(TS, K);
                } else {
                    int E = 1;  // entries per window
                    while (2 * E * TS <= k_cap && E < 1.5 * h.y_est) E *= 2;
                    K = E * TS;
                }
            } else if (wk) {
                if (c.mh_wide) {
                    K = std::min(c.mh_wide, k_cap);
                } else {
                    const double want = 1.5 * h.y_est * G / MH_WIDE_WAVES;
                    while (K < k_cap && K < want) K *= 2;
                }
            }
            P.team = TS;
            h.last_team = TS;
            P.start = pos;
            // a launch covers at least MH_CHUNK entries and four full windows
            P.stop_at = std::min(n, pos + std::max(MH_CHUNK, 4 * (MH_WIDE_WAVES / std::min(G, 8)) * K));
            SBPX_HIP(hipMemsetAsync(scan, 0, 5 * sizeof(double), c.stream));
            if (wk) wide_init_kernel<<<1, 256, 0, c.stream>>>(h.ctrl, h.team_arrivals, TS > 1 ? K / TS : 0);
            if (timed) c.timer_begin();
            if (trace) SBPX_HIP(hipEventRecord(tr_e0, c.stream));
            if (wk) {
                void* args[] = {&P};
                const void* kern = wide_kernel(G, learned, TS > 1);
                if (K > 1) {  // every workgroup resident at once: the grid barriers need them all
                    const hipError_t e =
                        hipLaunchCooperativeKernel(kern, dim3(K), dim3(MH_WIDE_WAVES * WAVE), args, 0, c.stream);
                    if (e != hipSuccess) {
                        // refused at launch (nothing ran), e.g. while another process on the device
                        // holds its cooperative-launch resources: this context goes on with one
                        // workgroup per window, which applies the same moves
                        (void)hipGetLastError();
                        h.k_max = h.k_max_l = 1;
                        K = 1;
                        TS = 1;
                        P.team = 1;
                        h.last_team = 1;
                        kern = wide_kernel(G, learned, false);
                        h.coop_refused++;
                    }
                }
                if (K == 1) SBPX_HIP(hipLaunchKernel(kern, dim3(1), dim3(MH_WIDE_WAVES * WAVE), args, 0, c.stream));
            } else {
                // the learned mixture runs in every chunk: the window-check kernel, or the group
                // kernels for wavefront groups
                if (learned && G == 1) throw StateError("learned sequential pass without the window-check kernel");
                switch (G) {
                    case 1: mh_kernel<<<1, MH_WAVES * WAVE, 0, c.stream>>>(P); break;
                    case 2: launch_group<2>(P, learned, c.stream); break;
                    case 4: launch_group<4>(P, learned, c.stream); break;
                    case 8: launch_group<8>(P, learned, c.stream); break;
                    default: launch_group<16>(P, learned, c.stream); break;
                }
            }
            SBPX_HIP(hipGetLastError());
            if (timed) c.timer_end(KERNEL_MH);
            if (trace) SBPX_HIP(hipEventRecord(tr_e1, c.stream));
            int stat[10];
            double sc[5];
            SBPX_HIP(hipMemcpyAsync(stat, h.status, sizeof(stat), hipMemcpyDeviceToHost, c.stream));
            SBPX_HIP(hipMemcpyAsync(sc, scan, sizeof(sc), hipMemcpyDeviceToHost, c.stream));
            SBPX_HIP(hipStreamSynchronize(c.stream));
            check_device_error(c);
            if (trace) {
                float ms = 0.0f;
                SBPX_HIP(hipEventElapsedTime(&ms, tr_e0, tr_e1));
                tr_ms += ms;
                tr_windows += sc[2];
                tr_scan += sc[0];
                tr_prop += sc[3];
                tr_evals += sc[1];
                tr_k += K * sc[2];
                tr_chunks++;
                tr_g1 += G == 1;
                tr_team += TS > 1;
            }
            const int advanced = stat[0] - pos;
            pos = stat[0];
            if (sc[1] > 0.0) {
                h.mean_scan = sc[0] / sc[1];
                h.mean_prop = sc[3] / sc[1];
                h.mean_wmax = sc[2] > 0.0 ? sc[4] / sc[2] : 0.0;
                h.last_B = B;
            }
            // Entries a window advanced. For the one-workgroup kernels (16 / G vertices per
            // window): near that cap the windows could be longer, so G halves; below it about 3 or
            // fewer standing entries take 4 wavefronts per vertex, up to 6 two, more one.
            if (sc[2] > 0.0 && advanced > 0) {
                const double y = (double)advanced / sc[2];
                // windows that ran full say only that longer ones may stand: ask for twice as many
                const double full = wk ? (double)(K / TS * (MH_WIDE_WAVES / G)) : (double)(MH_WAVES / G);
                h.y_est = y >= 0.75 * full ? 2.0 * full : y;
                if (!wk) {
                    if (y >= 0.75 * (double)(MH_WAVES / G)) h.g_yield = std::max(1, G / 2);
                    else h.g_yield = y < 3.0 ? 4 : (y < 6.0 ? 2 : 1);
                } else {
                    h.g_yield = 1;
                }
            }
            if (stat[9] > (1 << 30)) {  // restart the window serials well before they wrap
                SBPX_HIP(hipMemsetAsync(h.wstamp, 0, sizeof(unsigned long long) * c.g.N, c.stream));
                SBPX_HIP(hipMemsetAsync(h.status + 9, 0, sizeof(int), c.stream));
            }
            if (stat[4] > (1 << 30)) {  // restart the commit serials well before they wrap
                SBPX_HIP(hipMemsetAsync(h.touch, 0, sizeof(int) * c.g.N, c.stream));
                SBPX_HIP(hipMemsetAsync(h.status + 4, 0, sizeof(int), c.stream));
            }
            if (stat[1] == 1) {
                out_of_room = true;
                break;
            }
            if (pos < P.stop_at) throw StateError("Metropolis-Hastings pass stopped early");
        }
        // the matrix changed: bring the CSR snapshot up to date (dout/din exact, deg keeps drift)
        update_after_pass(c, d_list, n);
        if (out_of_room) h.headroom *= 2;
    }
    if (trace) {
        double ph[11];
        SBPX_HIP(hipMemcpyAsync(ph, c.red + 3000, sizeof(ph), hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipStreamSynchronize(c.stream));
        std::fprintf(stderr, "mh phases (wide kernel, ms): evaluate %.2f checks %.2f commits %.2f (%.0f) barrier %.2f\n",
                     ph[0] * 1e-5, ph[1] * 1e-5, ph[2] * 1e-5, ph[4], ph[3] * 1e-5);
        if (ph[9] > 0.0)
            std::fprintf(stderr,
                         "mh slowest evaluation per window: %.1f us (learned scoring %.1f us), delta-MDL slots %.0f, "
                         "proposal slots %.0f, degree %.0f (%.0f windows)\n",
                         ph[5] / ph[9] * 1e-2, ph[10] / ph[9] * 1e-2, ph[6] / ph[9], ph[7] / ph[9], ph[8] / ph[9],
                         ph[9]);
        std::fprintf(stderr,
                     "mh pass: B=%d n=%d chunks=%d (G=1 in %d, teams in %d) windows=%.0f entries/window=%.1f "
                     "workgroups/window=%.2f evals=%.0f slots/eval=%.0f proposal slots/eval=%.0f kernel=%.2f ms\n",
                     B, n, tr_chunks, tr_g1, tr_team, tr_windows, tr_windows > 0 ? n / tr_windows : 0.0,
                     tr_windows > 0 ? tr_k / tr_windows : 0.0, tr_evals, tr_evals > 0 ? tr_scan / tr_evals : 0.0,
                     tr_evals > 0 ? tr_prop / tr_evals : 0.0,
                     tr_ms);
        (void)hipEventDestroy(tr_e0);
        (void)hipEventDestroy(tr_e1);
    }
}

}  // namespace sbpx
