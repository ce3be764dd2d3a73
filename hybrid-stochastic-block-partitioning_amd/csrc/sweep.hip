// Batched vertex-move evaluation (one wavefront per vertex, asynchronous Gibbs semantics:
// every vertex of a batch is evaluated against the same snapshot) and batch application. The
// evaluation itself is eval_move in eval_core.hpp, read through the CSR snapshot view.
#include <algorithm>
#include <climits>

#include "eval_core.hpp"
#include "eval_group.hpp"
#include "internal.hpp"

namespace sbpx {

struct EvalParams {
    GraphView G;
    BlockState S;
    CsrView M;
    const int* __restrict__ list;
    Dyn n;
    // graph mode: the list starts at list + *list_off and the huge-table offsets at pool_off +
    // *pool_off_at (device words the host fills per round); persistent workgroups take list
    // entries from the work counter *queue until it passes n
    const int* __restrict__ list_off;
    const int* __restrict__ pool_off_at;
    int* __restrict__ queue;
    const int* __restrict__ given_s;
    const double* __restrict__ given_u;
    uint64_t seed;
    uint32_t call;
    uint32_t tag;
    int out_by_vertex;  // 1: outputs at index v; 0: at out_offset + position
    int out_offset;
    int propose_only;
    int* __restrict__ out_s;
    uint8_t* __restrict__ out_stat;  // 0 no move proposed, 1 evaluated and rejected, 2 accepted
    double* __restrict__ out_dS;
    double* __restrict__ out_H;
    double* __restrict__ out_u;
    long long* __restrict__ bytes;
    int* __restrict__ err;
    int* __restrict__ pool;
    const int* __restrict__ pool_off;
    // learned mixture (vertex-indexed): proposals, uniforms, net probabilities; lw < 0: plain
    const int* __restrict__ lps;
    const double* __restrict__ lpu;
    const float* __restrict__ lqf;
    const float* __restrict__ lqb;
    float lw;
};

// Stores one evaluation's outputs (one lane).
__device__ void write_result(const EvalParams& P, const MoveResult& R, int v, int idx) {
    if (R.status < 0) {
        atomicOr(P.err, DEV_ERR_TABLE);
        return;
    }
    if (R.status >= 1 && (__builtin_isnan(R.dS) || __builtin_isinf(R.dS))) atomicOr(P.err, DEV_ERR_NAN);
    const int o = P.out_by_vertex ? v : P.out_offset + idx;
    P.out_s[o] = R.s;
    if (P.out_stat) P.out_stat[o] = (uint8_t)R.status;
    if (P.out_dS) P.out_dS[o] = R.dS;
    if (P.out_H) P.out_H[o] = R.H;
    if (P.out_u) P.out_u[o] = R.u;
    if (P.bytes) {
        // algorithmic bytes per evaluation: CSR/CSC pointers, own block, neighbour ids and blocks,
        // proposal rows (row t + column t), delta-MDL rows (rows/columns r and s), degree
        // lookups and the Hastings cells, result record; an attempt with s == r stops after
        // the proposal.
        long long k = R.k_out + R.k_in;
        long long nnz_t = R.t >= 0 ? (long long)P.M.nnz_row(R.t) + P.M.nnz_col(R.t) : 0;
        long long b = 16 + 4 + 8 * k + 8 * nnz_t;
        if (R.status >= 1) b += 8 * R.nnz_rs + 12LL * (R.nuniq + 2) + 8LL * R.nuniq + 16;
        atomicAdd((unsigned long long*)P.bytes, (unsigned long long)b);
    }
}

// LRN: the learned mixture's draws and probabilities (P.lw >= 0), in instantiations of their own so
// that the plain ones carry none of the mixture's registers.
template <int U = 2, int W = WAVE, bool LRN = false>
__device__ __forceinline__ void eval_vertex(const EvalParams& P, const int* list, int* tab, int cap, int idx) {
    const int v = list[idx];
    int gs = P.given_s ? P.given_s[idx] : -1;
    double gu = P.given_u ? P.given_u[idx] : 0.0;
    double lqf = 0.0, lqb = 0.0;
    if constexpr (LRN) {
        gs = P.lps[v];
        gu = P.lpu[v];
        lqf = P.lqf[v];
        lqb = P.lqb[v];
    }
    MoveResult R = eval_move<CsrView, U, W>(P.G, P.S, P.M, v, tab, cap, P.seed, P.call, P.tag, gs, gu,
                                            P.propose_only != 0, LRN ? (double)P.lw : -1.0, lqf, lqb);
    if (group_lane<W>() == 0) write_result(P, R, v, idx);
}

__device__ __forceinline__ const int* eval_list(const EvalParams& P) {
    return P.list_off ? P.list + *P.list_off : P.list;
}

// The next `take` list entries of a persistent wavefront (graph mode): one atomic per wavefront
// on the launch's work counter; the first of them, the same in every lane.
__device__ __forceinline__ int next_entries(int* queue, int take) {
    int idx = 0;
    if (lane_id() == 0) idx = atomicAdd(queue, take);
    return __shfl(idx, 0, WAVE);
}

// One wavefront per vertex, its neighbour-block table in LDS (CAP slots); WAVES wavefronts per
// workgroup share nothing but the launch. U: row slots per lane per trip of the scans. PERS (graph
// mode): the workgroups stay resident and each wavefront takes vertices from the work counter
// until the list is done; a separate instantiation, since the loop's live state would cost the
// one-vertex launches registers (and occupancy).
template <int CAP, int WAVES, int U, bool PERS = false, bool LRN = false>
__global__ void __launch_bounds__(WAVES * WAVE) eval_bin_kernel(EvalParams P) {
    __shared__ int tab[WAVES][7 * CAP];
    const int wave = threadIdx.x / WAVE;
    if constexpr (PERS) {
        const int n = P.n.get();
        const int* list = eval_list(P);
        for (int idx = next_entries(P.queue, 1); idx < n; idx = next_entries(P.queue, 1))
            eval_vertex<U, WAVE, LRN>(P, list, tab[wave], CAP, idx);
    } else {
        const int idx = blockIdx.x * WAVES + wave;
        if (idx >= P.n.v) return;
        eval_vertex<U, WAVE, LRN>(P, P.list, tab[wave], CAP, idx);
    }
}

// Low-degree vertices packed WAVE / W to a wavefront: a group of W lanes per vertex, each with its
// own CAP-slot table in LDS. The evaluation is a chain of dependent gathers, so more vertices in
// flight per CU (not more lanes per vertex) is what shortens the launch.
template <int CAP, int WAVES, int U, int W, bool PERS = false, bool LRN = false>
__global__ void __launch_bounds__(WAVES * WAVE) eval_sub_kernel(EvalParams P) {
    constexpr int PER = WAVES * (WAVE / W);
    constexpr int SUBS = WAVE / W;  // vertices per wavefront
    __shared__ int tab[PER][7 * CAP];
    const int sub = threadIdx.x / W;
    if constexpr (PERS) {
        const int n = P.n.get();
        const int* list = eval_list(P);
        // the wavefront takes SUBS entries at a time, one per lane group; the groups run the same
        // trip count (the evaluation's wave-level steps need every lane)
        for (int base = next_entries(P.queue, SUBS); base < n; base = next_entries(P.queue, SUBS)) {
            const int idx = base + (lane_id() / W);
            if (idx < n) eval_vertex<U, W, LRN>(P, list, tab[sub], CAP, idx);
        }
    } else {
        const int idx = blockIdx.x * PER + sub;
        if (idx >= P.n.v) return;
        eval_vertex<U, W, LRN>(P, P.list, tab[sub], CAP, idx);
    }
}

// Vertices of the highest degrees: G wavefronts (the whole workgroup) share one vertex, its
// table in LDS (CAP slots) or, with CAP = 0, in global memory at a per-vertex offset. LRN: the
// learned mixture's draws and probabilities (P.lw >= 0) in a separate instantiation.
template <int G, int CAP, bool LRN>
__device__ void eval_group_entry(const EvalParams& P, const int* list, int idx, int* tab_lds, int* hist,
                                 double (*redd)[RED_D], int (*redi)[RED_I]) {
    const int v = list[idx];
    int* tab = tab_lds;
    int cap = CAP;
    if constexpr (CAP == 0) {
        const int* pool_off = P.pool_off_at ? P.pool_off + *P.pool_off_at : P.pool_off;
        const int base = pool_off[idx];
        cap = (pool_off[idx + 1] - base) / 7;
        tab = P.pool + base;
    }
    MoveResult R;
    if constexpr (LRN) {
        R = eval_group<G, CsrView, true>(P.G, P.S, P.M, v, tab, cap, P.seed, P.call, P.tag, P.lps[v], P.lpu[v],
                                         P.propose_only != 0, hist, redd, redi, (double)P.lw, nullptr, nullptr,
                                         (double)P.lqf[v], (double)P.lqb[v]);
    } else {
        const int gs = P.given_s ? P.given_s[idx] : -1;
        const double gu = P.given_u ? P.given_u[idx] : 0.0;
        R = eval_group<G, CsrView>(P.G, P.S, P.M, v, tab, cap, P.seed, P.call, P.tag, gs, gu, P.propose_only != 0,
                                   hist, redd, redi);
    }
    if (threadIdx.x == 0) write_result(P, R, v, idx);
}

// PERS (graph mode, an instantiation of its own): the workgroup stays resident and takes vertices
// from the work counter; the two barriers around the shared index keep the previous vertex's LDS
// in use until every wavefront has finished it.
template <int G, int CAP, bool LRN = false, bool PERS = false>
__global__ void __launch_bounds__(G * WAVE) eval_group_kernel(EvalParams P) {
    __shared__ int tab_lds[CAP > 0 ? 7 * CAP : 1];
    __shared__ int hist[256];
    __shared__ double redd[G][RED_D];
    __shared__ int redi[G][RED_I];
    if constexpr (PERS) {
        __shared__ int s_idx;
        const int n = P.n.get();
        const int* list = eval_list(P);
        for (;;) {
            if (threadIdx.x == 0) s_idx = atomicAdd(P.queue, 1);
            __syncthreads();
            const int idx = s_idx;
            __syncthreads();
            if (idx >= n) return;
            eval_group_entry<G, CAP, LRN>(P, list, idx, tab_lds, hist, redd, redi);
        }
    } else {
        const int idx = blockIdx.x;
        if (idx >= P.n.v) return;
        eval_group_entry<G, CAP, LRN>(P, P.list, idx, tab_lds, hist, redd, redi);
    }
}

// The group kernel for one launch: graph mode (a work counter) or one workgroup per vertex.
template <int G, int CAP, bool LRN = false>
static void launch_group_eval(const EvalParams& P, int grid, hipStream_t st) {
    if (P.queue)
        eval_group_kernel<G, CAP, LRN, true><<<grid, G * WAVE, 0, st>>>(P);
    else
        eval_group_kernel<G, CAP, LRN, false><<<grid, G * WAVE, 0, st>>>(P);
}

// Workgroups of an evaluation launch that covers `per` vertices each: every vertex its own
// workgroup slot for a host-known count; in graph mode at most PERSIST_GRID persistent workgroups.
constexpr int PERSIST_GRID = 1024;
static int eval_grid(const Dyn& n, int per) {
    const long g = ((long)n.v + per - 1) / per;
    return (int)std::max<long>(1, n.p ? std::min<long>(g, PERSIST_GRID) : g);
}

// PERS: graph mode (a work counter); LRN: the learned mixture. Graph mode runs batches of graphs,
// which the learned proposal does not, so the two are never combined.
template <int CAP, int WAVES, bool PERS, bool LRN>
static void launch_bin_as(const EvalParams& P, int variant, hipStream_t st) {
    if constexpr (CAP == BIN_CAP[0]) {
        if (variant == 2) {
            constexpr int W = 16, SW = 4, PER = SW * (WAVE / W);
            eval_sub_kernel<CAP, SW, 2, W, PERS, LRN><<<eval_grid(P.n, PER), SW * WAVE, 0, st>>>(P);
            return;
        }
    }
    if (variant == 1)
        eval_bin_kernel<CAP, WAVES, 4, PERS, LRN><<<eval_grid(P.n, WAVES), WAVES * WAVE, 0, st>>>(P);
    else
        eval_bin_kernel<CAP, WAVES, 2, PERS, LRN><<<eval_grid(P.n, WAVES), WAVES * WAVE, 0, st>>>(P);
}
template <int CAP, int WAVES>
static void launch_bin(const EvalParams& P, int variant, hipStream_t st) {
    if (P.queue) {
        if (P.lw >= 0.0f) throw StateError("graph-mode evaluation with the learned proposal");
        launch_bin_as<CAP, WAVES, true, false>(P, variant, st);
    } else if (P.lw >= 0.0f) {
        launch_bin_as<CAP, WAVES, false, true>(P, variant, st);
    } else {
        launch_bin_as<CAP, WAVES, false, false>(P, variant, st);
    }
}

static EvalParams make_params(Ctx& c) {
    EvalParams P{};
    P.G = GraphView{c.g.out_ptr, c.g.out_idx, c.g.in_ptr, c.g.in_idx, c.g.self_flag};
    P.S = BlockState{c.bm.assign, c.bm.dout, c.bm.din, c.bm.deg, c.bm.B};
    if (c.batch.G > 0) {
        const DevBatch& b = c.batch;
        P.S.vgraph = b.d_vgraph;
        P.S.gvo = b.d_vo;
        P.S.gbo = b.d_bo;
        P.S.gB = b.d_nb;
        P.S.gseed = b.d_seed;
        P.S.gcall = b.d_call;
    }
    P.M = CsrView{c.bm.row_ptr, c.bm.row_col, c.bm.row_val, c.bm.row_cum,
                  c.bm.col_ptr, c.bm.col_row, c.bm.col_val, c.bm.col_cum, c.bm.row_h, c.bm.col_h};
    P.err = c.dev_err;
    P.pool = c.table_pool;
    P.lw = -1.0f;
    return P;
}

void launch_eval(Ctx& c, const EvalLaunch& L) {
    if (L.n <= 0) return;
    EvalParams P = make_params(c);
    P.list = L.list;
    P.n = L.n_at ? Dyn(L.n, L.n_at) : Dyn(L.n);
    P.list_off = L.list_off;
    P.pool_off_at = L.huge_off_at;
    P.queue = L.queue;
    if ((L.n_at != nullptr) != (L.queue != nullptr)) throw StateError("a device-held count needs a work counter");
    P.given_s = L.given_s;
    P.given_u = L.given_u;
    P.seed = L.seed;
    P.call = L.call;
    P.tag = L.tag;
    P.out_by_vertex = L.out_by_vertex;
    P.out_offset = L.out_offset;
    P.propose_only = L.propose_only;
    P.out_s = c.ev_s;
    P.out_stat = c.ev_acc;
    P.out_dS = c.ev_dS;
    P.out_H = c.ev_H;
    P.out_u = L.out_u;
    P.bytes = L.bytes;
    P.pool_off = L.huge_offsets;
    if (L.learned) {
        if (L.bin >= NUM_LDS_BINS) throw StateError("learned proposals are evaluated by the LDS-table bins only");
        if (L.queue) throw StateError("graph-mode evaluation with the learned proposal");
        P.lps = c.mlp.ps;
        P.lpu = L.learned_u ? L.learned_u : c.ev_u;
        P.lqf = c.mlp.qf;
        P.lqb = c.mlp.qb;
        P.lw = LEARN_W_PLAIN;
    }
    static_assert(BIN_CAP[0] == 64 && BIN_CAP[1] == 128 && BIN_CAP[2] == 256 && BIN_CAP[3] == 512 &&
                      BIN_CAP[4] == 1024,
                  "bin table sizes");
    const int g1 = eval_grid(P.n, 1);
    switch (L.bin) {
        case 0: launch_bin<64, 8>(P, L.variant, c.stream); break;
        case 1: launch_bin<128, 4>(P, L.variant, c.stream); break;
        case 2: launch_bin<256, 4>(P, L.variant, c.stream); break;
        case 3:
            if (L.variant == 1 && L.learned) eval_group_kernel<4, 512, true><<<g1, 4 * WAVE, 0, c.stream>>>(P);
            else if (L.variant == 1) launch_group_eval<4, 512>(P, g1, c.stream);
            else launch_bin<512, 2>(P, 0, c.stream);
            break;
        case 4:
            if (L.variant == 1 && L.learned) eval_group_kernel<4, 1024, true><<<g1, 4 * WAVE, 0, c.stream>>>(P);
            else if (L.variant == 1) launch_group_eval<4, 1024>(P, g1, c.stream);
            else launch_bin<1024, 1>(P, 0, c.stream);
            break;
        case 5: launch_group_eval<4, 4096>(P, g1, c.stream); break;
        default: launch_group_eval<16, 0>(P, g1, c.stream); break;
    }
    SBPX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------------------
// Batch application (finetune.cpp:58-71): every accepted move takes effect; the block matrix
// and dout/din are rebuilt; for each block x the total degree keeps the drift that
// Blockmodel::move_vertex (blockmodel.cpp:344-360) leaves when the moves are applied one by
// one in ascending vertex id: deg[x] = exact - Delta_last[x,x], where Delta_last is the Delta of
// the last move (highest vertex id) touching x, evaluated with the blocks neighbours have at
// that point (moved earlier in the batch => new block).
constexpr int APPLY_TB = 256;

__global__ void mark_moves_kernel(VList L, const int* __restrict__ out_s, const uint8_t* __restrict__ out_stat,
                                  const int* __restrict__ assign, Dyn ep, int* __restrict__ moved_epoch,
                                  int* __restrict__ new_block, unsigned long long* __restrict__ last_touch,
                                  long long* __restrict__ counters) {
    __shared__ long long cnt[3];
    if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int* list = L.ptr();
    const int epoch = ep.get();
    long long a = 0, e = 0, m = 0;
    GRID_LOOP(i, L.n.get()) {
        int v = list[i];
        int st = out_stat[v];
        a += 1;
        e += st >= 1;
        m += st == 2;
        if (st == 2) {
            int s = out_s[v], r = assign[v];
            moved_epoch[v] = epoch;
            new_block[v] = s;
            unsigned long long key = ((unsigned long long)(unsigned)epoch << 32) | (unsigned)v;
            atomicMax(&last_touch[r], key);
            atomicMax(&last_touch[s], key);
        }
    }
    a = wave_sum_ll(a);
    e = wave_sum_ll(e);
    m = wave_sum_ll(m);
    if (lane_id() == 0) {
        atomicAdd((unsigned long long*)&cnt[0], (unsigned long long)a);
        atomicAdd((unsigned long long*)&cnt[1], (unsigned long long)e);
        atomicAdd((unsigned long long*)&cnt[2], (unsigned long long)m);
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicAdd((unsigned long long*)&counters[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

// One wavefront per batch vertex; only the last mover of a block computes its drift.
__global__ void drift_kernel(VList L, const uint8_t* __restrict__ out_stat,
                             const int* __restrict__ out_ptr, const int* __restrict__ out_idx,
                             const int* __restrict__ in_ptr, const int* __restrict__ in_idx,
                             const int* __restrict__ assign, Dyn ep, const int* __restrict__ moved_epoch,
                             const int* __restrict__ new_block, const unsigned long long* __restrict__ last_touch,
                             int* __restrict__ drift) {
    const int* list = L.ptr();
    const int n = L.n.get(), epoch = ep.get();
    const int wpb = blockDim.x / WAVE;
    for (int idx = blockIdx.x * wpb + threadIdx.x / WAVE; idx < n; idx += gridDim.x * wpb) {
    const int v = list[idx];
    if (out_stat[v] != 2) continue;
    const int lane = lane_id();
    const int r = assign[v], s = new_block[v];
    const unsigned long long key = ((unsigned long long)(unsigned)epoch << 32) | (unsigned)v;
    const bool do_r = last_touch[r] == key, do_s = last_touch[s] == key;
    if (!do_r && !do_s) continue;
    int cr = 0, cs = 0, self = 0;
    for (int j = out_ptr[v] + lane; j < out_ptr[v + 1]; j += WAVE) {
        int w = out_idx[j];
        if (w == v) {
            self = 1;
            continue;
        }
        int b = (w < v && moved_epoch[w] == epoch) ? new_block[w] : assign[w];
        cr += b == r;
        cs += b == s;
    }
    for (int j = in_ptr[v] + lane; j < in_ptr[v + 1]; j += WAVE) {
        int w = in_idx[j];
        if (w == v) continue;
        int b = (w < v && moved_epoch[w] == epoch) ? new_block[w] : assign[w];
        cr += b == r;
        cs += b == s;
    }
    cr = (int)wave_sum_ll(cr);
    cs = (int)wave_sum_ll(cs);
    self = wave_or(self);
    if (lane == 0) {
        if (do_r) drift[r] = -(cr + self);
        if (do_s) drift[s] = cs + self;
    }
    }
}

__global__ void commit_moves_kernel(VList L, const uint8_t* __restrict__ out_stat, const int* __restrict__ new_block,
                                    int* __restrict__ assign) {
    const int* list = L.ptr();
    GRID_LOOP(i, L.n.get()) {
        int v = list[i];
        if (out_stat[v] == 2) assign[v] = new_block[v];
    }
}

// Sums dS over the accepted moves of the batch (vertex-indexed outputs): SUM_BLOCKS partial
// sums over a grid-stride split, then one workgroup adds them in order (deterministic).
constexpr int SUM_BLOCKS = 256;
__global__ void sum_accepted_kernel(VList L, const uint8_t* __restrict__ out_stat, const double* __restrict__ out_dS,
                                    double* __restrict__ partial) {
    __shared__ double lds[APPLY_TB / WAVE];
    double acc = 0.0;
    const int* list = L.ptr();
    GRID_LOOP(i, L.n.get()) {
        int v = list[i];
        if (out_stat[v] == 2) acc += out_dS[v];
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < APPLY_TB / WAVE; ++w) s += lds[w];
        partial[blockIdx.x] = s;
    }
}

__global__ void add_partials_kernel(const double* __restrict__ partial, int n, double* __restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += WAVE) acc += partial[i];
    acc = wave_sum(acc);
    if (threadIdx.x == 0) out[0] += acc;
}

// Start of a round replayed from a graph: the sweep counters, the accepted-dS sum and the work
// counters of the persistent evaluation launches start at zero (kernel stores: no memset nodes).
__global__ void round_zero_kernel(long long* __restrict__ counters, double* __restrict__ dsum, int* __restrict__ queue,
                                  int nq) {
    for (int i = threadIdx.x; i < 8; i += blockDim.x) counters[i] = 0;
    if (threadIdx.x == 0) *dsum = 0.0;
    for (int i = threadIdx.x; i < nq; i += blockDim.x) queue[i] = 0;
}

void launch_round_zero(Ctx& c, int* queue, int nq) {
    round_zero_kernel<<<1, APPLY_TB, 0, c.stream>>>(c.counters, c.red + 2048, queue, nq);
}

// Marks the accepted moves, sums their dS and computes the degree drift (before the commit).
static void apply_marks(Ctx& c, const VList& L, Dyn epoch) {
    mark_moves_kernel<<<grid_for(L.n, APPLY_TB), APPLY_TB, 0, c.stream>>>(L, c.ev_s, c.ev_acc, c.bm.assign, epoch,
                                                                         c.moved_epoch, c.new_block, c.last_touch,
                                                                         c.counters);
    sum_accepted_kernel<<<SUM_BLOCKS, APPLY_TB, 0, c.stream>>>(L, c.ev_acc, c.ev_dS, c.red + 2560);
    add_partials_kernel<<<1, WAVE, 0, c.stream>>>(c.red + 2560, SUM_BLOCKS, c.red + 2048);
    const int wpb = APPLY_TB / WAVE;
    drift_kernel<<<grid_for(L.n, wpb), APPLY_TB, 0, c.stream>>>(L, c.ev_acc, c.g.out_ptr, c.g.out_idx, c.g.in_ptr,
                                                               c.g.in_idx, c.bm.assign, epoch, c.moved_epoch,
                                                               c.new_block, c.last_touch, c.bm.drift);
}

// list: the batch's vertices (any order); outputs are vertex-indexed.
void apply_accepted(Ctx& c, const int* list, int n) {
    if (n <= 0) return;
    c.epoch++;
    const VList L(list, n);
    apply_marks(c, L, Dyn(c.epoch));
    const int total = prepare_update(c, list, n);  // reads the assignment before the commit
    commit_moves_kernel<<<grid_for(L.n, APPLY_TB), APPLY_TB, 0, c.stream>>>(L, c.ev_acc, c.new_block, c.bm.assign);
    SBPX_HIP(hipGetLastError());
    if (total >= 0) finish_update(c, list, n, total);
    else rebuild_keep_drift(c);
}

// The same with the list's offset, length and epoch in device memory (graph mode): fixed grids,
// and the block-matrix update sized on the device.
void apply_accepted_dyn(Ctx& c, const VList& L, Dyn epoch) {
    apply_marks(c, L, epoch);
    prepare_update_dyn(c, L, epoch);
    commit_moves_kernel<<<grid_for(L.n, APPLY_TB), APPLY_TB, 0, c.stream>>>(L, c.ev_acc, c.new_block, c.bm.assign);
    finish_update_dyn(c);
}

}  // namespace sbpx
