// Block-merge proposals: one wavefront per block r evaluates its 10 proposals in order
// (block_merge.cpp:278-301, 335-368). A proposal draws t in proportion to row r + column r and
// s in proportion to row t + column t without index r (common.cpp:131-182, block_merge=true),
// both by wave-parallel searches of the prefix sums. Repeated proposals score DBL_MAX. The delta
// MDL of merging r into s follows entropy::block_merge_delta_mdl (entropy.cpp:127-180):
//   old cells of row r and column r (their new value is 0 and contributes nothing),
//   cells (s, c) and (c, s) that gain row/column r's counts, the diagonal (s, s),
//   and the untouched cells of row s and column s, whose degrees change.
#include <cfloat>

#include "eval_core.hpp"
#include "internal.hpp"
#include "merge_core.hpp"

namespace sbpx {

struct MergeParams {
    const int* __restrict__ row_ptr;
    const int* __restrict__ row_col;
    const int* __restrict__ row_val;
    const int* __restrict__ row_cum;
    const int* __restrict__ col_ptr;
    const int* __restrict__ col_row;
    const int* __restrict__ col_val;
    const int* __restrict__ col_cum;
    const int* __restrict__ dout;
    const int* __restrict__ din;
    int B;
    uint64_t seed;
    uint32_t call;
    double* __restrict__ out_dS;
    int* __restrict__ out_best;
    // a union of graphs (null for one graph): graph of each block id, its first block id, blocks
    // in use, seed and call, and whether it takes part in this merge phase
    const int* __restrict__ bgraph;
    const int* __restrict__ gbo;
    const int* __restrict__ gB;
    const uint64_t* __restrict__ gseed;
    const uint32_t* __restrict__ gcall;
    const uint8_t* __restrict__ active;
};

constexpr int MERGE_WAVES = 4;
constexpr int NUM_PROPOSALS = 10;  // NUM_AGG_PROPOSALS_PER_BLOCK (block_merge.hpp:24)

__global__ void __launch_bounds__(MERGE_WAVES * WAVE) merge_proposals_kernel(MergeParams P) {
    const int r = blockIdx.x * MERGE_WAVES + threadIdx.x / WAVE;
    if (r >= P.B) return;
    const int lane = lane_id();
    int B = P.B, base = 0;  // blocks of r's graph and its first block id
    uint64_t seed = P.seed;
    uint32_t call = P.call;
    if (P.bgraph != nullptr) {
        const int g = P.bgraph[r];
        if (!P.active[g]) return;
        base = P.gbo[g];
        B = P.gB[g];
        if (r - base >= B) return;
        seed = P.gseed[g];
        call = P.gcall[g];
    }
    const int rl = r - base;  // graph-local id: keys the randomness
    const int k_out = P.dout[r], k_in = P.din[r];
    const int k = k_out + k_in;
    int past[NUM_PROPOSALS];
    double best_dS = DBL_MAX;
    int best = -1;
    for (int i = 0; i < NUM_PROPOSALS; ++i) {
        Draw4 dr = philox4x32_10(seed, (uint32_t)rl, call, (uint32_t)i, TAG_MERGE);
        int s;
        if (k == 0) {
            s = base + random_block(rl, B, dr.x2);
        } else {
            int target = (int)mulhi32(dr.x0, (uint32_t)k);
            int t;
            if (target < k_out) {
                int p = wave_upper_bound(P.row_cum, P.row_ptr[r], P.row_ptr[r + 1], target);
                t = P.row_col[p];
            } else {
                int p = wave_upper_bound(P.col_cum, P.col_ptr[r], P.col_ptr[r + 1], target - k_out);
                t = P.col_row[p];
            }
            const int t0 = P.row_ptr[t], t1 = P.row_ptr[t + 1];
            const int u0 = P.col_ptr[t], u1 = P.col_ptr[t + 1];
            const int pr = lower_bound_lane(P.row_col, t0, t1, r);
            const int pc = lower_bound_lane(P.col_row, u0, u1, r);
            const int Mtr = (pr < t1 && P.row_col[pr] == r) ? P.row_val[pr] : 0;
            const int Mrt = (pc < u1 && P.col_row[pc] == r) ? P.col_val[pc] : 0;
            const int rowpart = P.dout[t] - Mtr;
            const int total = rowpart + P.din[t] - Mrt;
            if (total == 0) {
                s = base + random_block(rl, B, dr.x2);
            } else {
                int target2 = (int)mulhi32(dr.x1, (uint32_t)total);
                if (target2 < rowpart) {
                    int before = pr > t0 ? P.row_cum[pr - 1] : 0;
                    int tt = target2 < before ? target2 : target2 + Mtr;
                    int p = wave_upper_bound(P.row_cum, t0, t1, tt);
                    s = P.row_col[p];
                } else {
                    int target3 = target2 - rowpart;
                    int before = pc > u0 ? P.col_cum[pc - 1] : 0;
                    int tt = target3 < before ? target3 : target3 + Mrt;
                    int p = wave_upper_bound(P.col_cum, u0, u1, tt);
                    s = P.col_row[p];
                }
            }
        }
        bool repeat = false;
        for (int q = 0; q < i; ++q) repeat |= past[q] == s;
        past[i] = s;
        const CsrView M{P.row_ptr, P.row_col, P.row_val, P.row_cum, P.col_ptr, P.col_row, P.col_val, P.col_cum};
        double dS = repeat ? DBL_MAX : merge_delta_mdl_wave(M, P.dout, P.din, r, s, k_out, k_in);
        if (dS < best_dS) {
            best_dS = dS;
            best = s;
        }
    }
    if (lane == 0) {
        P.out_dS[r] = best_dS;
        P.out_best[r] = best;
    }
}

void launch_merge_proposals(Ctx& c, uint64_t seed, uint32_t call) {
    MergeParams P{};
    P.row_ptr = c.bm.row_ptr;
    P.row_col = c.bm.row_col;
    P.row_val = c.bm.row_val;
    P.row_cum = c.bm.row_cum;
    P.col_ptr = c.bm.col_ptr;
    P.col_row = c.bm.col_row;
    P.col_val = c.bm.col_val;
    P.col_cum = c.bm.col_cum;
    P.dout = c.bm.dout;
    P.din = c.bm.din;
    P.B = c.bm.B;
    P.seed = seed;
    P.call = call;
    P.out_dS = c.m_dS;
    P.out_best = c.m_best;
    if (c.batch.G > 0) {
        const DevBatch& b = c.batch;
        P.bgraph = b.d_bgraph;
        P.gbo = b.d_bo;
        P.gB = b.d_nb;
        P.gseed = b.d_seed;
        P.gcall = b.d_call;
        P.active = b.d_active;
    }
    int nb = (c.bm.B + MERGE_WAVES - 1) / MERGE_WAVES;
    merge_proposals_kernel<<<nb, MERGE_WAVES * WAVE, 0, c.stream>>>(P);
    SBPX_HIP(hipGetLastError());
}

__global__ void block_counts_kernel(const int* __restrict__ assign, int N, int* __restrict__ count) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N) atomicAdd(&count[assign[v]], 1);
}

__global__ void relabel_kernel(int* __restrict__ assign, int N, const int* __restrict__ map) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N) assign[v] = map[assign[v]];
}

__global__ void batch_relabel_kernel(int* __restrict__ assign, int N, const int* __restrict__ map, int lo, int hi) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int a = assign[v];
    if (a >= lo && a < hi) assign[v] = map[a - lo];
}

__global__ void zero_drift_active_kernel(int* __restrict__ drift, int range, const int* __restrict__ bgraph,
                                         const uint8_t* __restrict__ active) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < range && active[bgraph[x]]) drift[x] = 0;
}

// Request i (graph req[n + i]) holds the assignment values [pre[i], pre[i + 1]) of the staged
// list, for the vertices from vo[graph] on, in graph-local block ids.
__global__ void batch_scatter_kernel(const int* __restrict__ vals, int total, const int* __restrict__ req, int n,
                                     const int* __restrict__ vo, const int* __restrict__ bo,
                                     int* __restrict__ assign) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    int i = 0, hi = n;  // the last i with pre[i] <= p (req[0..n] = pre, pre[0] = 0, pre[n] = total)
    while (hi - i > 1) {
        const int mid = (i + hi) >> 1;
        if (req[mid] <= p) i = mid;
        else hi = mid;
    }
    const int g = req[n + 1 + i];
    assign[vo[g] + (p - req[i])] = bo[g] + vals[p];
}

__global__ void batch_dense_kernel(int range, const int* __restrict__ bgraph, const int* __restrict__ bo,
                                   const int* __restrict__ nb, const int* __restrict__ dbo, int* __restrict__ dense,
                                   int* __restrict__ inv) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= range) return;
    const int g = bgraph[x], l = x - bo[g];
    if (l < nb[g]) {
        dense[x] = dbo[g] + l;
        inv[dbo[g] + l] = x;
    } else {
        dense[x] = -1;
    }
}

// trace[g] = sum of diag over graph g's blocks in use; one wavefront per graph.
__global__ void batch_trace_kernel(const int* __restrict__ diag, int G, const int* __restrict__ bo,
                                   const int* __restrict__ nb, long long* __restrict__ trace) {
    const int g = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (g >= G) return;
    long long acc = 0;
    for (int x = bo[g] + lane_id(); x < bo[g] + nb[g]; x += WAVE) acc += diag[x];
    acc = wave_sum_ll(acc);
    if (lane_id() == 0) trace[g] = acc;
}

void launch_batch_trace(Ctx& c, long long* d_trace) {
    const DevBatch& b = c.batch;
    batch_trace_kernel<<<(b.G + 3) / 4, 4 * WAVE, 0, c.stream>>>(c.bm.diag, b.G, b.d_bo, b.d_nb, d_trace);
    SBPX_HIP(hipGetLastError());
}

void launch_batch_dense(Ctx& c) {
    const DevBatch& b = c.batch;
    batch_dense_kernel<<<(b.range + 255) / 256, 256, 0, c.stream>>>(b.range, b.d_bgraph, b.d_bo, b.d_nb, b.d_dbo,
                                                                    b.d_dense, b.d_inv);
    SBPX_HIP(hipGetLastError());
}

void launch_batch_relabel(Ctx& c, int lo, int hi) {
    batch_relabel_kernel<<<(c.g.N + 255) / 256, 256, 0, c.stream>>>(c.bm.assign, c.g.N, c.m_map, lo, hi);
    SBPX_HIP(hipGetLastError());
}

void launch_zero_drift_active(Ctx& c) {
    const int range = c.batch.range;
    zero_drift_active_kernel<<<(range + 255) / 256, 256, 0, c.stream>>>(c.bm.drift, range, c.batch.d_bgraph,
                                                                       c.batch.d_active);
    SBPX_HIP(hipGetLastError());
}

void launch_batch_scatter(Ctx& c, int n, int total) {
    if (total <= 0) return;
    batch_scatter_kernel<<<(total + 255) / 256, 256, 0, c.stream>>>(c.list_buf, total, c.batch.d_req, n,
                                                                   c.batch.d_vo, c.batch.d_bo, c.bm.assign);
    SBPX_HIP(hipGetLastError());
}

void launch_block_counts(Ctx& c) {
    SBPX_HIP(hipMemsetAsync(c.m_count, 0, sizeof(int) * c.bm.B, c.stream));
    block_counts_kernel<<<(c.g.N + 255) / 256, 256, 0, c.stream>>>(c.bm.assign, c.g.N, c.m_count);
    SBPX_HIP(hipGetLastError());
}

void launch_relabel(Ctx& c) {
    relabel_kernel<<<(c.g.N + 255) / 256, 256, 0, c.stream>>>(c.bm.assign, c.g.N, c.m_map);
    SBPX_HIP(hipGetLastError());
}

}  // namespace sbpx
