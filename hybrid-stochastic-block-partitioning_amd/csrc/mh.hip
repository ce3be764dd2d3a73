// Sequential Metropolis-Hastings (finetune.cpp:446-461, 488-501, used by metropolis_hastings
// :412-444 and by hybrid_mcmc for the high-degree vertices :358-363): each accepted move changes
// the block matrix before the next vertex is proposed.
//
// The block matrix is held, for the duration of the pass, as a per-row open-addressing table in
// both orientations (HashView), built from the CSR snapshot. One workgroup of 16 wavefronts runs
// the pass: in each round the wavefronts evaluate the next 16 vertices of the queue against the
// current state; the first accepted move in queue order is committed and the round restarts
// after it. Every vertex before it was rejected and so left the state unchanged, which makes the
// outcome identical to evaluating one vertex at a time (proposals and uniforms are keyed by
// vertex, not by evaluation order). Before a commit, every row and column it may extend is given
// room: a row that is too full moves to a larger segment taken from the free end of the pool
// (the other wavefronts wait at the barrier meanwhile). Only when the pool itself runs out does
// the kernel stop; the host then rebuilds the tables with a larger pool and the pass resumes at
// the same vertex.
#include <algorithm>
#include <cstdint>

#include "eval_core.hpp"
#include "internal.hpp"

namespace sbpx {

constexpr int MH_WAVES = 16;

// ------------------------------------------------------------------------------------------
// Table build: cap(x) = next power of two >= 2 nnz(x) + 32, filled from the CSR.
__device__ __forceinline__ int table_cap(int keys) {
    int want = 2 * keys + 32;
    int cap = 32;
    while (cap < want) cap <<= 1;
    return cap;
}

__global__ void hash_caps_kernel(const int* __restrict__ ptr, int B, int* __restrict__ caps) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= B) return;
    caps[x] = table_cap(ptr[x + 1] - ptr[x]);
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
    double* dS_sum;
    int* err;
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
__device__ bool make_room(const MhParams& P, const MoveResult& R, const int* tab, int cap) {
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

// Applies the accepted move: block matrix cells from (o_b, i_b, self), block degrees with the
// update rule of Blockmodel::move_vertex (blockmodel.cpp:344-360), the assignment.
__device__ void commit_move(const MhParams& P, const MoveResult& R, const int* tab, int cap, int serial) {
    const int lane = lane_id();
    const int r = R.r, s = R.s;
    {  // stamp the neighbours, whose neighbour-block tables this move changes
        const int v = R.v;
        const int ob = P.G.out_ptr[v], oe = P.G.out_ptr[v + 1];
        const int ib = P.G.in_ptr[v], ie = P.G.in_ptr[v + 1];
        for (int j = ob + lane; j < oe; j += WAVE) P.touch[P.G.out_idx[j]] = serial;
        for (int j = ib + lane; j < ie; j += WAVE) P.touch[P.G.in_idx[j]] = serial;
    }
    const int* to = tab + cap;
    const int* ti = tab + 2 * cap;
    for (int j = lane; j < cap; j += WAVE) {
        int b = tab[j];
        if (b < 0 || b == r || b == s) continue;
        int o = to[j], i = ti[j];
        if (o) {
            cell_add(P, s, b, o);
            cell_add(P, r, b, -o);
        }
        if (i) {
            cell_add(P, b, s, i);
            cell_add(P, b, r, -i);
        }
    }
    if (lane == 0) {
        if (R.d_ss) cell_add(P, s, s, R.d_ss);
        if (R.d_rr) cell_add(P, r, r, R.d_rr);
        if (R.d_sr) cell_add(P, s, r, R.d_sr);
        if (R.d_rs) cell_add(P, r, s, R.d_rs);
        P.assign[R.v] = s;
        const int nout_r = P.dout[r] - R.k_out, nout_s = P.dout[s] + R.k_out;
        const int nin_r = P.din[r] - (R.k_in + R.self), nin_s = P.din[s] + (R.k_in + R.self);
        P.dout[r] = nout_r;
        P.dout[s] = nout_s;
        P.din[r] = nin_r;
        P.din[s] = nin_s;
        // self edges read after the update, plus the Delta diagonal (blockmodel.cpp:348-359)
        P.deg[r] = nout_r + nin_r - ((R.m_rr + R.d_rr) + R.d_rr);
        P.deg[s] = nout_s + nin_s - ((R.m_ss + R.d_ss) + R.d_ss);
        P.drift[r] = R.d_rr;
        P.drift[s] = R.d_ss;
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
__device__ bool still_holds(const MhParams& P, const HashView& M, const MoveResult& R, const int* tab, int cap,
                            const int* ctab, int ccap, int cr, int cs, int serial) {
    const int lane = lane_id();
    int bad = 0;
    if (lane == 0) {
        bad = P.touch[R.v] == serial;  // a neighbour moved
    } else if (lane <= 3) {            // r, s or t is one of the blocks whose cells changed
        const int y = lane == 1 ? R.r : (lane == 2 ? R.s : R.t);
        if (y >= 0) bad = y == cr || y == cs || tab_find(ctab, ccap, y) >= 0;
    } else if (lane <= 11) {  // cr or cs keyed in rows / columns r and s
        const int q = lane - 4;
        const int x = (q & 1) ? cs : cr;
        const int row = (q & 2) ? R.s : R.r;
        bad = ((q & 4) ? M.clookup(row, x) : M.rlookup(row, x)) != 0;
    } else if (lane <= 13) {  // cr or cs a neighbour block
        const int x = lane == 12 ? cr : cs;
        bad = tab_find(tab, cap, x) >= 0;
    }
    return wave_or(bad) == 0;
}

__global__ void __launch_bounds__(MH_WAVES * WAVE) mh_kernel(MhParams P) {
    __shared__ int st[MH_WAVES];    // status of each wavefront's evaluation, 3 = idle
    __shared__ int holds[MH_WAVES]; // 1 while that evaluation still holds
    __shared__ double sdS[MH_WAVES];
    __shared__ int s_flag, s_cr, s_cs, s_ccap;
    __shared__ int* s_ctab;
    __shared__ int lds_tab[MH_WAVES][7 * MH_LDS_CAP];
    __shared__ int lds_hist[MH_WAVES][256];
    const int wave = threadIdx.x / WAVE, lane = lane_id();
    int* const gtab = P.tables + (size_t)wave * 7 * P.tcap;
    const GraphView G = P.G;
    const BlockState S{P.assign, P.dout, P.din, P.deg, P.B};
    const HashView M{P.rbase, P.rcap, P.rkeys, P.rvals, P.cbase, P.ccap, P.ckeys, P.cvals, P.B, lds_hist[wave]};
    long long att = 0, evs = 0, mvs = 0, bytes_local = 0;
    double dsum = 0.0;
    int pos = P.start;
    int stopped = 0;
    int serial = P.status[4];
    while (pos < P.n && stopped == 0) {
        // a window of up to 16 queue entries, each evaluated against the current state
        const int idx = pos + wave;
        MoveResult R{};
        const bool active = idx < P.n;
        int cap = 0;
        int* tab = gtab;
        if (active) {
            const int v = P.list[idx];
            cap = vertex_table_cap(G, v);
            if (cap <= MH_LDS_CAP) tab = lds_tab[wave];
            R = eval_move<HashView>(G, S, M, v, tab, cap, P.seed, P.call, P.tag, -1, 0.0, false);
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
            if (P.bytes != nullptr && lane == 0 && wave >= cur && wave < end && st[wave] != 3)
                bytes_local += eval_bytes(P, R);
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
            if (wave == j) {
                const bool ok = make_room(P, R, tab, cap);
                if (ok) commit_move(P, R, tab, cap, serial + 1);
                if (lane == 0) {
                    s_flag = ok ? 1 : 2;
                    s_cr = R.r;
                    s_cs = R.s;
                    s_ctab = tab;
                    s_ccap = cap;
                }
            }
            __syncthreads();
            if (s_flag == 2) {  // the pool is full: the host rebuilds it and resumes here
                stopped = 1;
                pos += j;
                break;
            }
            ++serial;
            if (wave == j && lane == 0 && P.bytes != nullptr) bytes_local += eval_bytes(P, R);
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
static void build_tables(Ctx& c) {
    const int B = c.bm.B;
    hipStream_t st = c.stream;
    MhState& h = c.mh;
    const int tb = 256;
    // capacities and bases, rows then columns
    hash_caps_kernel<<<(B + tb - 1) / tb, tb, 0, st>>>(c.bm.row_ptr, B, h.rcap);
    scan_exclusive(h.rcap, h.rbase, B, c.scan_tmp, st);
    hash_caps_kernel<<<(B + tb - 1) / tb, tb, 0, st>>>(c.bm.col_ptr, B, h.ccap);
    scan_exclusive(h.ccap, h.cbase, B, c.scan_tmp, st);
    int tot[2] = {0, 0};
    SBPX_HIP(hipMemcpyAsync(&tot[0], h.rbase + B, sizeof(int), hipMemcpyDeviceToHost, st));
    SBPX_HIP(hipMemcpyAsync(&tot[1], h.cbase + B, sizeof(int), hipMemcpyDeviceToHost, st));
    SBPX_HIP(hipStreamSynchronize(st));
    const size_t used = (size_t)std::max(tot[0], tot[1]);
    h.headroom = std::max(h.headroom, std::max(used, (size_t)1 << 20));
    // pool tops stay below 2^31 even after a failed claim
    const size_t need = std::min(used + h.headroom, (size_t)(INT32_MAX / 2));
    if (need < used + 64) throw StateError("block matrix too large for the Metropolis-Hastings tables");
    if (need > h.slots) {
        for (int** p : {&h.rkeys, &h.rvals, &h.ckeys, &h.cvals}) {
            if (*p) SBPX_HIP(hipFree(*p));
            SBPX_HIP(hipMalloc((void**)p, sizeof(int) * need));
        }
        h.slots = need;
    }
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

void mh_sequential(Ctx& c, const int* d_list, int n, int max_deg, uint64_t seed, uint32_t call, uint32_t tag,
                   long long* bytes, bool timed) {
    if (n <= 0) return;
    MhState& h = c.mh;
    const int B = c.bm.B;
    // max_deg counts a self-loop once; vertex_table_cap counts it in both lists
    int tcap = 16;
    while (tcap < 2 * (max_deg + 3)) tcap <<= 1;
    const size_t tslots = (size_t)MH_WAVES * 7 * tcap;
    if (tslots > h.table_slots) {
        if (h.tables) SBPX_HIP(hipFree(h.tables));
        SBPX_HIP(hipMalloc((void**)&h.tables, sizeof(int) * tslots));
        h.table_slots = tslots;
    }
    int pos = 0;
    int rounds = 0;
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
        P.pool = (int)h.slots;
        P.list = d_list;
        P.n = n;
        P.start = pos;
        P.seed = seed;
        P.call = call;
        P.tag = tag;
        P.tables = h.tables;
        P.tcap = tcap;
        P.status = h.status;
        P.counters = c.counters;
        P.bytes = bytes;
        P.touch = h.touch;
        P.dS_sum = c.red + 2048;
        P.err = c.dev_err;
        if (timed) c.timer_begin();
        mh_kernel<<<1, MH_WAVES * WAVE, 0, c.stream>>>(P);
        SBPX_HIP(hipGetLastError());
        if (timed) c.timer_end(KERNEL_MH);
        int stat[5];
        SBPX_HIP(hipMemcpyAsync(stat, h.status, sizeof(stat), hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipStreamSynchronize(c.stream));
        check_device_error(c);
        pos = stat[0];
        if (stat[4] > (1 << 30)) {  // restart the commit serials well before they wrap
            SBPX_HIP(hipMemsetAsync(h.touch, 0, sizeof(int) * c.g.N, c.stream));
            SBPX_HIP(hipMemsetAsync(h.status + 4, 0, sizeof(int), c.stream));
        }
        // the matrix changed: bring the CSR snapshot up to date (dout/din exact, deg keeps drift)
        rebuild_keep_drift(c);
        if (stat[1] == 0 && pos < n) throw StateError("Metropolis-Hastings pass stopped early");
        if (stat[1] == 1) h.headroom *= 2;
    }
}

}  // namespace sbpx
