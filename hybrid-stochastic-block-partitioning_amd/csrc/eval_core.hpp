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
    static constexpr bool kSlots = false;  // every entry in range is a nonzero
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
    // Block at unit `target` of row t (target < dout[t]) or of column t (target - dout[t]).
    __device__ int sample(int t, int target, int dout_t) const {
        if (target < dout_t) return row_col[wave_upper_bound(row_cum, row_ptr[t], row_ptr[t + 1], target)];
        return col_row[wave_upper_bound(col_cum, col_ptr[t], col_ptr[t + 1], target - dout_t)];
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
    static constexpr bool kSlots = true;  // skip key < 0 and value 0
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
    __device__ int sample(int t, int target, int dout_t) const {
        if (target < dout_t) return wave_weighted_select(rkeys, rvals, rbase[t], rbase[t] + rcap[t], target, B);
        return wave_weighted_select(ckeys, cvals, cbase[t], cbase[t] + ccap[t], target - dout_t, B);
    }
};

struct GraphView {
    const int* __restrict__ out_ptr;
    const int* __restrict__ out_idx;
    const int* __restrict__ in_ptr;
    const int* __restrict__ in_idx;
};

// No __restrict__: mutable during a Metropolis-Hastings pass.
struct BlockState {
    const int* assign;
    const int* dout;
    const int* din;
    const int* deg;
    int B;
};

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
};

__device__ __forceinline__ void fence_wave() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

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

// Table layout: 7 arrays of cap ints: key, o, i, M[s,b], M[b,s], M[r,b], M[b,r].
// given_s >= 0 fixes the proposal (with uniform given_u); propose_only stops after step 2.
template <class View>
__device__ __forceinline__ MoveResult eval_move(const GraphView& G, const BlockState& S, const View& M, int v,
                                                int* tab, int cap, uint64_t seed, uint32_t call, uint32_t tag,
                                                int given_s, double given_u, bool propose_only) {
    const int lane = lane_id();
    MoveResult R{};
    R.v = v;
    const int r = S.assign[v];
    R.r = r;
    R.t = -1;
    const int B = S.B;
    int* tkey = tab;
    int* to = tab + cap;
    int* ti = tab + 2 * cap;
    int* tmsb = tab + 3 * cap;
    int* tmbs = tab + 4 * cap;
    int* tmrb = tab + 5 * cap;
    int* tmbr = tab + 6 * cap;
    for (int j = lane; j < cap; j += WAVE) {
        tkey[j] = -1;
        to[j] = 0;
        ti[j] = 0;
        tmsb[j] = 0;
        tmbs[j] = 0;
        tmrb[j] = 0;
        tmbr[j] = 0;
    }
    fence_wave();

    // 1. neighbour blocks
    const int ob = G.out_ptr[v], oe = G.out_ptr[v + 1];
    const int ib = G.in_ptr[v], ie = G.in_ptr[v + 1];
    int self = 0, self_pos = INT_MAX, full = 0;
    for (int j = ob + lane; j < oe; j += WAVE) {
        int w = G.out_idx[j];
        if (w == v) {
            self = 1;
        } else {
            int slot = tab_insert(tkey, cap, S.assign[w]);
            if (slot < 0) full = 1;
            else atomicAdd(&to[slot], 1);
        }
    }
    for (int j = ib + lane; j < ie; j += WAVE) {
        int w = G.in_idx[j];
        if (w == v) {
            self_pos = j - ib;
        } else {
            int slot = tab_insert(tkey, cap, S.assign[w]);
            if (slot < 0) full = 1;
            else atomicAdd(&ti[slot], 1);
        }
    }
    self = wave_or(self);
    self_pos = wave_min(self_pos);
    if (self && lane == 0) {
        if (tab_insert(tkey, cap, r) < 0) full = 1;
    }
    full = wave_or(full);
    fence_wave();
    const int k_out = oe - ob;
    const int k_in = (ie - ib) - (self_pos != INT_MAX ? 1 : 0);
    const int k = k_out + k_in;
    R.k_out = k_out;
    R.k_in = k_in;
    R.self = self;
    if (full) {
        R.status = -1;
        return R;
    }

    // 2. proposal
    int s;
    double u;
    if (given_s >= 0) {
        s = given_s;
        u = given_u;
    } else {
        Draw4 dr = philox4x32_10(seed, (uint32_t)v, call, 0u, tag);
        u = unit53(dr.x2, dr.x3);
        if (k == 0) {
            s = random_block(r, B, dr.x0);
        } else {
            int slot = (int)mulhi32(dr.x0, (uint32_t)k);
            int w;
            if (slot < k_out) {
                w = G.out_idx[ob + slot];
            } else {
                int jj = slot - k_out;
                if (jj >= self_pos) jj++;
                w = G.in_idx[ib + jj];
            }
            const int t = S.assign[w];
            R.t = t;
            const int trow = S.dout[t];
            const int total = trow + S.din[t];
            const int target = (int)mulhi32(dr.x1, (uint32_t)total);
            s = M.sample(t, target, trow);
        }
    }
    R.s = s;
    R.u = u;
    R.H = 1.0;
    R.dS = 0.0;
    if (propose_only || s == r) {
        R.status = 0;
        return R;
    }

    // 3. one pass over rows r, s and columns r, s
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

    const int rs0 = M.rbeg(s), rs1 = M.rend(s);
    for (int j = rs0 + lane; j < rs1; j += WAVE) {
        int c = M.rkey(j), m = M.rval(j);
        if (View::kSlots && (c < 0 || m == 0)) continue;
        bool unchanged;
        if (c == s) {
            m_ss = m;
            unchanged = !t_ss;
        } else if (c == r) {
            m_sr = m;
            unchanged = !t_sr || d_sr == 0;
        } else {
            int sl = tab_find(tkey, cap, c);
            if (sl >= 0) {
                tmsb[sl] = m;
                unchanged = to[sl] == 0;
            } else {
                unchanged = true;
            }
        }
        if (unchanged) {
            int dc = S.din[c];
            int ndc = c == r ? ndin_r : (c == s ? ndin_s : dc);
            acc += (double)cell_entropy(fi(m), fi(dc), fi(dout_s));
            acc -= (double)cell_entropy(fi(m), fi(ndc), fi(ndout_s));
        }
    }
    const int rr0 = M.rbeg(r), rr1 = M.rend(r);
    for (int j = rr0 + lane; j < rr1; j += WAVE) {
        int c = M.rkey(j), m = M.rval(j);
        if (View::kSlots && (c < 0 || m == 0)) continue;
        bool unchanged;
        if (c == r) {
            m_rr = m;
            unchanged = !t_rr;
        } else if (c == s) {
            m_rs = m;
            unchanged = !t_rs || d_rs == 0;
        } else {
            int sl = tab_find(tkey, cap, c);
            if (sl >= 0) {
                tmrb[sl] = m;
                unchanged = to[sl] == 0;
            } else {
                unchanged = true;
            }
        }
        if (unchanged) {
            int dc = S.din[c];
            int ndc = c == r ? ndin_r : (c == s ? ndin_s : dc);
            acc += (double)cell_entropy(fi(m), fi(dc), fi(dout_r));
            acc -= (double)cell_entropy(fi(m), fi(ndc), fi(ndout_r));
        }
    }
    const int cs0 = M.cbeg(s), cs1 = M.cend(s);
    for (int j = cs0 + lane; j < cs1; j += WAVE) {
        int i = M.ckey(j);
        if (View::kSlots && i < 0) continue;
        if (i == r || i == s) continue;
        int m = M.cval(j);
        if (View::kSlots && m == 0) continue;
        int sl = tab_find(tkey, cap, i);
        bool unchanged = true;
        if (sl >= 0) {
            tmbs[sl] = m;
            unchanged = ti[sl] == 0;
        }
        if (unchanged) {
            int di = S.dout[i];
            acc += (double)cell_entropy(fi(m), fi(din_s), fi(di));
            acc -= (double)cell_entropy(fi(m), fi(ndin_s), fi(di));
        }
    }
    const int cr0 = M.cbeg(r), cr1 = M.cend(r);
    for (int j = cr0 + lane; j < cr1; j += WAVE) {
        int i = M.ckey(j);
        if (View::kSlots && i < 0) continue;
        if (i == r || i == s) continue;
        int m = M.cval(j);
        if (View::kSlots && m == 0) continue;
        int sl = tab_find(tkey, cap, i);
        bool unchanged = true;
        if (sl >= 0) {
            tmbr[sl] = m;
            unchanged = ti[sl] == 0;
        }
        if (unchanged) {
            int di = S.dout[i];
            acc += (double)cell_entropy(fi(m), fi(din_r), fi(di));
            acc -= (double)cell_entropy(fi(m), fi(ndin_r), fi(di));
        }
    }
    m_ss = wave_max(m_ss);
    m_rr = wave_max(m_rr);
    m_sr = wave_max(m_sr);
    m_rs = wave_max(m_rs);
    fence_wave();

    // 4. Delta keys and Hastings sums
    double pf = 0.0, pb = 0.0;
    int nuniq = 0;
    for (int j = lane; j < cap; j += WAVE) {
        int b = tkey[j];
        if (b < 0) continue;
        nuniq++;
        int ob_ = to[j], ib_ = ti[j];
        int Msb, Mbs, Mrb, Mbr, Drb, Dbr, nd;
        if (b == r) {
            Msb = m_sr; Mbs = m_rs; Mrb = m_rr; Mbr = m_rr; Drb = d_rr; Dbr = d_rr;
            nd = (dout_r - k_out) + (din_r - k_in) - (m_rr + d_rr);
        } else if (b == s) {
            Msb = m_ss; Mbs = m_ss; Mrb = m_rs; Mbr = m_sr; Drb = d_rs; Dbr = d_sr;
            nd = (dout_s + k_out) + (din_s + k_in) - (m_ss + d_ss);
        } else {
            Msb = tmsb[j]; Mbs = tmbs[j]; Mrb = tmrb[j]; Mbr = tmbr[j]; Drb = -ob_; Dbr = -ib_;
            nd = S.deg[b];
            const int db = S.din[b], ub = S.dout[b];
            if (ob_ > 0) {
                acc += (double)cell_entropy(fi(Msb), fi(db), fi(dout_s));
                acc -= (double)cell_entropy(fi(Msb + ob_), fi(db), fi(ndout_s));
                acc += (double)cell_entropy(fi(Mrb), fi(db), fi(dout_r));
                acc -= (double)cell_entropy(fi(Mrb - ob_), fi(db), fi(ndout_r));
            }
            if (ib_ > 0) {
                acc += (double)cell_entropy(fi(Mbs), fi(din_s), fi(ub));
                acc -= (double)cell_entropy(fi(Mbs + ib_), fi(ndin_s), fi(ub));
                acc += (double)cell_entropy(fi(Mbr), fi(din_r), fi(ub));
                acc -= (double)cell_entropy(fi(Mbr - ib_), fi(ndin_r), fi(ub));
            }
        }
        int nb = ob_ + ib_ + (b == r ? self : 0);
        if (nb == 0) continue;
        double count = (double)nb;
        double pw = (double)(Msb + Mbs) + 1.0;
        double bd = (double)(S.deg[b] + B);
        double bw = (double)(Mrb + Drb + Mbr + Dbr) + 1.0;
        double pd = (double)(nd + B);
        pf += count * pw / bd;
        pb += count * bw / pd;
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
    const double dS = wave_sum(acc);
    pf = wave_sum(pf);
    pb = wave_sum(pb);
    R.nuniq = (int)wave_sum_ll(nuniq);
    const double H = (k == 0) ? 1.0 : pb / pf;
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
    R.nnz_rs = (long long)(rr1 - rr0) + (rs1 - rs0) + (cr1 - cr0) + (cs1 - cs0);
    return R;
}

}  // namespace sbpx
