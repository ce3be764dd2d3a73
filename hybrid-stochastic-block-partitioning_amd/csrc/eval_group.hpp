// Vertex-move evaluation shared by G wavefronts of one workgroup (eval_move's steps, spread over
// G * 64 lanes): for vertices whose rows hold thousands of entries, or whose own degree is in
// the thousands, where one wavefront's chain of dependent loads would set the launch time.
#pragma once
#include "eval_core.hpp"

namespace sbpx {

// A workgroup may hold several groups (the Metropolis-Hastings kernel runs 16 / G of them side
// by side); every group runs the same sequence of workgroup barriers (idle groups, v < 0,
// included), so the barriers never depend on the data. The neighbour-block table is shared by
// the group (LDS or global memory) and partial results meet in LDS. Results equal those of
// eval_move: the same cells are visited and the same terms formed; only the order of the FP64
// sums differs. LRN with lw >= 0: the proposal was the learned mixture, as in eval_move (given_s /
// given_u carry the learned branch's draw); the net's forward and reverse probabilities come from
// the fixed scores lz / lcand, or with lz null as lqf_in / lqb_in.
constexpr int RED_D = 6;   // per-wavefront double partials: dS, pf, pb, and the mixture's pf0, pb0
constexpr int RED_I = 12;  // per-wavefront int partials

// One entry's rows and columns r and s for the workgroup-shared scan (eval_group with CG > 1): hash
// segment bases and lengths (row s, row r, column s, column r), r, s, its neighbour-block table,
// what the unchanged-cell test needs, the degrees before and after the move, and the cells among
// {r, s} the scan finds.
struct ScanShare {
    int b[4], L[4];
    int r, s, cap, flags, d_sr, d_rs;
    int dout_s, dout_r, ndout_s, ndout_r, din_s, din_r, ndin_s, ndin_r;
    int* tab;
    int m[4];  // m_ss, m_rr, m_sr, m_rs
};

// An entry evaluated by one workgroup: no team.
struct NoTeam {
    static constexpr bool on = false;
    int rank = 0, size = 1;
    int* hist = nullptr;     // unused
    double* part = nullptr;  // unused
    __device__ bool sync() const { return true; }
};
// A team of `size` workgroups evaluates one entry (the Metropolis-Hastings window kernel, for
// rows of tens of thousands of slots): every member forms the same neighbour table and draws the
// same proposal; the row scans of the proposal's radix selection and of the delta-MDL rows are
// split into `size` contiguous slices. The members' 256-bin histograms meet in `hist` ([2][size]
// [256], one half per selection pass in turn), their delta-MDL partial sums and scanned-slot
// counts in `part` ([size][2]), each exchange behind `sync()` (a barrier over the launch). Member
// 0 adds the terms of the changed cells and the Hastings sums, reading the neighbour blocks'
// cells by hash lookups (the scans no longer see every slot), and forms the result; the other
// members return status 3.

// GU: row slots per lane per trip of the row/column pass (the proposal's row scans take 8, or 2
// with the learned mixture).
// CG > 1: the workgroup holds CG groups (every wavefront runs eval_group), and their rows and
// columns r and s are scanned together: the union of the CG entries' slots is spread over all of
// the workgroup's lanes, so an entry with long rows does not hold the others at the barriers that
// follow. Each lane keeps one partial sum per entry; the partials are added in wavefront order, so
// the result does not depend on timing.
template <int G, class View, bool LRN = false, int GU = 2, class Team = NoTeam, int CG = 0>
__device__ MoveResult eval_group(const GraphView& Gr, const BlockState& S, const View& M, int v, int* tab, int cap,
                                 uint64_t seed, uint32_t call, uint32_t tag, int given_s, double given_u,
                                 bool propose_only, int* hist, double (*redd)[RED_D], int (*redi)[RED_I],
                                 double lw = -1.0, const float* lz = nullptr, const int* lcand = nullptr,
                                 double lqf_in = 0.0, double lqb_in = 0.0, const Team& team = Team{}) {
    static_assert(!Team::on || (View::kSlots && G == 8), "teams: hash rows, one 8-wavefront group per workgroup");
    static_assert(CG <= 1 || (View::kSlots && !Team::on && CG <= 4), "shared scans: hash rows, up to 4 groups");
    constexpr int GS = G * WAVE;
    const int lane = lane_id();
    const int wave = threadIdx.x / WAVE;
    const int wig = wave % G;
    const int w0 = wave - wig;
    const int gl = wig * WAVE + lane;
    const bool act = v >= 0;
    const LocalKey lk = act ? local_key(S, v, seed, call) : LocalKey{S.B, 0, 0, seed, call};
    const int B = lk.B;     // blocks of this vertex's graph
    const int Bids = S.B;   // range of block ids (radix descent)
    MoveResult R{};
    R.v = v;
    R.t = -1;
    R.status = 3;
    const int r = act ? S.assign[v] : 0;
    R.r = r;
    int* tkey = tab;
    int* to = tab + cap;
    int* ti = tab + 2 * cap;
    int* tmsb = tab + 3 * cap;
    int* tmbs = tab + 4 * cap;
    int* tmrb = tab + 5 * cap;
    int* tmbr = tab + 6 * cap;
    if (act) {
        for (int j = gl; j < cap; j += GS) {
            tkey[j] = -1;
            to[j] = 0;
            ti[j] = 0;
            tmsb[j] = 0;
            tmbs[j] = 0;
            tmrb[j] = 0;
            tmbr[j] = 0;
        }
    }
    const int ob = act ? Gr.out_ptr[v] : 0, dO = act ? Gr.out_ptr[v + 1] - ob : 0;
    const int ib = act ? Gr.in_ptr[v] : 0, dI = act ? Gr.in_ptr[v + 1] - ib : 0;
    const int self = act ? Gr.self[v] : 0;
    const int k_out = dO, k_in = dI - self, k = k_out + k_in;
    R.k_out = k_out;
    R.k_in = k_in;
    R.self = self;
    Draw4 dr{0u, 0u, 0u, 0u};
    double u = given_u;
    int pick0 = -1, pick1 = -1;
    if (act && given_s < 0) {
        dr = philox4x32_10(lk.seed, (uint32_t)lk.id, lk.call, 0u, tag);
        u = unit53(dr.x2, dr.x3);
        if (k > 0) {
            const int ps = (int)mulhi32(dr.x0, (uint32_t)k);
            if (ps < k_out) {
                pick0 = ps;
            } else {
                pick0 = dO + (ps - k_out);
                pick1 = pick0 + 1;
            }
        }
    }
    __syncthreads();  // table cleared

    // 1. neighbour blocks
    int self_pos = INT_MAX, full = 0, cand0 = -1, cand1 = -1;
    for (int j = gl; j < dO + dI; j += GS) {
        const bool is_out = j < dO;
        const int w = is_out ? Gr.out_idx[ob + j] : Gr.in_idx[ib + (j - dO)];
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
    if (act && self && gl == 0 && tab_insert(tkey, cap, r) < 0) full = 1;
    self_pos = wave_min(self_pos);
    cand0 = wave_max(cand0);
    cand1 = wave_max(cand1);
    full = wave_or(full);
    if (lane == 0) {
        redi[wave][0] = self_pos;
        redi[wave][1] = cand0;
        redi[wave][2] = cand1;
        redi[wave][3] = full;
    }
    __syncthreads();  // table built, partials visible
    for (int i = 0; i < G; ++i) {
        self_pos = min(self_pos, redi[w0 + i][0]);
        cand0 = max(cand0, redi[w0 + i][1]);
        cand1 = max(cand1, redi[w0 + i][2]);
        full |= redi[w0 + i][3];
    }

    // 2. proposal: s drawn from row t + column t (hash rows: a group-wide radix descent over the
    // block index; CSR rows: one wavefront's search of the prefix sums)
    int s = r;
    bool sample = false;
    int samp_t = 0, samp_target = 0, samp_trow = 0;
    if (act && !full) {
        if (given_s >= 0) {
            s = given_s;
        } else if (k == 0) {
            s = lk.base + random_block(r - lk.base, B, dr.x0);
        } else {
            int t = cand0;
            if (pick1 >= 0 && pick0 - dO >= self_pos) t = cand1;
            R.t = t;
            samp_t = t;
            samp_trow = S.dout[t];
            const int total = samp_trow + S.din[t];
            samp_target = (int)mulhi32(dr.x1, (uint32_t)total);
            sample = true;
        }
    }
    if constexpr (View::kSlots) {
        const bool row = samp_target < samp_trow;
        const int* skeys = row ? M.rkeys : M.ckeys;
        const int* svals = row ? M.rvals : M.cvals;
        int slo = sample ? (row ? M.rbase[samp_t] : M.cbase[samp_t]) : 0;
        int shi = sample ? slo + (row ? M.rcap[samp_t] : M.ccap[samp_t]) : 0;
        R.tslots = shi - slo;
        if constexpr (Team::on) {  // this member's slice of the row
            const int per = (shi - slo + team.size - 1) / team.size;
            const int lo = slo + team.rank * per;
            shi = min(shi, lo + per);
            slo = min(lo, shi);
        }
        [[maybe_unused]] int pass = 0;
        const int starget = row ? samp_target : samp_target - samp_trow;
        int nbits = 1;
        while (nbits < 31 && (1 << nbits) < Bids) ++nbits;
        int shift = ((nbits + 7) / 8 - 1) * 8;
        int prefix = 0;
        long long below = 0;
        for (;;) {
            for (int j = gl; j < 256; j += GS) hist[j] = 0;
            __syncthreads();
            if (sample) {
                const int up = shift + 8;
                // SU slots per lane per trip, their keys and counts loaded before any is used
                constexpr int SU = LRN ? 2 : 8;  // the learned scorer leaves fewer registers
                for (int j0 = slo + gl; j0 < shi; j0 += SU * GS) {
                    int kk[SU], w[SU];
#pragma unroll
                    for (int q = 0; q < SU; ++q) {
                        const int j = j0 + q * GS;
                        kk[q] = j < shi ? skeys[j] : -1;
                        w[q] = j < shi ? svals[j] : 0;
                    }
#pragma unroll
                    for (int q = 0; q < SU; ++q) {
                        if (kk[q] < 0 || w[q] == 0) continue;
                        const int hp = up >= 31 ? 0 : (kk[q] >> up);
                        if (hp == prefix) atomicAdd(&hist[(kk[q] >> shift) & 255], w[q]);
                    }
                }
            }
            __syncthreads();
            if constexpr (Team::on) {
                // publish this slice's counts, wait for the team, then every member sums the
                // team's counts in member order
                int* th = team.hist + (size_t)(pass & 1) * team.size * 256;
                for (int j = gl; j < 256; j += GS) th[team.rank * 256 + j] = hist[j];
                team.sync();
                for (int j = gl; j < 256; j += GS) {
                    int h = 0;
                    for (int q = 0; q < team.size; ++q) h += th[q * 256 + j];
                    hist[j] = h;
                }
                __syncthreads();
                ++pass;
            }
            if (sample) {
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
                    if (digit == INT_MAX && run + h[q] > starget) {
                        digit = lane * 4 + q;
                        before = run;
                    }
                    run += h[q];
                }
                const int d = wave_min(digit);
                below = __shfl(before, d == INT_MAX ? 0 : d / 4, WAVE);
                prefix = (prefix << 8) | (d == INT_MAX ? 255 : d);
            }
            __syncthreads();  // histogram read before the next pass clears it
            if (shift == 0) break;
            shift -= 8;
        }
        if (sample) s = prefix < Bids ? prefix : Bids - 1;
    } else {
        if (sample && wig == 0) {
            const int ss = M.sample(samp_t, samp_target, samp_trow);
            if (lane == 0) redi[w0][10] = ss;
        }
        __syncthreads();
        if (sample) s = redi[w0][10];
    }
    R.s = s;
    R.u = u;
    double lqf = lqf_in, lqb = lqb_in;
    if (LRN && act && lw >= 0.0 && lz != nullptr) {
        lqf = cand_prob(lz, lcand, r, s);
        lqb = cand_prob(lz, lcand, s, r);
    }
    R.H = 1.0;
    R.dS = 0.0;
    if (propose_only) {  // uniform over the launch
        R.status = act ? (full ? -1 : 0) : 3;
        return R;
    }
    const bool ev = act && !full && s != r;

    // 3. rows s, r and columns s, r
    int o_r = 0, i_r = 0, o_s = 0, i_s = 0;
    if (ev) {
        const int slot_r = tab_find(tkey, cap, r), slot_s = tab_find(tkey, cap, s);
        o_r = slot_r >= 0 ? to[slot_r] : 0;
        i_r = slot_r >= 0 ? ti[slot_r] : 0;
        o_s = slot_s >= 0 ? to[slot_s] : 0;
        i_s = slot_s >= 0 ? ti[slot_s] : 0;
    }
    const int dout_r = ev ? S.dout[r] : 0, dout_s = ev ? S.dout[s] : 0;
    const int din_r = ev ? S.din[r] : 0, din_s = ev ? S.din[s] : 0;
    const int ndout_r = dout_r - k_out, ndout_s = dout_s + k_out;
    const int ndin_r = din_r - (k_in + self), ndin_s = din_s + (k_in + self);
    const bool t_ss = o_s > 0 || i_s > 0 || self, t_rr = o_r > 0 || i_r > 0 || self;
    const bool t_sr = o_r > 0 || i_s > 0, t_rs = o_s > 0 || i_r > 0;
    const int d_ss = o_s + i_s + self, d_rr = -(o_r + i_r + self);
    const int d_sr = o_r - i_s, d_rs = i_r - o_s;
    int m_ss = 0, m_rr = 0, m_sr = 0, m_rs = 0;
    double acc = 0.0;
    long long live = 0;
    if constexpr (CG > 1) {
        __shared__ ScanShare sh[CG > 1 ? CG : 1];
        __shared__ double sh_acc[CG > 1 ? CG * G : 1][CG > 1 ? CG : 1];
        __shared__ int sh_live[CG > 1 ? CG * G : 1][CG > 1 ? CG : 1];
        const int grp = w0 / G;
        if (gl == 0) {
            ScanShare& E = sh[grp];
            if (ev) {
                E.b[0] = M.rbeg(s);
                E.L[0] = M.rend(s) - E.b[0];
                E.b[1] = M.rbeg(r);
                E.L[1] = M.rend(r) - E.b[1];
                E.b[2] = M.cbeg(s);
                E.L[2] = M.cend(s) - E.b[2];
                E.b[3] = M.cbeg(r);
                E.L[3] = M.cend(r) - E.b[3];
            } else {
                for (int q = 0; q < 4; ++q) E.b[q] = E.L[q] = 0;
            }
            E.r = r;
            E.s = s;
            E.cap = cap;
            E.tab = tab;
            E.flags = (t_ss ? 1 : 0) | (t_rr ? 2 : 0) | (t_sr ? 4 : 0) | (t_rs ? 8 : 0);
            E.d_sr = d_sr;
            E.d_rs = d_rs;
            E.dout_s = dout_s;
            E.dout_r = dout_r;
            E.ndout_s = ndout_s;
            E.ndout_r = ndout_r;
            E.din_s = din_s;
            E.din_r = din_r;
            E.ndin_s = ndin_s;
            E.ndin_r = ndin_r;
            for (int q = 0; q < 4; ++q) E.m[q] = 0;
        }
        __syncthreads();  // every entry's scan described
        int off[CG + 1];
        off[0] = 0;
#pragma unroll
        for (int e = 0; e < CG; ++e) off[e + 1] = off[e] + sh[e].L[0] + sh[e].L[1] + sh[e].L[2] + sh[e].L[3];
        const int TT = off[CG];
        constexpr int NT = CG * G * WAVE;
        double ae[CG];
        int le[CG];
#pragma unroll
        for (int e = 0; e < CG; ++e) {
            ae[e] = 0.0;
            le[e] = 0;
        }
        for (int i0 = (int)threadIdx.x; i0 < TT; i0 += GU * NT) {
            int cc[GU], mm[GU], sg[GU], en[GU], dsel[GU];
#pragma unroll
            for (int q = 0; q < GU; ++q) {
                const int i = i0 + q * NT;
                cc[q] = -1;
                mm[q] = 0;
                sg[q] = -1;
                en[q] = 0;
                if (i < TT) {
                    int e = 0;
#pragma unroll
                    for (int k = 1; k < CG; ++k) e += i >= off[k] ? 1 : 0;
                    en[q] = e;
                    const ScanShare& E = sh[e];
                    int j = i - off[e];
                    if (j < E.L[0]) {
                        sg[q] = 0;
                        cc[q] = M.rkey(E.b[0] + j);
                        mm[q] = M.rval(E.b[0] + j);
                    } else if ((j -= E.L[0]) < E.L[1]) {
                        sg[q] = 1;
                        cc[q] = M.rkey(E.b[1] + j);
                        mm[q] = M.rval(E.b[1] + j);
                    } else if ((j -= E.L[1]) < E.L[2]) {
                        sg[q] = 2;
                        cc[q] = M.ckey(E.b[2] + j);
                        mm[q] = M.cval(E.b[2] + j);
                    } else {
                        j -= E.L[2];
                        sg[q] = 3;
                        cc[q] = M.ckey(E.b[3] + j);
                        mm[q] = M.cval(E.b[3] + j);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < GU; ++q) {
                dsel[q] = -1;
                const int seg = sg[q], c = cc[q], m = mm[q];
                if (seg < 0 || c < 0 || m == 0) continue;
                ScanShare& E = sh[en[q]];
#pragma unroll
                for (int k = 0; k < CG; ++k) le[k] += k == en[q] ? 1 : 0;
                const int er = E.r, es = E.s, ecap = E.cap, fl = E.flags;
                int* const et = E.tab;
                bool unchanged = true;
                if (seg == 0) {
                    if (c == es) {
                        E.m[0] = m;
                        unchanged = (fl & 1) == 0;
                    } else if (c == er) {
                        E.m[2] = m;
                        unchanged = (fl & 4) == 0 || E.d_sr == 0;
                    } else {
                        const int sl = tab_find(et, ecap, c);
                        if (sl >= 0) {
                            et[3 * ecap + sl] = m;
                            unchanged = et[ecap + sl] == 0;
                        }
                    }
                } else if (seg == 1) {
                    if (c == er) {
                        E.m[1] = m;
                        unchanged = (fl & 2) == 0;
                    } else if (c == es) {
                        E.m[3] = m;
                        unchanged = (fl & 8) == 0 || E.d_rs == 0;
                    } else {
                        const int sl = tab_find(et, ecap, c);
                        if (sl >= 0) {
                            et[5 * ecap + sl] = m;
                            unchanged = et[ecap + sl] == 0;
                        }
                    }
                } else {
                    if (c == er || c == es) continue;
                    const int sl = tab_find(et, ecap, c);
                    if (sl >= 0) {
                        et[(seg == 2 ? 4 : 6) * ecap + sl] = m;
                        unchanged = et[2 * ecap + sl] == 0;
                    }
                }
                if (unchanged) dsel[q] = c;
            }
            int dv[GU];
#pragma unroll
            for (int q = 0; q < GU; ++q) dv[q] = dsel[q] >= 0 ? (sg[q] < 2 ? S.din[dsel[q]] : S.dout[dsel[q]]) : 0;
#pragma unroll
            for (int q = 0; q < GU; ++q) {
                if (dsel[q] < 0) continue;
                const ScanShare& E = sh[en[q]];
                const int seg = sg[q], c = dsel[q], d = dv[q];
                const float mf = fi(mm[q]);
                double term;
                if (seg < 2) {
                    const int xo = seg == 0 ? E.dout_s : E.dout_r, nxo = seg == 0 ? E.ndout_s : E.ndout_r;
                    const int nd = c == E.r ? E.ndin_r : (c == E.s ? E.ndin_s : d);
                    const double t0 = (double)cell_entropy(mf, fi(d), fi(xo));
                    term = t0 - (double)cell_entropy(mf, fi(nd), fi(nxo));
                } else {
                    const int xi = seg == 2 ? E.din_s : E.din_r, nxi = seg == 2 ? E.ndin_s : E.ndin_r;
                    const double t0 = (double)cell_entropy(mf, fi(xi), fi(d));
                    term = t0 - (double)cell_entropy(mf, fi(nxi), fi(d));
                }
#pragma unroll
                for (int k = 0; k < CG; ++k)
                    if (k == en[q]) ae[k] += term;
            }
        }
#pragma unroll
        for (int e = 0; e < CG; ++e) {
            const double a = wave_sum(ae[e]);
            const long long l = wave_sum_ll(le[e]);
            if (lane == 0) {
                sh_acc[wave][e] = a;
                sh_live[wave][e] = (int)l;
            }
        }
        __syncthreads();  // the entries' partials and cells among {r, s} visible
        double tot = 0.0;
        long long lv = 0;
        for (int w = 0; w < CG * G; ++w) {
            tot += sh_acc[w][grp];
            lv += sh_live[w][grp];
        }
        acc = gl == 0 ? tot : 0.0;
        live = gl == 0 ? lv : 0;
        m_ss = sh[grp].m[0];
        m_rr = sh[grp].m[1];
        m_sr = sh[grp].m[2];
        m_rs = sh[grp].m[3];
    } else if (ev) {
        const int b0 = M.rbeg(s), L0 = M.rend(s) - b0;
        const int b1 = M.rbeg(r), L1 = M.rend(r) - b1;
        const int b2 = M.cbeg(s), L2 = M.cend(s) - b2;
        const int b3 = M.cbeg(r), L3 = M.cend(r) - b3;
        const int T1 = L0 + L1, T2 = T1 + L2, T = T2 + L3;
        int lo = 0, hi = T;
        if constexpr (Team::on) {  // this member's slice of the four rows, taken end to end
            const int per = (T + team.size - 1) / team.size;
            lo = min(T, team.rank * per);
            hi = min(T, lo + per);
        }
        // GU slots per lane per trip: keys and counts, then the neighbour-block records, then the
        // degrees of the unchanged cells' other blocks, each stage's loads issued together
        for (int i0 = lo + gl; i0 < hi; i0 += GU * GS) {
            int cc[GU], mm[GU], sg[GU], dsel[GU];
#pragma unroll
            for (int q = 0; q < GU; ++q) {
                const int i = i0 + q * GS;
                sg[q] = i >= hi ? -1 : i < L0 ? 0 : i < T1 ? 1 : i < T2 ? 2 : 3;
                cc[q] = -1;
                mm[q] = 0;
                if (sg[q] == 0) {
                    cc[q] = M.rkey(b0 + i);
                    mm[q] = M.rval(b0 + i);
                } else if (sg[q] == 1) {
                    cc[q] = M.rkey(b1 + (i - L0));
                    mm[q] = M.rval(b1 + (i - L0));
                } else if (sg[q] == 2) {
                    cc[q] = M.ckey(b2 + (i - T1));
                    mm[q] = M.cval(b2 + (i - T1));
                } else if (sg[q] == 3) {
                    cc[q] = M.ckey(b3 + (i - T2));
                    mm[q] = M.cval(b3 + (i - T2));
                }
            }
#pragma unroll
            for (int q = 0; q < GU; ++q) {
                dsel[q] = -1;
                const int seg = sg[q], c = cc[q], m = mm[q];
                if (seg < 0 || (View::kSlots && (c < 0 || m == 0))) continue;
                ++live;
                bool unchanged = true;
                if (seg == 0) {
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
                        }
                    }
                } else if (seg == 1) {
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
                        }
                    }
                } else {
                    if (c == r || c == s) continue;
                    const int sl = tab_find(tkey, cap, c);
                    if (sl >= 0) {
                        (seg == 2 ? tmbs : tmbr)[sl] = m;
                        unchanged = ti[sl] == 0;
                    }
                }
                if (unchanged) dsel[q] = c;
            }
            int dv[GU];
#pragma unroll
            for (int q = 0; q < GU; ++q) dv[q] = dsel[q] >= 0 ? (sg[q] < 2 ? S.din[dsel[q]] : S.dout[dsel[q]]) : 0;
#pragma unroll
            for (int q = 0; q < GU; ++q) {
                if (dsel[q] < 0) continue;
                const int seg = sg[q], c = dsel[q], d = dv[q];
                const float mf = fi(mm[q]);
                if (seg < 2) {
                    const int xo = seg == 0 ? dout_s : dout_r, nxo = seg == 0 ? ndout_s : ndout_r;
                    const int nd = c == r ? ndin_r : (c == s ? ndin_s : d);
                    acc += (double)cell_entropy(mf, fi(d), fi(xo));
                    acc -= (double)cell_entropy(mf, fi(nd), fi(nxo));
                } else {
                    const int xi = seg == 2 ? din_s : din_r, nxi = seg == 2 ? ndin_s : ndin_r;
                    acc += (double)cell_entropy(mf, fi(xi), fi(d));
                    acc -= (double)cell_entropy(mf, fi(nxi), fi(d));
                }
            }
        }
    }
    m_ss = wave_max(m_ss);
    m_rr = wave_max(m_rr);
    m_sr = wave_max(m_sr);
    m_rs = wave_max(m_rs);
    live = wave_sum_ll(live);
    if (lane == 0) {
        redi[wave][4] = m_ss;
        redi[wave][5] = m_rr;
        redi[wave][6] = m_sr;
        redi[wave][7] = m_rs;
        redi[wave][8] = (int)live;
    }
    __syncthreads();  // cell records in the table and partials visible
    live = 0;
    for (int i = 0; i < G; ++i) {
        m_ss = max(m_ss, redi[w0 + i][4]);
        m_rr = max(m_rr, redi[w0 + i][5]);
        m_sr = max(m_sr, redi[w0 + i][6]);
        m_rs = max(m_rs, redi[w0 + i][7]);
        live += redi[w0 + i][8];
    }
    // a member saw only its slice of the rows: member 0 reads the cells of the neighbour blocks
    // and of r and s by lookups instead
    const bool lead4 = ev && (!Team::on || team.rank == 0);
    if constexpr (Team::on) {
        if (lead4) {
            m_ss = M.rlookup(s, s);
            m_rr = M.rlookup(r, r);
            m_sr = M.rlookup(s, r);
            m_rs = M.rlookup(r, s);
            for (int j = gl; j < cap; j += GS) {
                const int b = tkey[j];
                if (b < 0 || b == r || b == s) continue;
                tmsb[j] = M.rlookup(s, b);
                tmrb[j] = M.rlookup(r, b);
                tmbs[j] = M.clookup(s, b);
                tmbr[j] = M.clookup(r, b);
            }
        }
        __syncthreads();
    }

    // 4. Delta keys and Hastings sums; with the learned mixture also the plain sampler's own
    // probabilities (times k) of s now and of r after the move, as in eval_move
    const bool mix = LRN && lw >= 0.0;
    double pf = 0.0, pb = 0.0, pf0 = 0.0, pb0 = 0.0;
    int nuniq = 0;
    if (lead4) {
        for (int j = gl; j < cap; j += GS) {
            const int b = tkey[j];
            if (b < 0) continue;
            nuniq++;
            const int ob_ = to[j], ib_ = ti[j];
            const int degb = S.deg[b];
            int Msb, Mbs, Mrb, Mbr, Drb, Dbr, nd;
            if (b == r) {
                Msb = m_sr; Mbs = m_rs; Mrb = m_rr; Mbr = m_rr; Drb = d_rr; Dbr = d_rr;
                nd = (dout_r - k_out) + (din_r - k_in) - (m_rr + d_rr);
            } else if (b == s) {
                Msb = m_ss; Mbs = m_ss; Mrb = m_rs; Mbr = m_sr; Drb = d_rs; Dbr = d_sr;
                nd = (dout_s + k_out) + (din_s + k_in) - (m_ss + d_ss);
            } else {
                Msb = tmsb[j]; Mbs = tmbs[j]; Mrb = tmrb[j]; Mbr = tmbr[j]; Drb = -ob_; Dbr = -ib_;
                nd = degb;
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
            if (mix && ob_ + ib_ > 0) {  // v's neighbour slots in b (the self-loop's slot: below)
                const int tf = b == r ? dout_r + din_r : (b == s ? dout_s + din_s : S.dout[b] + S.din[b]);
                const int tb = b == r ? ndout_r + ndin_r : (b == s ? ndout_s + ndin_s : S.dout[b] + S.din[b]);
                if (tf > 0) pf0 += (double)(ob_ + ib_) * (double)(Msb + Mbs) / (double)tf;
                if (tb > 0) pb0 += (double)(ob_ + ib_) * (double)(Mrb + Drb + Mbr + Dbr) / (double)tb;
            }
            const int nb = ob_ + ib_ + (b == r ? self : 0);
            if (nb == 0) continue;
            const double count = (double)nb;
            pf += count * ((double)(Msb + Mbs) + 1.0) / (double)(degb + B);
            pb += count * ((double)(Mrb + Drb + Mbr + Dbr) + 1.0) / (double)(nd + B);
        }
        if (gl == 0 && mix && self) {  // the self-loop's slot: block r now, block s after the move
            const int tf = dout_r + din_r, tb = ndout_s + ndin_s;
            if (tf > 0) pf0 += (double)(m_sr + m_rs) / (double)tf;
            if (tb > 0) pb0 += (double)((m_rs + d_rs) + (m_sr + d_sr)) / (double)tb;
        }
        if (gl == 0) {
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
    }
    acc = wave_sum(acc);
    pf = wave_sum(pf);
    pb = wave_sum(pb);
    if (mix) {
        pf0 = wave_sum(pf0);
        pb0 = wave_sum(pb0);
    }
    nuniq = (int)wave_sum_ll(nuniq);
    if (lane == 0) {
        redd[wave][0] = acc;
        redd[wave][1] = pf;
        redd[wave][2] = pb;
        redd[wave][3] = pf0;
        redd[wave][4] = pb0;
        redi[wave][9] = nuniq;
    }
    __syncthreads();
    double dS = 0.0;
    pf = 0.0;
    pb = 0.0;
    pf0 = 0.0;
    pb0 = 0.0;
    nuniq = 0;
    for (int i = 0; i < G; ++i) {
        dS += redd[w0 + i][0];
        pf += redd[w0 + i][1];
        pb += redd[w0 + i][2];
        pf0 += redd[w0 + i][3];
        pb0 += redd[w0 + i][4];
        nuniq += redi[w0 + i][9];
    }
    if constexpr (Team::on) {
        // the members' partial sums, added in member order by member 0
        if (gl == 0) {
            team.part[2 * team.rank] = dS;
            team.part[2 * team.rank + 1] = (double)live;
        }
        team.sync();
        if (team.rank != 0) {
            R.status = 3;
            return R;
        }
        dS = 0.0;
        double lv = 0.0;
        for (int q = 0; q < team.size; ++q) {
            dS += team.part[2 * q];
            lv += team.part[2 * q + 1];
        }
        live = (long long)lv;
    }
    R.nuniq = nuniq;
    R.m_ss = m_ss;
    R.m_rr = m_rr;
    R.m_sr = m_sr;
    R.m_rs = m_rs;
    R.d_ss = d_ss;
    R.d_rr = d_rr;
    R.d_sr = d_sr;
    R.d_rs = d_rs;
    R.nnz_rs = live;
    if (!act) return R;
    if (full) {
        R.status = -1;
        return R;
    }
    if (!ev) {
        R.status = 0;
        return R;
    }
    double H = (k == 0) ? 1.0 : pb / pf;
    if (mix && k > 0) {
        // mixture proposal w q_plain + (1 - w) q_net, both the samplers' own probabilities, the
        // learned part from the fixed candidate scores
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
    return R;
}

}  // namespace sbpx
