// Delta MDL of merging block r into block s (entropy::block_merge_delta_mdl,
// entropy.cpp:127-180), read through a view of the block matrix; shared by the batched merge
// proposals (CSR snapshot) and the heap-ordered merge selection (mutable hash tables).
//   old cells of row r and column r (their new value is 0 and contributes nothing),
//   cells (s, c) and (c, s) that gain row/column r's counts, the diagonal (s, s),
//   and the untouched cells of row s and column s, whose degrees change.
#pragma once
#include "eval_core.hpp"

namespace sbpx {

__device__ __forceinline__ float fm(int x) { return (float)x; }

// Delta MDL of merging r into s, read through a view (CSR snapshot or mutable hash tables); the
// whole wavefront takes part. k_out / k_in: the row / column sums of r.
template <class View>
__device__ double merge_delta_mdl_wave(const View& M, const int* dout, const int* din, int r, int s, int k_out,
                                       int k_in) {
    const int lane = lane_id();
    const int r0 = M.rbeg(r), r1 = M.rend(r);
    const int c0 = M.cbeg(r), c1 = M.cend(r);
    const int s0 = M.rbeg(s), s1 = M.rend(s);
    const int cs0 = M.cbeg(s), cs1 = M.cend(s);
    const int dout_r = dout[r], din_r = din[r], dout_s = dout[s], din_s = din[s];
    const int ndout_s = dout_s + k_out, ndin_s = din_s + k_in;
    double acc = 0.0;
    int m_rr = 0, m_rs = 0, m_sr = 0;
    // row r: old (r, c); keys (s, c) for c not in {r, s}
    for (int j = r0 + lane; j < r1; j += WAVE) {
        int c = M.rkey(j), m = M.rval(j);
        if (View::kSlots && (c < 0 || m == 0)) continue;
        int dc = din[c];
        acc += (double)cell_entropy(fm(m), fm(dc), fm(dout_r));
        if (c == r) {
            m_rr = m;
        } else if (c == s) {
            m_rs = m;
        } else {
            int msc = M.rlookup(s, c);
            acc += (double)cell_entropy(fm(msc), fm(dc), fm(dout_s));
            acc -= (double)cell_entropy(fm(msc + m), fm(dc), fm(ndout_s));
        }
    }
    // column r without (r, r): old (i, r); keys (i, s) for i not in {r, s}
    for (int j = c0 + lane; j < c1; j += WAVE) {
        int i = M.ckey(j);
        if (View::kSlots && i < 0) continue;
        if (i == r) continue;
        int m = M.cval(j);
        if (View::kSlots && m == 0) continue;
        int di = dout[i];
        acc += (double)cell_entropy(fm(m), fm(din_r), fm(di));
        if (i == s) {
            m_sr = m;
        } else {
            int mis = M.clookup(s, i);
            acc += (double)cell_entropy(fm(mis), fm(din_s), fm(di));
            acc -= (double)cell_entropy(fm(mis + m), fm(ndin_s), fm(di));
        }
    }
    m_rr = wave_max(m_rr);
    m_rs = wave_max(m_rs);
    m_sr = wave_max(m_sr);
    const int d_ss = m_rr + m_rs + m_sr;
    // row s: untouched cells with the new degrees
    int m_ss = 0;
    for (int j = s0 + lane; j < s1; j += WAVE) {
        int c = M.rkey(j), m = M.rval(j);
        if (View::kSlots && (c < 0 || m == 0)) continue;
        if (c == s) m_ss = m;
        bool touched;
        if (c == s) touched = d_ss > 0;
        else if (c == r) touched = true;
        else touched = M.rlookup(r, c) != 0;
        if (!touched) {
            int dc = din[c];
            int ndc = (c == s) ? ndin_s : dc;
            acc += (double)cell_entropy(fm(m), fm(dc), fm(dout_s));
            acc -= (double)cell_entropy(fm(m), fm(ndc), fm(ndout_s));
        }
    }
    // column s, rows not in {r, s}: untouched cells
    for (int j = cs0 + lane; j < cs1; j += WAVE) {
        int i = M.ckey(j);
        if (View::kSlots && i < 0) continue;
        if (i == r || i == s) continue;
        int m = M.cval(j);
        if (View::kSlots && m == 0) continue;
        if (M.clookup(r, i) != 0) continue;
        int di = dout[i];
        acc += (double)cell_entropy(fm(m), fm(din_s), fm(di));
        acc -= (double)cell_entropy(fm(m), fm(ndin_s), fm(di));
    }
    m_ss = wave_max(m_ss);
    if (lane == 0 && d_ss > 0) {
        acc += (double)cell_entropy(fm(m_ss), fm(din_s), fm(dout_s));
        acc -= (double)cell_entropy(fm(m_ss + d_ss), fm(ndin_s), fm(ndout_s));
    }
    return wave_sum(acc);
}

}  // namespace sbpx
