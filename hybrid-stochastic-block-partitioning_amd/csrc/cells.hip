// The reference's finer entry points on one move or one merge, on the device's block matrix:
// delta MDL of a given Delta (entropy::delta_mdl(blockmodel, delta, proposal), entropy.cpp:296-401),
// delta MDL of merging block r into s (entropy::block_merge_delta_mdl, entropy.cpp:127-180), and
// cell lookups (blockmatrix()->get(row, col), dict_transpose_matrix.cpp:74-84). They serve the C++
// layer that restores those signatures; the sweeps and merge phases evaluate moves in bulk
// elsewhere.
#include <cfloat>

#include "internal.hpp"
#include "merge_core.hpp"

namespace sbpx {

constexpr int CELLS_TB = 256;

// The Delta's value at (row, col): its entries are sorted by (row, col), one per key; a key stored
// with a zero change reads 0, as Delta::get does (delta.hpp:85-95).
__device__ __forceinline__ int delta_get(const int* __restrict__ er, const int* __restrict__ ec,
                                         const int* __restrict__ ed, int n, int row, int col) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (er[mid] < row || (er[mid] == row && ec[mid] < col)) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && er[lo] == row && ec[lo] == col) ? ed[lo] : 0;
}

// One workgroup: the entries loop, then the untouched cells of rows r, s and columns r, s (the
// column loops leave rows r and s to the row loops). Keys stored with a zero change are counted
// in the entries loop and not skipped by the row/column loops, as in the reference.
__global__ void __launch_bounds__(CELLS_TB) delta_mdl_kernel(CsrView M, const int* __restrict__ dout,
                                                             const int* __restrict__ din, int r, int s, int n,
                                                             const int* __restrict__ er, const int* __restrict__ ec,
                                                             const int* __restrict__ ed, int k_out, int k_in,
                                                             int self_w, double* __restrict__ out) {
    __shared__ double part[CELLS_TB / WAVE];
    auto deg_in = [&](int x) {
        int v = din[x];
        if (x == r) v -= k_in + self_w;
        else if (x == s) v += k_in + self_w;
        return (float)v;
    };
    auto deg_out = [&](int x) {
        int v = dout[x];
        if (x == r) v -= k_out;
        else if (x == s) v += k_out;
        return (float)v;
    };
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += CELLS_TB) {
        const int row = er[i], col = ec[i];
        const int m = M.rlookup(row, col);
        acc += (double)cell_entropy((float)m, (float)din[col], (float)dout[row]);
        acc -= (double)cell_entropy((float)(m + ed[i]), deg_in(col), deg_out(row));
    }
    for (int q = 0; q < 2; ++q) {
        const int x = q == 0 ? r : s;
        for (int j = M.rbeg(x) + threadIdx.x; j < M.rend(x); j += CELLS_TB) {
            const int c = M.rkey(j), m = M.rval(j);
            if (delta_get(er, ec, ed, n, x, c) != 0) continue;
            acc += (double)cell_entropy((float)m, (float)din[c], (float)dout[x]);
            acc -= (double)cell_entropy((float)m, deg_in(c), deg_out(x));
        }
        for (int j = M.cbeg(x) + threadIdx.x; j < M.cend(x); j += CELLS_TB) {
            const int row = M.ckey(j), m = M.cval(j);
            if (delta_get(er, ec, ed, n, row, x) != 0 || row == r || row == s) continue;
            acc += (double)cell_entropy((float)m, (float)din[x], (float)dout[row]);
            acc -= (double)cell_entropy((float)m, deg_in(x), deg_out(row));
        }
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) part[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < CELLS_TB / WAVE; ++w) t += part[w];
        out[0] = t;
    }
}

// One wavefront per pair: delta MDL of merging r[i] into s[i] with r's whole row and column
// (proposal counts dout[r], din[r]).
__global__ void __launch_bounds__(CELLS_TB) merge_eval_kernel(CsrView M, const int* __restrict__ dout,
                                                              const int* __restrict__ din, int n,
                                                              const int* __restrict__ r, const int* __restrict__ s,
                                                              double* __restrict__ out) {
    const int i = blockIdx.x * (CELLS_TB / WAVE) + threadIdx.x / WAVE;
    if (i >= n) return;
    const int rr = r[i], ss = s[i];
    const double dS = merge_delta_mdl_wave(M, dout, din, rr, ss, dout[rr], din[rr]);
    if (lane_id() == 0) out[i] = dS;
}

__global__ void cells_kernel(CsrView M, int n, const int* __restrict__ rows, const int* __restrict__ cols,
                             int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = M.rlookup(rows[i], cols[i]);
}

static CsrView csr_view(Ctx& c) {
    return CsrView{c.bm.row_ptr, c.bm.row_col, c.bm.row_val, c.bm.row_cum,
                   c.bm.col_ptr, c.bm.col_row, c.bm.col_val, c.bm.col_cum};
}

void launch_delta_mdl(Ctx& c, int r, int s, int n, const int* er, const int* ec, const int* ed, int k_out, int k_in,
                      int self_w, double* d_out) {
    delta_mdl_kernel<<<1, CELLS_TB, 0, c.stream>>>(csr_view(c), c.bm.dout, c.bm.din, r, s, n, er, ec, ed, k_out, k_in,
                                                   self_w, d_out);
    SBPX_HIP(hipGetLastError());
}

void launch_merge_eval(Ctx& c, int n, const int* r, const int* s, double* d_out) {
    if (n <= 0) return;
    constexpr int per = CELLS_TB / WAVE;
    merge_eval_kernel<<<(n + per - 1) / per, CELLS_TB, 0, c.stream>>>(csr_view(c), c.bm.dout, c.bm.din, n, r, s,
                                                                      d_out);
    SBPX_HIP(hipGetLastError());
}

void launch_cells(Ctx& c, int n, const int* rows, const int* cols, int* d_out) {
    if (n <= 0) return;
    cells_kernel<<<(n + CELLS_TB - 1) / CELLS_TB, CELLS_TB, 0, c.stream>>>(csr_view(c), n, rows, cols, d_out);
    SBPX_HIP(hipGetLastError());
}

}  // namespace sbpx
