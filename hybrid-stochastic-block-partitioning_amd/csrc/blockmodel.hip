// Device Blockmodel: the block matrix M (M[b(u), b(v)] = number of edges u -> v) rebuilt from
// the assignment, as CSR and CSC with per-row / per-column inclusive prefix sums, plus
// dout = row sums, din = column sums, diag and the total degree deg = dout + din - diag - drift.
// Replaces Blockmodel::initialize_edge_counts (blockmodel.cpp:187-255) and the robin-map
// DictTransposeMatrix. The log posterior (blockmodel.cpp:267-285) is an FP64 reduction over
// the nonzeros with the FP32 log the reference uses.
#include <algorithm>

#include "internal.hpp"

namespace sbpx {

constexpr int TB = 256;

static inline int blocks_for(long n, int tb = TB) { return (int)std::max<long>(1, (n + tb - 1) / tb); }

static int bits_for(int B) {
    int bits = 1;
    while ((1L << bits) < B) ++bits;
    return bits;
}

// Sort keys use block ids, or, for a batch of graphs, their dense numbering among the blocks in
// use (dense[x], increasing with x, -1 for ids not in use; inv maps it back), so that the keys
// are as narrow as the blocks in use allow.
template <typename K>
__global__ void edge_keys_kernel(const int* __restrict__ edge_src, const int* __restrict__ out_idx,
                                 const int* __restrict__ assign, int E, int B, int shift, K* __restrict__ keys,
                                 int* __restrict__ err, const int* __restrict__ dense) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    int bu = assign[edge_src[i]];
    int bw = assign[out_idx[i]];
    if ((unsigned)bu >= (unsigned)B || (unsigned)bw >= (unsigned)B) {
        atomicOr(err, DEV_ERR_BLOCK_RANGE);
        bu = 0;
        bw = 0;
    }
    if (dense != nullptr) {
        bu = dense[bu];
        bw = dense[bw];
        if (bu < 0 || bw < 0) {
            atomicOr(err, DEV_ERR_BLOCK_RANGE);
            bu = 0;
            bw = 0;
        }
    }
    keys[i] = ((K)bu << shift) | (K)bw;
}

template <typename K>
__global__ void head_flags_kernel(const K* __restrict__ keys, int n, int* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Unique keys and the start index of each run; start[nnz] = n.
template <typename K>
__global__ void compact_runs_kernel(const K* __restrict__ keys, const int* __restrict__ flags,
                                    const int* __restrict__ pos, int n, K* __restrict__ ukeys,
                                    int* __restrict__ start) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flags[i]) {
        int j = pos[i];
        ukeys[j] = keys[i];
        start[j] = i;
    }
    if (i == n - 1) start[pos[n]] = n;
}

template <typename K>
__global__ void decode_csr_kernel(const K* __restrict__ ukeys, const int* __restrict__ start, int nnz,
                                  int shift, K mask, int* __restrict__ coo_row, int* __restrict__ row_col,
                                  int* __restrict__ row_val, const int* __restrict__ inv) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    K key = ukeys[j];
    int row = (int)(key >> shift), col = (int)(key & mask);
    if (inv != nullptr) {
        row = inv[row];
        col = inv[col];
    }
    coo_row[j] = row;
    row_col[j] = col;
    row_val[j] = start[j + 1] - start[j];
}

// ptr[x] = first index j with sorted_idx[j] >= x, for x in [0, B].
__global__ void segment_ptr_kernel(const int* __restrict__ sorted_idx, int nnz, int B, int* __restrict__ ptr) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x > B) return;
    ptr[x] = lower_bound_lane(sorted_idx, 0, nnz, x);
}

// cum[j] = inclusive prefix of val inside the segment of j; tot[x] = segment sum.
// ex is the exclusive scan of val over all nonzeros (nnz + 1 entries).
__global__ void segment_cum_kernel(const int* __restrict__ seg_of, const int* __restrict__ ptr,
                                   const int* __restrict__ ex, int nnz, int* __restrict__ cum) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    cum[j] = ex[j + 1] - ex[ptr[seg_of[j]]];
}
__global__ void segment_total_kernel(const int* __restrict__ ptr, const int* __restrict__ ex, int B,
                                     int* __restrict__ tot) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= B) return;
    tot[x] = ex[ptr[x + 1]] - ex[ptr[x]];
}

__global__ void diag_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col,
                            const int* __restrict__ row_val, int nnz, int* __restrict__ diag) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    if (coo_row[j] == row_col[j]) diag[coo_row[j]] = row_val[j];
}

template <typename K>
__global__ void csc_keys_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col, int nnz,
                                int shift, K* __restrict__ keys, const int* __restrict__ dense) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    int row = coo_row[j], col = row_col[j];
    if (dense != nullptr) {
        row = dense[row];
        col = dense[col];
    }
    keys[j] = ((K)col << shift) | (K)row;
}

template <typename K>
__global__ void decode_csc_kernel(const K* __restrict__ keys, const int* __restrict__ vals, int nnz, int shift,
                                  K mask, int* __restrict__ col_of, int* __restrict__ col_row,
                                  int* __restrict__ col_val, const int* __restrict__ inv) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    K key = keys[j];
    int col = (int)(key >> shift), row = (int)(key & mask);
    if (inv != nullptr) {
        col = inv[col];
        row = inv[row];
    }
    col_of[j] = col;
    col_row[j] = row;
    col_val[j] = vals[j];
}

// Current entropy of every nonzero, the "before" term of each cell an evaluation reads: CSR
// entry j is cell (coo_row[j], row_col[j]); CSC entry j is cell (col_row[j], col_of[j]).
__global__ void cell_entropy_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col,
                                    const int* __restrict__ row_val, const int* __restrict__ col_of,
                                    const int* __restrict__ col_row, const int* __restrict__ col_val,
                                    const int* __restrict__ dout, const int* __restrict__ din, int nnz,
                                    float* __restrict__ row_h, float* __restrict__ col_h) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnz) return;
    row_h[j] = cell_entropy((float)row_val[j], (float)din[row_col[j]], (float)dout[coo_row[j]]);
    col_h[j] = cell_entropy((float)col_val[j], (float)din[col_of[j]], (float)dout[col_row[j]]);
}

__global__ void degrees_kernel(const int* __restrict__ dout, const int* __restrict__ din,
                               const int* __restrict__ diag, const int* __restrict__ drift, int B,
                               int* __restrict__ deg) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= B) return;
    deg[x] = dout[x] + din[x] - diag[x] - drift[x];
}

template <typename K>
static void build_impl(Ctx& c) {
    DevGraph& g = c.g;
    DevBlockmodel& bm = c.bm;
    hipStream_t st = c.stream;
    const int B = bm.B;
    const bool dn = c.batch.G > 0;  // a batch: keys over the blocks in use
    const int* dense = dn ? c.batch.d_dense : nullptr;
    const int* inv = dn ? c.batch.d_inv : nullptr;
    const int bits = bits_for(dn ? c.batch.used : B);
    // key = row << bits | column: 2 * bits significant bits in 32- or 64-bit keys, sorted by
    // 8-bit digits from the lowest
    const int shift = bits;
    const K mask = (K)((((K)1) << bits) - 1);
    K* keys0 = (K*)c.keys0;
    K* keys1 = (K*)c.keys1;
    const int E = g.E;

    // 1. edge keys (block(u), block(v)) in CSR edge order, sorted by (row, col)
    int shifts[8];
    int np = 0;
    int nbytes = (bits + 7) / 8;
    for (int b = 0; b < 2 * bits; b += 8) shifts[np++] = b;
    int nnz = 0;
    if (E > 0) {
        edge_keys_kernel<K><<<blocks_for(E), TB, 0, st>>>(g.edge_src, g.out_idx, bm.assign, E, B, shift, keys0,
                                                          c.dev_err, dense);
        int which = radix_sort<K>(keys0, keys1, nullptr, nullptr, E, shifts, np, c.hist, c.scan_tmp, st);
        K* sorted = which ? keys1 : keys0;
        K* ukeys = which ? keys0 : keys1;
        // 2. run-length encode into unique cells with counts
        head_flags_kernel<K><<<blocks_for(E), TB, 0, st>>>(sorted, E, c.flags);
        scan_exclusive(c.flags, c.pos, E, c.scan_tmp, st);
        SBPX_HIP(hipMemcpyAsync(&nnz, c.pos + E, sizeof(int), hipMemcpyDeviceToHost, st));
        compact_runs_kernel<K><<<blocks_for(E), TB, 0, st>>>(sorted, c.flags, c.pos, E, ukeys, c.vals0);
        SBPX_HIP(hipStreamSynchronize(st));
        decode_csr_kernel<K><<<blocks_for(nnz), TB, 0, st>>>(ukeys, c.vals0, nnz, shift, mask, bm.coo_row,
                                                             bm.row_col, bm.row_val, inv);
    }
    bm.nnz = nnz;
    // 3. row pointers, row prefix sums, dout, diagonal
    segment_ptr_kernel<<<blocks_for(B + 1), TB, 0, st>>>(bm.coo_row, nnz, B, bm.row_ptr);
    scan_exclusive(bm.row_val, c.pos, nnz, c.scan_tmp, st);
    if (nnz > 0) segment_cum_kernel<<<blocks_for(nnz), TB, 0, st>>>(bm.coo_row, bm.row_ptr, c.pos, nnz, bm.row_cum);
    segment_total_kernel<<<blocks_for(B), TB, 0, st>>>(bm.row_ptr, c.pos, B, bm.dout);
    SBPX_HIP(hipMemsetAsync(bm.diag, 0, sizeof(int) * B, st));
    if (nnz > 0) diag_kernel<<<blocks_for(nnz), TB, 0, st>>>(bm.coo_row, bm.row_col, bm.row_val, nnz, bm.diag);
    // 4. CSC: stable sort of the unique cells by column (input already ordered by row)
    if (nnz > 0) {
        csc_keys_kernel<K><<<blocks_for(nnz), TB, 0, st>>>(bm.coo_row, bm.row_col, nnz, shift, keys0, dense);
        SBPX_HIP(hipMemcpyAsync(c.vals0, bm.row_val, sizeof(int) * nnz, hipMemcpyDeviceToDevice, st));
        int cshifts[4];
        int cnp = 0;
        for (int i = 0; i < nbytes; ++i) cshifts[cnp++] = shift + 8 * i;
        int which = radix_sort<K>(keys0, keys1, c.vals0, c.vals1, nnz, cshifts, cnp, c.hist, c.scan_tmp, st);
        K* ck = which ? keys1 : keys0;
        int* cv = which ? c.vals1 : c.vals0;
        int* col_of = c.flags;  // column index per CSC entry (scratch)
        decode_csc_kernel<K><<<blocks_for(nnz), TB, 0, st>>>(ck, cv, nnz, shift, mask, col_of, bm.col_row,
                                                             bm.col_val, inv);
        segment_ptr_kernel<<<blocks_for(B + 1), TB, 0, st>>>(col_of, nnz, B, bm.col_ptr);
        scan_exclusive(bm.col_val, c.pos, nnz, c.scan_tmp, st);
        segment_cum_kernel<<<blocks_for(nnz), TB, 0, st>>>(col_of, bm.col_ptr, c.pos, nnz, bm.col_cum);
        segment_total_kernel<<<blocks_for(B), TB, 0, st>>>(bm.col_ptr, c.pos, B, bm.din);
        cell_entropy_kernel<<<blocks_for(nnz), TB, 0, st>>>(bm.coo_row, bm.row_col, bm.row_val, col_of, bm.col_row,
                                                            bm.col_val, bm.dout, bm.din, nnz, bm.row_h, bm.col_h);
    } else {
        SBPX_HIP(hipMemsetAsync(bm.col_ptr, 0, sizeof(int) * (B + 1), st));
        SBPX_HIP(hipMemsetAsync(bm.din, 0, sizeof(int) * B, st));
    }
    degrees_kernel<<<blocks_for(B), TB, 0, st>>>(bm.dout, bm.din, bm.diag, bm.drift, B, bm.deg);
    bm.valid = true;
}

void rebuild_keep_drift(Ctx& c) {
    const int bits = bits_for(c.batch.G > 0 ? c.batch.used : c.bm.B);
    if (2 * bits <= 32) build_impl<uint32_t>(c);
    else build_impl<uint64_t>(c);
}

void build_blockmodel(Ctx& c) {
    SBPX_HIP(hipMemsetAsync(c.bm.drift, 0, sizeof(int) * c.bm.B, c.stream));
    rebuild_keep_drift(c);
}

// ------------------------------------------------------------------------------------------
// The log posterior is summed per graph (a one-graph context has one): the cells of graph g,
// [row_ptr[bo[g]], row_ptr[bo[g] + nb[g]]), fall in chunks of LP_CHUNK from its first cell; one
// workgroup sums a chunk in a fixed order, then one wavefront per graph adds its chunks in order.
// The result depends on the graph's own cells only, not on where they sit in a union.
constexpr int LP_CHUNK = 2048;

// seg = lo[G], hi[G], chunk offsets[G + 1]; one workgroup.
__global__ void lp_segments_kernel(const int* __restrict__ row_ptr, int G, const int* __restrict__ bo,
                                   const int* __restrict__ nb, int nnz, int* __restrict__ seg) {
    __shared__ int lds[TB];
    const int per = (G + TB - 1) / TB;
    const int g0 = threadIdx.x * per, g1 = min(G, g0 + per);
    int local = 0;
    for (int g = g0; g < g1; ++g) {
        const int lo = bo ? row_ptr[bo[g]] : 0;
        const int hi = bo ? row_ptr[bo[g] + nb[g]] : nnz;
        seg[g] = lo;
        seg[G + g] = hi;
        local += (hi - lo + LP_CHUNK - 1) / LP_CHUNK;
    }
    lds[threadIdx.x] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < TB; ++t) {
            const int x = lds[t];
            lds[t] = acc;
            acc += x;
        }
        seg[3 * G] = acc;
    }
    __syncthreads();
    int off = lds[threadIdx.x];
    for (int g = g0; g < g1; ++g) {
        seg[2 * G + g] = off;
        off += (seg[G + g] - seg[g] + LP_CHUNK - 1) / LP_CHUNK;
    }
}

__global__ void __launch_bounds__(TB) lp_chunk_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col,
                                                      const int* __restrict__ row_val, const int* __restrict__ dout,
                                                      const int* __restrict__ din, const int* __restrict__ seg, int G,
                                                      double* __restrict__ part, int* __restrict__ err) {
    __shared__ double lds[TB / WAVE];
    const int* choff = seg + 2 * G;
    const int w = blockIdx.x;
    if (w >= choff[G]) return;
    int g = 0, top = G;  // the last g with choff[g] <= w
    while (top - g > 1) {
        const int mid = (g + top) >> 1;
        if (choff[mid] <= w) g = mid;
        else top = mid;
    }
    const int lo = seg[g] + (w - choff[g]) * LP_CHUNK;
    const int end = min(seg[G + g], lo + LP_CHUNK);
    double acc = 0.0;
    for (int j = lo + (int)threadIdx.x; j < end; j += TB) {
        double value = (double)row_val[j];
        double d_out = (double)dout[coo_row[j]];
        double d_in = (double)din[row_col[j]];
        if (d_out == 0.0 || d_in == 0.0) atomicOr(err, DEV_ERR_ZERO_DEGREE);
        double ratio = value / (d_out * d_in);
        acc += value * (double)host_logf((float)ratio);
    }
    acc = wave_sum(acc);
    if (lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < TB / WAVE; ++q) s += lds[q];
        part[w] = s;
    }
}

__global__ void lp_graph_sum_kernel(const double* __restrict__ part, const int* __restrict__ seg, int G,
                                    double* __restrict__ out) {
    const int g = blockIdx.x * (TB / WAVE) + threadIdx.x / WAVE;
    if (g >= G) return;
    const int* choff = seg + 2 * G;
    double acc = 0.0;
    for (int i = choff[g] + lane_id(); i < choff[g + 1]; i += WAVE) acc += part[i];
    acc = wave_sum(acc);
    if (lane_id() == 0) out[g] = acc;
}

void log_posterior_segments(Ctx& c, double* out) {
    const int G = c.batch.G > 0 ? c.batch.G : 1;
    const int nnz = c.bm.nnz;
    const size_t chunks = (size_t)nnz / LP_CHUNK + G + 1;
    if (c.lp_part_slots < chunks + G) {
        SBPX_HIP(hipFree(c.lp_part));
        SBPX_HIP(hipFree(c.lp_seg));
        c.lp_part = nullptr;
        c.lp_seg = nullptr;
        c.lp_part_slots = 0;
        const size_t slots = chunks + G + 1024;
        SBPX_HIP(hipMalloc((void**)&c.lp_part, sizeof(double) * slots));
        SBPX_HIP(hipMalloc((void**)&c.lp_seg, sizeof(int) * (3 * (size_t)G + 1)));
        c.lp_part_slots = slots;
        c.lp_seg_graphs = G;
    } else if (c.lp_seg_graphs < G) {
        SBPX_HIP(hipFree(c.lp_seg));
        c.lp_seg = nullptr;
        SBPX_HIP(hipMalloc((void**)&c.lp_seg, sizeof(int) * (3 * (size_t)G + 1)));
        c.lp_seg_graphs = G;
    }
    const bool batch = c.batch.G > 0;
    lp_segments_kernel<<<1, TB, 0, c.stream>>>(c.bm.row_ptr, G, batch ? c.batch.d_bo : nullptr,
                                               batch ? c.batch.d_nb : nullptr, nnz, c.lp_seg);
    lp_chunk_kernel<<<(int)chunks, TB, 0, c.stream>>>(c.bm.coo_row, c.bm.row_col, c.bm.row_val, c.bm.dout,
                                                      c.bm.din, c.lp_seg, G, c.lp_part, c.dev_err);
    double* d_out = c.lp_part + chunks;
    lp_graph_sum_kernel<<<(G + TB / WAVE - 1) / (TB / WAVE), TB, 0, c.stream>>>(c.lp_part, c.lp_seg, G, d_out);
    SBPX_HIP(hipGetLastError());
    SBPX_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * G, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    check_device_error(c);
}

double log_posterior(Ctx& c) {
    if (c.batch.G > 0) throw StateError("the context holds a batch of graphs");
    double out = 0.0;
    log_posterior_segments(c, &out);
    return out;
}

__global__ void logf_table_kernel(uint32_t lo, uint32_t n, float* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = host_logf(__uint_as_float(lo + i));
}

void launch_logf_table(Ctx& c, uint32_t lo, uint32_t n, float* out) {
    logf_table_kernel<<<(n + TB - 1) / TB, TB, 0, c.stream>>>(lo, n, out);
    SBPX_HIP(hipGetLastError());
}

void check_device_error(Ctx& c) {
    int e = 0;
    SBPX_HIP(hipMemcpyAsync(&e, c.dev_err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (e) {
        SBPX_HIP(hipMemsetAsync(c.dev_err, 0, sizeof(int), c.stream));
        if (e & DEV_ERR_BLOCK_RANGE) throw ArgError("block id out of range [0, B)");
        if (e & DEV_ERR_ZERO_DEGREE) throw NumericError("nonzero block-matrix cell with a zero block degree");
        if (e & DEV_ERR_NAN) throw NumericError("NaN or Inf in a delta MDL");
        if (e & DEV_ERR_TABLE) throw StateError("neighbour-block table overflow");
        if (e & DEV_ERR_NO_PROPOSAL) throw StateError("block without a merge proposal");
        throw StateError("device error");
    }
}

}  // namespace sbpx
