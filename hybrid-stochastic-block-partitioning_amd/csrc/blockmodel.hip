// Device Blockmodel: the block matrix M (M[b(u), b(v)] = number of edges u -> v) rebuilt from
// the assignment, as CSR and CSC with per-row / per-column inclusive prefix sums, plus
// dout = row sums, din = column sums, diag and the total degree deg = dout + din - diag - drift.
// Replaces Blockmodel::initialize_edge_counts (blockmodel.cpp:187-255) and the robin-map
// DictTransposeMatrix. The log posterior (blockmodel.cpp:267-285) is an FP64 reduction over
// the nonzeros with the FP32 log the reference uses.
#include <algorithm>

#include <limits>

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
                                 const int* __restrict__ assign, Dyn E, int B, int shift, K* __restrict__ keys,
                                 int* __restrict__ err, const int* __restrict__ dense) {
    GRID_LOOP(i, E.get()) {
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
}

template <typename K>
__global__ void head_flags_kernel(const K* __restrict__ keys, Dyn n, int* __restrict__ flags) {
    GRID_LOOP(i, n.get()) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Unique keys and the start index of each run; start[nnz] = n.
template <typename K>
__global__ void compact_runs_kernel(const K* __restrict__ keys, const int* __restrict__ flags,
                                    const int* __restrict__ pos, Dyn cnt, K* __restrict__ ukeys,
                                    int* __restrict__ start) {
    const int n = cnt.get();
    GRID_LOOP(i, n) {
        if (flags[i]) {
            int j = pos[i];
            ukeys[j] = keys[i];
            start[j] = i;
        }
        if (i == n - 1) start[pos[n]] = n;
    }
}

template <typename K>
__global__ void decode_csr_kernel(const K* __restrict__ ukeys, const int* __restrict__ start, Dyn nnz,
                                  int shift, K mask, int* __restrict__ coo_row, int* __restrict__ row_col,
                                  int* __restrict__ row_val, const int* __restrict__ inv) {
    GRID_LOOP(j, nnz.get()) {
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
}

// ptr[x] = first index j with sorted_idx[j] >= x, for x in [0, B].
__global__ void segment_ptr_kernel(const int* __restrict__ sorted_idx, Dyn nnz, int B, int* __restrict__ ptr) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x > B) return;
    ptr[x] = lower_bound_lane(sorted_idx, 0, nnz.get(), x);
}

// cum[j] = inclusive prefix of val inside the segment of j; tot[x] = segment sum.
// ex is the exclusive scan of val over all nonzeros (nnz + 1 entries).
__global__ void segment_cum_kernel(const int* __restrict__ seg_of, const int* __restrict__ ptr,
                                   const int* __restrict__ ex, Dyn nnz, int* __restrict__ cum) {
    GRID_LOOP(j, nnz.get()) cum[j] = ex[j + 1] - ex[ptr[seg_of[j]]];
}
__global__ void segment_total_kernel(const int* __restrict__ ptr, const int* __restrict__ ex, int B,
                                     int* __restrict__ tot) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= B) return;
    tot[x] = ex[ptr[x + 1]] - ex[ptr[x]];
}

__global__ void diag_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col,
                            const int* __restrict__ row_val, Dyn nnz, int* __restrict__ diag) {
    GRID_LOOP(j, nnz.get()) if (coo_row[j] == row_col[j]) diag[coo_row[j]] = row_val[j];
}

template <typename K>
__global__ void csc_keys_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col, Dyn nnz,
                                int shift, K* __restrict__ keys, const int* __restrict__ dense) {
    GRID_LOOP(j, nnz.get()) {
        int row = coo_row[j], col = row_col[j];
        if (dense != nullptr) {
            row = dense[row];
            col = dense[col];
        }
        keys[j] = ((K)col << shift) | (K)row;
    }
}

template <typename K>
__global__ void decode_csc_kernel(const K* __restrict__ keys, const int* __restrict__ vals, Dyn nnz, int shift,
                                  K mask, int* __restrict__ col_of, int* __restrict__ col_row,
                                  int* __restrict__ col_val, const int* __restrict__ inv) {
    GRID_LOOP(j, nnz.get()) {
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
}

// Current entropy of every nonzero, the "before" term of each cell an evaluation reads: CSR
// entry j is cell (coo_row[j], row_col[j]); CSC entry j is cell (col_row[j], col_of[j]).
__global__ void cell_entropy_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col,
                                    const int* __restrict__ row_val, const int* __restrict__ col_of,
                                    const int* __restrict__ col_row, const int* __restrict__ col_val,
                                    const int* __restrict__ dout, const int* __restrict__ din, Dyn nnz,
                                    float* __restrict__ row_h, float* __restrict__ col_h) {
    GRID_LOOP(j, nnz.get()) {
        row_h[j] = cell_entropy((float)row_val[j], (float)din[row_col[j]], (float)dout[coo_row[j]]);
        col_h[j] = cell_entropy((float)col_val[j], (float)din[col_of[j]], (float)dout[col_row[j]]);
    }
}

__global__ void degrees_kernel(const int* __restrict__ dout, const int* __restrict__ din,
                               const int* __restrict__ diag, const int* __restrict__ drift, int B,
                               int* __restrict__ deg) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= B) return;
    deg[x] = dout[x] + din[x] - diag[x] - drift[x];
}

// Sort-key layout of one build: key = row << bits | column, over block ids or, for a batch of
// graphs, their dense numbering among the blocks in use.
template <typename K>
struct KeySpace {
    int bits, shift;
    K mask;
    const int* dense;
    const int* inv;
};

// The largest key of 2 * bits bits (all ones): padding that sorts after every real key. Shifting
// the maximum down avoids a shift by the key's full width when 2 * bits equals it.
template <typename K>
static inline K pad_key(int bits) {
    return std::numeric_limits<K>::max() >> (8 * (int)sizeof(K) - 2 * bits);
}

template <typename K>
static KeySpace<K> key_space(const Ctx& c) {
    const bool dn = c.batch.G > 0;  // a batch: keys over the blocks in use
    KeySpace<K> ks;
    ks.bits = bits_for(dn ? c.batch.used : c.bm.B);
    ks.shift = ks.bits;
    ks.mask = (K)((((K)1) << ks.bits) - 1);
    ks.dense = dn ? c.batch.d_dense : nullptr;
    ks.inv = dn ? c.batch.d_inv : nullptr;
    return ks;
}

// Small kernels of the graph-mode update: decisions and sizes in device words (DynWord).
__global__ void set_word_kernel(int* __restrict__ w, int value) { *w = value; }
__global__ void fill_int_kernel(int* __restrict__ p, Dyn n, int value) { GRID_LOOP(i, n.get()) p[i] = value; }
__global__ void copy_int_kernel(const int* __restrict__ src, int* __restrict__ dst, Dyn n) {
    GRID_LOOP(i, n.get()) dst[i] = src[i];
}
// After the change count is known: an incremental update when the changes are at most a quarter
// of the edges' worth (as update_blockmodel decides on the host), else a full rebuild.
__global__ void choose_update_kernel(int* __restrict__ w, const int* __restrict__ list_n, int E, int inc_max) {
    const int total = w[W_TOTAL];
    const bool inc = total <= inc_max;
    w[W_INC_TOTAL] = inc ? total : 0;
    w[W_EMIT_N] = inc ? *list_n : 0;
    w[W_FULL_E] = inc ? 0 : E;
    w[W_NOLD] = inc ? w[W_NNZ] : 0;
}
__global__ void add_words_kernel(int* __restrict__ w, int a, int b, int out) { w[out] = w[a] + w[b]; }
__global__ void pick_nnz_kernel(int* __restrict__ w) {
    w[W_NNZ] = w[W_FULL_E] > 0 ? w[W_NNZ_FULL] : w[W_NNZ_INC];
}

// Steps 3-4 of a build from the cells (coo_row, row_col, row_val; nnz of them, sorted by row
// then column): row pointers, prefix sums and totals, the diagonal, the CSC orientation, cell
// entropies and degrees. nnz is host-known, or device-held (graph mode: fixed grids, no memset
// or copy calls, the count stays in the device word it came from).
template <typename K>
static void finish_build(Ctx& c, const KeySpace<K>& ks, Dyn nnz) {
    DevBlockmodel& bm = c.bm;
    hipStream_t st = c.stream;
    const int B = bm.B;
    K* keys0 = (K*)c.keys0;
    K* keys1 = (K*)c.keys1;
    const int nbytes = (ks.bits + 7) / 8;
    const bool dyn = nnz.p != nullptr;
    if (!dyn) {
        bm.nnz = nnz.v;
        if (c.batch.G > 0) set_word_kernel<<<1, 1, 0, st>>>(c.words + W_NNZ, nnz.v);  // for graph-mode rounds
    }
    const bool cells = dyn || nnz.v > 0;
    const int gn = grid_for(nnz, TB);
    // 3. row pointers, row prefix sums, dout, diagonal
    segment_ptr_kernel<<<blocks_for(B + 1), TB, 0, st>>>(bm.coo_row, nnz, B, bm.row_ptr);
    scan_exclusive(bm.row_val, c.pos, nnz, c.scan_tmp, st);
    if (cells) segment_cum_kernel<<<gn, TB, 0, st>>>(bm.coo_row, bm.row_ptr, c.pos, nnz, bm.row_cum);
    segment_total_kernel<<<blocks_for(B), TB, 0, st>>>(bm.row_ptr, c.pos, B, bm.dout);
    if (dyn) fill_int_kernel<<<grid_for(Dyn(B), TB), TB, 0, st>>>(bm.diag, Dyn(B), 0);
    else SBPX_HIP(hipMemsetAsync(bm.diag, 0, sizeof(int) * B, st));
    if (cells) diag_kernel<<<gn, TB, 0, st>>>(bm.coo_row, bm.row_col, bm.row_val, nnz, bm.diag);
    // 4. CSC: stable sort of the unique cells by column (input already ordered by row)
    if (cells) {
        csc_keys_kernel<K><<<gn, TB, 0, st>>>(bm.coo_row, bm.row_col, nnz, ks.shift, keys0, ks.dense);
        if (dyn) copy_int_kernel<<<gn, TB, 0, st>>>(bm.row_val, c.vals0, nnz);
        else SBPX_HIP(hipMemcpyAsync(c.vals0, bm.row_val, sizeof(int) * nnz.v, hipMemcpyDeviceToDevice, st));
        int cshifts[4];
        int cnp = 0;
        for (int i = 0; i < nbytes; ++i) cshifts[cnp++] = ks.shift + 8 * i;
        int which = radix_sort<K>(keys0, keys1, c.vals0, c.vals1, nnz, cshifts, cnp, c.hist, c.scan_tmp, st);
        K* ck = which ? keys1 : keys0;
        int* cv = which ? c.vals1 : c.vals0;
        int* col_of = c.flags;  // column index per CSC entry (scratch)
        decode_csc_kernel<K><<<gn, TB, 0, st>>>(ck, cv, nnz, ks.shift, ks.mask, col_of, bm.col_row, bm.col_val,
                                                ks.inv);
        segment_ptr_kernel<<<blocks_for(B + 1), TB, 0, st>>>(col_of, nnz, B, bm.col_ptr);
        scan_exclusive(bm.col_val, c.pos, nnz, c.scan_tmp, st);
        segment_cum_kernel<<<gn, TB, 0, st>>>(col_of, bm.col_ptr, c.pos, nnz, bm.col_cum);
        segment_total_kernel<<<blocks_for(B), TB, 0, st>>>(bm.col_ptr, c.pos, B, bm.din);
        cell_entropy_kernel<<<gn, TB, 0, st>>>(bm.coo_row, bm.row_col, bm.row_val, col_of, bm.col_row, bm.col_val,
                                               bm.dout, bm.din, nnz, bm.row_h, bm.col_h);
    } else {
        SBPX_HIP(hipMemsetAsync(bm.col_ptr, 0, sizeof(int) * (B + 1), st));
        SBPX_HIP(hipMemsetAsync(bm.din, 0, sizeof(int) * B, st));
    }
    degrees_kernel<<<blocks_for(B), TB, 0, st>>>(bm.dout, bm.din, bm.diag, bm.drift, B, bm.deg);
    bm.valid = true;
}

// Steps 1-2 of a full build from the assignment: edge keys (block(u), block(v)) in CSR edge order,
// sorted by (row, col) in 8-bit digits, run-length encoded into cells with counts. E is the edge
// count, or in graph mode a device word that is E when the full build was chosen and 0 otherwise;
// then the cell count goes to the device word nnz_at.
template <typename K>
static void build_cells(Ctx& c, const KeySpace<K>& ks, Dyn E, int* nnz_at, int* nnz_host) {
    DevGraph& g = c.g;
    DevBlockmodel& bm = c.bm;
    hipStream_t st = c.stream;
    K* keys0 = (K*)c.keys0;
    K* keys1 = (K*)c.keys1;
    int shifts[8];
    int np = 0;
    for (int b = 0; b < 2 * ks.bits; b += 8) shifts[np++] = b;
    const int gE = grid_for(E, TB);
    edge_keys_kernel<K><<<gE, TB, 0, st>>>(g.edge_src, g.out_idx, bm.assign, E, bm.B, ks.shift, keys0, c.dev_err,
                                           ks.dense);
    int which = radix_sort<K>(keys0, keys1, nullptr, nullptr, E, shifts, np, c.hist, c.scan_tmp, st);
    K* sorted = which ? keys1 : keys0;
    K* ukeys = which ? keys0 : keys1;
    head_flags_kernel<K><<<gE, TB, 0, st>>>(sorted, E, c.flags);
    if (E.p) {
        scan_exclusive(c.flags, c.pos, E, c.scan_tmp, st, nnz_at);
        compact_runs_kernel<K><<<gE, TB, 0, st>>>(sorted, c.flags, c.pos, E, ukeys, c.vals0);
        const Dyn nnz(E.v, nnz_at);
        decode_csr_kernel<K><<<grid_for(nnz, TB), TB, 0, st>>>(ukeys, c.vals0, nnz, ks.shift, ks.mask, bm.coo_row,
                                                               bm.row_col, bm.row_val, ks.inv);
        return;
    }
    int nnz = 0;
    scan_exclusive(c.flags, c.pos, E.v, c.scan_tmp, st);
    SBPX_HIP(hipMemcpyAsync(&nnz, c.pos + E.v, sizeof(int), hipMemcpyDeviceToHost, st));
    compact_runs_kernel<K><<<gE, TB, 0, st>>>(sorted, c.flags, c.pos, E, ukeys, c.vals0);
    SBPX_HIP(hipStreamSynchronize(st));
    if (nnz > 0)
        decode_csr_kernel<K><<<blocks_for(nnz), TB, 0, st>>>(ukeys, c.vals0, nnz, ks.shift, ks.mask, bm.coo_row,
                                                             bm.row_col, bm.row_val, ks.inv);
    *nnz_host = nnz;
}

template <typename K>
static void build_impl(Ctx& c) {
    const KeySpace<K> ks = key_space<K>(c);
    int nnz = 0;
    if (c.g.E > 0) build_cells<K>(c, ks, Dyn(c.g.E), nullptr, &nnz);
    finish_build<K>(c, ks, Dyn(nnz));
}

// ------------------------------------------------------------------------------------------
// Incremental update after a batch of moves. Only the edges of moved vertices change cells:
// each contributes (old key, -1) and (new key, +1), an edge between two movers once (as its
// source's out-edge). The changes are sorted and summed per key, then merged into the previous
// cells (sorted keys with counts) by ranks: a change to an existing key adds to its count, a new
// key is placed among the old ones, and cells whose count drops to zero are removed. The cells
// then pass through the same finishing steps as a full build, so the result is the same arrays.

template <typename K>
__device__ __forceinline__ K cell_key(int row, int col, int shift, const int* dense) {
    if (dense != nullptr) {
        row = dense[row];
        col = dense[col];
    }
    return ((K)row << shift) | (K)col;
}

// Entries per listed vertex: 2 per incident edge when it moves in this batch, else 0.
__global__ void delta_count_kernel(VList L, const uint8_t* __restrict__ stat, const int* __restrict__ out_ptr,
                                   const int* __restrict__ in_ptr, int* __restrict__ cnt) {
    const int* list = L.ptr();
    GRID_LOOP(i, L.n.get()) {
        const int v = list[i];
        cnt[i] = stat[v] == 2 ? 2 * ((out_ptr[v + 1] - out_ptr[v]) + (in_ptr[v + 1] - in_ptr[v])) : 0;
    }
}

// One wavefront per listed vertex; reads the assignment before the batch is committed. Unused
// slots (in-edges from movers and the self-loop's in-list entry) get value 0 and the largest key
// of the sorted bit range, `pad` (the key of cell (2^bits - 1, 2^bits - 1)): the sort orders only
// those bits, so a key outside them would interleave with the cell whose low bits it shares.
template <typename K>
__global__ void delta_emit_kernel(VList L, const uint8_t* __restrict__ stat,
                                  const int* __restrict__ off, const int* __restrict__ out_ptr,
                                  const int* __restrict__ out_idx, const int* __restrict__ in_ptr,
                                  const int* __restrict__ in_idx, const int* __restrict__ assign,
                                  const int* __restrict__ new_block, const int* __restrict__ moved_epoch, Dyn ep,
                                  int shift, const int* __restrict__ dense, K pad, K* __restrict__ dkey,
                                  int* __restrict__ dval) {
    const int* list = L.ptr();
    const int n = L.n.get(), epoch = ep.get();
    const int wpb = blockDim.x / WAVE;
    for (int i = blockIdx.x * wpb + threadIdx.x / WAVE; i < n; i += gridDim.x * wpb) {
    const int v = list[i];
    if (stat[v] != 2) continue;
    const int lane = lane_id();
    const int r = assign[v], s = new_block[v];
    const int base = off[i];
    const int ob = out_ptr[v], dO = out_ptr[v + 1] - ob;
    const int ib = in_ptr[v], dI = in_ptr[v + 1] - ib;
    for (int j = lane; j < dO; j += WAVE) {
        const int w = out_idx[ob + j];
        const int bo = assign[w];
        const int bn = moved_epoch[w] == epoch ? new_block[w] : bo;
        dkey[base + 2 * j] = cell_key<K>(r, bo, shift, dense);
        dval[base + 2 * j] = -1;
        dkey[base + 2 * j + 1] = cell_key<K>(s, bn, shift, dense);
        dval[base + 2 * j + 1] = 1;
    }
    for (int j = lane; j < dI; j += WAVE) {
        const int u = in_idx[ib + j];
        const int q = base + 2 * dO + 2 * j;
        if (u == v || moved_epoch[u] == epoch) {
            dkey[q] = pad;
            dval[q] = 0;
            dkey[q + 1] = pad;
            dval[q + 1] = 0;
        } else {
            const int bu = assign[u];
            dkey[q] = cell_key<K>(bu, r, shift, dense);
            dval[q] = -1;
            dkey[q + 1] = cell_key<K>(bu, s, shift, dense);
            dval[q + 1] = 1;
        }
    }
    }
}

// Sum of the values of each run of equal sorted keys (run j: [start[j], start[j + 1])), kept
// when nonzero. Padding entries carry 0, so a run of padding alone is dropped, and one that
// shares its key with real changes (a full key space) sums those alone.
__global__ void run_sums_kernel(const int* __restrict__ start, const int* __restrict__ vex, Dyn nruns,
                                int* __restrict__ sum, int* __restrict__ keep) {
    GRID_LOOP(j, nruns.get()) {
        const int t = vex[start[j + 1]] - vex[start[j]];
        sum[j] = t;
        keep[j] = t != 0 ? 1 : 0;
    }
}

template <typename K>
__global__ void compact_pairs_kernel(const K* __restrict__ k, const int* __restrict__ v, const int* __restrict__ keep,
                                     const int* __restrict__ pos, Dyn n, K* __restrict__ ko, int* __restrict__ vo) {
    GRID_LOOP(j, n.get()) {
        if (!keep[j]) continue;
        ko[pos[j]] = k[j];
        vo[pos[j]] = v[j];
    }
}

template <typename K>
__device__ __forceinline__ int lower_bound_key(const K* __restrict__ a, int n, K x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename K>
__global__ void old_keys_kernel(const int* __restrict__ coo_row, const int* __restrict__ row_col,
                                const int* __restrict__ row_val, Dyn nnz, int shift, const int* __restrict__ dense,
                                K* __restrict__ okey, int* __restrict__ ocount) {
    GRID_LOOP(i, nnz.get()) {
        okey[i] = cell_key<K>(coo_row[i], row_col[i], shift, dense);
        ocount[i] = row_val[i];
    }
}

// A change to an existing cell adds to its count; others are flagged new, with their rank
// among the old keys.
template <typename K>
__global__ void merge_match_kernel(const K* __restrict__ dkey, const int* __restrict__ dsum, Dyn nd_,
                                   const K* __restrict__ okey, Dyn nold_, int* __restrict__ ocount,
                                   int* __restrict__ isnew, int* __restrict__ lb) {
    const int nold = nold_.get();
    GRID_LOOP(j, nd_.get()) {
        const int p = lower_bound_key<K>(okey, nold, dkey[j]);
        lb[j] = p;
        if (p < nold && okey[p] == dkey[j]) {
            ocount[p] += dsum[j];  // keys are unique on both sides
            isnew[j] = 0;
        } else {
            isnew[j] = 1;
        }
    }
}

template <typename K>
__global__ void merge_place_old_kernel(const K* __restrict__ okey, const int* __restrict__ ocount, Dyn nold,
                                       const K* __restrict__ dkey, Dyn nd_, const int* __restrict__ rank,
                                       K* __restrict__ mkey, int* __restrict__ mcount) {
    const int nd = nd_.get();
    GRID_LOOP(i, nold.get()) {
        const int p = i + rank[lower_bound_key<K>(dkey, nd, okey[i])];
        mkey[p] = okey[i];
        mcount[p] = ocount[i];
    }
}

template <typename K>
__global__ void merge_place_new_kernel(const K* __restrict__ dkey, const int* __restrict__ dsum, Dyn nd,
                                       const int* __restrict__ isnew, const int* __restrict__ lb,
                                       const int* __restrict__ rank, K* __restrict__ mkey, int* __restrict__ mcount,
                                       int* __restrict__ err) {
    GRID_LOOP(j, nd.get()) {
        if (!isnew[j]) continue;
        const int p = rank[j] + lb[j];
        mkey[p] = dkey[j];
        mcount[p] = dsum[j];
        if (dsum[j] < 0) atomicOr(err, DEV_ERR_BLOCK_RANGE);  // removes from a cell that is not there
    }
}

__global__ void nonzero_flags_kernel(const int* __restrict__ count, Dyn n, int* __restrict__ flag) {
    GRID_LOOP(i, n.get()) flag[i] = count[i] != 0 ? 1 : 0;
}

template <typename K>
__global__ void decode_cells_kernel(const K* __restrict__ keys, const int* __restrict__ count, Dyn nnz, int shift,
                                    K mask, int* __restrict__ coo_row, int* __restrict__ row_col,
                                    int* __restrict__ row_val, const int* __restrict__ inv) {
    GRID_LOOP(j, nnz.get()) {
    K key = keys[j];
    int row = (int)(key >> shift), col = (int)(key & mask);
    if (inv != nullptr) {
        row = inv[row];
        col = inv[col];
    }
    coo_row[j] = row;
    row_col[j] = col;
    row_val[j] = count[j];
    }
}

static void ensure_inc(Ctx& c) {
    const size_t E = (size_t)c.g.E + 2;
    const size_t need = std::max(E, (size_t)c.g.N + 2);
    if (c.inc.cap >= need) return;
    IncBuffers& b = c.inc;
    for (void** p : {&b.okey, &b.mkey, &b.dkey}) {
        if (*p) SBPX_HIP(hipFree(*p));
        *p = nullptr;
    }
    for (int** p : {&b.ocount, &b.mcount, &b.dsum, &b.lb, &b.rank, &b.flag, &b.tmp}) {
        if (*p) SBPX_HIP(hipFree(*p));
        *p = nullptr;
    }
    // merged list: old cells (<= E) plus new keys (<= a quarter of E, see update_blockmodel)
    SBPX_HIP(hipMalloc(&b.okey, sizeof(uint64_t) * need));
    SBPX_HIP(hipMalloc(&b.mkey, sizeof(uint64_t) * 2 * need));
    SBPX_HIP(hipMalloc(&b.dkey, sizeof(uint64_t) * need));
    SBPX_HIP(hipMalloc((void**)&b.ocount, sizeof(int) * need));
    SBPX_HIP(hipMalloc((void**)&b.mcount, sizeof(int) * 2 * need));
    SBPX_HIP(hipMalloc((void**)&b.dsum, sizeof(int) * need));
    SBPX_HIP(hipMalloc((void**)&b.lb, sizeof(int) * need));
    SBPX_HIP(hipMalloc((void**)&b.rank, sizeof(int) * (2 * need + 1)));
    SBPX_HIP(hipMalloc((void**)&b.flag, sizeof(int) * 2 * need));
    b.tmp_ints = scan_tmp_ints(2 * (long)need + 1);
    SBPX_HIP(hipMalloc((void**)&b.tmp, sizeof(int) * b.tmp_ints));
    b.cap = need;
}

// Steps 1-2 of an incremental update (the changes are in keys0 / vals0): the changes sorted and
// summed per key, merged into the previous cells by ranks, zero cells dropped. Host-known sizes
// (total changes, the previous cell count in bm.nnz), with the cell count returned; or in graph
// mode device words (W_INC_TOTAL, W_NOLD; 0 when the full build was chosen), the sizes of every
// step and the cell count (W_NNZ_INC) staying on the device.
template <typename K>
static int update_cells(Ctx& c, const KeySpace<K>& ks, Dyn total, bool dyn) {
    DevBlockmodel& bm = c.bm;
    hipStream_t st = c.stream;
    IncBuffers& ib = c.inc;
    int* w = c.words;
    const int E = c.g.E;
    K* keys0 = (K*)c.keys0;
    K* keys1 = (K*)c.keys1;
    // sort the changes by key (values follow), all 2 * bits of the key
    int shifts[8];
    int np = 0;
    for (int b = 0; b < 2 * ks.bits; b += 8) shifts[np++] = b;
    const int which = radix_sort<K>(keys0, keys1, c.vals0, c.vals1, total, shifts, np, c.hist, c.scan_tmp, st);
    K* sk = which ? keys1 : keys0;
    int* sv = which ? c.vals1 : c.vals0;
    K* uk = which ? keys0 : keys1;  // run keys
    // sum per key: runs of the sorted keys, value prefix sums
    const int gt = grid_for(total, TB);
    head_flags_kernel<K><<<gt, TB, 0, st>>>(sk, total, c.flags);
    int* start = ib.lb;  // run starts (nruns + 1), scratch until the merge
    K* dkey = (K*)ib.dkey;
    K* okey = (K*)ib.okey;
    K* mkey = (K*)ib.mkey;
    Dyn nruns, nd, nold, nnew_nm;
    if (dyn) {
        scan_exclusive(c.flags, c.pos, total, c.scan_tmp, st, w + W_NRUNS);
        nruns = Dyn(total.v, w + W_NRUNS);
        compact_runs_kernel<K><<<gt, TB, 0, st>>>(sk, c.flags, c.pos, total, uk, start);
        scan_exclusive(sv, ib.rank, total, ib.tmp, st);  // value prefix (rank as scratch)
        run_sums_kernel<<<grid_for(nruns, TB), TB, 0, st>>>(start, ib.rank, nruns, ib.mcount, ib.flag);
        scan_exclusive(ib.flag, c.pos, nruns, c.scan_tmp, st, w + W_ND);
        compact_pairs_kernel<K><<<grid_for(nruns, TB), TB, 0, st>>>(uk, ib.mcount, ib.flag, c.pos, nruns, dkey, ib.dsum);
        nd = Dyn(total.v, w + W_ND);
        nold = Dyn(E, w + W_NOLD);
        old_keys_kernel<K><<<grid_for(nold, TB), TB, 0, st>>>(bm.coo_row, bm.row_col, bm.row_val, nold, ks.shift,
                                                              ks.dense, okey, ib.ocount);
        merge_match_kernel<K><<<grid_for(nd, TB), TB, 0, st>>>(dkey, ib.dsum, nd, okey, nold, ib.ocount, ib.flag,
                                                               ib.lb);
        scan_exclusive(ib.flag, ib.rank, nd, ib.tmp, st, w + W_NNEW);
        merge_place_old_kernel<K><<<grid_for(nold, TB), TB, 0, st>>>(okey, ib.ocount, nold, dkey, nd, ib.rank, mkey,
                                                                     ib.mcount);
        merge_place_new_kernel<K><<<grid_for(nd, TB), TB, 0, st>>>(dkey, ib.dsum, nd, ib.flag, ib.lb, ib.rank, mkey,
                                                                   ib.mcount, c.dev_err);
        add_words_kernel<<<1, 1, 0, st>>>(w, W_NOLD, W_NNEW, W_NM);
        const Dyn nm(E + total.v, w + W_NM);
        nonzero_flags_kernel<<<grid_for(nm, TB), TB, 0, st>>>(ib.mcount, nm, ib.flag);
        scan_exclusive(ib.flag, ib.rank, nm, ib.tmp, st, w + W_NNZ_INC);
        compact_pairs_kernel<K><<<grid_for(nm, TB), TB, 0, st>>>(mkey, ib.mcount, ib.flag, ib.rank, nm, okey,
                                                                 ib.ocount);
        const Dyn nnz(E, w + W_NNZ_INC);
        decode_cells_kernel<K><<<grid_for(nnz, TB), TB, 0, st>>>(okey, ib.ocount, nnz, ks.shift, ks.mask, bm.coo_row,
                                                                 bm.row_col, bm.row_val, ks.inv);
        return -1;
    }
    const int tot = total.v;
    scan_exclusive(c.flags, c.pos, tot, c.scan_tmp, st);
    int h_nruns = 0;
    SBPX_HIP(hipMemcpyAsync(&h_nruns, c.pos + tot, sizeof(int), hipMemcpyDeviceToHost, st));
    compact_runs_kernel<K><<<gt, TB, 0, st>>>(sk, c.flags, c.pos, total, uk, start);
    scan_exclusive(sv, ib.rank, tot, ib.tmp, st);  // value prefix (rank as scratch)
    SBPX_HIP(hipStreamSynchronize(st));
    run_sums_kernel<<<blocks_for(h_nruns), TB, 0, st>>>(start, ib.rank, h_nruns, ib.mcount, ib.flag);
    scan_exclusive(ib.flag, c.pos, h_nruns, c.scan_tmp, st);
    int h_nd = 0;
    SBPX_HIP(hipMemcpyAsync(&h_nd, c.pos + h_nruns, sizeof(int), hipMemcpyDeviceToHost, st));
    compact_pairs_kernel<K><<<blocks_for(h_nruns), TB, 0, st>>>(uk, ib.mcount, ib.flag, c.pos, h_nruns, dkey, ib.dsum);
    // the previous cells as sorted keys with counts
    const int h_nold = bm.nnz;
    if (h_nold > 0)
        old_keys_kernel<K><<<blocks_for(h_nold), TB, 0, st>>>(bm.coo_row, bm.row_col, bm.row_val, h_nold, ks.shift,
                                                              ks.dense, okey, ib.ocount);
    SBPX_HIP(hipStreamSynchronize(st));
    // merge by ranks
    int h_nnew = 0;
    if (h_nd > 0) {
        merge_match_kernel<K><<<blocks_for(h_nd), TB, 0, st>>>(dkey, ib.dsum, h_nd, okey, h_nold, ib.ocount, ib.flag,
                                                               ib.lb);
        scan_exclusive(ib.flag, ib.rank, h_nd, ib.tmp, st);
        SBPX_HIP(hipMemcpyAsync(&h_nnew, ib.rank + h_nd, sizeof(int), hipMemcpyDeviceToHost, st));
    } else {
        SBPX_HIP(hipMemsetAsync(ib.rank, 0, sizeof(int), st));
    }
    if (h_nold > 0)
        merge_place_old_kernel<K><<<blocks_for(h_nold), TB, 0, st>>>(okey, ib.ocount, h_nold, dkey, h_nd, ib.rank,
                                                                     mkey, ib.mcount);
    SBPX_HIP(hipStreamSynchronize(st));
    if (h_nnew > 0)
        merge_place_new_kernel<K><<<blocks_for(h_nd), TB, 0, st>>>(dkey, ib.dsum, h_nd, ib.flag, ib.lb, ib.rank, mkey,
                                                                   ib.mcount, c.dev_err);
    // drop the cells that reached zero
    const int nm = h_nold + h_nnew;
    int nnz = 0;
    if (nm > 0) {
        nonzero_flags_kernel<<<blocks_for(nm), TB, 0, st>>>(ib.mcount, nm, ib.flag);
        scan_exclusive(ib.flag, ib.rank, nm, ib.tmp, st);
        SBPX_HIP(hipMemcpyAsync(&nnz, ib.rank + nm, sizeof(int), hipMemcpyDeviceToHost, st));
        compact_pairs_kernel<K><<<blocks_for(nm), TB, 0, st>>>(mkey, ib.mcount, ib.flag, ib.rank, nm, okey, ib.ocount);
        SBPX_HIP(hipStreamSynchronize(st));
        decode_cells_kernel<K><<<blocks_for(nnz), TB, 0, st>>>(okey, ib.ocount, nnz, ks.shift, ks.mask, bm.coo_row,
                                                               bm.row_col, bm.row_val, ks.inv);
    }
    return nnz;
}

template <typename K>
static void update_impl(Ctx& c, int total) {
    const KeySpace<K> ks = key_space<K>(c);
    const int nnz = update_cells<K>(c, ks, Dyn(total), false);
    finish_build<K>(c, ks, Dyn(nnz));
}

void rebuild_keep_drift(Ctx& c) {
    const int bits = bits_for(c.batch.G > 0 ? c.batch.used : c.bm.B);
    if (2 * bits <= 32) build_impl<uint32_t>(c);
    else build_impl<uint64_t>(c);
}

// Called before the batch's moves are committed to the assignment (moved_epoch / new_block of
// this epoch mark them). Emits the changes of the moved vertices' edges when they are at most a
// quarter of the edges' worth (incremental update after the commit), else returns -1 (full
// rebuild). Returns the number of change entries.
template <typename K>
static int emit_delta(Ctx& c, const int* list, int n, const int* old_assign) {
    ensure_inc(c);
    hipStream_t st = c.stream;
    const KeySpace<K> ks = key_space<K>(c);
    int* cnt = c.inc.flag;
    const VList L(list, n);
    delta_count_kernel<<<blocks_for(n), TB, 0, st>>>(L, c.ev_acc, c.g.out_ptr, c.g.in_ptr, cnt);
    scan_exclusive(cnt, c.inc.rank, n, c.inc.tmp, st);
    int total = 0;
    SBPX_HIP(hipMemcpyAsync(&total, c.inc.rank + n, sizeof(int), hipMemcpyDeviceToHost, st));
    SBPX_HIP(hipStreamSynchronize(st));
    if (total > c.g.E / 2) return -1;  // more than a quarter of the edges changed
    if (total == 0) return 0;
    const int wpb = TB / WAVE;
    delta_emit_kernel<K><<<(n + wpb - 1) / wpb, TB, 0, st>>>(L, c.ev_acc, c.inc.rank, c.g.out_ptr, c.g.out_idx,
                                                            c.g.in_ptr, c.g.in_idx, old_assign, c.new_block,
                                                            c.moved_epoch, Dyn(c.epoch), ks.shift, ks.dense,
                                                            pad_key<K>(ks.bits), (K*)c.keys0, c.vals0);
    SBPX_HIP(hipGetLastError());
    return total;
}

int prepare_update(Ctx& c, const int* list, int n, const int* old_assign) {
    if (!c.bm.valid || c.g.E <= 0 || c.incremental == 0) return -1;
    if (old_assign == nullptr) old_assign = c.bm.assign;
    const int bits = bits_for(c.batch.G > 0 ? c.batch.used : c.bm.B);
    return 2 * bits <= 32 ? emit_delta<uint32_t>(c, list, n, old_assign) : emit_delta<uint64_t>(c, list, n, old_assign);
}

void finish_update(Ctx& c, const int* list, int n, int total) {
    if (total == 0) {  // nothing moved: the cells stand; degrees take the new drift
        degrees_kernel<<<blocks_for(c.bm.B), TB, 0, c.stream>>>(c.bm.dout, c.bm.din, c.bm.diag, c.bm.drift, c.bm.B,
                                                                c.bm.deg);
        SBPX_HIP(hipGetLastError());
        return;
    }
    const int bits = bits_for(c.batch.G > 0 ? c.batch.used : c.bm.B);
    if (2 * bits <= 32) update_impl<uint32_t>(c, total);
    else update_impl<uint64_t>(c, total);
}

// Graph mode: the same update with every size in device words, for launches captured once and
// replayed. Before the commit: the change count of the listed movers (W_TOTAL), the choice between
// an incremental update and a full rebuild, and the changes when incremental.
template <typename K>
static void emit_delta_dyn(Ctx& c, const VList& L, Dyn epoch) {
    hipStream_t st = c.stream;
    const KeySpace<K> ks = key_space<K>(c);
    int* cnt = c.inc.flag;
    delta_count_kernel<<<grid_for(L.n, TB), TB, 0, st>>>(L, c.ev_acc, c.g.out_ptr, c.g.in_ptr, cnt);
    scan_exclusive(cnt, c.inc.rank, L.n, c.inc.tmp, st, c.words + W_TOTAL);
    choose_update_kernel<<<1, 1, 0, st>>>(c.words, L.n.p, c.g.E, c.incremental ? c.g.E / 2 : -1);
    const VList LE(L.base, L.off, Dyn(L.n.v, c.words + W_EMIT_N));
    delta_emit_kernel<K><<<grid_for(LE.n, TB / WAVE), TB, 0, st>>>(LE, c.ev_acc, c.inc.rank, c.g.out_ptr, c.g.out_idx,
                                                                  c.g.in_ptr, c.g.in_idx, c.bm.assign, c.new_block,
                                                                  c.moved_epoch, epoch, ks.shift, ks.dense,
                                                                  pad_key<K>(ks.bits), (K*)c.keys0, c.vals0);
}

// After the commit: the incremental update (sizes 0 when the full rebuild was chosen), the full
// rebuild (edge count 0 when not), the chosen cell count, and the finishing steps.
template <typename K>
static void finish_update_dyn(Ctx& c) {
    const KeySpace<K> ks = key_space<K>(c);
    const int E = c.g.E;
    update_cells<K>(c, ks, Dyn(E / 2, c.words + W_INC_TOTAL), true);
    build_cells<K>(c, ks, Dyn(E, c.words + W_FULL_E), c.words + W_NNZ_FULL, nullptr);
    pick_nnz_kernel<<<1, 1, 0, c.stream>>>(c.words);
    finish_build<K>(c, ks, Dyn(E, c.words + W_NNZ));
}

void prepare_update_dyn(Ctx& c, const VList& L, Dyn epoch) {
    const int bits = bits_for(c.batch.G > 0 ? c.batch.used : c.bm.B);
    if (2 * bits <= 32) emit_delta_dyn<uint32_t>(c, L, epoch);
    else emit_delta_dyn<uint64_t>(c, L, epoch);
}

void finish_update_dyn(Ctx& c) {
    const int bits = bits_for(c.batch.G > 0 ? c.batch.used : c.bm.B);
    if (2 * bits <= 32) finish_update_dyn<uint32_t>(c);
    else finish_update_dyn<uint64_t>(c);
}

void ensure_update_buffers(Ctx& c) { ensure_inc(c); }

int sort_key_bits(const Ctx& c) { return bits_for(c.batch.G > 0 ? c.batch.used : c.bm.B); }

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

// ------------------------------------------------------------------------------------------
// The log posterior without allocations (sbpx_log_posterior_noalloc): the kernels above, with the
// segment words, the error word and the chunk partial sums in a fixed device array. Each device
// holds its own copy of the array until the process ends; every call on a device shares it.
constexpr long long NOALLOC_MAX_CELLS = 1LL << 28;
constexpr size_t NOALLOC_PARTS = (size_t)(NOALLOC_MAX_CELLS / LP_CHUNK) + 8;
constexpr size_t NOALLOC_HEAD = 256;  // bytes before the partial sums: segment words, error word
static __device__ __align__(256) unsigned char scratch[NOALLOC_HEAD + sizeof(double) * NOALLOC_PARTS];

// The current device's address of the array (host code must not use the symbol itself).
void* noalloc_scratch_address() {
    void* p = nullptr;
    SBPX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    return p;
}

__global__ void zero_word_kernel(int* w) { *w = 0; }

void log_posterior_noalloc(void* base, long long nnz, const int* coo_row, const int* col, const int* val,
                           const int* dout, const int* din, double* out, hipStream_t st) {
    if (nnz < 0 || nnz > NOALLOC_MAX_CELLS) throw ArgError("cell count must be in [0, 2^28]");
    int* seg = (int*)base;  // lo, hi and the chunk offsets of one graph: 4 words
    int* err = seg + 8;
    double* part = (double*)((unsigned char*)base + NOALLOC_HEAD);
    const int chunks = (int)((nnz + LP_CHUNK - 1) / LP_CHUNK);
    zero_word_kernel<<<1, 1, 0, st>>>(err);
    lp_segments_kernel<<<1, TB, 0, st>>>(nullptr, 1, nullptr, nullptr, (int)nnz, seg);
    if (chunks > 0) lp_chunk_kernel<<<chunks, TB, 0, st>>>(coo_row, col, val, dout, din, seg, 1, part, err);
    lp_graph_sum_kernel<<<1, TB, 0, st>>>(part, seg, 1, out);
    SBPX_HIP(hipGetLastError());
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
        if (e & DEV_ERR_BARRIER) throw StateError("a workgroup of the Metropolis-Hastings pass missed a barrier");
        throw StateError("device error");
    }
}

}  // namespace sbpx
