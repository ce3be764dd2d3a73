// Prefix sums and a stable LSD radix sort, written for 64-wide wavefronts. They build the
// block matrix from the assignment: edge keys (block(u), block(v)) are sorted, run-length
// encoded into CSR, then the unique cells are sorted again by (column, row) for CSC.
#include "internal.hpp"

namespace sbpx {

// ------------------------------------------------------------------------------------------
// Exclusive scan of int32 with the total appended: out[0..n] with out[n] = sum(in).
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread; returns the prefix, *total gets the sum.
__device__ __forceinline__ int block_exclusive_scan(int x, int* lds_waves, int* total) {
    const int lane = lane_id(), wave = threadIdx.x / WAVE, nwaves = blockDim.x / WAVE;
    int incl = x;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        int y = __shfl_up(incl, off, WAVE);
        if (lane >= off) incl += y;
    }
    if (lane == WAVE - 1) lds_waves[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < nwaves; ++w) {
            int t = lds_waves[w];
            lds_waves[w] = acc;
            acc += t;
        }
        lds_waves[nwaves] = acc;
    }
    __syncthreads();
    int res = lds_waves[wave] + incl - x;
    *total = lds_waves[nwaves];
    __syncthreads();
    return res;
}

__global__ void scan_reduce_kernel(const int* __restrict__ in, int n, int* __restrict__ partial) {
    __shared__ int lds[SCAN_THREADS / WAVE + 1];
    const long base = (long)blockIdx.x * SCAN_TILE;
    int sum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        long idx = base + (long)i * SCAN_THREADS + threadIdx.x;
        if (idx < n) sum += in[idx];
    }
    int total;
    block_exclusive_scan(sum, lds, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Each thread owns SCAN_ITEMS consecutive elements.
__global__ void scan_apply_kernel(const int* __restrict__ in, int n, int* __restrict__ out,
                                  const int* __restrict__ block_offset) {
    __shared__ int lds[SCAN_THREADS / WAVE + 1];
    const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int sum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        long idx = base + i;
        v[i] = idx < n ? in[idx] : 0;
        sum += v[i];
    }
    int total;
    int pre = block_exclusive_scan(sum, lds, &total);
    int acc = pre + (block_offset ? block_offset[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        long idx = base + i;
        if (idx < n) out[idx] = acc;
        acc += v[i];
    }
    if (base <= n - 1 && n - 1 < base + SCAN_ITEMS) out[n] = acc;  // owner of the last element
}

// tmp must hold scan_tmp_ints(n) ints. in and out may alias.
size_t scan_tmp_ints(long n) {
    size_t total = 0;
    long m = (n + SCAN_TILE - 1) / SCAN_TILE;
    while (m > 1) {
        total += 2 * (size_t)m + 1;
        m = (m + SCAN_TILE - 1) / SCAN_TILE;
    }
    return total + 8;
}

void scan_exclusive(const int* in, int* out, int n, int* tmp, hipStream_t st) {
    if (n <= 0) {
        (void)hipMemsetAsync(out, 0, sizeof(int), st);
        return;
    }
    int nblk = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nblk == 1) {
        scan_apply_kernel<<<1, SCAN_THREADS, 0, st>>>(in, n, out, nullptr);
        return;
    }
    int* partial = tmp;
    int* offsets = tmp + nblk;
    scan_reduce_kernel<<<nblk, SCAN_THREADS, 0, st>>>(in, n, partial);
    scan_exclusive(partial, offsets, nblk, tmp + 2 * nblk + 1, st);
    scan_apply_kernel<<<nblk, SCAN_THREADS, 0, st>>>(in, n, out, offsets);
}

// Scan over a count held in device memory: DYN_SCAN_BLOCKS (or fewer) workgroups, each owning a
// contiguous run of whole tiles; its sum, one scan of those sums, then each workgroup rescans its
// run tile by tile, carrying the offset.
__device__ __forceinline__ void dyn_scan_range(int n, int* first_tile, int* end_tile) {
    const int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    const int per = (tiles + (int)gridDim.x - 1) / (int)gridDim.x;
    const int a = (int)blockIdx.x * per, b = a + per;
    *first_tile = a < tiles ? a : tiles;
    *end_tile = b < tiles ? b : tiles;
}

__global__ void dyn_scan_reduce_kernel(const int* __restrict__ in, Dyn cnt, int* __restrict__ partial) {
    __shared__ int lds[SCAN_THREADS / WAVE + 1];
    const int n = cnt.get();
    int t0, t1;
    dyn_scan_range(n, &t0, &t1);
    int sum = 0;
    const long end = (long)t1 * SCAN_TILE < (long)n ? (long)t1 * SCAN_TILE : (long)n;
    for (long idx = (long)t0 * SCAN_TILE + threadIdx.x; idx < end; idx += SCAN_THREADS) sum += in[idx];
    int total;
    block_exclusive_scan(sum, lds, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ void dyn_scan_apply_kernel(const int* __restrict__ in, Dyn cnt, int* __restrict__ out,
                                      const int* __restrict__ block_offset, int nblocks, int* __restrict__ total_at) {
    __shared__ int lds[SCAN_THREADS / WAVE + 1];
    const int n = cnt.get();
    int t0, t1;
    dyn_scan_range(n, &t0, &t1);
    int carry = block_offset[blockIdx.x];
    for (int t = t0; t < t1; ++t) {
        const long base = (long)t * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
        int v[SCAN_ITEMS];
        int sum = 0;
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const long idx = base + i;
            v[i] = idx < n ? in[idx] : 0;
            sum += v[i];
        }
        int total;
        int acc = carry + block_exclusive_scan(sum, lds, &total);
#pragma unroll
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const long idx = base + i;
            if (idx < n) out[idx] = acc;
            acc += v[i];
        }
        carry += total;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // every workgroup has read in[0, n) before it writes; out[n] is past the input
        out[n] = block_offset[nblocks];
        if (total_at) *total_at = block_offset[nblocks];
    }
}

void scan_exclusive(const int* in, int* out, Dyn n, int* tmp, hipStream_t st, int* total_at) {
    if (!n.p) {
        scan_exclusive(in, out, n.v, tmp, st);
        if (total_at) (void)hipMemcpyAsync(total_at, out + std::max(n.v, 0), sizeof(int), hipMemcpyDeviceToDevice, st);
        return;
    }
    const int nblk = (int)std::max<long>(1, std::min<long>(DYN_SCAN_BLOCKS, ((long)n.v + SCAN_TILE - 1) / SCAN_TILE));
    int* partial = tmp;
    int* offsets = tmp + nblk;
    dyn_scan_reduce_kernel<<<nblk, SCAN_THREADS, 0, st>>>(in, n, partial);
    scan_apply_kernel<<<1, SCAN_THREADS, 0, st>>>(partial, nblk, offsets, nullptr);  // nblk <= one tile
    dyn_scan_apply_kernel<<<nblk, SCAN_THREADS, 0, st>>>(in, n, out, offsets, nblk, total_at);
}

// ------------------------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits. Tiles of 4096 keys; per pass: tile histograms
// (digit-major), one exclusive scan, then a stable scatter that ranks keys inside a tile round
// by round (256 keys per round) with wavefront ballots.
constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
constexpr int RS_WAVES = RS_THREADS / WAVE;

template <typename K>
__global__ void radix_hist_kernel(const K* __restrict__ keys, long n, int shift, int ntiles,
                                  int* __restrict__ hist) {
    __shared__ int cnt[RS_WAVES][256];
    const int wave = threadIdx.x / WAVE;
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ROUNDS; ++j) {
        long idx = base + (long)j * RS_THREADS + threadIdx.x;
        if (idx < n) {
            int d = (int)((keys[idx] >> shift) & 0xFF);
            atomicAdd(&cnt[wave][d], 1);
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += RS_THREADS) {
        int s = 0;
        for (int w = 0; w < RS_WAVES; ++w) s += cnt[w][d];
        hist[(long)d * ntiles + blockIdx.x] = s;
    }
}

template <typename K, bool HAS_VAL>
__global__ void radix_scatter_kernel(const K* __restrict__ keys, const int* __restrict__ vals, long n,
                                     int shift, int ntiles, const int* __restrict__ offsets,
                                     K* __restrict__ keys_out, int* __restrict__ vals_out) {
    __shared__ int tile_off[256];
    __shared__ int run[256];
    __shared__ int wc[RS_WAVES][256];
    const int lane = lane_id(), wave = threadIdx.x / WAVE;
    for (int d = threadIdx.x; d < 256; d += RS_THREADS) {
        tile_off[d] = offsets[(long)d * ntiles + blockIdx.x];
        run[d] = 0;
    }
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&wc[0][0])[i] = 0;
    __syncthreads();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (WAVE - lane));
    const long base = (long)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ROUNDS; ++j) {
        long idx = base + (long)j * RS_THREADS + threadIdx.x;
        bool valid = idx < n;
        K key = valid ? keys[idx] : (K)0;
        int val = 0;
        if (HAS_VAL && valid) val = vals[idx];
        int d = (int)((key >> shift) & 0xFF);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1;
            unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        int rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wc[wave][d] = __popcll(peers);
        __syncthreads();
        for (int dd = threadIdx.x; dd < 256; dd += RS_THREADS) {
            int acc = run[dd];
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) {
                int t = wc[w][dd];
                wc[w][dd] = acc;
                acc += t;
            }
            run[dd] = acc;
        }
        __syncthreads();
        if (valid) {
            long pos = (long)tile_off[d] + wc[wave][d] + rank;
            keys_out[pos] = key;
            if (HAS_VAL) vals_out[pos] = val;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&wc[0][0])[i] = 0;
        __syncthreads();
    }
}

size_t radix_hist_ints(long n) {
    long ntiles = (n + RS_TILE - 1) / RS_TILE;
    return (size_t)(256 * ntiles + 1);
}

// Sorts keys (and vals) over the 8-bit digits starting at the given bit shifts. Returns the
// buffer index (0 or 1) holding the result: 0 = keys/vals, 1 = keys_alt/vals_alt.
template <typename K>
int radix_sort(K* keys, K* keys_alt, int* vals, int* vals_alt, long n, const int* shifts, int npass,
               int* hist, int* scan_tmp, hipStream_t st) {
    if (n <= 1 || npass == 0) return 0;
    int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    K* src = keys;
    K* dst = keys_alt;
    int* vsrc = vals;
    int* vdst = vals_alt;
    int cur = 0;
    for (int p = 0; p < npass; ++p) {
        radix_hist_kernel<K><<<ntiles, RS_THREADS, 0, st>>>(src, n, shifts[p], ntiles, hist);
        scan_exclusive(hist, hist, 256 * ntiles, scan_tmp, st);
        if (vals)
            radix_scatter_kernel<K, true><<<ntiles, RS_THREADS, 0, st>>>(src, vsrc, n, shifts[p], ntiles, hist,
                                                                          dst, vdst);
        else
            radix_scatter_kernel<K, false><<<ntiles, RS_THREADS, 0, st>>>(src, nullptr, n, shifts[p], ntiles,
                                                                           hist, dst, nullptr);
        std::swap(src, dst);
        std::swap(vsrc, vdst);
        cur ^= 1;
    }
    return cur;
}

// The same passes over a count held in device memory: nblk fixed workgroups, each owning a
// contiguous run of the keys (a multiple of RS_THREADS long), ranked round by round as above.
__device__ __forceinline__ void dyn_sort_range(long n, long* lo, long* hi) {
    long per = (n + gridDim.x - 1) / gridDim.x;
    per = (per + RS_THREADS - 1) / RS_THREADS * RS_THREADS;
    const long a = (long)blockIdx.x * per, b = a + per;
    *lo = a < n ? a : n;
    *hi = b < n ? b : n;
}

template <typename K>
__global__ void dyn_radix_hist_kernel(const K* __restrict__ keys, Dyn cnt, int shift, int* __restrict__ hist) {
    __shared__ int c[RS_WAVES][256];
    const int wave = threadIdx.x / WAVE;
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&c[0][0])[i] = 0;
    __syncthreads();
    long lo, hi;
    dyn_sort_range(cnt.get(), &lo, &hi);
    for (long idx = lo + threadIdx.x; idx < hi; idx += RS_THREADS) atomicAdd(&c[wave][(int)((keys[idx] >> shift) & 0xFF)], 1);
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += RS_THREADS) {
        int s = 0;
        for (int w = 0; w < RS_WAVES; ++w) s += c[w][d];
        hist[(long)d * gridDim.x + blockIdx.x] = s;
    }
}

template <typename K, bool HAS_VAL>
__global__ void dyn_radix_scatter_kernel(const K* __restrict__ keys, const int* __restrict__ vals, Dyn cnt, int shift,
                                         const int* __restrict__ offsets, K* __restrict__ keys_out,
                                         int* __restrict__ vals_out) {
    __shared__ int tile_off[256];
    __shared__ int run[256];
    __shared__ int wc[RS_WAVES][256];
    const int lane = lane_id(), wave = threadIdx.x / WAVE;
    for (int d = threadIdx.x; d < 256; d += RS_THREADS) {
        tile_off[d] = offsets[(long)d * gridDim.x + blockIdx.x];
        run[d] = 0;
    }
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&wc[0][0])[i] = 0;
    __syncthreads();
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (WAVE - lane));
    long lo, hi;
    dyn_sort_range(cnt.get(), &lo, &hi);
    for (long r0 = lo; r0 < hi; r0 += RS_THREADS) {  // the same trip count in every thread
        const long idx = r0 + threadIdx.x;
        const bool valid = idx < hi;
        K key = valid ? keys[idx] : (K)0;
        int val = 0;
        if (HAS_VAL && valid) val = vals[idx];
        const int d = (int)((key >> shift) & 0xFF);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt_mask);
        if (valid && rank == 0) wc[wave][d] = __popcll(peers);
        __syncthreads();
        for (int dd = threadIdx.x; dd < 256; dd += RS_THREADS) {
            int acc = run[dd];
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) {
                const int t = wc[w][dd];
                wc[w][dd] = acc;
                acc += t;
            }
            run[dd] = acc;
        }
        __syncthreads();
        if (valid) {
            const long pos = (long)tile_off[d] + wc[wave][d] + rank;
            keys_out[pos] = key;
            if (HAS_VAL) vals_out[pos] = val;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&wc[0][0])[i] = 0;
        __syncthreads();
    }
}

template <typename K>
int radix_sort(K* keys, K* keys_alt, int* vals, int* vals_alt, Dyn n, const int* shifts, int npass, int* hist,
               int* scan_tmp, hipStream_t st) {
    if (!n.p) return radix_sort<K>(keys, keys_alt, vals, vals_alt, (long)n.v, shifts, npass, hist, scan_tmp, st);
    if (npass == 0) return 0;
    const int nblk = (int)std::max<long>(1, std::min<long>(DYN_SORT_BLOCKS, ((long)n.v + RS_TILE - 1) / RS_TILE));
    K* src = keys;
    K* dst = keys_alt;
    int* vsrc = vals;
    int* vdst = vals_alt;
    int cur = 0;
    for (int p = 0; p < npass; ++p) {
        dyn_radix_hist_kernel<K><<<nblk, RS_THREADS, 0, st>>>(src, n, shifts[p], hist);
        scan_exclusive(hist, hist, 256 * nblk, scan_tmp, st);
        if (vals)
            dyn_radix_scatter_kernel<K, true><<<nblk, RS_THREADS, 0, st>>>(src, vsrc, n, shifts[p], hist, dst, vdst);
        else
            dyn_radix_scatter_kernel<K, false><<<nblk, RS_THREADS, 0, st>>>(src, nullptr, n, shifts[p], hist, dst,
                                                                            nullptr);
        std::swap(src, dst);
        std::swap(vsrc, vdst);
        cur ^= 1;
    }
    return cur;
}

template int radix_sort<uint32_t>(uint32_t*, uint32_t*, int*, int*, long, const int*, int, int*, int*,
                                  hipStream_t);
template int radix_sort<uint64_t>(uint64_t*, uint64_t*, int*, int*, long, const int*, int, int*, int*,
                                  hipStream_t);
template int radix_sort<uint32_t>(uint32_t*, uint32_t*, int*, int*, Dyn, const int*, int, int*, int*, hipStream_t);
template int radix_sort<uint64_t>(uint64_t*, uint64_t*, int*, int*, Dyn, const int*, int, int*, int*, hipStream_t);

}  // namespace sbpx
