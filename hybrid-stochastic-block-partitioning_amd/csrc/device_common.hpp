// Device helpers shared by the SBP kernels: counter-based randomness, the FP32 log used by the
// objective, the FP32 cell entropy, and wavefront (64-lane) primitives for gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sbpx {

constexpr int WAVE = 64;

// ------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). Counter layout shared with the CPU oracle:
// (vertex or block, call id, proposal index, tag); key = 64-bit seed.
struct Draw4 { uint32_t x0, x1, x2, x3; };

__device__ __forceinline__ Draw4 philox4x32_10(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Draw4{c0, c1, c2, c3};
}

constexpr uint32_t TAG_ASYNC = 0x4D430000u;
constexpr uint32_t TAG_MH = 0x4D480000u;
constexpr uint32_t TAG_MERGE = 0x4D520000u;
constexpr uint32_t TAG_LEARN = 0x4D4C0000u;  // learned-mixture branch and categorical draws

__device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }

__device__ __forceinline__ double unit53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// A uniform block other than r (common.cpp:184-192).
__device__ __forceinline__ int random_block(int r, int B, uint32_t x) {
    int p = (int)mulhi32(x, (uint32_t)(B - 1));
    return p >= r ? p + 1 : p;
}

// ------------------------------------------------------------------------------------------
// logf as the host C library computes it (glibc >= 2.28): the table-driven method of
// Szabolcs Nagy's optimized-routines logf. The 16-entry table and the polynomial coefficients
// below are that routine's (Arm optimized-routines, Copyright (c) 2017-2018 Arm Limited,
// SPDX-License-Identifier: MIT OR Apache-2.0 WITH LLVM-exception; glibc ships the same code). x = 2^k z with z in [OFF, 2 OFF); z is split into
// 16 subintervals, each with a reciprocal c^-1 and log(c); log(x) = k ln2 + log(c) +
// log1p(z/c - 1), with log1p approximated by a degree-4 polynomial evaluated in double and
// rounded once to float. The objective depends on this function bit for bit, so the device
// evaluates the same sequence of double operations; the test suite compares it with the host
// logf over every positive float.
struct LogfEntry { double invc, logc; };
__device__ __constant__ LogfEntry kLogfTable[16] = {
    {0x1.661ec79f8f3bep+0, -0x1.57bf7808caadep-2}, {0x1.571ed4aaf883dp+0, -0x1.2bef0a7c06ddbp-2},
    {0x1.49539f0f010bp+0, -0x1.01eae7f513a67p-2},  {0x1.3c995b0b80385p+0, -0x1.b31d8a68224e9p-3},
    {0x1.30d190c8864a5p+0, -0x1.6574f0ac07758p-3}, {0x1.25e227b0b8eap+0, -0x1.1aa2bc79c81p-3},
    {0x1.1bb4a4a1a343fp+0, -0x1.a4e76ce8c0e5ep-4}, {0x1.12358f08ae5bap+0, -0x1.1973c5a611cccp-4},
    {0x1.0953f419900a7p+0, -0x1.252f438e10c1ep-5}, {0x1p+0, 0x0p+0},
    {0x1.e608cfd9a47acp-1, 0x1.aa5aa5df25984p-5},  {0x1.ca4b31f026aap-1, 0x1.c5e53aa362eb4p-4},
    {0x1.b2036576afce6p-1, 0x1.526e57720db08p-3},  {0x1.9c2d163a1aa2dp-1, 0x1.bc2860d22477p-3},
    {0x1.886e6037841edp-1, 0x1.1058bc8a07ee1p-2},  {0x1.767dcf5534862p-1, 0x1.4043057b6ee09p-2},
};

__device__ __forceinline__ float host_logf(float x) {
    const double A0 = -0x1.00ea348b88334p-2, A1 = 0x1.5575b0be00b6ap-2, A2 = -0x1.ffffef20a4123p-2;
    const double Ln2 = 0x1.62e42fefa39efp-1;
    uint32_t ix = __float_as_uint(x);
    if (ix == 0x3f800000u) return 0.0f;
    if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
        if ((ix << 1) == 0) return -__builtin_inff();
        if (ix == 0x7f800000u) return x;
        if ((ix & 0x80000000u) || (ix << 1) >= 0xff000000u) return __builtin_nanf("");
        ix = __float_as_uint(x * 0x1p23f);  // subnormal: normalise
        ix -= 23u << 23;
    }
    uint32_t tmp = ix - 0x3f330000u;
    int i = (int)((tmp >> 19) & 15u);
    int k = (int32_t)tmp >> 23;
    uint32_t iz = ix - (tmp & 0xff800000u);
    double invc = kLogfTable[i].invc, logc = kLogfTable[i].logc;
    double z = (double)__uint_as_float(iz);
    double r = z * invc - 1.0;
    double y0 = logc + (double)k * Ln2;
    double r2 = r * r;
    double y = A1 * r + A2;
    y = A0 * r2 + y;
    y = y * r2 + (y0 + r);
    return (float)y;
}

// common::cell_entropy (common.hpp:35-44): value * logf(value / (din * dout)), all FP32.
__device__ __forceinline__ float cell_entropy(float value, float din, float dout) {
    if (value == 0.0f) return 0.0f;
    return value * host_logf(value / (din * dout));
}

// ------------------------------------------------------------------------------------------
// Wavefront primitives.
__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }

// Orders this wavefront's earlier memory writes before its later reads (LDS and global).
__device__ __forceinline__ void fence_wave() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, WAVE));
    return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, WAVE));
    return v;
}
__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ int first_lane(unsigned long long mask) { return __ffsll((long long)mask) - 1; }

// First index j in [lo, hi) with cum[j] > target, for a non-decreasing cum with
// cum[hi-1] > target. All 64 lanes take part; each round narrows the range 64-fold.
__device__ __forceinline__ int wave_upper_bound(const int* __restrict__ cum, int lo, int hi, int target) {
    const int lane = lane_id();
    while (hi - lo > WAVE) {
        int len = hi - lo;
        int step = (len + WAVE - 1) / WAVE;
        int end = min((lane + 1) * step, len);
        bool gt = (lane * step < len) && (cum[lo + end - 1] > target);
        unsigned long long m = __ballot(gt);
        int L = first_lane(m);
        int nlo = lo + L * step;
        hi = min(lo + (L + 1) * step, hi);
        lo = nlo;
    }
    int p = lo + lane;
    bool gt = (p < hi) && (cum[p] > target);
    unsigned long long m = __ballot(gt);
    return lo + first_lane(m);
}

// Lane groups: W consecutive lanes (W a power of two dividing 64) evaluate one vertex together.
// Shuffles stay inside the group (offsets below W) and ballots are read at the group's bits, so
// groups of one wavefront may follow different paths; every lane of a group follows the same.
template <int W>
__device__ __forceinline__ int group_lane() { return threadIdx.x & (W - 1); }
template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
    return v;
}
template <int W>
__device__ __forceinline__ long long group_sum_ll(long long v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
    return v;
}
template <int W>
__device__ __forceinline__ int group_max(int v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, W));
    return v;
}
template <int W>
__device__ __forceinline__ int group_min(int v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, W));
    return v;
}
template <int W>
__device__ __forceinline__ int group_or(int v) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) v |= __shfl_xor(v, off, W);
    return v;
}
// The group's bits of a wavefront ballot.
template <int W>
__device__ __forceinline__ unsigned long long group_ballot(bool p) {
    const unsigned long long m = __ballot(p);
    if constexpr (W == WAVE) return m;
    const int base = (threadIdx.x & (WAVE - 1)) & ~(W - 1);
    return (m >> base) & ((1ull << W) - 1ull);
}

// wave_upper_bound for a group of W lanes: each round narrows the range W-fold.
template <int W>
__device__ __forceinline__ int group_upper_bound(const int* __restrict__ cum, int lo, int hi, int target) {
    if constexpr (W == WAVE) return wave_upper_bound(cum, lo, hi, target);
    const int lane = group_lane<W>();
    while (hi - lo > W) {
        int len = hi - lo;
        int step = (len + W - 1) / W;
        int end = min((lane + 1) * step, len);
        bool gt = (lane * step < len) && (cum[lo + end - 1] > target);
        int L = first_lane(group_ballot<W>(gt));
        int nlo = lo + L * step;
        hi = min(lo + (L + 1) * step, hi);
        lo = nlo;
    }
    int p = lo + lane;
    bool gt = (p < hi) && (cum[p] > target);
    return lo + first_lane(group_ballot<W>(gt));
}

// Lower bound of key in the sorted array keys[lo, hi) (one lane).
__device__ __forceinline__ int lower_bound_lane(const int* __restrict__ keys, int lo, int hi, int key) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Value of the entry with key in sorted keys[lo, hi), or 0.
__device__ __forceinline__ int lookup_sorted(const int* __restrict__ keys, const int* __restrict__ vals,
                                             int lo, int hi, int key) {
    int p = lower_bound_lane(keys, lo, hi, key);
    return (p < hi && keys[p] == key) ? vals[p] : 0;
}

__device__ __forceinline__ uint32_t hash_block(int b) { return (uint32_t)b * 0x9E3779B1u; }

}  // namespace sbpx
