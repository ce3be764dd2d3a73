// Shapes and device state of the proposal network (64 -> 256 -> 256 -> 32, ReLU).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sbpx {

constexpr int MLP_IN = 64;     // features per vertex
constexpr int MLP_H = 256;     // hidden width (two layers)
constexpr int MLP_OUT = 32;    // output scores
constexpr int MLP_CAND = 32;   // candidate blocks scored (one output each)
constexpr int MLP_BATCH = 8192;  // training rows per step (upper bound)

// Weights are stored output-major (W_l[out][in]), so a GEMM reads eight consecutive inputs of one
// output as one vector.
inline long mlp_weight_count(int l) { return l == 0 ? (long)MLP_H * MLP_IN : (l == 1 ? (long)MLP_H * MLP_H : (long)MLP_OUT * MLP_H); }
inline long mlp_bias_count(int l) { return l == 2 ? MLP_OUT : MLP_H; }
// Flat parameter order of the C ABI: W1, b1, W2, b2, W3, b3.
inline long mlp_param_count() {
    long n = 0;
    for (int l = 0; l < 3; ++l) n += mlp_weight_count(l) + mlp_bias_count(l);
    return n;
}

struct MlpState {
    bool ready = false;       // weights allocated and initialised
    float* w[3] = {};         // FP32 master weights
    __bf16* w16[3] = {};      // BF16 copies read by the GEMMs
    float* b[3] = {};         // FP32 biases
    float* gw[3] = {};        // gradients
    float* gb[3] = {};
    float* m1w[3] = {};       // Adam moments
    float* m2w[3] = {};
    float* m1b[3] = {};
    float* m2b[3] = {};
    int adam_t = 0;
    // training batch buffers [MLP_BATCH rows]
    __bf16 *h1 = nullptr, *h2 = nullptr, *dz = nullptr, *dh1 = nullptr, *dh2 = nullptr, *ones = nullptr;
    float* z = nullptr;
    float* loss = nullptr;
    __bf16* tx = nullptr;     // training batch rows gathered from the logged samples
    uint32_t* tmask = nullptr;
    int* ty = nullptr;
    int* tidx = nullptr;
    // logged training samples
    __bf16* ds_x = nullptr;   // [cap][MLP_IN]
    uint32_t* ds_mask = nullptr;
    int* ds_y = nullptr;
    int* ds_count = nullptr;  // device counter
    int *log_flag = nullptr, *log_rank = nullptr;  // [log_slots] rows of a batch and their ranks
    size_t log_slots = 0;
    int ds_cap = 0;
    bool logging = false;
    // sweep-time buffers, [N] rows
    int rows = 0;
    __bf16 *x = nullptr, *sh1 = nullptr, *sh2 = nullptr;
    float* sz = nullptr;
    int* cand = nullptr;      // [N][MLP_CAND] candidate blocks (-1: none)
    int* cnt = nullptr;       // [N][MLP_CAND] edges of the vertex to each candidate
    float* qf = nullptr;      // [N] learned forward proposal probability (< 0: no learned mass)
    float* qb = nullptr;      // [N] learned reverse proposal probability
    int* ps = nullptr;        // [N] proposal of the mixture
    double* pu = nullptr;     // [N] accept uniform
};

}  // namespace sbpx
