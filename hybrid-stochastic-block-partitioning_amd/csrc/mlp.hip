// The proposal network's arithmetic: a 64 -> 256 -> 256 -> 32 perceptron with ReLU, BF16
// weights and activations on the matrix cores (V_MFMA_F32_16X16X32_BF16, FP32 accumulation),
// trained by softmax cross-entropy with Adam on FP32 master weights.
//
// One GEMM kernel serves the forward pass and every backward product:
//   C[m][n] = epilogue( sum_k A(m, k) * B(k, n) ),  A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn]
// each wavefront owns a 16 x 16 tile of C and walks K in steps of 32; a lane holds eight
// consecutive k of one row of A (row = lane % 16) and of one column of B (column = lane % 16),
// k = 8 * (lane / 16) + i, and receives rows 4 * (lane / 16) + i of column lane % 16 of the tile.
// Out-of-range rows, columns and k are read as zero, so any M, N, K work.
#include <cstdint>

#include "internal.hpp"
#include "mlp.hpp"

namespace sbpx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float f) {
    // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    const unsigned short h = (unsigned short)(u >> 16);
    return __builtin_bit_cast(__bf16, h);
}
__device__ __forceinline__ float from_bf16(__bf16 b) {
    const unsigned short h = __builtin_bit_cast(unsigned short, b);
    return __uint_as_float((uint32_t)h << 16);
}

constexpr int GEMM_WAVES = 4;  // a workgroup covers 16 rows x 64 columns

struct GemmArgs {
    const __bf16* A;
    long sam, sak;
    const __bf16* B;
    long sbk, sbn;
    int M, N, K;
    const float* bias;     // [N] or null
    const __bf16* mask;    // [M][ldc]: output zero where mask <= 0 (ReLU backward), or null
    int relu;              // apply max(0, x) after the bias
    float* C32;            // FP32 output [M][ldc], or null
    __bf16* C16;           // BF16 output [M][ldc], or null
    int ldc;
};

__global__ void __launch_bounds__(GEMM_WAVES * WAVE) gemm_bf16_kernel(GemmArgs g) {
    const int lane = lane_id();
    const int wave = threadIdx.x / WAVE;
    const int m0 = blockIdx.y * 16;
    const int n0 = (blockIdx.x * GEMM_WAVES + wave) * 16;
    if (n0 >= g.N) return;
    const int am = m0 + (lane & 15);
    const int bn = n0 + (lane & 15);
    const int kq = 8 * (lane >> 4);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const __bf16 zero = to_bf16(0.0f);
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = k0 + kq + i;
            a[i] = (am < g.M && k < g.K) ? g.A[am * g.sam + k * g.sak] : zero;
            b[i] = (bn < g.N && k < g.K) ? g.B[k * g.sbk + bn * g.sbn] : zero;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int cn = n0 + (lane & 15);
    if (cn >= g.N) return;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int cm = m0 + 4 * (lane >> 4) + i;
        if (cm >= g.M) continue;
        float x = acc[i];
        if (g.bias) x += g.bias[cn];
        if (g.relu) x = x > 0.0f ? x : 0.0f;
        if (g.mask && !(from_bf16(g.mask[(long)cm * g.ldc + cn]) > 0.0f)) x = 0.0f;
        if (g.C32) g.C32[(long)cm * g.ldc + cn] = x;
        if (g.C16) g.C16[(long)cm * g.ldc + cn] = to_bf16(x);
    }
}

// The forward products (A and B both contiguous in k, K a multiple of 32, 16-byte aligned rows):
// each wavefront computes 16 rows x 64 columns, four 16 x 16 tiles that share one A fragment per
// k step, and each fragment is one 16-byte load per lane. Every tile runs the k steps and the
// MFMA of gemm_bf16_kernel in the same order, so the results are the same bits.
constexpr int GEMM_NT = 4;  // 16-column tiles per wavefront

__global__ void __launch_bounds__(GEMM_WAVES * WAVE) gemm_bf16_kcontig_kernel(GemmArgs g) {
    const int lane = lane_id();
    const int wave = threadIdx.x / WAVE;
    const int m0 = (blockIdx.y * GEMM_WAVES + wave) * 16;
    const int n0 = blockIdx.x * 16 * GEMM_NT;
    if (m0 >= g.M) return;
    const int am = m0 + (lane & 15);
    const int kq = 8 * (lane >> 4);
    const bool arow = am < g.M;
    const __bf16* Ap = g.A + (long)(arow ? am : 0) * g.sam + kq;
    const __bf16* Bp[GEMM_NT];
    bool bcol[GEMM_NT];
#pragma unroll
    for (int j = 0; j < GEMM_NT; ++j) {
        const int bn = n0 + 16 * j + (lane & 15);
        bcol[j] = bn < g.N;
        Bp[j] = g.B + (long)(bcol[j] ? bn : 0) * g.sbn + kq;
    }
    const bf16x8 zero8 = {};
    f32x4 acc[GEMM_NT];
#pragma unroll
    for (int j = 0; j < GEMM_NT; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        const bf16x8 a = arow ? *(const bf16x8*)(Ap + k0) : zero8;
#pragma unroll
        for (int j = 0; j < GEMM_NT; ++j) {
            if (n0 + 16 * j >= g.N) break;  // uniform over the wavefront
            const bf16x8 b = bcol[j] ? *(const bf16x8*)(Bp[j] + k0) : zero8;
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int j = 0; j < GEMM_NT; ++j) {
        const int cn = n0 + 16 * j + (lane & 15);
        if (cn >= g.N) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cm = m0 + 4 * (lane >> 4) + i;
            if (cm >= g.M) continue;
            float x = acc[j][i];
            if (g.bias) x += g.bias[cn];
            if (g.relu) x = x > 0.0f ? x : 0.0f;
            if (g.mask && !(from_bf16(g.mask[(long)cm * g.ldc + cn]) > 0.0f)) x = 0.0f;
            if (g.C32) g.C32[(long)cm * g.ldc + cn] = x;
            if (g.C16) g.C16[(long)cm * g.ldc + cn] = to_bf16(x);
        }
    }
}

static void gemm(hipStream_t st, const GemmArgs& g) {
    if (g.M <= 0 || g.N <= 0) return;
    const bool kcontig = g.sak == 1 && g.sbk == 1 && g.K % 32 == 0 && g.sam % 8 == 0 && g.sbn % 8 == 0 &&
                         (uintptr_t)g.A % 16 == 0 && (uintptr_t)g.B % 16 == 0;
    if (kcontig) {
        dim3 grid((g.N + 16 * GEMM_NT - 1) / (16 * GEMM_NT), (g.M + 16 * GEMM_WAVES - 1) / (16 * GEMM_WAVES));
        gemm_bf16_kcontig_kernel<<<grid, GEMM_WAVES * WAVE, 0, st>>>(g);
    } else {
        dim3 grid((g.N + 16 * GEMM_WAVES - 1) / (16 * GEMM_WAVES), (g.M + 15) / 16);
        gemm_bf16_kernel<<<grid, GEMM_WAVES * WAVE, 0, st>>>(g);
    }
    SBPX_HIP(hipGetLastError());
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ x, long n, __bf16* __restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = to_bf16(x[i]);
}
__global__ void bf16_to_f32_kernel(const __bf16* __restrict__ x, long n, float* __restrict__ y) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = from_bf16(x[i]);
}

static void to_bf16(hipStream_t st, const float* x, long n, __bf16* y) {
    if (n <= 0) return;
    f32_to_bf16_kernel<<<(int)std::min<long>((n + 255) / 256, 4096), 256, 0, st>>>(x, n, y);
    SBPX_HIP(hipGetLastError());
}

// Column sums of a BF16 [M][N] matrix (bias gradients) as a product with a row of ones.
static void colsum(Ctx& c, const __bf16* x, int M, int N, float* out) {
    GemmArgs g{c.mlp.ones, 0, 1, x, N, 1, 1, N, M, nullptr, nullptr, 0, out, nullptr, N};
    gemm(c.stream, g);
}

// Softmax cross-entropy over the first `valid[m]`-masked logits of each row: logits Z[M][ZW]
// (FP32), candidate mask bits cmask[m] (bit j: candidate j exists and may be proposed), label
// y[m] in [0, 16). dZ = (softmax - onehot) / M as BF16; loss[m] = -log softmax[y].
__global__ void xent_kernel(const float* __restrict__ Z, const uint32_t* __restrict__ cmask,
                            const int* __restrict__ y, int M, float* __restrict__ loss, __bf16* __restrict__ dZ) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const float* z = Z + (long)m * MLP_OUT;
    const uint32_t mk = cmask[m];
    float mx = -1e30f;
    for (int j = 0; j < MLP_CAND; ++j)
        if (mk >> j & 1u) mx = fmaxf(mx, z[j]);
    float den = 0.0f;
    for (int j = 0; j < MLP_CAND; ++j)
        if (mk >> j & 1u) den += expf(z[j] - mx);
    const float inv = 1.0f / (den * (float)M);
    for (int j = 0; j < MLP_OUT; ++j) {
        float g = 0.0f;
        if (j < MLP_CAND && (mk >> j & 1u)) g = expf(z[j] - mx) * inv;
        if (j == y[m]) g -= 1.0f / (float)M;
        dZ[(long)m * MLP_OUT + j] = to_bf16(g);
    }
    loss[m] = -(z[y[m]] - mx - logf(den));
}

// Training rows idx[0..n) of the logged samples into the batch buffers.
__global__ void gather_rows_kernel(const int* __restrict__ idx, int n, const __bf16* __restrict__ ds_x,
                                   const uint32_t* __restrict__ ds_mask, const int* __restrict__ ds_y,
                                   __bf16* __restrict__ x, uint32_t* __restrict__ mask, int* __restrict__ y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = t / MLP_IN, j = t % MLP_IN;
    if (row >= n) return;
    const int src = idx[row];
    x[(long)row * MLP_IN + j] = ds_x[(long)src * MLP_IN + j];
    if (j == 0) {
        mask[row] = ds_mask[src];
        y[row] = ds_y[src];
    }
}

void mlp_gather(Ctx& c, const int* idx, int n) {
    MlpState& m = c.mlp;
    gather_rows_kernel<<<(n * MLP_IN + 255) / 256, 256, 0, c.stream>>>(idx, n, m.ds_x, m.ds_mask, m.ds_y, m.tx,
                                                                      m.tmask, m.ty);
    SBPX_HIP(hipGetLastError());
}

// Adam (Kingma and Ba, 2015) on FP32 master weights; writes the BF16 copy the GEMMs read.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m1,
                            float* __restrict__ m2, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2, __bf16* __restrict__ w16) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float a = b1 * m1[i] + (1.0f - b1) * g;
        const float b = b2 * m2[i] + (1.0f - b2) * g * g;
        m1[i] = a;
        m2[i] = b;
        const float upd = lr * (a / c1) / (sqrtf(b / c2) + eps);
        const float nw = w[i] - upd;
        w[i] = nw;
        if (w16) w16[i] = to_bf16(nw);
    }
}

// ------------------------------------------------------------------------------------------
void mlp_forward(Ctx& c, const __bf16* X, int V, __bf16* H1, __bf16* H2, float* Z) {
    MlpState& s = c.mlp;
    GemmArgs g{};
    // H1 = relu(X W1 + b1): A = X [V][64], B(k, n) = W1t[n][k]
    g = GemmArgs{X, MLP_IN, 1, s.w16[0], 1, MLP_IN, V, MLP_H, MLP_IN, s.b[0], nullptr, 1, nullptr, H1, MLP_H};
    gemm(c.stream, g);
    g = GemmArgs{H1, MLP_H, 1, s.w16[1], 1, MLP_H, V, MLP_H, MLP_H, s.b[1], nullptr, 1, nullptr, H2, MLP_H};
    gemm(c.stream, g);
    g = GemmArgs{H2, MLP_H, 1, s.w16[2], 1, MLP_H, V, MLP_OUT, MLP_H, s.b[2], nullptr, 0, Z, nullptr, MLP_OUT};
    gemm(c.stream, g);
}

// Gradients of the mean cross-entropy for one batch (forward recomputed here); grads[l] and
// gbias[l] receive dW_l (out x in, like the weights) and db_l.
void mlp_backward(Ctx& c, const __bf16* X, const uint32_t* cmask, const int* y, int V, float* loss, float* grads[3],
                  float* gbias[3]) {
    MlpState& s = c.mlp;
    __bf16* H1 = s.h1;
    __bf16* H2 = s.h2;
    float* Z = s.z;
    mlp_forward(c, X, V, H1, H2, Z);
    __bf16* dZ = s.dz;
    xent_kernel<<<(V + 255) / 256, 256, 0, c.stream>>>(Z, cmask, y, V, loss, dZ);
    SBPX_HIP(hipGetLastError());
    GemmArgs g{};
    // dW3t[o][i] = sum_v dZ[v][o] H2[v][i]
    g = GemmArgs{dZ, 1, MLP_OUT, H2, MLP_H, 1, MLP_OUT, MLP_H, V, nullptr, nullptr, 0, grads[2], nullptr, MLP_H};
    gemm(c.stream, g);
    colsum(c, dZ, V, MLP_OUT, gbias[2]);
    // dH2 = (dZ W3) * [H2 > 0]: A(v, o) = dZ[v][o], B(o, i) = W3t[o][i]
    __bf16* dH2 = s.dh2;
    g = GemmArgs{dZ, MLP_OUT, 1, s.w16[2], MLP_H, 1, V, MLP_H, MLP_OUT, nullptr, H2, 0, nullptr, dH2, MLP_H};
    gemm(c.stream, g);
    g = GemmArgs{dH2, 1, MLP_H, H1, MLP_H, 1, MLP_H, MLP_H, V, nullptr, nullptr, 0, grads[1], nullptr, MLP_H};
    gemm(c.stream, g);
    colsum(c, dH2, V, MLP_H, gbias[1]);
    __bf16* dH1 = s.dh1;
    g = GemmArgs{dH2, MLP_H, 1, s.w16[1], MLP_H, 1, V, MLP_H, MLP_H, nullptr, H1, 0, nullptr, dH1, MLP_H};
    gemm(c.stream, g);
    g = GemmArgs{dH1, 1, MLP_H, X, MLP_IN, 1, MLP_H, MLP_IN, V, nullptr, nullptr, 0, grads[0], nullptr, MLP_IN};
    gemm(c.stream, g);
    colsum(c, dH1, V, MLP_H, gbias[0]);
}

void mlp_adam_step(Ctx& c, float lr) {
    MlpState& s = c.mlp;
    s.adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)s.adam_t), c2 = 1.0f - powf(b2, (float)s.adam_t);
    for (int l = 0; l < 3; ++l) {
        const long n = mlp_weight_count(l);
        adam_kernel<<<(int)std::min<long>((n + 255) / 256, 1024), 256, 0, c.stream>>>(
            s.w[l], s.gw[l], s.m1w[l], s.m2w[l], n, lr, b1, b2, eps, c1, c2, s.w16[l]);
        const long nb = mlp_bias_count(l);
        adam_kernel<<<1, 256, 0, c.stream>>>(s.b[l], s.gb[l], s.m1b[l], s.m2b[l], nb, lr, b1, b2, eps, c1, c2, nullptr);
    }
    SBPX_HIP(hipGetLastError());
}

void mlp_sync_bf16(Ctx& c) {
    for (int l = 0; l < 3; ++l) to_bf16(c.stream, c.mlp.w[l], mlp_weight_count(l), c.mlp.w16[l]);
}

void mlp_convert_to_bf16(Ctx& c, const float* x, long n, __bf16* y) { to_bf16(c.stream, x, n, y); }

void mlp_convert_to_f32(Ctx& c, const __bf16* x, long n, float* y) {
    if (n <= 0) return;
    bf16_to_f32_kernel<<<(int)std::min<long>((n + 255) / 256, 4096), 256, 0, c.stream>>>(x, n, y);
    SBPX_HIP(hipGetLastError());
}

}  // namespace sbpx
