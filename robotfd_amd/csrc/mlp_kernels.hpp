// Bracket-position network: BF16 MLP on MFMA with FP32 accumulation.
//
//   x (400 = 386 features + zero pad) -> 256 -> 256 -> 256 -> 256 -> 1
//   hidden layers: a = relu(x W + b) stored in BF16; head: p = sigmoid(a w + b) in FP32.
//
// The hidden layers' forward and backward products run through one GEMM
// kernel on gfx950's v_mfma_f32_16x16x32_bf16: a 64x64 output tile per
// workgroup of four waves, each wave a 32x32 block of 2x2 MFMA tiles, with
// the operands of each 32-deep K step staged through LDS (each element read
// from memory once per workgroup, each wave's fragments one 16-byte LDS read
// apiece).  The three operand layouts training needs (X W, X^T dZ, dZ W^T)
// are strides of the same kernel.  Epilogues fuse bias + ReLU (forward) and
// the ReLU mask of the layer below (backward).  Master weights and Adam
// moments are FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "mlp_layout.hpp"

namespace rfd {

typedef __bf16 mfma_bf8 __attribute__((ext_vector_type(8)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int mfma_u4 __attribute__((ext_vector_type(4)));

enum GemmEpi { EPI_F32 = 0, EPI_BIAS_RELU_BF16 = 1, EPI_MASK_BF16 = 2 };

struct GemmArgs {
    const uint16_t *A;   // BF16, element (i, k) at A[i*sai + k*sak]
    const uint16_t *B;   // BF16, element (k, j) at B[k*sbk + j*sbj]
    int M, N, K;
    long long sai, sak, sbk, sbj;
    float *Cf;           // EPI_F32: C[i*N + j]
    uint16_t *Cb;        // BF16 epilogues: C[i*N + j]
    const float *bias;   // EPI_BIAS_RELU_BF16: bias[j]
    const uint16_t *mask;  // EPI_MASK_BF16: keep where mask[i*N + j] > 0
};

// K step of the staged tiles: A as [64 rows][k], B transposed as [64
// columns][k], k contiguous, rows 16 bytes longer than the staged K (row
// starts 16-byte aligned).  KSTEPS = 1 streams K through one step's tiles;
// KSTEPS > 1 stages the whole K (at most 32 KSTEPS) at once, so that every
// load of the launch is issued before the first multiply: the layers' K is
// 400 or 256, where a streamed launch waits on memory once per step.
constexpr int kGemmK = 32;
constexpr int kGemmStepsAll = 13;  // the network's widest K, 400, in 32-deep steps

// MFMA 16x16x32 BF16 operand layout: lane l supplies A[l%16][8*(l/16) + t]
// and B[8*(l/16) + t][l%16], t = 0..7; it receives D[4*(l/16) + t][l%16],
// t = 0..3.
template <int EPI, int KSTEPS>
__global__ __launch_bounds__(256) void mlp_gemm(GemmArgs g) {
    constexpr int kGemmLd = KSTEPS * kGemmK + 8;
    __shared__ __attribute__((aligned(16))) uint16_t As[64 * kGemmLd];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * kGemmLd];
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int tr0 = (int)blockIdx.y * 64, tc0 = (int)blockIdx.x * 64;
    const int row0 = tr0 + (wave >> 1) * 32;
    const int col0 = tc0 + (wave & 1) * 32;
    const int r16 = lane & 15, kq = (lane >> 4) * 4, k8 = (lane >> 4) * 8;
    // Staging: thread t loads tile row (and tile column) t / 4, k t % 4 * 8 .. + 7.
    const int lr = tid >> 2, lk = (tid & 3) * 8;
    const int li = tr0 + lr, lj = tc0 + lr;
    // One thread's eight A and eight B elements of the K step at k0, packed
    // two to a word; zero outside the matrices.  A row whose k runs
    // contiguously (sak = 1, the forward's X) with 16-byte aligned rows is
    // one 16-byte load.
    const bool avec = g.sak == 1 && (g.sai & 7) == 0 && ((uintptr_t)g.A & 15) == 0;
    auto load = [&](const int k0, mfma_u4 &va, mfma_u4 &vb) __attribute__((always_inline)) {
        const int kb = k0 + lk;
        unsigned a2[4] = {0, 0, 0, 0}, b2[4] = {0, 0, 0, 0};
        if (avec && kb + 8 <= g.K && li < g.M) {
            const mfma_u4 v = *(const mfma_u4 *)(g.A + (long long)li * g.sai + kb);
            a2[0] = v[0], a2[1] = v[1], a2[2] = v[2], a2[3] = v[3];
        } else if (li < g.M) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = kb + u;
                const unsigned e = k < g.K ? g.A[(long long)li * g.sai + (long long)k * g.sak] : 0u;
                a2[u >> 1] |= e << (16 * (u & 1));
            }
        }
        if (lj < g.N) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = kb + u;
                const unsigned e = k < g.K ? g.B[(long long)k * g.sbk + (long long)lj * g.sbj] : 0u;
                b2[u >> 1] |= e << (16 * (u & 1));
            }
        }
        va = mfma_u4{a2[0], a2[1], a2[2], a2[3]};
        vb = mfma_u4{b2[0], b2[1], b2[2], b2[3]};
    };
    mfma_f4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    // The fragments of the K step staged at column kk of the tiles.
    auto multiply = [&](const int kk) __attribute__((always_inline)) {
        mfma_bf8 fa[2], fb[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const mfma_u4 v = *(const mfma_u4 *)&As[((wave >> 1) * 32 + a * 16 + r16) * kGemmLd + kk + k8];
            fa[a] = __builtin_bit_cast(mfma_bf8, v);
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const mfma_u4 v = *(const mfma_u4 *)&Bs[((wave & 1) * 32 + b * 16 + r16) * kGemmLd + kk + k8];
            fb[b] = __builtin_bit_cast(mfma_bf8, v);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    };
    if constexpr (KSTEPS > 1) {
        // The whole K at once (the host launches this form for K <= 32 KSTEPS).
        mfma_u4 va[KSTEPS], vb[KSTEPS];
#pragma unroll
        for (int st = 0; st < KSTEPS; ++st) load(st * kGemmK, va[st], vb[st]);
#pragma unroll
        for (int st = 0; st < KSTEPS; ++st) {
            *(mfma_u4 *)&As[lr * kGemmLd + st * kGemmK + lk] = va[st];
            *(mfma_u4 *)&Bs[lr * kGemmLd + st * kGemmK + lk] = vb[st];
        }
        __syncthreads();
        for (int k0 = 0; k0 < g.K; k0 += kGemmK) multiply(k0);
    } else {
        mfma_u4 va, vb;
        load(0, va, vb);
        for (int k0 = 0; k0 < g.K; k0 += kGemmK) {
            __syncthreads();  // the previous step's fragments are read
            *(mfma_u4 *)&As[lr * kGemmLd + lk] = va;
            *(mfma_u4 *)&Bs[lr * kGemmLd + lk] = vb;
            __syncthreads();
            // The next step's elements are in flight while this one's multiply.
            if (k0 + kGemmK < g.K) load(k0 + kGemmK, va, vb);
            multiply(0);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int i = row0 + a * 16 + kq + t;
                const int j = col0 + b * 16 + r16;
                if (i >= g.M || j >= g.N) continue;
                const float v = acc[a][b][t];
                const long long o = (long long)i * g.N + j;
                if (EPI == EPI_F32) {
                    g.Cf[o] = v;
                } else if (EPI == EPI_BIAS_RELU_BF16) {
                    const float z = v + g.bias[j];
                    g.Cb[o] = f2bf(z > 0.f ? z : 0.f);
                } else {
                    g.Cb[o] = g.mask[o] > 0 ? f2bf(v) : (uint16_t)0;  // BF16 > 0 as raw bits: sign clear, nonzero
                }
            }
}

// Head: z_i = a4_i . w5 + b5, p_i = sigmoid(z_i).  Training also forms the
// binary cross-entropy gradient dz_i = (p_i - y_i) / count over valid rows.
__global__ void mlp_head(const uint16_t *A4, const float *params, int rows, const float *labels,
                         int valid, float *p_out, float *dz, float *loss_rows) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= rows) return;
    const float *w5 = params + MLP_W5;
    float z = params[MLP_B5];
    for (int k = 0; k < MLP_H; ++k) z += bf2f(A4[(long long)i * MLP_H + k]) * w5[k];
    const float p = 1.f / (1.f + __expf(-z));
    if (p_out) p_out[i] = p;
    if (dz) {
        const bool ok = i < valid;
        const float y = ok ? labels[i] : 0.f;
        dz[i] = ok ? (p - y) / (float)valid : 0.f;
        if (loss_rows) {
            // log(1 + e^-|z|) form of the cross-entropy, stable for large |z|.
            const float l = fmaxf(z, 0.f) - z * y + __logf(1.f + __expf(-fabsf(z)));
            loss_rows[i] = ok ? l / (float)valid : 0.f;
        }
    }
}

// dA4 = dz w5^T masked by a4 > 0 (BF16, the next GEMMs' operand); dw5 and db5.
__global__ void mlp_head_back(const uint16_t *A4, const float *params, const float *dz, int rows,
                              uint16_t *dZ4, float *grads) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long long)rows * MLP_H) {
        const int i = (int)(t / MLP_H), k = (int)(t % MLP_H);
        const float g = dz[i] * params[MLP_W5 + k];
        dZ4[t] = A4[t] > 0 ? f2bf(g) : (uint16_t)0;
    }
    if (t < MLP_H + 1) {
        float s = 0.f;
        if (t < MLP_H) {
            for (int i = 0; i < rows; ++i) s += dz[i] * bf2f(A4[(long long)i * MLP_H + t]);
            grads[MLP_W5 + t] = s;
        } else {
            for (int i = 0; i < rows; ++i) s += dz[i];
            grads[MLP_B5] = s;
        }
    }
}

// Column sums of a BF16 [rows x cols] matrix (bias gradients).
__global__ void mlp_colsum(const uint16_t *X, int rows, int cols, float *out) {
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= cols) return;
    float s = 0.f;
    for (int i = 0; i < rows; ++i) s += bf2f(X[(long long)i * cols + j]);
    out[j] = s;
}

__global__ void mlp_adam(float *params, const float *grads, float *m1, float *m2, long long n,
                         float lr, float b1, float b2, float eps, float c1, float c2) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float g = grads[t];
    const float a = b1 * m1[t] + (1.f - b1) * g;
    const float v = b2 * m2[t] + (1.f - b2) * g * g;
    m1[t] = a;
    m2[t] = v;
    params[t] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// BF16 copies of the hidden layers' weights (the MFMA operands).
__global__ void mlp_cast_weights(const float *params, uint16_t *wb) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n0 = (long long)MLP_IN * MLP_H, nh = (long long)MLP_H * MLP_H;
    if (t >= n0 + 3 * nh) return;
    long long src;
    if (t < n0) src = t;
    else {
        const long long l = 1 + (t - n0) / nh, r = (t - n0) % nh;
        src = mlp_w_off((int)l) + r;
    }
    wb[t] = f2bf(params[src]);
}

__global__ void mlp_rows_to_bf16(const float *x, int rows, uint16_t *xb) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)rows * MLP_IN) return;
    const int i = (int)(t / MLP_IN), c = (int)(t % MLP_IN);
    xb[t] = c < MLP_FEAT ? f2bf(x[(long long)i * MLP_FEAT + c]) : (uint16_t)0;
}

}  // namespace rfd
