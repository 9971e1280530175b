// Parameter layout and BF16 conversion of the bracket-position network.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rfd {

constexpr int MLP_IN = 400;      // padded input width (multiple of 16)
constexpr int MLP_FEAT = 386;    // 2 x 64 x 3 curve samples + 2 bounds
constexpr int MLP_H = 256;
constexpr int MLP_LAYERS = 4;    // hidden layers

// Parameter layout (FP32): W1 [400x256], b1, W2 [256x256], b2, W3, b3, W4, b4, w5 [256], b5.
__host__ __device__ constexpr long long mlp_w_off(int l) {  // l = 0..3
    return l == 0 ? 0 : (long long)MLP_IN * MLP_H + MLP_H + (long long)(l - 1) * (MLP_H * MLP_H + MLP_H);
}
__host__ __device__ constexpr long long mlp_b_off(int l) {
    return mlp_w_off(l) + (long long)(l == 0 ? MLP_IN : MLP_H) * MLP_H;
}
constexpr long long MLP_W5 = mlp_w_off(3) + (long long)MLP_H * MLP_H + MLP_H;
constexpr long long MLP_B5 = MLP_W5 + MLP_H;
constexpr long long MLP_NPARAMS = MLP_B5 + 1;

// FP32 -> BF16 with round-to-nearest-even (finite inputs).
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

}  // namespace rfd
