// mlp_common.h -- the learned surrogate 16 -> 256 -> 256 -> 1 (GELU) shared by
// the training kernels (mlp_device.hip) and the MLP screen (screen_device.hip).
// No counterpart in the reference: codec-arena scores candidates with the
// quadratic surface only (src/screen.zig:97-112).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace csf {

constexpr int MLP_IN = 16;   // the 16 legacy effector genes
constexpr int MLP_H = 256;
// flat fp32 parameter vector: W1 [16][256], b1, W2 [256][256], b2, w3, b3
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + MLP_IN * MLP_H;
constexpr int P_W2 = P_B1 + MLP_H;
constexpr int P_B2 = P_W2 + MLP_H * MLP_H;
constexpr int P_W3 = P_B2 + MLP_H;
constexpr int P_B3 = P_W3 + MLP_H;
constexpr int P_TOTAL = P_B3 + 1;  // 70401

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// GELU, tanh form (Hendrycks & Gimpel 2016):
//   gelu(x) = 0.5 x (1 + tanh(u)),  u = sqrt(2/pi) (x + 0.044715 x^3)
//           = x * sigmoid(2u)
// Evaluated with the hardware base-2 exponential and reciprocal (about 1 ulp
// each): sigmoid(2u) = 1 / (1 + 2^(-2u log2 e)).
constexpr float GELU_K = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float GELU_C = 0.044715f;
constexpr float GELU_E = -2.0f * GELU_K * 1.4426950408889634f;  // -2 sqrt(2/pi) log2(e)
__device__ __forceinline__ float gelu_sig(float x) {
    const float inner = __builtin_fmaf(GELU_C * x * x, x, x);  // x + c x^3
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(GELU_E * inner));
}
__device__ __forceinline__ float gelu_f(float x) { return x * gelu_sig(x); }
// d gelu / dx = s + x s (1 - s) du2/dx, s = sigmoid(2u)
__device__ __forceinline__ float gelu_grad_f(float x) {
    const float s = gelu_sig(x);
    const float du2 = 2.0f * GELU_K * (1.0f + 3.0f * GELU_C * x * x);
    return s + x * s * (1.0f - s) * du2;
}

// Layouts prepared for the screen (one wavefront scores 64 candidates):
//   w1f  [16 m-tiles][64 lanes] x 4 bf16: A fragments of W1^T for
//        v_mfma_f32_16x16x16_bf16 (lane l: hidden 16 m + l%16, genes 4(l/16)..+3)
//   w2f  [8 k-blocks][16 n-tiles][64 lanes] x 8: B fragments of W2 for
//        v_mfma_f32_16x16x32_{bf16,fp8} in the permuted K order that matches the
//        layer-1 accumulator layout: slot e of lane group q in k-block kb is
//        hidden unit 32 kb + (e < 4 ? 4 q + e : 16 + 4 q + e - 4)
__host__ __device__ constexpr int w2_k_of(int kb, int q, int e) {
    return 32 * kb + (e < 4 ? 4 * q + e : 16 + 4 * q + (e - 4));
}

struct MlpScreenW {
    const bf16x4 *w1f;     // 16 * 64
    const bf16x8 *w2f;     // 8 * 16 * 64 (bf16)
    const uint64_t *w2f8;  // 8 * 16 * 64 (fp8 x 8)
    const float *b1, *b2, *w3;
    float b3;
    float h_scale;   // fp8: H1 is multiplied by this before conversion
    float w_scale;   // fp8: W2 was multiplied by this before conversion
    float descale;   // 1 / (h_scale * w_scale)
};

}  // namespace csf
