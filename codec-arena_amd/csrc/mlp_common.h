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
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef unsigned long u64x4 __attribute__((ext_vector_type(4)));
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

// Packed-f16 GELU of the MLP screen (two activations per lane and
// instruction, no transcendentals).  gelu(x) = x S(x) with
//   x16 = f16(x) (toward zero), v = x16 q, w = v v - 1,
//   S = clamp(1/2 + v R(w), 0, 1), R of degree 6 in Horner form;
// q = f16(sqrt 2) / 4 maps |x| <= 4.0004 onto w in [-1, 1], the fitted range.
// Beyond it R (even degree, positive leading coefficient) stays above
// 1 / (2 |v|), up to f16 overflow to +inf, so the FMA's clamp bit saturates
// S to 0 or 1 without clamping w first.  A power-of-two scale carried by x
// (FP8 screen) is absorbed into q.  Coefficients and error bounds:
// gelu_poly_fit.py (max |S - S_exact| 6.8e-4 with every step in f16, checked
// out to the f16 range of x).
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
constexpr float GELU_Q = 0.353515625f;  // f16(sqrt 2) / 4, exact in f16
__device__ __forceinline__ h2 h2s(_Float16 c) { return h2{c, c}; }
__device__ __forceinline__ h2 pk_rtz(float a, float b) { return __builtin_bit_cast(h2, __builtin_amdgcn_cvt_pkrtz(a, b)); }
// V is a vector of f16 (2 or 8 lanes): an 8-lane vector is four independent
// packed chains, so each Horner step issues four v_pk_fma_f16 back to back
// instead of waiting on one dependent chain.
template <typename V>
__device__ __forceinline__ V gelu_sig_scaled(V v) {  // S from v = x q
    const V one = (V)(_Float16)1.0f;
    const V w = __builtin_elementwise_fma(v, v, -one);
    V p = (V)(_Float16)0.022354126f;
    p = __builtin_elementwise_fma(p, w, (V)(_Float16)-0.0514831543f);
    p = __builtin_elementwise_fma(p, w, (V)(_Float16)0.0627441406f);
    p = __builtin_elementwise_fma(p, w, (V)(_Float16)-0.0956420898f);
    p = __builtin_elementwise_fma(p, w, (V)(_Float16)0.157470703f);
    p = __builtin_elementwise_fma(p, w, (V)(_Float16)-0.239624023f);
    p = __builtin_elementwise_fma(p, w, (V)(_Float16)0.497802734f);
    const V sg = __builtin_elementwise_fma(v, p, (V)(_Float16)0.5f);
    return __builtin_elementwise_min(__builtin_elementwise_max(sg, (V)(_Float16)0.0f), one);
}
template <typename V>
__device__ __forceinline__ V gelu_sig_v(V x, V q) {
    return gelu_sig_scaled<V>(x * q);
}
__device__ __forceinline__ h2 gelu_sig_h2(h2 x, h2 q) { return gelu_sig_v<h2>(x, q); }
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// Layouts prepared for the screen (one wavefront scores 64 candidates):
//   w1f  [16 m-tiles][64 lanes] x 4 bf16: A fragments of W1^T for
//        v_mfma_f32_16x16x16_bf16 (lane l: hidden 16 m + l%16, genes 4(l/16)..+3)
//   w2f  [8 k-blocks][16 n-tiles][64 lanes] x 8: B fragments of W2 for
//        v_mfma_f32_16x16x32_bf16 (and, as 8 E4M3 bytes, of the FP8 screen's
//        v_mfma_scale_f32_16x16x128_f8f6f4, four k-blocks per instruction) in
//        the permuted K order that matches the layer-1 accumulator layout:
//        slot e of lane group q in k-block kb is hidden unit
//        32 kb + (e < 4 ? 4 q + e : 16 + 4 q + e - 4)
__host__ __device__ constexpr int w2_k_of(int kb, int q, int e) {
    return 32 * kb + (e < 4 ? 4 * q + e : 16 + 4 * q + (e - 4));
}

struct MlpScreenW {
    const bf16x4 *w1f;     // 16 * 64
    const bf16x4 *w1fh;    // 16 * 64, times h_scale (fp8 screen)
    const bf16x8 *w2f;     // 8 * 16 * 64 (bf16)
    const uint64_t *w2f8;  // 8 * 16 * 64 (fp8 x 8): GELU_Q W2 in per-block E8M0 scale units
    const uint32_t *w2s8;  // [8 n-tile pairs][64 lanes]: the E8M0 bytes of those blocks
    const float *b1, *b2, *w3;
    float b3;
    float h_scale;   // fp8: power of two; W1, b1 (so H1) are multiplied by it
    int h_e8m0;      // fp8: E8M0 byte of 1 / h_scale, the layer-2 MFMA's H1 scale
    float q1;        // v = x q1 in the layer-1 GELU: GELU_Q / h_scale (fp8) or GELU_Q
    // FP64 scorer: W1 n-major [256][16], W2 as v_mfma_f64_16x16x4f64 A
    // fragments [n-step 64][m-tile 16][lane 64], b1 | b2 | w3, b3
    const double *w1d, *w2d, *vd;
    double b3d;
};

}  // namespace csf
