// mlp_device.hip -- training of the learned surrogate 16 -> 256 -> 256 -> 1
// with hand-written BF16 MFMA kernels (FP32 accumulation, FP32 master
// weights, AdamW).  No counterpart in the reference.
//
// Forward   Z1 = X W1 + b1, H1 = gelu(Z1), Z2 = H1 W2 + b2, H2 = gelu(Z2),
//           out = H2 w3 + b3;  loss = mean (out - y)^2
// Backward  dout = 2 (out - y) / B
//           dw3 = H2^T dout, db3 = sum dout, dZ2 = (dout w3^T) * gelu'(Z2)
//           dW2 = H1^T dZ2, db2 = colsum dZ2, dH1 = dZ2 W2^T
//           dZ1 = dH1 * gelu'(Z1), dW1 = X^T dZ1, db1 = colsum dZ1
// GEMMs run on v_mfma_f32_16x16x32_bf16 with BF16 operands (X, W, H, dZ) and
// FP32 accumulation; elementwise work and the optimizer are FP32.
#include <cmath>
#include <cstring>
#include <vector>

#include "device_ctx.h"
#include "mlp_common.h"

namespace csf {

// ------------------------------------------------------------------ GEMM

constexpr int KSPLIT = 16;  // K layers of the weight-gradient GEMMs

struct GemmArgs {
    int M, N, K;
    const __bf16 *A;
    long lda_m, lda_k;
    const __bf16 *B;
    long ldb_k, ldb_n;
    const float *bias;  // [N] or null
    float *Z;           // pre-activation (f32) or null
    float *C;           // output (f32) or null
    __bf16 *Cb;         // output (bf16) or null
    int ldc;
    int act;  // 0 identity, 1 gelu
    int kchunk;         // split-K: K per workgroup layer (blockIdx.z), 0 = all of K
    long split_stride;  // split-K: elements between the partial outputs of consecutive layers
};

// 64 x 64 output tile per workgroup, 2 x 2 waves of 32 x 32 (2 x 2 MFMA tiles);
// K advances 32 per step through LDS (B stored transposed, [n][k]).  With
// kchunk > 0 the workgroups of layer z = blockIdx.z cover K range
// [z kchunk, (z + 1) kchunk) and write their partial tile to C + z split_stride
// (the weight gradients, whose K is the batch and whose output is one or a few
// tiles, would otherwise run on a handful of workgroups).
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __bf16 As[64][40];
    __shared__ __bf16 Bs[64][40];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int tm = blockIdx.y * 64, tn = blockIdx.x * 64;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kbeg = g.kchunk ? (int)blockIdx.z * g.kchunk : 0;
    const int kend = g.kchunk ? min(g.K, kbeg + g.kchunk) : g.K;
    for (int k0 = kbeg; k0 < kend; k0 += 32) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid * 8 + e, r = idx >> 5, kk = idx & 31;
            const int gm = tm + r, gn = tn + r, gk = k0 + kk;
            As[r][kk] = (gm < g.M && gk < kend) ? g.A[(long)gm * g.lda_m + (long)gk * g.lda_k] : (__bf16)0.0f;
            Bs[r][kk] = (gn < g.N && gk < kend) ? g.B[(long)gk * g.ldb_k + (long)gn * g.ldb_n] : (__bf16)0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + mi * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + ni * 16 + (lane & 15)][8 * (lane >> 4)]);
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[mi][ni], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = tm + wm * 32 + mi * 16 + 4 * (lane >> 4) + r;
                const int col = tn + wn * 32 + ni * 16 + (lane & 15);
                if (row < g.M && col < g.N) {
                    float z = acc[mi][ni][r] + (g.bias ? g.bias[col] : 0.0f);
                    const long o = (long)row * g.ldc + col + (long)blockIdx.z * g.split_stride;
                    if (g.Z) g.Z[o] = z;
                    const float h = g.act ? gelu_f(z) : z;
                    if (g.C) g.C[o] = h;
                    if (g.Cb) g.Cb[o] = (__bf16)h;
                }
            }
}

// C[row][col] = sum over z = 0..nz-1 of part[z][row][col] (fixed order)
__global__ void k_sum_splits(const float *__restrict__ part, int nz, int M, int N, int ldc, long stride,
                             float *__restrict__ C) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)M * N) return;
    const long o = (i / N) * ldc + (i % N);
    float s = 0.0f;
    for (int z = 0; z < nz; ++z) s += part[z * stride + o];
    C[o] = s;
}

// ------------------------------------------------------------ elementwise

__global__ void k_pack_x(const float *__restrict__ genes, int ldg, int n, __bf16 *__restrict__ xb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * MLP_IN) xb[i] = (__bf16)genes[(long)(i / MLP_IN) * ldg + (i % MLP_IN)];
}

// out = H2 w3 + b3; with targets: dout = 2 (out - y) / B and per-block sums of (out - y)^2
__global__ void k_head(const __bf16 *__restrict__ h2, const float *__restrict__ p, int n, const float *__restrict__ y,
                       float *__restrict__ out, float *__restrict__ dout, float *__restrict__ loss_part) {
    __shared__ float red[256];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float r2 = 0.0f;
    if (b < n) {
        float s = 0.0f;
        for (int k = 0; k < MLP_H; ++k) s += (float)h2[(long)b * MLP_H + k] * p[P_W3 + k];
        s += p[P_B3];
        out[b] = s;
        if (y) {
            const float r = s - y[b];
            dout[b] = 2.0f * r / (float)n;
            r2 = r * r;
        }
    }
    red[threadIdx.x] = r2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss_part) loss_part[blockIdx.x] = red[0];
}

// dZ = dH * gelu'(Z); for layer 2 dH = dout w3^T
__global__ void k_gelu_back(const float *__restrict__ z, const float *__restrict__ dh, const float *__restrict__ dout,
                            const float *__restrict__ w3, int n, float *__restrict__ dzf, __bf16 *__restrict__ dzb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * MLP_H) return;
    const int b = (int)(i / MLP_H), k = (int)(i % MLP_H);
    const float up = dh ? dh[i] : dout[b] * w3[k];
    const float d = up * gelu_grad_f(z[i]);
    dzf[i] = d;
    dzb[i] = (__bf16)d;
}

// column sums of an [n][256] f32 matrix (one workgroup per column)
__global__ void k_colsum(const float *__restrict__ src, int n, float *__restrict__ dst) {
    __shared__ float red[256];
    const int col = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < n; b += blockDim.x) s += src[(long)b * MLP_H + col];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[col] = red[0];
}

// dw3[k] = sum_b H2[b][k] dout[b] (workgroup per k); block 256 computes db3
__global__ void k_dw3(const __bf16 *__restrict__ h2, const float *__restrict__ dout, int n, float *__restrict__ g) {
    __shared__ float red[256];
    const int k = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < n; b += blockDim.x) s += (k < MLP_H ? (float)h2[(long)b * MLP_H + k] : 1.0f) * dout[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) g[k < MLP_H ? P_W3 + k : P_B3] = red[0];
}

// torch.optim.AdamW (decoupled decay, bias corrections as in torch 2.x)
__global__ void k_adamw(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                        float *__restrict__ v, int n, float lr, float wd, float b1, float b2, float eps,
                        float step_size, float bc2_sqrt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float w = p[i] * (1.0f - lr * wd);
    const float gi = g[i];
    const float mi = m[i] + (gi - m[i]) * (1.0f - b1);
    const float vi = v[i] * b2 + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float denom = sqrtf(vi) / bc2_sqrt + eps;
    p[i] = w - step_size * (mi / denom);
}

// BF16 copies for training and the fragment layouts for the screen
__global__ void k_refresh(const float *__restrict__ p, __bf16 *__restrict__ w1b, __bf16 *__restrict__ w2b,
                          bf16x4 *__restrict__ w1f, bf16x4 *__restrict__ w1fh, bf16x8 *__restrict__ w2f,
                          uint64_t *__restrict__ w2f8, uint32_t *__restrict__ w2s8, float h_scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < MLP_IN * MLP_H) w1b[i] = (__bf16)p[P_W1 + i];
    if (i < MLP_H * MLP_H) w2b[i] = (__bf16)p[P_W2 + i];
    if (i < 16 * 64) {  // w1f[mt][lane]: hidden 16 mt + lane%16, genes 4 (lane/16) + e
        const int mt = i / 64, lane = i % 64, h = 16 * mt + (lane & 15), k0 = 4 * (lane >> 4);
        bf16x4 v, vh;
        for (int e = 0; e < 4; ++e) {
            const float w1 = p[P_W1 + (k0 + e) * MLP_H + h];
            v[e] = (__bf16)(w1 * GELU_Q);  // the BF16 screen's layer-1 GELU input scale, folded in
            vh[e] = (__bf16)((float)(__bf16)w1 * h_scale);  // exact: h_scale is a power of two
        }
        w1f[i] = v;
        w1fh[i] = vh;
    }
    if (i < 8 * 16 * 64) {  // w2f[kb][nt][lane]: out unit 16 nt + lane%16, permuted k slots
        const int kb = i / 1024, nt = (i / 64) % 16, lane = i % 64, q = lane >> 4;
        const int n = 16 * nt + (lane & 15);
        bf16x8 v;
        for (int e = 0; e < 8; ++e) v[e] = (__bf16)p[P_W2 + w2_k_of(kb, q, e) * MLP_H + n];
        w2f[i] = v;
    }
    if (i < 2 * 16 * 64) {
        // FP8 screen: GELU_Q W2 (the layer-2 GELU input scale, which the BF16
        // screen's H1 carries instead) in E4M3 with one E8M0 scale per block
        // of 32 k values of an output unit.  In v_mfma_scale_f32_16x16x128_f8f6f4 the
        // A bytes of lane group g sit at k = 16 g .. 16 g + 15 (its first four
        // VGPRs) and 64 + 16 g .. (its last four), so scale block b (k = 32 b
        // .. 32 b + 31, scale byte from lane 16 b + row) holds the 8-byte
        // pieces j of the k-quad kq with j >> 1 = b >> 1 (pieces j = 0, 1 are
        // the first four VGPRs) of lane groups q with q >> 1 = b & 1
        // (measured on the device with one nonzero VGPR of one lane group at
        // a time).  Thread (kq, nt, 16 b + r) scales block b of out unit
        // 16 nt + r: the block holds its values divided by 2^e, e the smallest
        // exponent that keeps the largest magnitude within E4M3's 448, and the
        // lane's scale byte is e + 127.  w2s8[np][lane] packs the bytes of the
        // four instructions of n-tile pair np: byte 2 kq + (nt & 1).
        const int kq = i / 1024, nt = (i / 64) % 16, lane = i % 64, b = lane >> 4, r = lane & 15;
        const int n = 16 * nt + r;
        float w[32];  // [q - q0][j - j0][slot]
        float mx = 0.0f;
        const int q0 = 2 * (b & 1), j0 = 2 * (b >> 1);
        for (int u = 0; u < 32; ++u) {
            const int q = q0 + (u >> 4), j = j0 + ((u >> 3) & 1), e = u & 7;
            w[u] = GELU_Q * p[P_W2 + w2_k_of(4 * kq + j, q, e) * MLP_H + n];
            mx = fmaxf(mx, fabsf(w[u]));
        }
        int e = -127;
        if (mx > 0.0f) {
            int ex;
            (void)frexpf(mx / 448.0f, &ex);  // mx / 448 = f 2^ex, f in [0.5, 1)
            e = ex;
            if (ldexpf(mx, -(e - 1)) <= 448.0f) e -= 1;
            e = max(-127, min(127, e));
        }
        for (int pc = 0; pc < 4; ++pc) {  // the block's four 8-byte pieces
            const int q = q0 + (pc >> 1), j = j0 + (pc & 1);
            uint64_t f8 = 0;
            for (int e8 = 0; e8 < 8; ++e8) {
                const float x = ldexpf(w[8 * pc + e8], -e);
                const uint32_t pk = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false);
                f8 |= (uint64_t)(pk & 0xffu) << (8 * e8);
            }
            w2f8[((4 * kq + j) * 16 + nt) * 64 + 16 * q + r] = f8;
        }
        reinterpret_cast<uint8_t *>(w2s8)[((nt >> 1) * 64 + lane) * 4 + 2 * kq + (nt & 1)] = (uint8_t)(e + 127);
    }
}

}  // namespace csf

namespace csf {

static int mlp_alloc(cs_mlp *m, int n) {
    if (n <= m->cap) return CS_OK;
    int rc;
    const size_t B = (size_t)n;
    if ((rc = ensure(m->xb, B * MLP_IN * 2)) || (rc = ensure(m->z1, B * MLP_H * 4)) ||
        (rc = ensure(m->h1b, B * MLP_H * 2)) || (rc = ensure(m->z2, B * MLP_H * 4)) ||
        (rc = ensure(m->h2b, B * MLP_H * 2)) || (rc = ensure(m->out, B * 4)) || (rc = ensure(m->dout, B * 4)) ||
        (rc = ensure(m->dz2f, B * MLP_H * 4)) || (rc = ensure(m->dz2b, B * MLP_H * 2)) ||
        (rc = ensure(m->dh1, B * MLP_H * 4)) || (rc = ensure(m->dz1f, B * MLP_H * 4)) ||
        (rc = ensure(m->dz1b, B * MLP_H * 2)) || (rc = ensure(m->loss_part, (B / 256 + 2) * 4)) ||
        (rc = ensure(m->gws, (size_t)KSPLIT * MLP_H * MLP_H * 4)))
        return rc;
    m->cap = n;
    return CS_OK;
}

static void launch_gemm(hipStream_t s, int M, int N, int K, const __bf16 *A, long lam, long lak, const __bf16 *B,
                        long lbk, long lbn, const float *bias, float *Z, float *C, __bf16 *Cb, int act) {
    GemmArgs g{M, N, K, A, lam, lak, B, lbk, lbn, bias, Z, C, Cb, MLP_H, act, 0, 0};
    hipLaunchKernelGGL(k_gemm, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, g);
}

// weight gradient C = A^T B over the batch (K = n): KSPLIT layers of K into
// the workspace, then a fixed-order sum into C (row pitch MLP_H)
static void launch_gemm_wgrad(hipStream_t s, int M, int N, int K, const __bf16 *A, long lak, const __bf16 *B,
                              long lbk, float *ws, float *C) {
    const int kchunk = ((K + KSPLIT * 32 - 1) / (KSPLIT * 32)) * 32;
    const long stride = (long)M * MLP_H;
    GemmArgs g{M, N, K, A, 1, lak, B, lbk, 1, nullptr, nullptr, ws, nullptr, MLP_H, 0, kchunk, stride};
    hipLaunchKernelGGL(k_gemm, dim3((N + 63) / 64, (M + 63) / 64, KSPLIT), dim3(256), 0, s, g);
    hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)(((long)M * N + 255) / 256)), dim3(256), 0, s, (const float *)ws,
                       KSPLIT, M, N, MLP_H, stride, C);
}

static int refresh(cs_mlp *m) {
    cs_ctx *ctx = m->ctx;
    hipLaunchKernelGGL(k_refresh, dim3(MLP_H * MLP_H / 256), dim3(256), 0, ctx->stream, (const float *)m->params.p,
                       (__bf16 *)m->w1b.p, (__bf16 *)m->w2b.p, (bf16x4 *)m->w1f.p, (bf16x4 *)m->w1fh.p,
                       (bf16x8 *)m->w2f.p, (uint64_t *)m->w2f8.p, (uint32_t *)m->w2s8.p, m->h_scale);
    HIPCHK(hipGetLastError());
    return CS_OK;
}

// forward on n <= cap rows; y != null also computes dout and the loss partials
static int forward(cs_mlp *m, const float *genes, int ldg, int n, const float *y) {
    cs_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const float *p = (const float *)m->params.p;
    hipLaunchKernelGGL(k_pack_x, dim3((n * MLP_IN + 255) / 256), dim3(256), 0, s, genes, ldg, n, (__bf16 *)m->xb.p);
    launch_gemm(s, n, MLP_H, MLP_IN, (const __bf16 *)m->xb.p, MLP_IN, 1, (const __bf16 *)m->w1b.p, MLP_H, 1,
                p + P_B1, (float *)m->z1.p, nullptr, (__bf16 *)m->h1b.p, 1);
    launch_gemm(s, n, MLP_H, MLP_H, (const __bf16 *)m->h1b.p, MLP_H, 1, (const __bf16 *)m->w2b.p, MLP_H, 1,
                p + P_B2, (float *)m->z2.p, nullptr, (__bf16 *)m->h2b.p, 1);
    hipLaunchKernelGGL(k_head, dim3((n + 255) / 256), dim3(256), 0, s, (const __bf16 *)m->h2b.p, p, n, y,
                       (float *)m->out.p, (float *)m->dout.p, (float *)m->loss_part.p);
    HIPCHK(hipGetLastError());
    return CS_OK;
}

// gradients of the mean squared error into m->grads (needs a forward with y)
static int backward(cs_mlp *m, int n) {
    cs_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    float *g = (float *)m->grads.p;
    const float *p = (const float *)m->params.p;
    const long nh = (long)n * MLP_H;
    hipLaunchKernelGGL(k_dw3, dim3(MLP_H + 1), dim3(256), 0, s, (const __bf16 *)m->h2b.p, (const float *)m->dout.p,
                       n, g);
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, (const float *)m->z2.p,
                       (const float *)nullptr, (const float *)m->dout.p, p + P_W3, n, (float *)m->dz2f.p,
                       (__bf16 *)m->dz2b.p);
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, s, (const float *)m->dz2f.p, n, g + P_B2);
    // dW2 = H1^T dZ2
    launch_gemm_wgrad(s, MLP_H, MLP_H, n, (const __bf16 *)m->h1b.p, MLP_H, (const __bf16 *)m->dz2b.p, MLP_H,
                      (float *)m->gws.p, g + P_W2);
    // dH1 = dZ2 W2^T
    launch_gemm(s, n, MLP_H, MLP_H, (const __bf16 *)m->dz2b.p, MLP_H, 1, (const __bf16 *)m->w2b.p, 1, MLP_H, nullptr,
                nullptr, (float *)m->dh1.p, nullptr, 0);
    hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, (const float *)m->z1.p,
                       (const float *)m->dh1.p, (const float *)nullptr, (const float *)nullptr, n,
                       (float *)m->dz1f.p, (__bf16 *)m->dz1b.p);
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, s, (const float *)m->dz1f.p, n, g + P_B1);
    // dW1 = X^T dZ1
    launch_gemm_wgrad(s, MLP_IN, MLP_H, n, (const __bf16 *)m->xb.p, MLP_IN, (const __bf16 *)m->dz1b.p, MLP_H,
                      (float *)m->gws.p, g + P_W1);
    HIPCHK(hipGetLastError());
    return CS_OK;
}

static int read_loss(cs_mlp *m, int n, float *loss) {
    const int nb = (n + 255) / 256;
    std::vector<float> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), m->loss_part.p, 4 * nb, hipMemcpyDeviceToHost, m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    double s = 0;
    for (float x : part) s += x;
    *loss = (float)(s / n);
    return CS_OK;
}

// torch.nn.Linear's default initialization: U(-1/sqrt(fan_in), 1/sqrt(fan_in))
// for weights and biases, drawn here from splitmix64(seed).
static void init_params(uint64_t seed, std::vector<float> &p) {
    p.assign(P_TOTAL, 0.0f);
    uint64_t s = seed;
    auto uni = [&](float bound) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * 0x1p-53;
        return (float)((2.0 * u - 1.0) * bound);
    };
    const float b1 = 1.0f / std::sqrt((float)MLP_IN), b2 = 1.0f / std::sqrt((float)MLP_H);
    for (int i = 0; i < MLP_IN * MLP_H; ++i) p[P_W1 + i] = uni(b1);
    for (int i = 0; i < MLP_H; ++i) p[P_B1 + i] = uni(b1);
    for (int i = 0; i < MLP_H * MLP_H; ++i) p[P_W2 + i] = uni(b2);
    for (int i = 0; i < MLP_H; ++i) p[P_B2 + i] = uni(b2);
    for (int i = 0; i < MLP_H; ++i) p[P_W3 + i] = uni(b2);
    p[P_B3] = uni(b2);
}

}  // namespace csf

using namespace csf;

extern "C" {

namespace {
int mlp_setup(cs_mlp *m, uint64_t seed) {
    cs_ctx *ctx = m->ctx;
    int rc;
    if ((rc = ensure(m->params, 4 * P_TOTAL)) || (rc = ensure(m->grads, 4 * P_TOTAL)) ||
        (rc = ensure(m->m, 4 * P_TOTAL)) || (rc = ensure(m->v, 4 * P_TOTAL)) ||
        (rc = ensure(m->w1b, 2 * MLP_IN * MLP_H)) || (rc = ensure(m->w2b, 2 * MLP_H * MLP_H)) ||
        (rc = ensure(m->w1f, 8 * 16 * 64)) || (rc = ensure(m->w1fh, 8 * 16 * 64)) ||
        (rc = ensure(m->w2f, 16 * 8 * 16 * 64)) ||
        (rc = ensure(m->w2f8, 8 * 8 * 16 * 64)) || (rc = ensure(m->w2s8, 4 * 8 * 64)))
        return rc;
    std::vector<float> p;
    init_params(seed, p);
    HIPCHK(hipMemcpyAsync(m->params.p, p.data(), 4 * P_TOTAL, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(m->m.p, 0, 4 * P_TOTAL, ctx->stream));
    HIPCHK(hipMemsetAsync(m->v.p, 0, 4 * P_TOTAL, ctx->stream));
    HIPCHK(hipMemsetAsync(m->grads.p, 0, 4 * P_TOTAL, ctx->stream));
    if ((rc = refresh(m))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CS_OK;
}
}  // namespace

int cs_mlp_create(cs_ctx *ctx, uint64_t seed, cs_mlp **out) {
    if (!ctx || !out) return fail(CS_E_BADARGS, "null argument");
    HIPCHK(hipSetDevice(ctx->device));
    cs_mlp *m = new cs_mlp();
    m->ctx = ctx;
    ctx->n_mlp++;
    if (int rc = mlp_setup(m, seed)) {
        cs_mlp_destroy(m);  // keeps the setup error: destroy does not touch it
        return rc;
    }
    *out = m;
    return CS_OK;
}

int cs_mlp_destroy(cs_mlp *m) {
    if (!m) return CS_OK;
    (void)hipStreamSynchronize(m->ctx->stream);
    csf::DevBuf *bufs[] = {&m->params, &m->grads, &m->m, &m->v, &m->w1b, &m->w2b, &m->w1f, &m->w1fh, &m->w2f, &m->w2f8, &m->w2s8, &m->f64pack,
                           &m->xb, &m->z1, &m->h1b, &m->z2, &m->h2b, &m->out, &m->dout, &m->dz2f, &m->dz2b,
                           &m->dh1, &m->dz1f, &m->dz1b, &m->loss_part, &m->gws};
    for (auto *b : bufs)
        if (b->p) (void)hipFree(b->p);
    m->ctx->n_mlp--;
    delete m;
    return CS_OK;
}

int cs_mlp_get_params(cs_mlp *m, float *host) {
    if (!m || !host) return fail(CS_E_BADARGS, "null argument");
    HIPCHK(hipMemcpyAsync(host, m->params.p, 4 * P_TOTAL, hipMemcpyDeviceToHost, m->ctx->stream));
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    return CS_OK;
}

int cs_mlp_set_params(cs_mlp *m, const float *host) {
    if (!m || !host) return fail(CS_E_BADARGS, "null argument");
    HIPCHK(hipMemcpyAsync(m->params.p, host, 4 * P_TOTAL, hipMemcpyHostToDevice, m->ctx->stream));
    int rc;
    if ((rc = refresh(m))) return rc;
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    return CS_OK;
}

int cs_mlp_set_fp8_scale(cs_mlp *m, float h_scale) {
    int eh = 0;
    if (!m || !(h_scale > 0) || std::frexp(h_scale, &eh) != 0.5f || eh < -14 || eh > 16)
        return fail(CS_E_BADARGS, "the fp8 H1 scale must be a power of two in [2^-15, 2^15]");
    m->h_scale = h_scale;
    int rc;
    if ((rc = refresh(m))) return rc;
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    return CS_OK;
}

int cs_mlp_forward(cs_mlp *m, const float *genes, int ldg, uint64_t n, float *out, float *h1max) {
    if (!m || !genes || !out || ldg < MLP_IN) return fail(CS_E_BADARGS, "bad argument");
    HIPCHK(hipSetDevice(m->ctx->device));
    const int chunk = 65536;
    int rc;
    if ((rc = mlp_alloc(m, chunk))) return rc;
    float hmax = 0.0f;
    std::vector<uint16_t> hb;
    for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
        const int c = (int)std::min<uint64_t>(chunk, n - r0);
        if ((rc = forward(m, genes + r0 * ldg, ldg, c, nullptr))) return rc;
        HIPCHK(hipMemcpyAsync(out + r0, m->out.p, 4 * (size_t)c, hipMemcpyDeviceToDevice, m->ctx->stream));
        if (h1max) {
            hb.resize((size_t)c * MLP_H);
            HIPCHK(hipMemcpyAsync(hb.data(), m->h1b.p, 2 * hb.size(), hipMemcpyDeviceToHost, m->ctx->stream));
            HIPCHK(hipStreamSynchronize(m->ctx->stream));
            for (uint16_t v : hb) {
                uint32_t bits = (uint32_t)v << 16;
                float f;
                std::memcpy(&f, &bits, 4);
                hmax = std::max(hmax, std::fabs(f));
            }
        }
    }
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    if (h1max) *h1max = hmax;
    return CS_OK;
}

int cs_mlp_grad(cs_mlp *m, const float *genes, int ldg, const float *y, int n, float *loss, float *grads_host) {
    if (!m || !genes || !y || !loss || n < 1 || ldg < MLP_IN) return fail(CS_E_BADARGS, "bad argument");
    HIPCHK(hipSetDevice(m->ctx->device));
    int rc;
    if ((rc = mlp_alloc(m, n)) || (rc = forward(m, genes, ldg, n, y)) || (rc = backward(m, n)) ||
        (rc = read_loss(m, n, loss)))
        return rc;
    if (grads_host) {
        HIPCHK(hipMemcpyAsync(grads_host, m->grads.p, 4 * P_TOTAL, hipMemcpyDeviceToHost, m->ctx->stream));
        HIPCHK(hipStreamSynchronize(m->ctx->stream));
    }
    return CS_OK;
}

int cs_mlp_train_step(cs_mlp *m, const float *genes, int ldg, const float *y, int n, float lr, float wd,
                      float *loss) {
    if (!m || !genes || !y || !loss || n < 1 || ldg < MLP_IN) return fail(CS_E_BADARGS, "bad argument");
    HIPCHK(hipSetDevice(m->ctx->device));
    int rc;
    if ((rc = mlp_alloc(m, n)) || (rc = forward(m, genes, ldg, n, y)) || (rc = backward(m, n))) return rc;
    m->step += 1;
    const double b1 = 0.9, b2 = 0.999;
    const double bc1 = 1.0 - std::pow(b1, (double)m->step), bc2 = 1.0 - std::pow(b2, (double)m->step);
    hipLaunchKernelGGL(k_adamw, dim3((P_TOTAL + 255) / 256), dim3(256), 0, m->ctx->stream, (float *)m->params.p,
                       (const float *)m->grads.p, (float *)m->m.p, (float *)m->v.p, P_TOTAL, lr, wd, (float)b1,
                       (float)b2, 1e-8f, (float)(lr / bc1), (float)std::sqrt(bc2));
    HIPCHK(hipGetLastError());
    if ((rc = refresh(m))) return rc;
    return read_loss(m, n, loss);
}

}  // extern "C"
