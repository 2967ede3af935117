// ridge_device.hip -- FP64 ridge refit of the screen's quadratic surface on
// gfx950, and the loss gradient used by the differentiable surrogate.
//
// Reference (codec-arena snapshot):
//   features        brain/evolve.py:211-220  [1, g_0..g_{ng-1}, g_i g_j (i<16, j<=i)]
//   refit_surrogate brain/evolve.py:250-268  w = solve(X'X + a I, X'y + a w_prior)
// With ng = 16 the feature vector is the 153-term layout of Model.fromCompact
// (src/screen.zig:36-95); with ng = 29 it is the 166-term layout the brain
// refits every fourth generation.
//
// Kernels:
//   k_gram        X'X, X'y and y'y in one pass: features are built per row in
//                 LDS (one multiply per column from the extended row
//                 [g, y, 1, 0]; the next rows are fetched into registers
//                 meanwhile), [phi | y] (padded to 176 columns = 11 tiles of
//                 16) is multiplied with v_mfma_f64_16x16x4f64 into the 66
//                 upper tiles; each workgroup writes its partial tiles.
//   k_gram_reduce fixed-order sum of the partial tiles (deterministic).
//   k_chol        Cholesky of X'X + a I in LDS (packed lower triangle) and the
//                 two triangular solves, one workgroup.
//   k_ridge_grad  residuals r = phi.w - y and the partials of sum r^2 and
//                 X'r, per workgroup; k_ridge_grad_reduce finishes
//                 L = (|Xw - y|^2 + a |w - w_p|^2) / n and dL/dw.
#include "device_ctx.h"

namespace csf {

constexpr int RP = 176;  // padded columns of [phi | y]
constexpr int RT = RP / 16;
constexpr int NTILE = RT * (RT + 1) / 2;  // 66 upper-triangle tiles
constexpr int TPW = (NTILE + 3) / 4;      // tiles per wave (17)
constexpr int GSB = 32;                   // rows per LDS sub-batch
constexpr int GCH = 512;                  // rows per workgroup

typedef double dbl4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int tile_i(int t) {
    int i = 0;
    while (t >= RT - i) {
        t -= RT - i;
        ++i;
    }
    return i;
}
__host__ __device__ constexpr int tile_j(int t) {
    int i = 0;
    while (t >= RT - i) {
        t -= RT - i;
        ++i;
    }
    return i + t;
}

// Column e of [phi | y] as (kind, a, b): 0 -> 0, 1 -> 1, 2 -> g[a], 3 -> g[a] g[b], 4 -> y.
__device__ __forceinline__ uint32_t feature_code(int e, int ng, int p) {
    if (e == 0) return 1u;
    if (e <= ng) return 2u | ((uint32_t)(e - 1) << 8);
    if (e < p) {
        int q = e - 1 - ng, i = 0;
        while (q > i) {
            q -= i + 1;
            ++i;
        }
        return 3u | ((uint32_t)i << 8) | ((uint32_t)q << 16);
    }
    if (e == p) return 4u;
    return 0u;
}

__device__ __forceinline__ double feature_value(uint32_t code, const double *g, double y) {
    const uint32_t kind = code & 255u, a = (code >> 8) & 255u, b = code >> 16;
    switch (kind) {
    case 1: return 1.0;
    case 2: return g[a];
    case 3: return g[a] * g[b];
    case 4: return y;
    default: return 0.0;
    }
}

template <int W>
__device__ __forceinline__ void gram_mma(dbl4 (&acc)[TPW], const double (*phi)[RP], int lane) {
#pragma unroll
    for (int kk = 0; kk < GSB / 4; ++kk) {
        double v[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) v[t] = phi[kk * 4 + (lane >> 4)][t * 16 + (lane & 15)];
#pragma unroll
        for (int q = 0; q < TPW; ++q) {
            const int T = 4 * q + W;
            if (T < NTILE) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[tile_i(T)], v[tile_j(T)], acc[q], 0, 0, 0);
        }
    }
}

template <int W>
__device__ __forceinline__ void gram_store(const dbl4 (&acc)[TPW], double *part, int lane) {
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
        const int T = 4 * q + W;
        if (T < NTILE) {
#pragma unroll
            for (int r = 0; r < 4; ++r) part[(size_t)T * 256 + lane * 4 + r] = acc[q][r];
        }
    }
}

// Column e of [phi | y] as the product of two entries a | b << 8 of the
// extended row [g_0 .. g_{ng-1}, y, 1, 0]: every column is one multiply and
// the values are those of feature_value (x * 1 and 0 * 0 are exact).
__device__ __forceinline__ uint32_t feature_pair(int e, int ng, int p) {
    const uint32_t Y = (uint32_t)ng, ONE = Y + 1u, ZERO = Y + 2u;
    const uint32_t c = feature_code(e, ng, p), kind = c & 255u, a = (c >> 8) & 255u, b = c >> 16;
    switch (kind) {
    case 1: return ONE | (ONE << 8);
    case 2: return a | (ONE << 8);
    case 3: return a | (b << 8);
    case 4: return Y | (ONE << 8);
    default: return ZERO | (ZERO << 8);
    }
}

constexpr int GXW = 32;                          // extended row width (ng + 3 <= 32)
constexpr int GPF = (GSB * GXW + 255) / 256;     // prefetched values per thread

__global__ __launch_bounds__(256) void k_gram(const double *__restrict__ genes, int ng, const double *__restrict__ y,
                                              uint64_t n, int p, double *__restrict__ part) {
    __shared__ double phi[GSB][RP];
    __shared__ uint32_t pair[RP];
    __shared__ double xrow[GSB][GXW];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int e = tid; e < RP; e += 256) pair[e] = feature_pair(e, ng, p);
    dbl4 acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q) acc[q] = dbl4{0.0, 0.0, 0.0, 0.0};
    const uint64_t r_begin = (uint64_t)blockIdx.x * GCH;
    const uint64_t r_end = min<uint64_t>(n, r_begin + GCH);
    // the next sub-batch's genes, y and the constant 1 travel in registers
    // while the current one is multiplied; rows past the end are all zero
    double pf[GPF];
    auto fetch = [&](uint64_t r0) {
#pragma unroll
        for (int k = 0; k < GPF; ++k) {
            const int idx = tid + 256 * k, row = idx / GXW, col = idx % GXW;
            const uint64_t r = r0 + (uint64_t)row;
            double v = 0.0;
            if (row < GSB && r < r_end) {
                if (col < ng) v = genes[r * (uint64_t)ng + (uint64_t)col];
                else if (col == ng) v = y[r];
                else if (col == ng + 1) v = 1.0;
            }
            pf[k] = v;
        }
    };
    fetch(r_begin);
    for (uint64_t r0 = r_begin; r0 < r_end; r0 += GSB) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < GPF; ++k) {
            const int idx = tid + 256 * k;
            if (idx < GSB * GXW) xrow[idx / GXW][idx % GXW] = pf[k];
        }
        __syncthreads();
        if (r0 + GSB < r_end) fetch(r0 + GSB);
        for (int idx = tid; idx < GSB * RP; idx += 256) {
            const int row = idx / RP, e = idx % RP;
            const uint32_t pr = pair[e];
            phi[row][e] = xrow[row][pr & 255u] * xrow[row][pr >> 8];
        }
        __syncthreads();
        switch (wave) {
        case 0: gram_mma<0>(acc, phi, lane); break;
        case 1: gram_mma<1>(acc, phi, lane); break;
        case 2: gram_mma<2>(acc, phi, lane); break;
        default: gram_mma<3>(acc, phi, lane); break;
        }
    }
    double *mine = part + (size_t)blockIdx.x * NTILE * 256;
    switch (wave) {
    case 0: gram_store<0>(acc, mine, lane); break;
    case 1: gram_store<1>(acc, mine, lane); break;
    case 2: gram_store<2>(acc, mine, lane); break;
    default: gram_store<3>(acc, mine, lane); break;
    }
}

// one thread per tile element; sums the workgroup partials in order.
// v_mfma_f64_16x16x4f64 leaves D[i][j] in accumulator register r of lane l
// with i = 4 r + l / 16 and j = l % 16, so element (lane, r) of tile (ti, tj)
// is entry (16 ti + 4 r + lane / 16, 16 tj + lane % 16).
__global__ void k_gram_reduce(const double *__restrict__ part, int nwg, int p, double *__restrict__ gram,
                              double *__restrict__ xty, double *__restrict__ yy) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= NTILE * 256) return;
    const int T = idx / 256, el = idx % 256, lane = el / 4, r = el % 4;
    double s = 0.0;
    for (int w = 0; w < nwg; ++w) s += part[(size_t)w * NTILE * 256 + idx];
    const int i = tile_i(T) * 16 + 4 * r + (lane >> 4);
    const int j = tile_j(T) * 16 + (lane & 15);
    // diagonal tiles hold both (i, j) and (j, i): take the upper one only, so
    // every output has exactly one writer
    if (i > j) return;
    if (j < p) {
        gram[(size_t)i * p + j] = s;
        gram[(size_t)j * p + i] = s;
    } else if (i < p && j == p) {
        xty[i] = s;
    } else if (i == p && j == p) {
        yy[0] = s;
    }
}

// Solve (gram + alpha I) w = xty + alpha prior by Cholesky.  p <= 176.
// status[0] = 0 on success, k+1 if the pivot of column k was not positive.
__global__ __launch_bounds__(256) void k_chol(const double *__restrict__ gram, const double *__restrict__ xty, int p,
                                              double alpha, const double *__restrict__ prior,
                                              double *__restrict__ w, int *__restrict__ status) {
    extern __shared__ double sm[];
    double *M = sm;                      // packed lower triangle, (i, j<=i) at i(i+1)/2 + j
    double *rhs = sm + p * (p + 1) / 2;  // right-hand side, then forward solution
    __shared__ int bad;
    const int tid = threadIdx.x;
    for (int i = 0; i < p; ++i)
        for (int j = tid; j <= i; j += 256) M[i * (i + 1) / 2 + j] = gram[(size_t)i * p + j] + (i == j ? alpha : 0.0);
    for (int i = tid; i < p; i += 256) rhs[i] = xty[i] + (prior ? alpha * prior[i] : 0.0);
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < p; ++k) {
        if (tid == 0) {
            const double d = M[k * (k + 1) / 2 + k];
            if (!(d > 0.0) || !isfinite(d)) bad = k + 1;
            else M[k * (k + 1) / 2 + k] = sqrt(d);
        }
        __syncthreads();
        if (bad) break;
        const double dk = M[k * (k + 1) / 2 + k];
        for (int i = k + 1 + tid; i < p; i += 256) M[i * (i + 1) / 2 + k] /= dk;
        __syncthreads();
        for (int i = k + 1; i < p; ++i) {
            const double lik = M[i * (i + 1) / 2 + k];
            for (int j = k + 1 + tid; j <= i; j += 256) M[i * (i + 1) / 2 + j] -= lik * M[j * (j + 1) / 2 + k];
        }
        __syncthreads();
    }
    if (bad) {
        if (tid == 0) status[0] = bad;
        return;
    }
    for (int j = 0; j < p; ++j) {  // L z = rhs
        if (tid == 0) rhs[j] /= M[j * (j + 1) / 2 + j];
        __syncthreads();
        const double zj = rhs[j];
        for (int i = j + 1 + tid; i < p; i += 256) rhs[i] -= M[i * (i + 1) / 2 + j] * zj;
        __syncthreads();
    }
    for (int j = p - 1; j >= 0; --j) {  // L' w = z
        if (tid == 0) rhs[j] /= M[j * (j + 1) / 2 + j];
        __syncthreads();
        const double wj = rhs[j];
        for (int i = tid; i < j; i += 256) rhs[i] -= M[j * (j + 1) / 2 + i] * wj;
        __syncthreads();
    }
    for (int i = tid; i < p; i += 256) w[i] = rhs[i];
    if (tid == 0) status[0] = 0;
}

constexpr int GRAD_ROWS = 256;

// Per workgroup: partial sum r^2 and partial X'r over its rows (grid-stride
// over 256-row chunks; the partition is fixed by the grid size, so the
// result is deterministic).
__global__ __launch_bounds__(256) void k_ridge_grad(const double *__restrict__ genes, int ng,
                                                    const double *__restrict__ y, uint64_t n, int p,
                                                    const double *__restrict__ w, double *__restrict__ part) {
    __shared__ double wl[RP];
    __shared__ uint32_t code[RP];
    __shared__ double gl[GRAD_ROWS][30];
    __shared__ double rl[GRAD_ROWS];
    __shared__ double red[256];
    const int tid = threadIdx.x;
    for (int e = tid; e < RP; e += 256) {
        code[e] = feature_code(e, ng, p);
        wl[e] = e < p ? w[e] : 0.0;
    }
    double acc = 0.0, r2 = 0.0;
    for (uint64_t r0 = (uint64_t)blockIdx.x * GRAD_ROWS; r0 < n; r0 += (uint64_t)gridDim.x * GRAD_ROWS) {
        __syncthreads();
        const uint64_t row = r0 + tid;
        const bool live = row < n;
        for (int k = 0; k < ng; ++k) gl[tid][k] = live ? genes[row * ng + k] : 0.0;
        double dot = 0.0;
        for (int e = 0; e < p; ++e) dot += wl[e] * feature_value(code[e], gl[tid], 0.0);
        const double r = live ? dot - y[row] : 0.0;
        rl[tid] = r;
        r2 += r * r;
        __syncthreads();
        if (tid < p) {
            const uint32_t c = code[tid];
            for (int k = 0; k < GRAD_ROWS; ++k) acc += rl[k] * feature_value(c, gl[k], 0.0);
        }
    }
    double *mine = part + (size_t)blockIdx.x * (RP + 1);
    if (tid < p) mine[tid] = acc;
    red[tid] = r2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) mine[RP] = red[0];
}

__global__ void k_ridge_grad_reduce(const double *__restrict__ part, int nwg, int p, uint64_t n, double alpha,
                                    const double *__restrict__ w, const double *__restrict__ prior,
                                    double *__restrict__ loss, double *__restrict__ grad) {
    __shared__ double pen[256];
    const int tid = threadIdx.x;
    double pn = 0.0;
    for (int f = tid; f < p; f += blockDim.x) {
        double s = 0.0;
        for (int wg = 0; wg < nwg; ++wg) s += part[(size_t)wg * (RP + 1) + f];
        const double d = w[f] - (prior ? prior[f] : 0.0);
        grad[f] = 2.0 * (s + alpha * d) / (double)n;
        pn += d * d;
    }
    pen[tid] = pn;
    __syncthreads();
    if (tid == 0) {
        double sr = 0.0, sp = 0.0;
        for (int wg = 0; wg < nwg; ++wg) sr += part[(size_t)wg * (RP + 1) + RP];
        for (int t = 0; t < (int)blockDim.x; ++t) sp += pen[t];
        loss[0] = (sr + alpha * sp) / (double)n;
    }
}

static int check_ridge_dims(int ng, int *p) {
    if (ng != 16 && ng != 29) return fail(CS_E_BADARGS, "ng must be 16 (153 features) or 29 (166 features)");
    *p = 1 + ng + 136;
    return CS_OK;
}

}  // namespace csf

using namespace csf;

extern "C" {

int cs_ridge_gram_f64(cs_ctx *ctx, const double *genes, int ng, const double *y, uint64_t n, double *gram,
                      double *xty, double *yy) {
    if (!ctx || !genes || !y || !gram || !xty || !yy) return fail(CS_E_BADARGS, "null argument");
    int p, rc;
    if ((rc = check_ridge_dims(ng, &p))) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const uint64_t nwg64 = (n + GCH - 1) / GCH;
    if (nwg64 == 0 || nwg64 > (1u << 20)) return fail(CS_E_BADARGS, "n must be in [1, 2^30]");
    const int nwg = (int)nwg64;
    if ((rc = ensure(ctx->gram_part, sizeof(double) * NTILE * 256 * (size_t)nwg))) return rc;
    hipLaunchKernelGGL(k_gram, dim3(nwg), dim3(256), 0, ctx->stream, genes, ng, y, n, p, (double *)ctx->gram_part.p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_gram_reduce, dim3((NTILE * 256 + 255) / 256), dim3(256), 0, ctx->stream,
                       (const double *)ctx->gram_part.p, nwg, p, gram, xty, yy);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CS_OK;
}

int cs_ridge_solve_f64(cs_ctx *ctx, const double *gram, const double *xty, int p, double alpha, const double *prior,
                       double *w) {
    if (!ctx || !gram || !xty || !w) return fail(CS_E_BADARGS, "null argument");
    if (p < 1 || p > RP) return fail(CS_E_BADARGS, "p must be in [1, 176]");
    HIPCHK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure(ctx->ridge_scalar, 64))) return rc;
    const size_t smem = sizeof(double) * ((size_t)p * (p + 1) / 2 + p);
    HIPCHK(hipFuncSetAttribute((const void *)k_chol, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    hipLaunchKernelGGL(k_chol, dim3(1), dim3(256), smem, ctx->stream, gram, xty, p, alpha, prior, w,
                       (int *)ctx->ridge_scalar.p);
    HIPCHK(hipGetLastError());
    int status = -1;
    HIPCHK(hipMemcpyAsync(&status, ctx->ridge_scalar.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (status != 0)
        return fail(CS_E_INTERNAL, "matrix is not positive definite (pivot " + std::to_string(status - 1) + ")");
    return CS_OK;
}

int cs_ridge_loss_grad_f64(cs_ctx *ctx, const double *genes, int ng, const double *y, uint64_t n, const double *w,
                           const double *prior, double alpha, double *loss, double *grad) {
    if (!ctx || !genes || !y || !w || !loss || !grad) return fail(CS_E_BADARGS, "null argument");
    int p, rc;
    if ((rc = check_ridge_dims(ng, &p))) return rc;
    if (n == 0) return fail(CS_E_BADARGS, "n must be >= 1");
    HIPCHK(hipSetDevice(ctx->device));
    const int nwg = 512;
    if ((rc = ensure(ctx->ridge_part, sizeof(double) * (RP + 1) * nwg))) return rc;
    hipLaunchKernelGGL(k_ridge_grad, dim3(nwg), dim3(256), 0, ctx->stream, genes, ng, y, n, p, w,
                       (double *)ctx->ridge_part.p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_ridge_grad_reduce, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->ridge_part.p,
                       nwg, p, n, alpha, w, prior, loss, grad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CS_OK;
}

}  // extern "C"
