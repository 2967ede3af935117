// VALU instruction-rate probe: the issue cost of the instruction classes the
// screen kernels are made of, measured on the device the screen runs on.  The
// quadratic screen is bound by VALU issue, and its instruction mix is about
// one part in six 32-bit integer multiplies (the splitmix64 mixing), which
// the SIMD does not issue at the rate of an add; the probe measures each
// class so that the screen's achieved fraction of the VALU bound can be
// stated from measured rates rather than from one nominal rate.
//
// Each thread runs `iters` iterations of eight independent instructions of
// one class (inline asm, so the compiler neither folds nor reorders them), on
// a grid of 8 waves per SIMD; two classes time the matrix core alone and
// with packed-f16 work beside it.  No memory is touched in the loop; one
// store per thread keeps the results live, and each wave's s_memtime count
// gives the cost in that counter's ticks beside the event-timed ns.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "funnel_internal.h"

namespace csf {

enum RateOp {
    RATE_ADD_U32 = 0,
    RATE_MUL_LO_U32 = 1,
    RATE_MUL_HI_U32 = 2,
    RATE_MAD_U64_U32 = 3,
    RATE_MUL_F32 = 4,
    RATE_ADD_U64 = 5,
    RATE_ALIGNBIT = 6,
    RATE_PK_FMA_F16 = 7,
    RATE_MFMA_BF16 = 8,     // v_mfma_f32_16x16x32_bf16
    RATE_MFMA_PK_F16 = 9,   // the same MFMAs with four independent v_pk_fma_f16 after each
    RATE_MFMA_F8 = 10,      // v_mfma_f32_16x16x128_f8f6f4 (E4M3 x E4M3)
    RATE_SHR_U64 = 11,      // v_lshrrev_b64
    RATE_NOPS = 12
};
typedef __bf16 pbf16x8 __attribute__((ext_vector_type(8)));
typedef float pf32x4 __attribute__((ext_vector_type(4)));
typedef int pi32x8 __attribute__((ext_vector_type(8)));

template <int OP>
__global__ __launch_bounds__(256) void k_rate(uint32_t *out, uint64_t *cyc, uint32_t iters) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t t0 = __builtin_amdgcn_s_memtime();
    uint32_t a[8];
    uint64_t d[8];
    float f[8];
    pf32x4 acc[8];
    pbf16x8 ma, mb;
    pi32x8 fa, fb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        ma[i] = (__bf16)(1.0f + (float)(tid & 7) * 0.125f);
        mb[i] = (__bf16)(0.5f - (float)i * 0.0625f);
        acc[i] = pf32x4{0.0f, 0.0f, 0.0f, 0.0f};
        fa[i] = 0x38383838 ^ (int)(tid & 7);  // E4M3 values near 0.5
        fb[i] = 0x30303030 + i;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = tid * 2654435761u + (uint32_t)i;
        d[i] = ((uint64_t)a[i] << 32) | (a[i] ^ 0x9e3779b9u);
        f[i] = 1.0f + (float)i * 1e-3f;
    }
    const uint32_t c = 0x9e3779b9u ^ tid;
    const uint64_t c64 = 0xbf58476d1ce4e5b9ull;
    const float cf = 1.0000001f;
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (OP == RATE_ADD_U32) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[i]) : "v"(c));
            if constexpr (OP == RATE_MUL_LO_U32) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[i]) : "v"(c));
            if constexpr (OP == RATE_MUL_HI_U32) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a[i]) : "v"(c));
            if constexpr (OP == RATE_MAD_U64_U32) {
                uint64_t cc;
                asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(d[i]), "=s"(cc) : "v"(a[i]), "v"(c));
            }
            if constexpr (OP == RATE_MUL_F32) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(f[i]) : "v"(cf));
            if constexpr (OP == RATE_ADD_U64) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(d[i]) : "v"(c64));
            if constexpr (OP == RATE_SHR_U64) asm volatile("v_lshrrev_b64 %0, 30, %0" : "+v"(d[i]));
            if constexpr (OP == RATE_ALIGNBIT) asm volatile("v_alignbit_b32 %0, %0, %1, 7" : "+v"(a[i]) : "v"(c));
            if constexpr (OP == RATE_PK_FMA_F16) asm volatile("v_pk_fma_f16 %0, %0, %1, %1" : "+v"(a[i]) : "v"(c));
            if constexpr (OP == RATE_MFMA_BF16 || OP == RATE_MFMA_PK_F16)
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ma, mb, acc[i], 0, 0, 0);
            if constexpr (OP == RATE_MFMA_F8)
                acc[i] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(fa, fb, acc[i], 0, 0, 0, 0, 0, 0);
            if constexpr (OP == RATE_MFMA_PK_F16) {
                asm volatile("v_pk_fma_f16 %0, %0, %1, %1" : "+v"(a[i]) : "v"(c));
                asm volatile("v_pk_fma_f16 %0, %0, %1, %1" : "+v"(a[(i + 1) & 7]) : "v"(c));
                asm volatile("v_pk_fma_f16 %0, %0, %1, %1" : "+v"(a[(i + 2) & 7]) : "v"(c));
                asm volatile("v_pk_fma_f16 %0, %0, %1, %1" : "+v"(a[(i + 3) & 7]) : "v"(c));
            }
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        r ^= a[i] ^ (uint32_t)d[i] ^ (uint32_t)(d[i] >> 32) ^ __float_as_uint(f[i]) ^ __float_as_uint(acc[i][0]) ^
             __float_as_uint(acc[i][3]);
    out[tid] = r;
    const uint64_t t1 = __builtin_amdgcn_s_memtime();
    if ((threadIdx.x & 63) == 0) cyc[tid >> 6] = t1 - t0;
}

}  // namespace csf

using namespace csf;

extern "C" int cs_valu_rate_probe(int device, int op, uint32_t iters, double *ns_per_iter, double *cycles_per_iter,
                                  uint32_t *simds) {
    if (!ns_per_iter || op < 0 || op >= RATE_NOPS || iters == 0) return fail(CS_E_BADARGS, "rate probe args");
    if (hipSetDevice(device) != hipSuccess) return fail(CS_E_NODEV, "hipSetDevice");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(CS_E_HIP, "hipGetDeviceProperties");
    const uint32_t nsimd = (uint32_t)prop.multiProcessorCount * 4u;
    const uint32_t blocks = nsimd * 2u;  // 256 threads = 4 waves: 8 waves per SIMD
    const uint32_t nwaves = blocks * 4u;
    uint32_t *out = nullptr;
    uint64_t *cyc = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)) != hipSuccess) return fail(CS_E_OOM, "rate probe");
    if (hipMalloc(&cyc, (size_t)nwaves * sizeof(uint64_t)) != hipSuccess) {
        hipFree(out);
        return fail(CS_E_OOM, "rate probe");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        hipFree(out);
        hipFree(cyc);
        return fail(CS_E_HIP, "rate probe events");
    }
    typedef void (*Kern)(uint32_t *, uint64_t *, uint32_t);
    static const Kern kerns[RATE_NOPS] = {k_rate<0>, k_rate<1>, k_rate<2>, k_rate<3>, k_rate<4>, k_rate<5>,
                                          k_rate<6>, k_rate<7>, k_rate<8>, k_rate<9>, k_rate<10>, k_rate<11>};
    int rc = CS_OK;
    float ms = 0.0f;
    // one warm-up launch, then the timed one
    for (int rep = 0; rep < 2 && rc == CS_OK; ++rep) {
        if (rep == 1) hipEventRecord(e0, 0);
        hipLaunchKernelGGL(kerns[op], dim3(blocks), dim3(256), 0, 0, out, cyc, iters);
        if (hipGetLastError() != hipSuccess) rc = fail(CS_E_HIP, "rate probe launch");
        if (rep == 1) hipEventRecord(e1, 0);
    }
    if (rc == CS_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(CS_E_HIP, "rate probe sync");
    if (rc == CS_OK) hipEventElapsedTime(&ms, e0, e1);
    std::vector<uint64_t> hc(nwaves);
    if (rc == CS_OK && hipMemcpy(hc.data(), cyc, nwaves * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(CS_E_HIP, "rate probe copy");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(out);
    hipFree(cyc);
    if (rc != CS_OK) return rc;
    // every SIMD runs 8 waves of `iters` iterations
    *ns_per_iter = (double)ms * 1e6 / (8.0 * (double)iters);
    if (cycles_per_iter) {
        std::sort(hc.begin(), hc.end());
        *cycles_per_iter = (double)hc[nwaves / 2] / (8.0 * (double)iters);
    }
    if (simds) *simds = nsimd;
    return CS_OK;
}
