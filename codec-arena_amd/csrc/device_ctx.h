// device_ctx.h -- the cs_ctx device context shared by the screen and the
// ridge/MLP halves of libcsfunnel (HIP runtime types; device code only).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "funnel_internal.h"

namespace csf {

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            return fail(CS_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));             \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

inline int ensure(DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes) return CS_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    size_t want = bytes + bytes / 4 + 4096;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        // the runtime keeps the error as its last one: clear it, or the next
        // launch check of this context would report it again
        (void)hipGetLastError();
        return fail(CS_E_OOM, "hipMalloc(" + std::to_string(want) + "): " + hipGetErrorString(e));
    }
    b.bytes = want;
    return CS_OK;
}

// Device buffers of one screen call.  A context owns one set and grows it
// with hipMalloc; the allocation-free entry point carves a set out of the
// library's static scratch array instead.
struct ScreenBufs {
    DevBuf streams, blk_first, maps, entries, upper_maps, upper_entries, level_first, elites, segs, seg_first, rec,
        counters, model, seg_first_buf, segout, seg_base;
};

}  // namespace csf

struct cs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    csf::ScreenBufs sb;
    csf::DevBuf gram_part, ridge_part, ridge_scalar;
    std::vector<csf::Record> host_rec;
    cs_screen_stats stats{};
    uint32_t rec_cap = 1u << 20;
    bool force_maps = false;  // testing hook: always use the transfer maps
    int walk_seg_log = 0;     // speculative walk: 0 = balanced segments, else 2^walk_seg_log draws each
    int n_cu = 256;
    int n_mlp = 0;  // live cs_mlp handles bound to this context
    void *scratch = nullptr;  // this device's copy of the static scratch array
};

// Learned surrogate state (mlp_device.hip): FP32 master parameters, AdamW
// moments, BF16 copies, the fragment layouts the MLP screen reads, and the
// activations of the last training batch.
struct cs_mlp {
    cs_ctx *ctx = nullptr;
    csf::DevBuf params, grads, m, v, w1b, w2b, w1f, w1fh, w2f, w2f8, w2s8, f64pack;
    csf::DevBuf xb, z1, h1b, z2, h2b, out, dout, dz2f, dz2b, dh1, dz1f, dz1b, loss_part, gws;
    int cap = 0;
    long step = 0;
    float h_scale = 1.0f;  // FP8 screen: power-of-two scale of H1 (W2 has per-block E8M0 scales)
};
