/*
 * cscodec.h -- C ABI of the codec label source: a from-scratch CPU
 * restatement of codec-arena's genome-driven encoder/decoder (src/codec.zig)
 * and its metrics (src/yuv.zig psnrPlane, src/msssim.zig planeMsssim), used
 * to label (genome, score) pairs for the learned surrogate and to re-score
 * screen survivors.  The reference exposes this path only as the `arena eval`
 * subcommand (src/main.zig:219-351 evalAll -> src/codec.zig:2537 evalClip).
 *
 * Coverage: every tool that is off when its gene says "legacy" is restated,
 * plus the in-loop deblocking filter (g[17], src/codec.zig:1258-1310) and the
 * adaptive binary range coder (g[18] > 0.5, src/codec.zig:239-525).  Genomes
 * that switch on intra prediction (g[16] >= 0.25), chroma-from-luma
 * (g[19] > 0.5), CDEF (g[20] > 0.5), quarter-pel MC (g[21] > 0.5) or the 16x16
 * superblock layer (g[22] > 0.5) are rejected with CSC_E_UNSUPPORTED.
 *
 * All functions are thread-safe (no shared mutable state); buffers are
 * caller-owned.
 */
#ifndef CSCODEC_H
#define CSCODEC_H
#include <stdint.h>

#include "csfunnel.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CSC_OK 0
#define CSC_E_UNSUPPORTED (-20) /* genome needs a codec tool not restated here */
#define CSC_E_ARGS (-21)        /* bad clip geometry or null pointer          */
#define CSC_E_OVERFLOW (-22)    /* bitstream larger than the 8 MiB clip buffer */
#define CSC_E_CORRUPT (-23)     /* decoder hit malformed data (never expected) */
#define CSC_E_DESYNC (-24)      /* decoder output != encoder recon (CSC_CLOSED_LOOP) */

#define CSC_FAST_SAD 1u   /* EvalOpts.fast_sad (src/codec.zig:17), the default */
#define CSC_CLOSED_LOOP 2u /* fail with CSC_E_DESYNC unless decoder output == encoder recon */
#define CSC_MSSSIM 4u     /* compute MS-SSIM of the Y plane per frame           */

/* ClipStats (src/codec.zig:23-39) plus the raw counters behind the mops. */
typedef struct csc_clip_stats {
    uint64_t bytes;
    float psnr_y, psnr_u, psnr_v;
    float msssim_y;   /* mean over frames; 0 unless CSC_MSSSIM */
    float enc_mops;   /* (dct + idct calls) * 64 / 1e6 */
    float dec_mops;   /* idct calls * 64 / 1e6 (encoder and decoder calls) */
    uint64_t dct_calls, idct_calls;
    int32_t closed_loop; /* 1 when the decoder output equals the encoder recon */
    int32_t _pad;
} csc_clip_stats;

/* 1 when every tool the genome switches on is restated here, else 0. */
int csc_genome_supported(const cs_genome* g);

/* evalClip (src/codec.zig:2537-2670) on one planar YUV420p clip laid out as
 * all Y frames, then all U, then all V (src/yuv.zig:33-66): encode, decode
 * the bitstream independently, measure on the decoded frames.  The byte-varint
 * stream (entropy gene <= 0.5) decodes a negative value x as x + 1
 * (src/codec.zig:215-220 against :165-168), so its decoder output drifts from
 * the encoder recon; that is reproduced and reported in closed_loop. */
int csc_eval_clip(const uint8_t* yuv, uint32_t w, uint32_t h, uint32_t frames, const cs_genome* g,
                  uint32_t flags, csc_clip_stats* out);

/* planeMsssim (src/msssim.zig:89-175) of two u8 planes. */
int csc_plane_msssim(const uint8_t* src, const uint8_t* rec, uint32_t w, uint32_t h, float* out);

/* psnrPlane (src/yuv.zig:68-80). */
float csc_psnr_plane(const uint8_t* a, const uint8_t* b, uint64_t n);

/* putSvarint then getSvarint (src/codec.zig:165-168, 215-220): v for v >= 0,
 * v + 1 for v < 0 (the decode quirk described at csc_eval_clip). */
int64_t csc_svarint_roundtrip(int64_t v);

/* Adaptive range coder round trip over a symbol stream (src/codec.zig:
 * 279-358): codes `n` bits with contexts ctx[i] (< 64) into buf (capacity
 * cap), decodes them back into bits_out and returns the coded length, or a
 * negative status. */
int64_t csc_rc_roundtrip(const uint8_t* bits, const uint8_t* ctx, uint64_t n, uint8_t* buf, uint64_t cap,
                         uint8_t* bits_out);

#ifdef __cplusplus
}
#endif
#endif
