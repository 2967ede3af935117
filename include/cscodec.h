/*
 * cscodec.h -- C ABI of the codec label source: a from-scratch CPU
 * restatement of codec-arena's genome-driven encoder/decoder (src/codec.zig)
 * and its metrics (src/yuv.zig psnrPlane, src/msssim.zig planeMsssim), used
 * to label (genome, score) pairs for the learned surrogate and to re-score
 * screen survivors.  The reference exposes this path only as the `arena eval`
 * subcommand (src/main.zig:219-351 evalAll -> src/codec.zig:2537 evalClip).
 *
 * Coverage: the 8x8 codec with every gene-gated tool -- intra prediction
 * (g[16], src/codec.zig:1009-1256), deblocking (g[17], :1258-1310), the
 * adaptive binary range coder (g[18], :239-525), chroma-from-luma (g[19]),
 * CDEF-lite (g[20], :1320-1400), quarter-pel MC (g[21], :813-1006) -- and
 * the 16x16 superblock layer (g[22] > 0.5, :1640-1925, :2195-2226,
 * :2350-2447).  Two events of that layer leave a bitstream whose decode the
 * reference does not define, and a clip that meets one fails with
 * CSC_E_UNSUPPORTED: the unified trial of a row pair's last superblock
 * winning in a plane more than two blocks wide (the rollback to its
 * snapshot drops the nbx - 2 blocks coded since), and a committed unified
 * superblock whose range-coded pair count or zero run exceeds the 6-bit
 * field (68 or more).
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
#define CSC_E_UNSUPPORTED (-20) /* the clip's stream is one the reference leaves undefined */
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
    float enc_ms;        /* wall time of the encoder (ms); not reproducible */
    uint32_t sb_unified[3]; /* unified 16x16 superblocks in the decoded stream:
                               intra16, skip16, inter16 (0 unless g[22] > 0.5) */
} csc_clip_stats;

/* 1 for any genome: every gene-gated tool is restated (kept for callers
 * that screened genomes with it; CSC_E_UNSUPPORTED is now per clip). */
int csc_genome_supported(const cs_genome* g);

/* evalClip (src/codec.zig:2537-2670) on one planar YUV420p clip laid out as
 * all Y frames, then all U, then all V (src/yuv.zig:33-66): encode, decode
 * the bitstream independently, measure on the decoded frames.  The byte-varint
 * stream (entropy gene <= 0.5) decodes a negative value x as x + 1
 * (src/codec.zig:215-220 against :165-168), so its decoder output drifts from
 * the encoder recon; that is reproduced and reported in closed_loop. */
int csc_eval_clip(const uint8_t* yuv, uint32_t w, uint32_t h, uint32_t frames, const cs_genome* g,
                  uint32_t flags, csc_clip_stats* out);

/* csc_eval_clip with the `arena eval` options of src/main.zig:74-93:
 * `qscale` is the uniform quantizer override (--qscale, codec.qscale_override
 * at src/codec.zig:750-754; 1 = off, bit-identical to csc_eval_clip), and a
 * non-null `recon` (frames * w * h * 3 / 2 bytes) receives the decoded clip in
 * the input's global-planar layout (--dump-recon, src/codec.zig:2649-2658). */
int csc_eval_clip_ex(const uint8_t* yuv, uint32_t w, uint32_t h, uint32_t frames, const cs_genome* g,
                     uint32_t flags, float qscale, uint8_t* recon, csc_clip_stats* out);

/* planeMsssim (src/msssim.zig:89-175) of two u8 planes. */
int csc_plane_msssim(const uint8_t* src, const uint8_t* rec, uint32_t w, uint32_t h, float* out);

/* psnrPlane (src/yuv.zig:68-80). */
float csc_psnr_plane(const uint8_t* a, const uint8_t* b, uint64_t n);

/* putSvarint then getSvarint (src/codec.zig:165-168, 215-220): v for v >= 0,
 * v + 1 for v < 0 (the decode quirk described at csc_eval_clip). */
int64_t csc_svarint_roundtrip(int64_t v);

/* intraPredict (src/codec.zig:1118-1246): the 8x8 prediction of `mode`
 * (0 NONE ... 13 CFL +2/2) at (sx, sy) of an 8-bit plane; `luma` (row pitch
 * lpw) is the co-located luma plane, needed by the chroma-from-luma modes. */
int csc_intra_predict(int mode, const uint8_t* plane, int pw, int ph, int sx, int sy, const uint8_t* luma, int lpw,
                      uint8_t* out);

/* Quarter-pel prediction at (ox + qx/4, oy + qy/4), qx, qy in [-2, 2]: from
 * prebuilt half planes (direct == 0, mcBlockQ, src/codec.zig:959-1006, the
 * path both codec sides use) or by the direct filters (direct != 0,
 * interpBlockQ, :866-921). */
int csc_interp_q(const uint8_t* prev, int pw, int ph, int ox, int oy, int qx, int qy, int direct, uint8_t* out);

/* deblockPair (src/codec.zig:1258-1272) on one edge pair. */
int csc_deblock_pair(uint8_t* p0, uint8_t* q0, uint8_t p1, uint8_t q1, float alpha, float beta);

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
