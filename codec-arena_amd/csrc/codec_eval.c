/*
 * codec_eval.c -- the codec label source (CPU): codec-arena's genome-driven
 * hybrid video codec restated from its semantics, with the decoder run as an
 * independent path and PSNR / MS-SSIM measured on the decoded frames.
 *
 * Followed, by part:
 *   phenotype and regulators ........ src/genome.zig:205-290
 *   DCT-II tables, dct8 / idct8 ...... src/codec.zig:41-64, 660-701
 *   zigzag and bands ................. src/codec.zig:67-80
 *   byte stream, varints ............. src/codec.zig:120-237
 *   adaptive range coder, contexts ... src/codec.zig:239-525
 *   quantizer, levelQ ................ src/codec.zig:752-775, 1446-1483
 *   bilinear half-pel MC ............. src/codec.zig:788-812
 *   quarter-pel MC, half planes ...... src/codec.zig:813-1006
 *   intra prediction, mode sets ...... src/codec.zig:1009-1256
 *   deblocking ....................... src/codec.zig:1258-1310
 *   CDEF-lite ........................ src/codec.zig:1320-1400
 *   frame / plane encode ............. src/codec.zig:1505-2237 (8x8 syntax; the
 *                                      16x16 superblock layer is not restated)
 *   SAD .............................. src/codec.zig:2239-2261
 *   frame / plane decode ............. src/codec.zig:2263-2535
 *   evalClip ......................... src/codec.zig:2537-2670
 *   psnrPlane ........................ src/yuv.zig:68-80
 *   planeMsssim ...................... src/msssim.zig:17-183
 *
 * Arithmetic follows the reference's f32/f64 types and operation order with
 * no contraction (built with -ffp-contract=off); Zig's @round is C's round
 * (half away from zero) and @intFromFloat truncates.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "cscodec.h"

/* ------------------------------------------------------------------ tables */

static float DCT8[64]; /* t[u*8 + x] = c(u)/2 cos((2x+1) u pi / 16) */
static float DCT16[256]; /* t[u*16 + x] = c(u)/2 cos((2x+1) u pi / 32) (src/codec.zig:56-62) */
static uint16_t ZZ16[256]; /* the 16x16 diagonal snake, raster index of the k-th slot (src/codec.zig:85-112) */
static float GAUSS[11];
static pthread_once_t tables_once = PTHREAD_ONCE_INIT;

static const uint8_t ZZ[64] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
};

static inline int zz_band(int k) { return k < 8 ? 0 : (k < 16 ? 1 : 2); }
/* the 8x8 band proportions on the 256 slots (src/codec.zig:114-116) */
static inline int zz_band16(int k) { return k < 32 ? 0 : (k < 64 ? 1 : 2); }

static void init_tables(void) {
    const float pi = 3.14159265358979f;
    for (int u = 0; u < 8; u++) {
        const float cu = u == 0 ? sqrtf(0.5f) : 1.0f;
        for (int x = 0; x < 8; x++) {
            const float ang = pi * (2.0f * (float)x + 1.0f) * (float)u / 16.0f;
            DCT8[u * 8 + x] = 0.5f * cu * (float)cos((double)ang);
        }
    }
    for (int u = 0; u < 16; u++) {
        const float cu = u == 0 ? sqrtf(0.5f) : 1.0f;
        for (int x = 0; x < 16; x++) {
            const float ang = pi * (2.0f * (float)x + 1.0f) * (float)u / 32.0f;
            DCT16[u * 16 + x] = 0.5f * cu * (float)cos((double)ang);
        }
    }
    /* anti-diagonal s: even s run up from the bottom-left, odd s down */
    int zk = 0;
    for (int d = 0; d < 31; d++) {
        if (d % 2 == 0)
            for (int r = d < 16 ? d : 15, c = d - r; r >= 0 && c < 16; r--, c++) ZZ16[zk++] = (uint16_t)(r * 16 + c);
        else
            for (int c = d < 16 ? d : 15, r = d - c; c >= 0 && r < 16; c--, r++) ZZ16[zk++] = (uint16_t)(r * 16 + c);
    }
    float sum = 0.0f;
    for (int i = 0; i < 11; i++) {
        const float d = (float)i - 5.0f;
        GAUSS[i] = (float)exp((double)(-(d * d) / 4.5f));
        sum += GAUSS[i];
    }
    for (int i = 0; i < 11; i++) GAUSS[i] /= sum;
}

static inline float clamp01(float x) { return fmaxf(0.0f, fminf(1.0f, x)); }
static inline float clampf(float x, float lo, float hi) { return fmaxf(lo, fminf(hi, x)); }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* --------------------------------------------------------------- phenotype */

typedef struct pheno {
    uint32_t key_int, search_r, refresh;
    float dz_dc, base_q, band[3], deadzone, skip_th, p_qmul, intra_b0, dither, adapt_q, chroma_scale;
    int halfpel;
    float gain[6];
    float deblock;
    int entropy;
    float intra_gene; /* g[16]: intra mode set tier */
    int cfl;          /* g[19] > 0.5: chroma-from-luma modes join the chroma set */
    int cdef;         /* g[20] > 0.5: CDEF-lite after deblocking */
    int qpel;         /* g[21] > 0.5: quarter-pel MC supersedes the bilinear half-pel path */
    int sb16;         /* g[22] > 0.5: the 16x16 superblock layer (src/genome.zig:72-74) */
    float sb16_gene;  /* clamp01(g[22]): the superblock RD lambda factor (src/genome.zig:69-71) */
    float qscale;     /* uniform quantizer override (src/codec.zig:750), 1 = off */
} pheno;

static void phenotype(const cs_genome* g, pheno* p) {
    p->key_int = 1u + (uint32_t)(clamp01(g->g[0]) * 47.0f);
    p->dz_dc = 0.5f + clamp01(g->g[1]) * 1.5f;
    p->search_r = (uint32_t)(clamp01(g->g[2]) * 12.0f);
    p->base_q = 2.0f + clamp01(g->g[3]) * 46.0f;
    p->band[0] = 0.5f + clamp01(g->g[4]) * 1.5f;
    p->band[1] = 0.5f + clamp01(g->g[5]) * 2.5f;
    p->band[2] = 0.5f + clamp01(g->g[6]) * 4.0f;
    p->deadzone = clamp01(g->g[7]) * 14.0f;
    p->skip_th = clamp01(g->g[8]) * 8.0f;
    p->p_qmul = 0.3f + clamp01(g->g[9]) * 1.7f;
    p->intra_b0 = 0.5f + clamp01(g->g[10]) * 1.5f;
    p->dither = (clamp01(g->g[11]) - 0.5f) * 1.0f;
    p->halfpel = g->g[12] > 0.6f;
    p->refresh = 1u + (uint32_t)(clamp01(g->g[13]) * 8.0f);
    p->adapt_q = clamp01(g->g[14]) * 10.0f;
    p->chroma_scale = 0.5f + clamp01(g->g[15]) * 1.5f;
    for (int i = 0; i < 6; i++) p->gain[i] = (clamp01(g->reg[i]) - 0.5f) * 0.5f;
    p->deblock = clamp01(g->g[17]);
    p->entropy = clamp01(g->g[18]) > 0.5f;
    p->intra_gene = clamp01(g->g[16]);
    p->cfl = clamp01(g->g[19]) > 0.5f;
    p->cdef = clamp01(g->g[20]) > 0.5f;
    p->qpel = g->g[21] > 0.5f;
    p->sb16 = g->g[22] > 0.5f;
    p->sb16_gene = clamp01(g->g[22]);
    p->qscale = 1.0f;
}

/* Every gene-gated tool is restated: 1 for any genome.  The superblock
 * layer's one ill-defined event is reported per clip by csc_eval_clip. */
int csc_genome_supported(const cs_genome* g) { return g != NULL; }

typedef struct regctx {
    float mv_n, by_n, gop;
    int intra;
} regctx;

static inline float reg_qmod(const pheno* p, const regctx* rc) {
    const float mv_term = 1.0f + p->gain[1] * (2.0f * rc->mv_n - 1.0f);
    const float row_term = 1.0f + p->gain[3] * (2.0f * rc->by_n - 1.0f);
    float m = mv_term * row_term;
    if (rc->intra) m *= 1.0f + p->gain[5] * (2.0f * rc->gop - 1.0f);
    return clampf(m, 0.5f, 2.0f);
}

static inline float reg_mod(float gain, float mv_n) { return clampf(1.0f + gain * (2.0f * mv_n - 1.0f), 0.5f, 2.0f); }

/* levelQ for the three zigzag bands of one block (src/codec.zig:768-775). */
static void band_steps(const pheno* p, float qp, float base_scale, const regctx* rc, float q[3]) {
    const float m = reg_qmod(p, rc);
    for (int b = 0; b < 3; b++) {
        float s = base_scale * m * p->band[b];
        if (rc->intra && b == 0) s *= p->intra_b0;
        q[b] = fmaxf(1.0f, qp * s * p->qscale);
    }
}

/* -------------------------------------------------------------- transforms */

static void dct8(const uint8_t* src, float out[64]) {
    float tmp[64];
    for (int r = 0; r < 8; r++) {
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int x = 0; x < 8; x++) {
            const float px = (float)src[r * 8 + x] - 128.0f;
            for (int u = 0; u < 8; u++) s[u] = s[u] + px * DCT8[u * 8 + x];
        }
        for (int u = 0; u < 8; u++) tmp[r * 8 + u] = s[u];
    }
    for (int v = 0; v < 8; v++) {
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int r = 0; r < 8; r++) {
            const float t = DCT8[v * 8 + r];
            for (int u = 0; u < 8; u++) s[u] = s[u] + tmp[r * 8 + u] * t;
        }
        for (int u = 0; u < 8; u++) out[v * 8 + u] = s[u];
    }
}

static void idct8(const float coef[64], uint8_t* dst, size_t stride) {
    float tmp[64];
    for (int v = 0; v < 8; v++) {
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int u = 0; u < 8; u++) {
            const float c = coef[v * 8 + u];
            for (int x = 0; x < 8; x++) s[x] = s[x] + c * DCT8[u * 8 + x];
        }
        for (int x = 0; x < 8; x++) tmp[x * 8 + v] = s[x];
    }
    for (int r = 0; r < 8; r++) {
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int v = 0; v < 8; v++) {
            const float t = tmp[r * 8 + v];
            for (int c = 0; c < 8; c++) s[c] = s[c] + t * DCT8[v * 8 + c];
        }
        for (int c = 0; c < 8; c++) {
            const int px = (int)roundf(s[c] + 128.0f);
            dst[r * stride + c] = (uint8_t)clampi(px, 0, 255);
        }
    }
}

/* 16x16 transform of the superblock layer (src/codec.zig:701-741): the same
 * separable products and summation order as dct8 / idct8 */
static void dct16(const uint8_t* src, float out[256]) {
    float tmp[256];
    for (int r = 0; r < 16; r++)
        for (int u = 0; u < 16; u++) {
            float sm = 0.0f;
            for (int x = 0; x < 16; x++) sm = sm + ((float)src[r * 16 + x] - 128.0f) * DCT16[u * 16 + x];
            tmp[r * 16 + u] = sm;
        }
    for (int u = 0; u < 16; u++)
        for (int v = 0; v < 16; v++) {
            float sm = 0.0f;
            for (int r = 0; r < 16; r++) sm = sm + tmp[r * 16 + u] * DCT16[v * 16 + r];
            out[v * 16 + u] = sm;
        }
}

static void idct16(const float coef[256], uint8_t* dst, size_t stride) {
    float tmp[256];
    for (int v = 0; v < 16; v++)
        for (int x = 0; x < 16; x++) {
            float sm = 0.0f;
            for (int u = 0; u < 16; u++) sm = sm + coef[v * 16 + u] * DCT16[u * 16 + x];
            tmp[x * 16 + v] = sm;
        }
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) {
            float sm = 0.0f;
            for (int v = 0; v < 16; v++) sm = sm + tmp[r * 16 + v] * DCT16[v * 16 + c];
            dst[r * stride + c] = (uint8_t)clampi((int)roundf(sm + 128.0f), 0, 255);
        }
}

/* --------------------------------------------------------------- quantizer */

static inline int32_t quant_level(float c, float q, float dz, float dith) {
    const float a = fabsf(c) + dith;
    const float lvl = roundf(fmaxf(0.0f, (a - dz * 0.5f) / q));
    const int32_t li = (int32_t)lvl;
    return c < 0.0f ? -li : li;
}

typedef struct qblock {
    float q[3];   /* band steps */
    float dz0;    /* motion-regulated deadzone */
} qblock;

static void quantize_block(const pheno* p, const qblock* qb, const uint8_t* src, int32_t lv[64]) {
    float coef[64];
    dct8(src, coef);
    for (int k = 0; k < 64; k++) {
        const float dz = k == 0 ? qb->dz0 * p->dz_dc : qb->dz0;
        lv[k] = quant_level(coef[ZZ[k]], qb->q[zz_band(k)], dz, p->dither);
    }
}

static void dequant(const float q[3], const int32_t lv[64], float out[64]) {
    for (int k = 0; k < 64; k++) out[ZZ[k]] = (float)lv[k] * q[zz_band(k)];
}

/* quantizeBlock16 / dequantLevels16 (src/codec.zig:1485-1501): the 8x8
 * quantizer's deadzone and band steps over the 256 slots */
static void quantize_block16(const pheno* p, const qblock* qb, const uint8_t* src, int32_t lv[256]) {
    float coef[256];
    dct16(src, coef);
    for (int k = 0; k < 256; k++) {
        const float dz = k == 0 ? qb->dz0 * p->dz_dc : qb->dz0;
        lv[k] = quant_level(coef[ZZ16[k]], qb->q[zz_band16(k)], dz, p->dither);
    }
}

static void dequant16(const float q[3], const int32_t lv[256], float out[256]) {
    for (int k = 0; k < 256; k++) out[ZZ16[k]] = (float)lv[k] * q[zz_band16(k)];
}

/* ------------------------------------------------------------ byte streams */

typedef struct bytes_out {
    uint8_t* buf;
    size_t n, cap;
    int overflow;
} bytes_out;

static inline void put_u8(bytes_out* b, uint8_t v) {
    if (b->n < b->cap) b->buf[b->n] = v;
    else b->overflow = 1;
    b->n++;
}

static void put_uvarint(bytes_out* b, uint64_t v) {
    for (;;) {
        const uint8_t byte = (uint8_t)(v & 0x7F);
        v >>= 7;
        if (v) put_u8(b, byte | 0x80);
        else {
            put_u8(b, byte);
            return;
        }
    }
}

static void put_svarint(bytes_out* b, int64_t v) {
    put_uvarint(b, ((uint64_t)v << 1) ^ (uint64_t)(v >> 63));
}

typedef struct bytes_in {
    const uint8_t* buf;
    size_t p, n;
    int err;
} bytes_in;

static inline uint8_t get_u8(bytes_in* b) {
    if (b->p >= b->n) {
        b->err = 1;
        return 0;
    }
    return b->buf[b->p++];
}

static uint64_t get_uvarint(bytes_in* b) {
    uint64_t v = 0;
    unsigned shift = 0;
    for (;;) {
        const uint8_t byte = get_u8(b);
        if (b->err) return 0;
        v |= (uint64_t)(byte & 0x7F) << shift;
        if (!(byte & 0x80)) return v;
        shift += 7;
        if (shift > 56) {
            b->err = 1;
            return 0;
        }
    }
}

/* the reference's decode of its zigzag varint (src/codec.zig:215-220): an odd
 * code z yields -(z >> 1), so a negative value x written by put_svarint reads
 * back as x + 1.  Reproduced as is; it only affects the byte-varint stream. */
static int64_t get_svarint(bytes_in* b) {
    const uint64_t z = get_uvarint(b);
    const int64_t v = (int64_t)(z >> 1);
    return (z & 1) ? -v : v;
}

int64_t csc_svarint_roundtrip(int64_t v) {
    uint8_t buf[16];
    bytes_out o = {buf, 0, sizeof buf, 0};
    put_svarint(&o, v);
    bytes_in in = {buf, 0, o.n, 0};
    return get_svarint(&in);
}

/* ------------------------------------------------------ adaptive range coder */

enum {
    PROB_BITS = 11, PROB_ONE = 1 << PROB_BITS, PROB_MOVE = 5,
    CX_MARK_A = 0, CX_MARK_B = 1, CX_MV_SIGN = 2, CX_MV_U0 = 3, CX_MV_B0 = 6, CX_SP_B0 = 10,
    CX_NP_U0 = 14, CX_NP_B0 = 18, CX_RN_U0 = 22, CX_RN_B0 = 26, CX_LV_U0 = 30, CX_LV_ESC = 33,
    CX_LV_B0 = 34, CX_LV_SIGN = 46, CX_MARK_C = 62,
};
#define RC_TOP (1u << 24)

static inline void prob_update(uint16_t* p, int bit) {
    const int32_t target = bit ? 0 : PROB_ONE;
    const int32_t d = (target - (int32_t)*p) / (1 << PROB_MOVE); /* C division truncates */
    int32_t np = (int32_t)*p + d;
    if (np < 1) np = 1;
    if (np > PROB_ONE - 1) np = PROB_ONE - 1;
    *p = (uint16_t)np;
}

typedef struct rc_enc {
    bytes_out* out;
    uint64_t low;
    uint32_t range;
    uint8_t cache;
    uint32_t cache_size;
    uint16_t ps[64];
} rc_enc;

static void rc_enc_init(rc_enc* e, bytes_out* out) {
    e->out = out;
    e->low = 0;
    e->range = 0xFFFFFFFFu;
    e->cache = 0;
    e->cache_size = 1;
    for (int i = 0; i < 64; i++) e->ps[i] = PROB_ONE / 2;
}

/* delayed-carry byte output (the LZMA scheme) */
static void rc_shift_low(rc_enc* e) {
    const uint8_t carry = (uint8_t)(e->low >> 32);
    if (e->low < 0xFF000000ull || carry) {
        uint8_t temp = e->cache;
        do {
            put_u8(e->out, (uint8_t)(temp + carry));
            temp = 0xFF;
        } while (--e->cache_size);
        e->cache = (uint8_t)(e->low >> 24);
    }
    e->cache_size++;
    e->low = (e->low << 8) & 0xFFFFFFFFull;
}

static void rc_bit(rc_enc* e, int ci, int bit) {
    uint16_t* p = &e->ps[ci];
    const uint32_t bound = (e->range >> PROB_BITS) * (uint32_t)*p;
    if (!bit) e->range = bound;
    else {
        e->low += bound;
        e->range -= bound;
    }
    while (e->range < RC_TOP) {
        e->range <<= 8;
        rc_shift_low(e);
    }
    prob_update(p, bit);
}

static void rc_flush(rc_enc* e) {
    for (int i = 0; i < 5; i++) rc_shift_low(e);
}

static void rc_ubits(rc_enc* e, int base, int nb, uint32_t v) {
    for (int k = 0; k < nb; k++) rc_bit(e, base + k, (v >> (nb - 1 - k)) & 1);
}

/* truncated unary: v ones capped at n_units, then a zero unless v == n_units */
static void rc_unary(rc_enc* e, int base, uint32_t v, uint32_t n_units) {
    for (uint32_t k = 0; k < n_units; k++) {
        const int one = v > k;
        rc_bit(e, base + (int)k, one);
        if (!one) break;
    }
}

typedef struct rc_dec {
    bytes_in* in;
    uint32_t range, code;
    uint16_t ps[64];
} rc_dec;

static void rc_dec_init(rc_dec* d, bytes_in* in) {
    d->in = in;
    d->range = 0xFFFFFFFFu;
    d->code = 0;
    (void)get_u8(in); /* the encoder's initial cache byte */
    for (int i = 0; i < 4; i++) d->code = (d->code << 8) | get_u8(in);
    for (int i = 0; i < 64; i++) d->ps[i] = PROB_ONE / 2;
}

static int rc_dbit(rc_dec* d, int ci) {
    uint16_t* p = &d->ps[ci];
    const uint32_t bound = (d->range >> PROB_BITS) * (uint32_t)*p;
    int bit = 0;
    if (d->code < bound) d->range = bound;
    else {
        d->code -= bound;
        d->range -= bound;
        bit = 1;
    }
    while (d->range < RC_TOP) {
        d->range <<= 8;
        d->code = (d->code << 8) | get_u8(d->in);
    }
    prob_update(p, bit);
    return bit;
}

static uint32_t rc_dubits(rc_dec* d, int base, int nb) {
    uint32_t v = 0;
    for (int k = 0; k < nb; k++) v = (v << 1) | (uint32_t)rc_dbit(d, base + k);
    return v;
}

static uint32_t rc_dunary(rc_dec* d, int base, uint32_t n_units) {
    uint32_t v = 0;
    while (v < n_units && rc_dbit(d, base + (int)v)) v++;
    return v;
}

int64_t csc_rc_roundtrip(const uint8_t* bits, const uint8_t* ctx, uint64_t n, uint8_t* buf, uint64_t cap,
                         uint8_t* bits_out) {
    if ((!bits || !ctx || !bits_out) && n) return CSC_E_ARGS;
    for (uint64_t i = 0; i < n; i++)
        if (ctx[i] >= 64) return CSC_E_ARGS;
    bytes_out o = {buf, 0, cap, 0};
    rc_enc e;
    rc_enc_init(&e, &o);
    for (uint64_t i = 0; i < n; i++) rc_bit(&e, ctx[i], bits[i] & 1);
    rc_flush(&e);
    if (o.overflow) return CSC_E_OVERFLOW;
    bytes_in in = {buf, 0, o.n, 0};
    rc_dec d;
    rc_dec_init(&d, &in);
    for (uint64_t i = 0; i < n; i++) bits_out[i] = (uint8_t)rc_dbit(&d, ctx[i]);
    if (in.err || in.p != o.n) return CSC_E_CORRUPT;
    return (int64_t)o.n;
}

/* ---------------------------------------------------- block stream (writer) */

typedef struct blockw {
    int ac;
    bytes_out* out;
    rc_enc e;
} blockw;

static void bw_marker(blockw* w, int m) {
    if (!w->ac) {
        put_u8(w->out, (uint8_t)m);
        return;
    }
    if (m == 1) {
        rc_bit(&w->e, CX_MARK_A, 1);
        return;
    }
    rc_bit(&w->e, CX_MARK_A, 0);
    rc_bit(&w->e, CX_MARK_B, m == 2);
}

/* the superblock flag (src/codec.zig:407-425, 556-563): with `wide` a
 * non-skip marker carries a third context bit, 1 = unified (value 3) */
static void bw_marker_sb(blockw* w, int m, int wide) {
    if (!w->ac) {
        put_u8(w->out, (uint8_t)m);
        return;
    }
    if (!wide) {
        bw_marker(w, m);
        return;
    }
    if (m == 1) {
        rc_bit(&w->e, CX_MARK_A, 1);
        return;
    }
    rc_bit(&w->e, CX_MARK_A, 0);
    rc_bit(&w->e, CX_MARK_B, m == 2);
    rc_bit(&w->e, CX_MARK_C, m == 3);
}

static void rc_mv(rc_enc* e, int32_t d) {
    rc_bit(e, CX_MV_SIGN, d < 0);
    const uint32_t mag = (uint32_t)(d < 0 ? -d : d);
    rc_unary(e, CX_MV_U0, mag, 3);
    if (mag > 2) rc_ubits(e, CX_MV_B0, 4, mag - 3);
}

static void bw_mv(blockw* w, int32_t dx, int32_t dy) {
    if (w->ac) {
        rc_mv(&w->e, dx);
        rc_mv(&w->e, dy);
    } else {
        put_svarint(w->out, dx);
        put_svarint(w->out, dy);
    }
}

static void bw_subpel(blockw* w, uint32_t code, int bits) {
    if (w->ac) rc_ubits(&w->e, CX_SP_B0, bits, code);
    else put_u8(w->out, (uint8_t)code);
}

static void rc_level(rc_enc* e, int32_t l) {
    rc_bit(e, CX_LV_SIGN, l < 0);
    const uint32_t v = (uint32_t)(l < 0 ? -l : l) - 1;
    rc_unary(e, CX_LV_U0, v, 3);
    if (v > 2) {
        const uint32_t x = v - 3;
        if (x >= 4095) {
            rc_bit(e, CX_LV_ESC, 1);
            rc_ubits(e, CX_LV_B0, 12, 4095);
            rc_ubits(e, CX_LV_B0, 16, x - 4095);
        } else {
            rc_bit(e, CX_LV_ESC, 0);
            rc_ubits(e, CX_LV_B0, 12, x);
        }
    }
}

/* (run, level) pairs up to the last nonzero zigzag slot */
static void bw_levels(blockw* w, const int32_t* lv, int last) {
    uint32_t npairs = 0;
    for (int k = 0; k <= last; k++) npairs += lv[k] != 0;
    if (!w->ac) {
        put_uvarint(w->out, npairs);
        uint64_t run = 0;
        for (int k = 0; k <= last; k++) {
            if (!lv[k]) run++;
            else {
                put_uvarint(w->out, run);
                put_svarint(w->out, lv[k]);
                run = 0;
            }
        }
        return;
    }
    rc_unary(&w->e, CX_NP_U0, npairs, 4);
    if (npairs > 3) rc_ubits(&w->e, CX_NP_B0, 6, npairs - 4);
    uint32_t run = 0;
    for (int k = 0; k <= last; k++) {
        if (!lv[k]) run++;
        else {
            rc_unary(&w->e, CX_RN_U0, run, 4);
            if (run > 3) rc_ubits(&w->e, CX_RN_B0, 6, run - 4);
            rc_level(&w->e, lv[k]);
            run = 0;
        }
    }
}

/* ---------------------------------------------------- block stream (reader) */

typedef struct blockr {
    int ac;
    bytes_in* in;
    rc_dec d;
} blockr;

static int br_marker(blockr* r) {
    if (!r->ac) return get_u8(r->in);
    if (rc_dbit(&r->d, CX_MARK_A)) return 1;
    return rc_dbit(&r->d, CX_MARK_B) ? 2 : 0;
}

static int br_marker_sb(blockr* r, int wide) {
    if (!r->ac) return get_u8(r->in);
    if (!wide) return br_marker(r);
    if (rc_dbit(&r->d, CX_MARK_A)) return 1;
    const int b = rc_dbit(&r->d, CX_MARK_B);
    if (rc_dbit(&r->d, CX_MARK_C)) return 3;
    return b ? 2 : 0;
}

static int32_t rc_dmv(rc_dec* d) {
    const int s = rc_dbit(d, CX_MV_SIGN);
    uint32_t mag = rc_dunary(d, CX_MV_U0, 3);
    if (mag > 2) mag = 3 + rc_dubits(d, CX_MV_B0, 4);
    return s ? -(int32_t)mag : (int32_t)mag;
}

static int br_mv(blockr* r, int32_t* dx, int32_t* dy) {
    if (r->ac) {
        *dx = rc_dmv(&r->d);
        *dy = rc_dmv(&r->d);
        return 0;
    }
    const int64_t a = get_svarint(r->in), b = get_svarint(r->in);
    if (a < INT32_MIN || a > INT32_MAX || b < INT32_MIN || b > INT32_MAX) return -1;
    *dx = (int32_t)a;
    *dy = (int32_t)b;
    return 0;
}

static uint32_t br_subpel(blockr* r, int bits) {
    if (r->ac) return rc_dubits(&r->d, CX_SP_B0, bits);
    return get_u8(r->in);
}

static int32_t rc_dlevel(rc_dec* d) {
    const int s = rc_dbit(d, CX_LV_SIGN);
    uint32_t v = rc_dunary(d, CX_LV_U0, 3);
    if (v > 2) {
        const int esc = rc_dbit(d, CX_LV_ESC);
        uint32_t x = rc_dubits(d, CX_LV_B0, 12);
        if (esc) x = 4095 + rc_dubits(d, CX_LV_B0, 16);
        v = 3 + x;
    }
    const int32_t m = (int32_t)(v + 1);
    return s ? -m : m;
}

/* levels -> dequantized coefficients in raster order */
static int br_levels(blockr* r, const float q[3], float coef[64]) {
    memset(coef, 0, 64 * sizeof(float));
    uint64_t k = 0;
    if (!r->ac) {
        const uint64_t npairs = get_uvarint(r->in);
        for (uint64_t i = 0; i < npairs; i++) {
            const uint64_t run = get_uvarint(r->in);
            const int64_t lvl = get_svarint(r->in);
            if (r->in->err) return -1;
            k += run;
            if (k >= 64) return -1;
            coef[ZZ[k]] = (float)lvl * q[zz_band((int)k)];
            k++;
        }
        return 0;
    }
    uint32_t npairs = rc_dunary(&r->d, CX_NP_U0, 4);
    if (npairs > 3) npairs = 4 + rc_dubits(&r->d, CX_NP_B0, 6);
    for (uint32_t i = 0; i < npairs; i++) {
        uint32_t run = rc_dunary(&r->d, CX_RN_U0, 4);
        if (run > 3) run = 4 + rc_dubits(&r->d, CX_RN_B0, 6);
        k += run;
        if (k >= 64) return -1;
        const int32_t lvl = rc_dlevel(&r->d);
        coef[ZZ[k]] = (float)lvl * q[zz_band((int)k)];
        k++;
    }
    return 0;
}

/* the big-TX twin (src/codec.zig:507-522, 635-655): 256 slots, ZZ16 */
static int br_levels16(blockr* r, const float q[3], float coef[256]) {
    memset(coef, 0, 256 * sizeof(float));
    uint64_t k = 0;
    if (!r->ac) {
        const uint64_t npairs = get_uvarint(r->in);
        for (uint64_t i = 0; i < npairs; i++) {
            const uint64_t run = get_uvarint(r->in);
            const int64_t lvl = get_svarint(r->in);
            if (r->in->err) return -1;
            k += run;
            if (k >= 256) return -1;
            coef[ZZ16[k]] = (float)lvl * q[zz_band16((int)k)];
            k++;
        }
        return 0;
    }
    uint32_t npairs = rc_dunary(&r->d, CX_NP_U0, 4);
    if (npairs > 3) npairs = 4 + rc_dubits(&r->d, CX_NP_B0, 6);
    for (uint32_t i = 0; i < npairs; i++) {
        uint32_t run = rc_dunary(&r->d, CX_RN_U0, 4);
        if (run > 3) run = 4 + rc_dubits(&r->d, CX_RN_B0, 6);
        k += run;
        if (k >= 256) return -1;
        const int32_t lvl = rc_dlevel(&r->d);
        coef[ZZ16[k]] = (float)lvl * q[zz_band16((int)k)];
        k++;
    }
    return 0;
}

/* ------------------------------------------------------- motion compensation */

/* bilinear half-pel block at (ox + hx/2, oy + hy/2), hx, hy in {-1, 0, 1}, edge-clamped */
static void interp_block(const uint8_t* prev, int pw, int ph, int ox, int oy, int hx, int hy, uint8_t out[64]) {
    for (int r = 0; r < 8; r++) {
        const int yn = 2 * (oy + r) + hy;
        const int yi = yn >> 1, yf = yn - 2 * yi; /* floor division by 2 */
        const int y0 = clampi(yi, 0, ph - 1), y1 = clampi(yi + yf, 0, ph - 1);
        for (int c = 0; c < 8; c++) {
            const int xn = 2 * (ox + c) + hx;
            const int xi = xn >> 1, xf = xn - 2 * xi;
            const int x0 = clampi(xi, 0, pw - 1), x1 = clampi(xi + xf, 0, pw - 1);
            const int a = prev[y0 * pw + x0], b = prev[y0 * pw + x1];
            const int c2 = prev[y1 * pw + x0], d = prev[y1 * pw + x1];
            const int v = (a * (2 - xf) * (2 - yf) + b * xf * (2 - yf) + c2 * (2 - xf) * yf + d * xf * yf + 2) / 4;
            out[r * 8 + c] = (uint8_t)v;
        }
    }
}

/* ---- quarter-pel MC (g[21]): 4-tap (-1, 9, 9, -1)/16 halves, clamped to
 * [0, 255], bilinear quarters; the half planes are built once per reference
 * plane and read by both sides (src/codec.zig:828-1006) */

static inline int half_t4(int p0, int p1, int p2, int p3) {
    const int v = (-p0 + 9 * p1 + 9 * p2 - p3 + 8) / 16; /* truncating division */
    return v < 0 ? 0 : (v > 255 ? 255 : v);
}

static inline int at_clamp(const uint8_t* pl, int pw, int ph, int x, int y) {
    return pl[(size_t)clampi(y, 0, ph - 1) * pw + clampi(x, 0, pw - 1)];
}

static inline int half_h(const uint8_t* pl, int pw, int ph, int x, int y) {
    return half_t4(at_clamp(pl, pw, ph, x - 1, y), at_clamp(pl, pw, ph, x, y), at_clamp(pl, pw, ph, x + 1, y),
                   at_clamp(pl, pw, ph, x + 2, y));
}

static inline int half_v(const uint8_t* pl, int pw, int ph, int x, int y) {
    return half_t4(at_clamp(pl, pw, ph, x, y - 1), at_clamp(pl, pw, ph, x, y), at_clamp(pl, pw, ph, x, y + 1),
                   at_clamp(pl, pw, ph, x, y + 2));
}

static void build_half_planes(const uint8_t* prev, int pw, int ph, uint8_t* hH, uint8_t* hV, uint8_t* hD) {
    for (int y = 0; y < ph; y++)
        for (int x = 0; x < pw; x++) hH[(size_t)y * pw + x] = (uint8_t)half_h(prev, pw, ph, x, y);
    for (int y = 0; y < ph; y++)
        for (int x = 0; x < pw; x++) hV[(size_t)y * pw + x] = (uint8_t)half_v(prev, pw, ph, x, y);
    for (int y = 0; y < ph; y++) {
        const int ym = y == 0 ? 0 : y - 1, yp = y + 1 < ph ? y + 1 : ph - 1, y2 = y + 2 < ph ? y + 2 : ph - 1;
        for (int x = 0; x < pw; x++)
            hD[(size_t)y * pw + x] = (uint8_t)half_t4(hH[(size_t)ym * pw + x], hH[(size_t)y * pw + x],
                                                      hH[(size_t)yp * pw + x], hH[(size_t)y2 * pw + x]);
    }
}

/* prediction at (ox + qx/4, oy + qy/4), qx, qy in [-2, 2], from the half planes */
static void mc_block_q(const uint8_t* prev, const uint8_t* hH, const uint8_t* hV, const uint8_t* hD, int pw, int ph,
                       int ox, int oy, int qx, int qy, uint8_t out[64]) {
    const int sx = qx < 0 ? -1 : 0, sy = qy < 0 ? -1 : 0;
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
            const int x = ox + c, y = oy + r;
            int v;
            if (!qx && !qy) v = at_clamp(prev, pw, ph, x, y);
            else if (!qy && abs(qx) == 2) v = at_clamp(hH, pw, ph, x + (qx < 0 ? -1 : 0), y);
            else if (!qx && abs(qy) == 2) v = at_clamp(hV, pw, ph, x, y + (qy < 0 ? -1 : 0));
            else if (abs(qx) == 2 && abs(qy) == 2) v = at_clamp(hD, pw, ph, x + (qx < 0 ? -1 : 0), y + (qy < 0 ? -1 : 0));
            else if (!qy) v = (at_clamp(prev, pw, ph, x, y) + at_clamp(hH, pw, ph, qx == -1 ? x - 1 : x, y) + 1) / 2;
            else if (!qx) v = (at_clamp(prev, pw, ph, x, y) + at_clamp(hV, pw, ph, x, qy == -1 ? y - 1 : y) + 1) / 2;
            else
                v = (at_clamp(prev, pw, ph, x, y) + at_clamp(hH, pw, ph, x + sx, y) + at_clamp(hV, pw, ph, x, y + sy) +
                     at_clamp(hD, pw, ph, x + sx, y + sy) + 2) / 4;
            out[r * 8 + c] = (uint8_t)clampi(v, 0, 255);
        }
}

/* the same prediction computed directly from the reference plane (the
 * definition the half planes must reproduce) */
int csc_interp_q(const uint8_t* prev, int pw, int ph, int ox, int oy, int qx, int qy, int direct, uint8_t* out) {
    if (!prev || !out || pw < 1 || ph < 1 || qx < -2 || qx > 2 || qy < -2 || qy > 2) return CSC_E_ARGS;
    if (!direct) {
        uint8_t* hp = (uint8_t*)malloc((size_t)3 * pw * ph);
        if (!hp) return CSC_E_ARGS;
        build_half_planes(prev, pw, ph, hp, hp + (size_t)pw * ph, hp + (size_t)2 * pw * ph);
        mc_block_q(prev, hp, hp + (size_t)pw * ph, hp + (size_t)2 * pw * ph, pw, ph, ox, oy, qx, qy, out);
        free(hp);
        return CSC_OK;
    }
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
            const int x = ox + c, y = oy + r;
            const int sx = qx < 0 ? -1 : 0, sy = qy < 0 ? -1 : 0;
            int v;
            if (!qx && !qy) v = at_clamp(prev, pw, ph, x, y);
            else if (!qy && abs(qx) == 2) v = half_h(prev, pw, ph, x + sx, y);
            else if (!qx && abs(qy) == 2) v = half_v(prev, pw, ph, x, y + sy);
            else if (abs(qx) == 2 && abs(qy) == 2)
                v = half_t4(half_h(prev, pw, ph, x + sx, y + sy - 1), half_h(prev, pw, ph, x + sx, y + sy),
                            half_h(prev, pw, ph, x + sx, y + sy + 1), half_h(prev, pw, ph, x + sx, y + sy + 2));
            else if (!qy) v = (at_clamp(prev, pw, ph, x, y) + half_h(prev, pw, ph, qx == -1 ? x - 1 : x, y) + 1) / 2;
            else if (!qx) v = (at_clamp(prev, pw, ph, x, y) + half_v(prev, pw, ph, x, qy == -1 ? y - 1 : y) + 1) / 2;
            else
                v = (at_clamp(prev, pw, ph, x, y) + half_h(prev, pw, ph, x + sx, y) + half_v(prev, pw, ph, x, y + sy) +
                     half_t4(half_h(prev, pw, ph, x + sx, y + sy - 1), half_h(prev, pw, ph, x + sx, y + sy),
                             half_h(prev, pw, ph, x + sx, y + sy + 1), half_h(prev, pw, ph, x + sx, y + sy + 2)) + 2) / 4;
            out[r * 8 + c] = (uint8_t)clampi(v, 0, 255);
        }
    return CSC_OK;
}

/* ---- intra prediction (g[16], g[19]): src/codec.zig:1009-1256 */

enum {
    INTRA_NONE = 0, INTRA_DC, INTRA_PLANAR, INTRA_VERTICAL, INTRA_HORIZONTAL, INTRA_DIAG_DR, INTRA_DIAG_DL,
    INTRA_HOR_DOWN, INTRA_SMOOTH, INTRA_PAETH, INTRA_CFL_N2, INTRA_CFL_N1, INTRA_CFL_P1, INTRA_CFL_P2,
};

/* the enabled modes for a plane kind: NONE first, then the tier's angular
 * set, then the four chroma-from-luma variants when that gene is on */
static int intra_modes(const pheno* p, int chroma, uint8_t modes[14]) {
    static const int tier_n[5] = {1, 2, 5, 8, 10};
    const float v = p->intra_gene;
    const int tier = v < 0.25f ? 0 : v < 0.55f ? 1 : v < 0.85f ? 2 : v < 0.925f ? 3 : 4;
    int n = 0;
    for (int m = 0; m < tier_n[tier]; m++) modes[n++] = (uint8_t)m;
    if (chroma && p->cfl)
        for (int m = INTRA_CFL_N2; m <= INTRA_CFL_P2; m++) modes[n++] = (uint8_t)m;
    return n;
}

static inline int mode_bits(int n) {
    if (n <= 1) return 0;
    int b = 0;
    while ((1 << b) < n) b++;
    return b; /* ceil(log2 n) */
}

static uint8_t dc_pred8(const uint8_t* rc, int pw, int sx, int sy) {
    int dc = 128 * 16;
    int s = 0;
    if (sy > 0 && sx > 0) {
        for (int i = 0; i < 8; i++) s += rc[(size_t)(sy - 1) * pw + sx + i] + rc[(size_t)(sy + i) * pw + sx - 1];
        dc = (s + 8) / 16;
    } else if (sy > 0) {
        for (int i = 0; i < 8; i++) s += rc[(size_t)(sy - 1) * pw + sx + i];
        dc = (s + 4) / 8;
    } else if (sx > 0) {
        for (int i = 0; i < 8; i++) s += rc[(size_t)(sy + i) * pw + sx - 1];
        dc = (s + 4) / 8;
    }
    return (uint8_t)clampi(dc, 0, 255);
}

static void intra_predict(int mode, const uint8_t* rc, int pw, int ph, int sx, int sy, uint8_t out[64],
                          const uint8_t* luma, int lpw) {
    uint8_t top[8], left[8], tl = 128, tr = 128, bl = 128;
    memset(top, 128, 8);
    memset(left, 128, 8);
    if (sy > 0) {
        for (int c = 0; c < 8; c++) top[c] = rc[(size_t)(sy - 1) * pw + sx + c];
        tl = sx > 0 ? rc[(size_t)(sy - 1) * pw + sx - 1] : 128;
        tr = sx + 8 < pw ? rc[(size_t)(sy - 1) * pw + sx + 8] : top[7];
    }
    if (sx > 0) {
        for (int r = 0; r < 8; r++) left[r] = rc[(size_t)(sy + r) * pw + sx - 1];
        bl = sy + 8 < ph ? rc[(size_t)(sy + 8) * pw + sx - 1] : left[7];
    }
    switch (mode) {
    case INTRA_NONE: memset(out, 128, 64); break;
    case INTRA_VERTICAL:
        for (int r = 0; r < 8; r++) memcpy(out + r * 8, top, 8);
        break;
    case INTRA_HORIZONTAL:
        for (int r = 0; r < 8; r++) memset(out + r * 8, left[r], 8);
        break;
    case INTRA_DC: memset(out, dc_pred8(rc, pw, sx, sy), 64); break;
    case INTRA_PLANAR:
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++)
                out[r * 8 + c] = (uint8_t)clampi(((7 - c) * left[r] + (c + 1) * tr + (7 - r) * top[c] + (r + 1) * bl + 8) >> 4,
                                                 0, 255);
        break;
    case INTRA_DIAG_DR: {
        uint8_t ref[16];
        memset(ref, 128, 16);
        ref[0] = tl;
        memcpy(ref + 1, top, 8);
        memcpy(ref + 9, left, 7);
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) out[r * 8 + c] = ref[r + c + 1 < 15 ? r + c + 1 : 15];
        break;
    }
    case INTRA_DIAG_DL:
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) {
                const int i = c + (r + 1) / 2;
                out[r * 8 + c] = top[i < 7 ? i : 7];
            }
        break;
    case INTRA_HOR_DOWN:
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) {
                const int i = r + (c + 1) / 2;
                out[r * 8 + c] = left[i < 7 ? i : 7];
            }
        break;
    case INTRA_SMOOTH:
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) {
                const int hz = ((7 - c) * left[r] + (c + 1) * tr + 4) / 8;
                const int vt = ((7 - r) * top[c] + (r + 1) * bl + 4) / 8;
                out[r * 8 + c] = (uint8_t)clampi((hz + vt + 1) >> 1, 0, 255);
            }
        break;
    case INTRA_PAETH:
        for (int r = 0; r < 8; r++)
            for (int c = 0; c < 8; c++) {
                const int pl = left[r] + top[c] - tl;
                const int dl = abs(pl - left[r]), dt = abs(pl - top[c]), dtl = abs(pl - tl);
                out[r * 8 + c] = (dl <= dt && dl <= dtl) ? left[r] : (dt <= dtl ? top[c] : tl);
            }
        break;
    default: { /* chroma from luma: 2x2-mean luma minus its mean, times alpha, on the chroma DC */
        int ld[64];
        int64_t mean = 0;
        for (int r = 0; r < 8; r++) {
            const uint8_t* l0 = luma + (size_t)(2 * sy + 2 * r) * lpw + 2 * sx;
            const uint8_t* l1 = l0 + lpw;
            for (int c = 0; c < 8; c++) {
                ld[r * 8 + c] = (l0[2 * c] + l0[2 * c + 1] + l1[2 * c] + l1[2 * c + 1]) >> 2;
                mean += ld[r * 8 + c];
            }
        }
        const int lm = (int)((mean + 32) / 64);
        const int dc = dc_pred8(rc, pw, sx, sy);
        const int a2 = mode == INTRA_CFL_N2 ? -2 : mode == INTRA_CFL_N1 ? -1 : mode == INTRA_CFL_P1 ? 1 : 2;
        for (int k = 0; k < 64; k++) out[k] = (uint8_t)clampi(dc + a2 * (ld[k] - lm) / 2, 0, 255);
        break;
    }
    }
}

int csc_intra_predict(int mode, const uint8_t* plane, int pw, int ph, int sx, int sy, const uint8_t* luma, int lpw,
                      uint8_t* out) {
    if (!plane || !out || mode < 0 || mode > INTRA_CFL_P2 || sx < 0 || sy < 0 || sx + 8 > pw || sy + 8 > ph ||
        (mode >= INTRA_CFL_N2 && !luma))
        return CSC_E_ARGS;
    intra_predict(mode, plane, pw, ph, sx, sy, out, luma, lpw);
    return CSC_OK;
}

/* ---- CDEF-lite (g[20]): src/codec.zig:1320-1400 */

static void cdef_plane(uint8_t* p, int pw, int ph, float qp, uint8_t* snap) {
    static const int dirs[4][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}};
    memcpy(snap, p, (size_t)pw * ph);
    const int damp = 2 + (int)qp / 24;
    for (int by = 0; by < ph / 8; by++)
        for (int bx = 0; bx < pw / 8; bx++) {
            const int sx = bx * 8, sy = by * 8;
            int best_dir = 0;
            int64_t best_cost = INT64_MAX;
            for (int di = 0; di < 4; di++) {
                int64_t cost = 0;
                for (int r = 2; r < 6; r++)
                    for (int c = 2; c < 6; c++) {
                        const int x = sx + c, y = sy + r, ctr = snap[(size_t)y * pw + x];
                        for (int k = 1; k <= 2; k++) {
                            const int smp = at_clamp(snap, pw, ph, x + dirs[di][0] * k, y + dirs[di][1] * k);
                            cost += abs(smp - ctr) / 4;
                        }
                    }
                if (cost < best_cost) {
                    best_cost = cost;
                    best_dir = di;
                }
            }
            if (best_cost > 512) continue;
            const int dx = dirs[best_dir][0], dy = dirs[best_dir][1];
            for (int r = 0; r < 8; r++)
                for (int c = 0; c < 8; c++) {
                    const int x = sx + c, y = sy + r, ctr = snap[(size_t)y * pw + x];
                    int64_t sum = ctr;
                    for (int k = 1; k <= 2; k++) {
                        const int kk = k < damp ? k : damp;
                        for (int sg = 1; sg >= -1; sg -= 2) {
                            const int smp = at_clamp(snap, pw, ph, x + dx * kk * sg, y + dy * kk * sg);
                            sum += ctr + clampi(smp - ctr, -6, 6);
                        }
                    }
                    p[(size_t)y * pw + x] = (uint8_t)clampi((int)((sum + 2) / 5), 0, 255);
                }
        }
}

/* SAD of an 8x8 block; the fast form samples every 2nd row and column and
 * doubles the sum (src/codec.zig:2239-2251) */
static inline uint64_t sad_block(const uint8_t* a, const uint8_t* b, size_t stride, int fast) {
    uint64_t acc = 0;
    if (fast) {
        for (int r = 0; r < 8; r += 2)
            for (int c = 0; c < 8; c += 2) acc += (uint64_t)abs((int)a[r * 8 + c] - (int)b[r * stride + c]);
        return acc * 2;
    }
    for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) acc += (uint64_t)abs((int)a[r * 8 + c] - (int)b[r * stride + c]);
    return acc;
}

/* ------------------------------------------------------------------- deblock */

static inline void deblock_pair(uint8_t* p0, uint8_t* q0, uint8_t p1, uint8_t q1, float alpha, float beta) {
    const float fp0 = (float)*p0, fq0 = (float)*q0;
    const float d = fq0 - fp0;
    if (fabsf(d) >= alpha) return;
    if (fabsf((float)p1 - fp0) >= beta) return;
    if (fabsf((float)q1 - fq0) >= beta) return;
    const float filt = (3.0f * d + ((float)p1 - (float)q1)) / 8.0f;
    const float maxd = fmaxf(1.0f, fabsf(d) * 0.5f);
    const float delta = clampf(filt, -maxd, maxd);
    *p0 = (uint8_t)clampf(roundf(fp0 + delta), 0.0f, 255.0f);
    *q0 = (uint8_t)clampf(roundf(fq0 - delta), 0.0f, 255.0f);
}

int csc_deblock_pair(uint8_t* p0, uint8_t* q0, uint8_t p1, uint8_t q1, float alpha, float beta) {
    if (!p0 || !q0) return CSC_E_ARGS;
    deblock_pair(p0, q0, p1, q1, alpha, beta);
    return CSC_OK;
}

static void deblock_plane(uint8_t* pl, int pw, int ph, float strength, float qp) {
    const float alpha = strength * (6.0f + qp * 0.8f);
    const float beta = strength * (3.0f + qp * 0.4f);
    for (int c = 8; c < pw; c += 8)
        for (int r = 0; r < ph; r++) {
            uint8_t* row = pl + (size_t)r * pw;
            deblock_pair(&row[c - 1], &row[c], row[c - 2], row[c + 1], alpha, beta);
        }
    for (int r = 8; r < ph; r += 8)
        for (int c = 0; c < pw; c++)
            deblock_pair(&pl[(size_t)(r - 1) * pw + c], &pl[(size_t)r * pw + c], pl[(size_t)(r - 2) * pw + c],
                         pl[(size_t)(r + 1) * pw + c], alpha, beta);
}

static void deblock_frame(const pheno* p, uint8_t* pl[3], int w, int h, float qp) {
    if (p->deblock <= 0.0f) return;
    deblock_plane(pl[0], w, h, p->deblock, qp);
    deblock_plane(pl[1], w / 2, h / 2, p->deblock, qp);
    deblock_plane(pl[2], w / 2, h / 2, p->deblock, qp);
}

/* ------------------------------------------------------------------- encoder */

typedef struct codec {
    const pheno* p;
    int w, h, frames, fast;
    size_t ysz, csz;
    uint8_t* enc[3];  /* frame being coded (starts as a copy of the reference) */
    uint8_t* prev[3]; /* reference = previous recon */
    bytes_out main;
    bytes_out scratch;
    uint64_t dct_calls, idct_calls;
    float qp_adj, cur_gop;
    size_t prev_frame_bytes, last_frame_bytes;
    uint8_t *mcH, *mcV, *mcD; /* quarter-pel half planes of the current reference plane */
    uint8_t* snap;            /* CDEF snapshot */
    const uint8_t* dec_cur_y; /* decoder: the frame's luma recon (chroma-from-luma source) */
    int unsupported;          /* the superblock layer's ill-defined rollback fired */
    uint32_t sb_dec[3];       /* decoder: unified superblocks read, by subtype */
} codec;

/* in-loop filters of a finished frame, identical on both sides: deblocking,
 * then CDEF-lite (src/codec.zig:1543-1549, 2281-2284) */
static void post_filters(const pheno* p, uint8_t* pl[3], int w, int h, float qp, uint8_t* snap) {
    deblock_frame(p, pl, w, h, qp);
    if (p->cdef) {
        cdef_plane(pl[0], w, h, qp, snap);
        cdef_plane(pl[1], w / 2, h / 2, qp, snap);
        cdef_plane(pl[2], w / 2, h / 2, qp, snap);
    }
}

static inline float frame_gop(uint32_t fi, uint32_t key_int) { return (float)(fi % key_int) / (float)key_int; }

/* ---------------------------------------------------- 16x16 superblock layer */

/* dcPred16 (src/codec.zig:1088-1109), recon neighbours; with none the
 * default 128 * 32 clamps to 255, as dcPred8's 128 * 16 does */
static uint8_t dc_pred16(const uint8_t* rc, int pw, int sx, int sy) {
    int dc = 128 * 32, sm = 0;
    if (sy > 0 && sx > 0) {
        for (int i = 0; i < 16; i++) sm += rc[(size_t)(sy - 1) * pw + sx + i] + rc[(size_t)(sy + i) * pw + sx - 1];
        dc = (sm + 16) / 32;
    } else if (sy > 0) {
        for (int i = 0; i < 16; i++) sm += rc[(size_t)(sy - 1) * pw + sx + i];
        dc = (sm + 8) / 16;
    } else if (sx > 0) {
        for (int i = 0; i < 16; i++) sm += rc[(size_t)(sy + i) * pw + sx - 1];
        dc = (sm + 8) / 16;
    }
    return (uint8_t)clampi(dc, 0, 255);
}

/* The block-stream and mode-map state a superblock trial rolls back to
 * (src/codec.zig:1651-1697): bytes written, range-coder state and
 * probabilities, mode-map accumulator. */
typedef struct sb_state {
    size_t n;
    uint64_t low;
    uint32_t range, cache_size;
    uint8_t cache;
    uint16_t ps[64];
    uint32_t mbacc;
    int mbn;
    size_t mmap_n;
} sb_state;

static void sb_snap(const blockw* w, uint32_t mbacc, int mbn, size_t mmap_n, sb_state* st) {
    st->n = w->out->n;
    if (w->ac) {
        st->low = w->e.low;
        st->range = w->e.range;
        st->cache = w->e.cache;
        st->cache_size = w->e.cache_size;
        memcpy(st->ps, w->e.ps, sizeof st->ps);
    }
    st->mbacc = mbacc;
    st->mbn = mbn;
    st->mmap_n = mmap_n;
}

static void sb_restore(blockw* w, const sb_state* st, uint32_t* mbacc, int* mbn, size_t* mmap_n) {
    w->out->n = st->n;
    if (w->ac) {
        w->e.low = st->low;
        w->e.range = st->range;
        w->e.cache = st->cache;
        w->e.cache_size = st->cache_size;
        memcpy(w->e.ps, st->ps, sizeof st->ps);
    }
    *mbacc = st->mbacc;
    *mbn = st->mbn;
    *mmap_n = st->mmap_n;
}

/* SbU.encode (src/codec.zig:1701-1860): one full superblock in unified mode
 * -- flag 3, then subtype 0 (I frames: DC prediction + 16x16 transform),
 * 1 (skip16: one integer MV, the superblock copied) or 2 (inter16: one MV +
 * 16x16 residual) -- with its recon written into enc.  The integer search's
 * SAD compares every source row with the candidate's first reference row,
 * as the reference's loop does.  Returns 1 when the range coder's 6-bit
 * pair-count or run field (src/codec.zig:468-490) cannot hold a value of
 * these 256 slots, so the decoder would read a different syntax. */
static int sb_unified(codec* cx, blockw* w, const uint8_t* src, uint8_t* enc, const uint8_t* prev, int pw, int ph,
                       int is_i, int chroma, float qp, float iq_scale, float pq_scale, int sx, int sy) {
    const pheno* p = cx->p;
    bw_marker_sb(w, 3, 1);
    uint8_t sbs[256], pred[256], res[256], rec[256];
    int32_t lv[256];
    float deq[256];
    for (int r = 0; r < 16; r++) memcpy(sbs + r * 16, src + (size_t)(sy + r) * pw + sx, 16);
    const float by_n = ((float)(sy / 8) + 0.5f) / (float)(ph / 8);
    regctx rc = {0.0f, by_n, cx->cur_gop, is_i};
    qblock qb;
    float scale = iq_scale;
    if (is_i) {
        bw_marker_sb(w, 0, 0);
        memset(pred, dc_pred16(enc, pw, sx, sy), 256);
    } else {
        const int sr = (int)p->search_r;
        const int rmax = chroma ? (sr / 2 > 1 ? sr / 2 : 1) : sr;
        const int ymin = -(sr < sy ? sr : sy), ymax = sr < ph - 16 - sy ? sr : ph - 16 - sy;
        const int xmin = -(sr < sx ? sr : sx), xmax = sr < pw - 16 - sx ? sr : pw - 16 - sx;
        uint64_t best = UINT64_MAX;
        int bdx = 0, bdy = 0;
        for (int dy = ymin; dy <= ymax; dy++) {
            if (dy < -rmax || dy > rmax) continue;
            for (int dx = xmin; dx <= xmax; dx++) {
                if (dx < -rmax || dx > rmax) continue;
                const uint8_t* ref = prev + (size_t)(sy + dy) * pw + (sx + dx);
                uint64_t acc = 0;
                if (cx->fast) {
                    for (int fr = 0; fr < 16; fr += 2)
                        for (int fc = 0; fc < 16; fc += 2) acc += (uint64_t)abs((int)sbs[fr * 16 + fc] - (int)ref[fc]);
                    acc *= 4;
                } else {
                    for (int fr = 0; fr < 16; fr++)
                        for (int fc = 0; fc < 16; fc++) acc += (uint64_t)abs((int)sbs[fr * 16 + fc] - (int)ref[fc]);
                }
                if (acc < best) {
                    best = acc;
                    bdx = dx;
                    bdy = dy;
                }
            }
        }
        rc.mv_n = fminf(1.0f, ((float)abs(bdx) + (float)abs(bdy)) / 8.0f);
        const float perpx = (float)best / 256.0f;
        const float skip_th = p->skip_th * reg_mod(p->gain[4], rc.mv_n);
        const int ox = sx + bdx, oy = sy + bdy;
        for (int r = 0; r < 16; r++) memcpy(pred + r * 16, prev + (size_t)(oy + r) * pw + ox, 16);
        if (perpx < skip_th) {
            bw_marker_sb(w, 1, 0);
            bw_mv(w, bdx, bdy);
            for (int r = 0; r < 16; r++) memcpy(enc + (size_t)(sy + r) * pw + sx, pred + r * 16, 16);
            return 0;
        }
        bw_marker_sb(w, 2, 0);
        bw_mv(w, bdx, bdy);
        scale = pq_scale;
    }
    for (int k = 0; k < 256; k++) res[k] = (uint8_t)clampi((int)sbs[k] - (int)pred[k] + 128, 0, 255);
    band_steps(p, qp, scale, &rc, qb.q);
    qb.dz0 = p->deadzone * reg_mod(p->gain[0], rc.mv_n);
    quantize_block16(p, &qb, res, lv);
    cx->dct_calls++;
    int last = 0, npairs = 0, run = 0, wide = 0;
    for (int k = 0; k < 256; k++)
        if (lv[k]) last = k;
    for (int k = 0; k <= last; k++) {
        if (!lv[k]) run++;
        else {
            wide |= run > 67;
            npairs++;
            run = 0;
        }
    }
    bw_levels(w, lv, last);
    dequant16(qb.q, lv, deq);
    idct16(deq, rec, 16);
    cx->idct_calls++;
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++)
            enc[(size_t)(sy + r) * pw + sx + c] = (uint8_t)clampi((int)pred[r * 16 + c] + (int)rec[r * 16 + c] - 128, 0, 255);
    return w->ac && (wide || npairs > 67);
}

static uint64_t sb_sse(const uint8_t* a, const uint8_t* b, int pw, int sx, int sy) {
    uint64_t e = 0;
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) {
            const int d = (int)a[(size_t)(sy + r) * pw + sx + c] - (int)b[(size_t)(sy + r) * pw + sx + c];
            e += (uint64_t)(d * d);
        }
    return e;
}

static void encode_plane(codec* cx, uint32_t fi, int is_i, float qp, const uint8_t* src, uint8_t* enc,
                         const uint8_t* prev, int pw, int ph, int chroma) {
    const pheno* p = cx->p;
    const float iq_scale = chroma ? p->chroma_scale : 1.0f;
    const float pq_scale = chroma ? p->p_qmul * p->chroma_scale : p->p_qmul;
    const int nbx = pw / 8, nby = ph / 8;
    const uint32_t rrow = (fi / p->refresh) % (uint32_t)(nby > 1 ? nby : 1);
    const int do_refresh = !is_i && !chroma && (fi / p->refresh) > 0;
    const int sr = (int)p->search_r;
    const int rmax = chroma ? (sr / 2 > 1 ? sr / 2 : 1) : sr;
    cx->scratch.n = 0;
    blockw w = {p->entropy, &cx->scratch, {0}};
    if (w.ac) rc_enc_init(&w.e, &cx->scratch);
    if (p->qpel && !is_i) build_half_planes(prev, pw, ph, cx->mcH, cx->mcV, cx->mcD);
    /* intra modes: a bit-packed map of the chosen mode per intra block, framed
     * before the block data; only whole bytes are kept, the map's length is
     * fixed by the block count (src/codec.zig:1627-1636, 2227-2231) */
    uint8_t modes[14];
    const int nmodes = intra_modes(p, chroma, modes);
    const int mwidth = mode_bits(nmodes);
    const size_t map_len = ((size_t)nbx * nby * mwidth + 7) / 8;
    uint8_t mmap[1024];
    memset(mmap, 0, sizeof mmap);
    uint32_t mbacc = 0;
    int mbn = 0;
    size_t mmap_n = 0;
    const uint8_t* luma = chroma ? cx->enc[0] : NULL; /* the frame's luma recon before deblocking */

    /* superblock layer (src/codec.zig:1640-1700, 1862-1918, 2195-2226): at
     * the top-left block of each full, non-mixed 16x16 superblock the unified
     * coding is tried and rolled back, a legacy flag is written, and the four
     * blocks follow.  The trial is kept in ONE pending slot, so only the last
     * superblock of a row pair reaches its decision at its bottom-right
     * block; when unified wins there the stream is rolled back to the trial's
     * snapshot, which also drops the nbx - 2 blocks coded in between. */
    typedef struct {
        sb_state s0;
        uint8_t recon0[256];
        size_t rate0, rate_b;
        uint64_t err_b;
        int sx0, sy0;
    } sb_pend_t;
    sb_pend_t pend;
    int have_pend = 0;

    uint8_t blk[64], pred[64], res[64], rec[64], hp[64], trial[64];
    int32_t lv[64];
    float deq[64];
    for (int by = 0; by < nby; by++) {
        const float by_n = ((float)by + 0.5f) / (float)nby;
        for (int bx = 0; bx < nbx; bx++) {
            const int sx = bx * 8, sy = by * 8;
            const int sb_full = (by % 2 == 0) && (bx % 2 == 0) && by + 2 <= nby && bx + 2 <= nbx;
            const int sb_mixed = !is_i && do_refresh && ((uint32_t)by == rrow || (uint32_t)by + 1 == rrow);
            if (p->sb16 && sb_full && !sb_mixed) {
                sb_snap(&w, mbacc, mbn, mmap_n, &pend.s0);
                pend.rate0 = pend.s0.n;
                for (int r = 0; r < 16; r++) memcpy(pend.recon0 + r * 16, enc + (size_t)(sy + r) * pw + sx, 16);
                (void)sb_unified(cx, &w, src, enc, prev, pw, ph, is_i, chroma, qp, iq_scale, pq_scale, sx, sy);
                pend.rate_b = w.out->n - pend.rate0;
                pend.err_b = sb_sse(src, enc, pw, sx, sy);
                sb_restore(&w, &pend.s0, &mbacc, &mbn, &mmap_n);
                for (int r = 0; r < 16; r++) memcpy(enc + (size_t)(sy + r) * pw + sx, pend.recon0 + r * 16, 16);
                bw_marker_sb(&w, 0, 1);
                pend.sx0 = bx;
                pend.sy0 = by;
                have_pend = 1;
            }
            for (int r = 0; r < 8; r++) memcpy(blk + r * 8, src + (size_t)(sy + r) * pw + sx, 8);
            const int force_intra = is_i || (do_refresh && (uint32_t)by == rrow);
            regctx rc = {0.0f, by_n, cx->cur_gop, force_intra};
            qblock qb;
            int last = 0;
            if (force_intra) {
                if (!is_i) bw_marker(&w, 2);
                band_steps(p, qp, iq_scale, &rc, qb.q);
                qb.dz0 = p->deadzone * reg_mod(p->gain[0], rc.mv_n);
                if (nmodes == 1) { /* raw-DCT intra, no mode code */
                    quantize_block(p, &qb, blk, lv);
                    cx->dct_calls++;
                    for (int k = 0; k < 64; k++)
                        if (lv[k]) last = k;
                    bw_levels(&w, lv, last);
                    dequant(qb.q, lv, deq);
                    idct8(deq, enc + (size_t)sy * pw + sx, (size_t)pw);
                    cx->idct_calls++;
                    goto block_end;
                }
                /* mode by the true pre-quantization error of the u8 residual */
                int mode_i = 0;
                uint64_t best_err = UINT64_MAX;
                for (int mi = 0; mi < nmodes; mi++) {
                    intra_predict(modes[mi], enc, pw, ph, sx, sy, trial, luma, cx->w);
                    uint64_t err = 0;
                    for (int k = 0; k < 64; k++) {
                        const int sv = blk[k], pv = trial[k];
                        err += (uint64_t)abs(pv + clampi(sv - pv, -128, 127) - sv);
                    }
                    if (err < best_err) {
                        best_err = err;
                        mode_i = mi;
                        memcpy(pred, trial, 64);
                    }
                }
                mbacc = (mbacc << mwidth) | (uint32_t)mode_i;
                mbn += mwidth;
                while (mbn >= 8) {
                    mbn -= 8;
                    if (mmap_n < sizeof mmap) mmap[mmap_n] = (uint8_t)(mbacc >> mbn);
                    mmap_n++;
                }
                for (int k = 0; k < 64; k++) res[k] = (uint8_t)clampi((int)blk[k] - (int)pred[k] + 128, 0, 255);
                quantize_block(p, &qb, res, lv);
                cx->dct_calls++;
                for (int k = 0; k < 64; k++)
                    if (lv[k]) last = k;
                bw_levels(&w, lv, last);
                dequant(qb.q, lv, deq);
                idct8(deq, rec, 8);
                cx->idct_calls++;
                uint8_t* dst = enc + (size_t)sy * pw + sx;
                for (int r = 0; r < 8; r++)
                    for (int c = 0; c < 8; c++)
                        dst[(size_t)r * pw + c] = (uint8_t)clampi((int)pred[r * 8 + c] + (int)rec[r * 8 + c] - 128, 0, 255);
                goto block_end;
            }
            /* integer motion search against the reference, raster order, first minimum */
            uint64_t best = UINT64_MAX;
            int bdx = 0, bdy = 0;
            const int ymin = -(sr < sy ? sr : sy), ymax = sr < ph - 8 - sy ? sr : ph - 8 - sy;
            const int xmin = -(sr < sx ? sr : sx), xmax = sr < pw - 8 - sx ? sr : pw - 8 - sx;
            for (int dy = ymin; dy <= ymax; dy++) {
                if (dy < -rmax || dy > rmax) continue;
                for (int dx = xmin; dx <= xmax; dx++) {
                    if (dx < -rmax || dx > rmax) continue;
                    const uint64_t s = sad_block(blk, prev + (size_t)(sy + dy) * pw + (sx + dx), (size_t)pw, cx->fast);
                    if (s < best) {
                        best = s;
                        bdx = dx;
                        bdy = dy;
                    }
                }
            }
            const float perpx = (float)best / 64.0f;
            rc.mv_n = fminf(1.0f, ((float)abs(bdx) + (float)abs(bdy)) / 8.0f);
            const float skip_th = p->skip_th * reg_mod(p->gain[4], rc.mv_n);
            const uint32_t center = p->qpel ? 12u : 4u;
            uint32_t subpel = center;
            const int ox = sx + bdx, oy = sy + bdy;
            if (p->qpel) {
                /* the 8 half positions, then the quarter ring around the winning half */
                uint64_t hb = best;
                int bqx = 0, bqy = 0;
                for (int qy = -2; qy <= 2; qy += 2)
                    for (int qx = -2; qx <= 2; qx += 2) {
                        if (!qx && !qy) continue;
                        mc_block_q(prev, cx->mcH, cx->mcV, cx->mcD, pw, ph, ox, oy, qx, qy, hp);
                        const uint64_t s2 = sad_block(blk, hp, 8, cx->fast);
                        if (s2 < hb) {
                            hb = s2;
                            bqx = qx;
                            bqy = qy;
                        }
                    }
                if (bqx || bqy) {
                    /* the ring moves with the running best, as in src/codec.zig:2075-2095 */
                    for (int qy = -1; qy <= 1; qy++)
                        for (int qx = -1; qx <= 1; qx++) {
                            if (!qx && !qy) continue;
                            const int cxq = bqx + qx, cyq = bqy + qy;
                            if (abs(cxq) > 2 || abs(cyq) > 2) continue;
                            if (abs(cxq) != 1 && abs(cyq) != 1) continue;
                            mc_block_q(prev, cx->mcH, cx->mcV, cx->mcD, pw, ph, ox, oy, cxq, cyq, hp);
                            const uint64_t s2 = sad_block(blk, hp, 8, cx->fast);
                            if (s2 < hb) {
                                hb = s2;
                                bqx = cxq;
                                bqy = cyq;
                            }
                        }
                    subpel = (uint32_t)((bqx + 2) * 5 + (bqy + 2));
                    best = hb;
                    mc_block_q(prev, cx->mcH, cx->mcV, cx->mcD, pw, ph, ox, oy, bqx, bqy, pred);
                }
            } else if (p->halfpel) {
                uint64_t hb = best;
                int bhx = 0, bhy = 0;
                for (int hy = -1; hy <= 1; hy++)
                    for (int hx = -1; hx <= 1; hx++) {
                        if (!hx && !hy) continue;
                        interp_block(prev, pw, ph, ox, oy, hx, hy, hp);
                        const uint64_t s2 = sad_block(blk, hp, 8, cx->fast);
                        if (s2 < hb) {
                            hb = s2;
                            bhx = hx;
                            bhy = hy;
                        }
                    }
                if (bhx || bhy) {
                    subpel = (uint32_t)((bhx + 1) * 3 + (bhy + 1));
                    best = hb;
                    interp_block(prev, pw, ph, ox, oy, bhx, bhy, pred);
                }
            }
            if (perpx < skip_th) {
                bw_marker(&w, 1); /* skip: the block keeps the co-located reference pixels */
                goto block_end;
            }
            bw_marker(&w, subpel != center ? 2 : 0);
            bw_mv(&w, bdx, bdy);
            if (subpel != center) bw_subpel(&w, subpel, p->qpel ? 5 : 4);
            else
                for (int r = 0; r < 8; r++) memcpy(pred + r * 8, prev + (size_t)(oy + r) * pw + ox, 8);
            for (int k = 0; k < 64; k++) res[k] = (uint8_t)clampi((int)blk[k] - (int)pred[k] + 128, 0, 255);
            band_steps(p, qp, pq_scale, &rc, qb.q);
            qb.dz0 = p->deadzone * reg_mod(p->gain[0], rc.mv_n);
            quantize_block(p, &qb, res, lv);
            cx->dct_calls++;
            for (int k = 0; k < 64; k++)
                if (lv[k]) last = k;
            bw_levels(&w, lv, last);
            dequant(qb.q, lv, deq);
            idct8(deq, rec, 8);
            cx->idct_calls++;
            {
                uint8_t* dst = enc + (size_t)sy * pw + sx;
                for (int r = 0; r < 8; r++)
                    for (int c = 0; c < 8; c++)
                        dst[(size_t)r * pw + c] = (uint8_t)clampi((int)pred[r * 8 + c] + (int)rec[r * 8 + c] - 128, 0, 255);
            }
        block_end:
            if (have_pend && by == pend.sy0 + 1 && bx == pend.sx0 + 1) {
                have_pend = 0;
                const uint64_t err_a = sb_sse(src, enc, pw, pend.sx0 * 8, pend.sy0 * 8);
                const size_t rate_a = w.out->n - pend.rate0;
                const float lam = 0.85f * qp * qp * (0.5f + (p->sb16_gene - 0.5f) * 2.0f);
                const double cost_a = (double)err_a + (double)lam * 8.0 * (double)rate_a;
                const double cost_b = (double)pend.err_b + (double)lam * 8.0 * (double)pend.rate_b;
                if (cost_b < cost_a) {
                    /* the rollback drops the blocks between the superblock's
                     * two rows unless it is the plane's only column pair */
                    if (nbx != 2) {
                        cx->unsupported = 1;
                        return;
                    }
                    sb_restore(&w, &pend.s0, &mbacc, &mbn, &mmap_n);
                    for (int r = 0; r < 16; r++)
                        memcpy(enc + (size_t)(pend.sy0 * 8 + r) * pw + pend.sx0 * 8, pend.recon0 + r * 16, 16);
                    if (sb_unified(cx, &w, src, enc, prev, pw, ph, is_i, chroma, qp, iq_scale, pq_scale, pend.sx0 * 8,
                                   pend.sy0 * 8)) {
                        cx->unsupported = 1;
                        return;
                    }
                    bx += 1;
                }
            }
        }
    }
    if (mwidth > 0) {
        if (map_len > sizeof mmap) cx->main.overflow = 1;
        for (size_t i = 0; i < map_len && i < sizeof mmap; i++) put_u8(&cx->main, mmap[i]);
    }
    if (w.ac) rc_flush(&w.e);
    for (size_t i = 0; i < cx->scratch.n && i < cx->scratch.cap; i++) put_u8(&cx->main, cx->scratch.buf[i]);
    if (cx->scratch.overflow) cx->main.overflow = 1;
}

static void encode_frame(codec* cx, const uint8_t* yuv, uint32_t fi, uint8_t* recon[3]) {
    const pheno* p = cx->p;
    const int is_i = (fi % p->key_int) == 0;
    const float gop = frame_gop(fi, p->key_int);
    const float qp_f0 = clampf(p->base_q + cx->qp_adj, 1.0f, 120.0f);
    const float qp_f = clampf(qp_f0 * (1.0f + p->gain[2] * (2.0f * gop - 1.0f)), 1.0f, 120.0f);
    const float qp = roundf(qp_f);
    cx->cur_gop = gop;
    put_u8(&cx->main, is_i ? 1 : 0);
    put_u8(&cx->main, (uint8_t)qp);
    for (int c = 0; c < 3; c++) memcpy(cx->enc[c], cx->prev[c], c ? cx->csz : cx->ysz);
    const uint8_t* sy = yuv + (size_t)fi * cx->ysz;
    const uint8_t* su = yuv + (size_t)cx->frames * cx->ysz + (size_t)fi * cx->csz;
    const uint8_t* sv = su + (size_t)cx->frames * cx->csz;
    encode_plane(cx, fi, is_i, qp, sy, cx->enc[0], cx->prev[0], cx->w, cx->h, 0);
    if (cx->unsupported) return;
    encode_plane(cx, fi, is_i, qp, su, cx->enc[1], cx->prev[1], cx->w / 2, cx->h / 2, 1);
    if (cx->unsupported) return;
    encode_plane(cx, fi, is_i, qp, sv, cx->enc[2], cx->prev[2], cx->w / 2, cx->h / 2, 1);
    if (cx->unsupported) return;
    post_filters(p, cx->enc, cx->w, cx->h, qp, cx->snap);
    for (int c = 0; c < 3; c++) {
        uint8_t* t = cx->prev[c];
        cx->prev[c] = cx->enc[c];
        cx->enc[c] = t;
        memcpy(recon[c], cx->prev[c], c ? cx->csz : cx->ysz);
    }
    /* open-loop adaptive q: compares the two most recently finished frames */
    const size_t fb = cx->last_frame_bytes;
    if (cx->prev_frame_bytes > 0 && p->adapt_q > 0.01f) {
        if (fb > cx->prev_frame_bytes) cx->qp_adj += 0.05f * p->adapt_q;
        else cx->qp_adj -= 0.025f * p->adapt_q;
        cx->qp_adj = clampf(cx->qp_adj, -2.0f, p->adapt_q);
    }
    cx->prev_frame_bytes = fb;
}

/* ------------------------------------------------------------------- decoder */

static int decode_plane(codec* cx, bytes_in* in, uint32_t fi, int is_i, float qp, uint8_t* cur,
                        const uint8_t* prev, int pw, int ph, int chroma) {
    const pheno* p = cx->p;
    const float iq_scale = chroma ? p->chroma_scale : 1.0f;
    const float pq_scale = chroma ? p->p_qmul * p->chroma_scale : p->p_qmul;
    const int nbx = pw / 8, nby = ph / 8;
    const uint32_t rrow = (fi / p->refresh) % (uint32_t)(nby > 1 ? nby : 1);
    const int do_refresh = !is_i && !chroma && (fi / p->refresh) > 0;
    if (p->qpel && !is_i) build_half_planes(prev, pw, ph, cx->mcH, cx->mcV, cx->mcD);
    uint8_t modes[14], mmap[1024];
    const int nmodes = intra_modes(p, chroma, modes);
    const int mwidth = mode_bits(nmodes);
    size_t map_len = 0;
    if (mwidth > 0) {
        map_len = ((size_t)nbx * nby * mwidth + 7) / 8;
        if (map_len > sizeof mmap) return CSC_E_CORRUPT;
        for (size_t i = 0; i < map_len; i++) mmap[i] = get_u8(in);
    }
    size_t map_p = 0;
    uint32_t map_acc = 0;
    int map_bn = 0;
    float q16[3];
    const uint8_t* luma = chroma ? cx->dec_cur_y : NULL;
    blockr r = {p->entropy, in, {0}};
    if (r.ac) rc_dec_init(&r.d, in);
    uint8_t pred[64], rec[64];
    float coef[64];
    /* odd-row columns already covered by a unified superblock decoded from
     * the even row above (src/codec.zig:2350-2372) */
    uint8_t skipmap[256];
    if (p->sb16 && nbx > 256) return CSC_E_CORRUPT;
    for (int by = 0; by < nby; by++) {
        const float by_n = ((float)by + 0.5f) / (float)nby;
        if (p->sb16 && by % 2 == 0) memset(skipmap, 0, sizeof skipmap);
        for (int bx = 0; bx < nbx; bx++) {
            const int sx = bx * 8, sy = by * 8;
            if (p->sb16 && by % 2 == 1 && skipmap[bx]) continue;
            const int sb_full = (by % 2 == 0) && (bx % 2 == 0) && by + 2 <= nby && bx + 2 <= nbx;
            const int sb_mixed = !is_i && do_refresh && ((uint32_t)by == rrow || (uint32_t)by + 1 == rrow);
            if (p->sb16 && sb_full && !sb_mixed) {
                /* src/codec.zig:2377-2447: flag 3 = unified, 0 = four legacy blocks follow */
                const int flag = br_marker_sb(&r, 1);
                if (flag == 3) {
                    const int sub = br_marker_sb(&r, 0);
                    uint8_t sp[256], srec[256];
                    float scoef[256];
                    if (sub == 0) {
                        regctx rc = {0.0f, by_n, cx->cur_gop, 1};
                        memset(sp, dc_pred16(cur, pw, sx, sy), 256);
                        band_steps(p, qp, iq_scale, &rc, q16);
                        if (br_levels16(&r, q16, scoef)) return CSC_E_CORRUPT;
                    } else {
                        if (sub > 2) return CSC_E_CORRUPT;
                        int32_t dx, dy;
                        if (br_mv(&r, &dx, &dy)) return CSC_E_CORRUPT;
                        const int ox = sx + dx, oy = sy + dy;
                        if (ox < 0 || oy < 0 || ox + 16 > pw || oy + 16 > ph) return CSC_E_CORRUPT;
                        regctx rc = {fminf(1.0f, ((float)abs(dx) + (float)abs(dy)) / 8.0f), by_n, cx->cur_gop, 0};
                        for (int rr = 0; rr < 16; rr++) memcpy(sp + rr * 16, prev + (size_t)(oy + rr) * pw + ox, 16);
                        if (sub == 2) {
                            band_steps(p, qp, pq_scale, &rc, q16);
                            if (br_levels16(&r, q16, scoef)) return CSC_E_CORRUPT;
                        }
                    }
                    uint8_t* dst = cur + (size_t)sy * pw + sx;
                    if (sub == 1) {
                        for (int rr = 0; rr < 16; rr++) memcpy(dst + (size_t)rr * pw, sp + rr * 16, 16);
                    } else {
                        idct16(scoef, srec, 16);
                        cx->idct_calls++;
                        for (int rr = 0; rr < 16; rr++)
                            for (int c = 0; c < 16; c++)
                                dst[(size_t)rr * pw + c] =
                                    (uint8_t)clampi((int)sp[rr * 16 + c] + (int)srec[rr * 16 + c] - 128, 0, 255);
                    }
                    cx->sb_dec[sub]++;
                    skipmap[bx] = 1;
                    skipmap[bx + 1] = 1;
                    bx += 1;
                    continue;
                }
                if (flag != 0) return CSC_E_CORRUPT;
            }
            const int force_intra = is_i || (do_refresh && (uint32_t)by == rrow);
            regctx rc = {0.0f, by_n, cx->cur_gop, force_intra};
            float q[3];
            if (force_intra) {
                if (!is_i) (void)br_marker(&r);
                band_steps(p, qp, iq_scale, &rc, q);
                if (nmodes == 1) {
                    if (br_levels(&r, q, coef)) return CSC_E_CORRUPT;
                    idct8(coef, cur + (size_t)sy * pw + sx, (size_t)pw);
                    cx->idct_calls++;
                    continue;
                }
                while (map_bn < mwidth) { /* MSB-first bits of the mode map */
                    if (map_p >= map_len) return CSC_E_CORRUPT;
                    map_acc = (map_acc << 8) | mmap[map_p++];
                    map_bn += 8;
                }
                map_bn -= mwidth;
                const uint32_t midx = (map_acc >> map_bn) & ((1u << mwidth) - 1u);
                if ((int)midx >= nmodes) return CSC_E_CORRUPT;
                intra_predict(modes[midx], cur, pw, ph, sx, sy, pred, luma, cx->w);
                if (br_levels(&r, q, coef)) return CSC_E_CORRUPT;
                idct8(coef, rec, 8);
                cx->idct_calls++;
                uint8_t* dst = cur + (size_t)sy * pw + sx;
                for (int rr = 0; rr < 8; rr++)
                    for (int c = 0; c < 8; c++)
                        dst[(size_t)rr * pw + c] = (uint8_t)clampi((int)pred[rr * 8 + c] + (int)rec[rr * 8 + c] - 128, 0, 255);
                continue;
            }
            const int flags = br_marker(&r);
            if (flags == 1) continue; /* skip: cur already holds the reference copy */
            int32_t dx, dy;
            if (br_mv(&r, &dx, &dy)) return CSC_E_CORRUPT;
            rc.mv_n = fminf(1.0f, ((float)abs(dx) + (float)abs(dy)) / 8.0f);
            const int ox = sx + dx, oy = sy + dy;
            if (flags == 2) {
                if (p->qpel) {
                    const uint32_t code = br_subpel(&r, 5);
                    if (code >= 25) return CSC_E_CORRUPT;
                    mc_block_q(prev, cx->mcH, cx->mcV, cx->mcD, pw, ph, ox, oy, (int)(code / 5) - 2, (int)(code % 5) - 2,
                               pred);
                } else {
                    const uint32_t code = br_subpel(&r, 4);
                    if (code >= 9) return CSC_E_CORRUPT;
                    interp_block(prev, pw, ph, ox, oy, (int)(code / 3) - 1, (int)(code % 3) - 1, pred);
                }
            } else {
                if (ox < 0 || oy < 0 || ox + 8 > pw || oy + 8 > ph) return CSC_E_CORRUPT;
                for (int rr = 0; rr < 8; rr++) memcpy(pred + rr * 8, prev + (size_t)(oy + rr) * pw + ox, 8);
            }
            band_steps(p, qp, pq_scale, &rc, q);
            if (br_levels(&r, q, coef)) return CSC_E_CORRUPT;
            idct8(coef, rec, 8);
            cx->idct_calls++;
            uint8_t* dst = cur + (size_t)sy * pw + sx;
            for (int rr = 0; rr < 8; rr++)
                for (int c = 0; c < 8; c++)
                    dst[(size_t)rr * pw + c] = (uint8_t)clampi((int)pred[rr * 8 + c] + (int)rec[rr * 8 + c] - 128, 0, 255);
        }
    }
    return in->err ? CSC_E_CORRUPT : CSC_OK;
}

/* ------------------------------------------------------------------- metrics */

float csc_psnr_plane(const uint8_t* a, const uint8_t* b, uint64_t n) {
    double acc = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        const int32_t d = (int32_t)a[i] - (int32_t)b[i];
        acc += (double)(d * d);
    }
    const double mse = acc / (double)n;
    if (mse <= 0.0000001) return 99.0f;
    const double val = 10.0 * log10(255.0 * 255.0 / mse);
    return (float)(val < 99.0 ? val : 99.0);
}

/* valid-mode separable 11-tap gaussian: (w x h) -> (w-10) x (h-10) */
static void gauss_valid(const float* src, float* dst, int w, int h, float* tmp) {
    const int ow = w - 10, oh = h - 10;
    for (int y = 0; y < h; y++) {
        const float* row = src + (size_t)y * w;
        float* o = tmp + (size_t)y * ow;
        for (int x = 0; x < ow; x++) {
            float acc = 0.0f;
            for (int i = 0; i < 11; i++) acc += GAUSS[i] * row[x + i];
            o[x] = acc;
        }
    }
    for (int y = 0; y < oh; y++) {
        float* o = dst + (size_t)y * ow;
        for (int x = 0; x < ow; x++) {
            float acc = 0.0f;
            for (int i = 0; i < 11; i++) acc += GAUSS[i] * tmp[(size_t)(y + i) * ow + x];
            o[x] = acc;
        }
    }
}

static void downsample2(float* buf, int w, int h) {
    const int nw = w / 2, nh = h / 2;
    for (int r = 0; r < nh; r++)
        for (int c = 0; c < nw; c++) {
            const float* a = buf + (size_t)(2 * r) * w + 2 * c;
            const float* b = a + w;
            buf[(size_t)r * nw + c] = (a[0] + a[1] + b[0] + b[1]) * 0.25f;
        }
}

int csc_plane_msssim(const uint8_t* src, const uint8_t* rec, uint32_t w, uint32_t h, float* out) {
    static const double WEIGHTS[5] = {0.0448, 0.2856, 0.3001, 0.2363, 0.1333};
    const float C1 = 6.5025f, C2 = 58.5225f;
    if (!src || !rec || !out || w < 11 || h < 11) return CSC_E_ARGS;
    pthread_once(&tables_once, init_tables);
    const size_t n0 = (size_t)w * h;
    float* mem = (float*)malloc(sizeof(float) * (n0 * 3 + (size_t)(w - 10) * (h - 10) * 5 + (size_t)(w - 10) * h));
    if (!mem) return CSC_E_ARGS;
    float *x = mem, *y = x + n0, *sq = y + n0, *fx = sq + n0;
    const size_t nv0 = (size_t)(w - 10) * (h - 10);
    float *fy = fx + nv0, *fxx = fy + nv0, *fyy = fxx + nv0, *fxy = fyy + nv0, *tmp = fxy + nv0;
    for (size_t i = 0; i < n0; i++) {
        x[i] = (float)src[i];
        y[i] = (float)rec[i];
    }
    int lw = (int)w, lh = (int)h;
    double log_ss = 0.0;
    for (int scale = 0; scale < 5; scale++) {
        if (lw < 11 || lh < 11) break;
        const int vw = lw - 10, vh = lh - 10;
        const size_t n = (size_t)vw * vh, cur = (size_t)lw * lh;
        gauss_valid(x, fx, lw, lh, tmp);
        gauss_valid(y, fy, lw, lh, tmp);
        for (size_t i = 0; i < cur; i++) sq[i] = x[i] * x[i];
        gauss_valid(sq, fxx, lw, lh, tmp);
        for (size_t i = 0; i < cur; i++) sq[i] = y[i] * y[i];
        gauss_valid(sq, fyy, lw, lh, tmp);
        for (size_t i = 0; i < cur; i++) sq[i] = x[i] * y[i];
        gauss_valid(sq, fxy, lw, lh, tmp);
        double acc_l = 0.0, acc_cs = 0.0;
        for (size_t i = 0; i < n; i++) {
            const float mx = fx[i], my = fy[i];
            const float vx2 = fxx[i] - mx * mx, vy2 = fyy[i] - my * my, vxy = fxy[i] - mx * my;
            acc_l += (double)((2.0f * mx * my + C1) / (mx * mx + my * my + C1));
            acc_cs += (double)((2.0f * vxy + C2) / (vx2 + vy2 + C2));
        }
        const double m_l = acc_l / (double)n, m_cs = acc_cs / (double)n;
        if (scale == 4) log_ss += WEIGHTS[scale] * log(m_l > 1e-6 ? m_l : 1e-6);
        log_ss += WEIGHTS[scale] * log(m_cs > 1e-6 ? m_cs : 1e-6);
        if (scale + 1 < 5) {
            downsample2(x, lw, lh);
            downsample2(y, lw, lh);
            lw /= 2;
            lh /= 2;
        }
    }
    free(mem);
    *out = (float)exp(log_ss);
    return CSC_OK;
}

/* ------------------------------------------------------------------ evalClip */

#define CLIP_STREAM_CAP (8u << 20) /* bitbuf of src/codec.zig:2563 */
#define PLANE_SCRATCH_CAP (1u << 20)

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

int csc_eval_clip(const uint8_t* yuv, uint32_t w, uint32_t h, uint32_t frames, const cs_genome* g,
                  uint32_t flags, csc_clip_stats* out) {
    return csc_eval_clip_ex(yuv, w, h, frames, g, flags, 1.0f, NULL, out);
}

int csc_eval_clip_ex(const uint8_t* yuv, uint32_t w, uint32_t h, uint32_t frames, const cs_genome* g,
                     uint32_t flags, float qscale, uint8_t* recon, csc_clip_stats* out) {
    if (!yuv || !g || !out || w < 16 || h < 16 || (w % 16) || (h % 16) || !frames) return CSC_E_ARGS;
    if (!isfinite(qscale)) return CSC_E_ARGS;
    pthread_once(&tables_once, init_tables);
    pheno p;
    phenotype(g, &p);
    p.qscale = qscale;
    codec cx;
    memset(&cx, 0, sizeof cx);
    cx.p = &p;
    cx.w = (int)w;
    cx.h = (int)h;
    cx.frames = (int)frames;
    cx.fast = (flags & CSC_FAST_SAD) != 0;
    cx.ysz = (size_t)w * h;
    cx.csz = cx.ysz / 4;
    const size_t fsz = cx.ysz + 2 * cx.csz;
    uint8_t* planes = (uint8_t*)malloc(4 * fsz + 4 * cx.ysz);
    uint8_t* stream = (uint8_t*)malloc(CLIP_STREAM_CAP + PLANE_SCRATCH_CAP);
    uint8_t* encout = (uint8_t*)malloc((size_t)frames * fsz); /* encoder recon, frame-major */
    uint8_t* decout = (uint8_t*)malloc((size_t)frames * fsz);
    int rc = CSC_OK;
    if (!planes || !stream || !encout || !decout) {
        rc = CSC_E_ARGS;
        goto done;
    }
    memset(planes, 128, 4 * fsz);
    cx.mcH = planes + 4 * fsz;
    cx.mcV = cx.mcH + cx.ysz;
    cx.mcD = cx.mcV + cx.ysz;
    cx.snap = cx.mcD + cx.ysz;
    cx.enc[0] = planes;
    cx.enc[1] = planes + cx.ysz;
    cx.enc[2] = planes + cx.ysz + cx.csz;
    cx.prev[0] = planes + fsz;
    cx.prev[1] = cx.prev[0] + cx.ysz;
    cx.prev[2] = cx.prev[1] + cx.csz;
    cx.main = (bytes_out){stream, 0, CLIP_STREAM_CAP, 0};
    cx.scratch = (bytes_out){stream + CLIP_STREAM_CAP, 0, PLANE_SCRATCH_CAP, 0};
    const double t_enc = now_ms();
    for (uint32_t fi = 0; fi < frames; fi++) {
        const size_t fb_start = cx.main.n;
        uint8_t* rec[3] = {encout + fi * fsz, encout + fi * fsz + cx.ysz, encout + fi * fsz + cx.ysz + cx.csz};
        encode_frame(&cx, yuv, fi, rec);
        if (cx.unsupported) {
            rc = CSC_E_UNSUPPORTED;
            goto done;
        }
        cx.last_frame_bytes = cx.main.n - fb_start;
    }
    const float enc_ms = (float)(now_ms() - t_enc);
    if (cx.main.overflow) {
        rc = CSC_E_OVERFLOW;
        goto done;
    }
    const uint64_t enc_dct = cx.dct_calls;
    /* the independent decode of the bitstream into fresh planes; the metrics
     * are measured on its output, as the reference does */
    memset(planes, 128, 4 * fsz);
    uint8_t* cur[3] = {planes, planes + cx.ysz, planes + cx.ysz + cx.csz};
    uint8_t* prv[3] = {planes + fsz, planes + fsz + cx.ysz, planes + fsz + cx.ysz + cx.csz};
    bytes_in in = {stream, 0, cx.main.n, 0};
    for (uint32_t fi = 0; fi < frames && rc == CSC_OK; fi++) {
        for (int c = 0; c < 3; c++) memcpy(cur[c], prv[c], c ? cx.csz : cx.ysz);
        const int ftype = get_u8(&in);
        const float qp = (float)get_u8(&in);
        const int is_i = ftype == 1;
        cx.cur_gop = frame_gop(fi, p.key_int);
        cx.dec_cur_y = cur[0];
        rc = decode_plane(&cx, &in, fi, is_i, qp, cur[0], prv[0], cx.w, cx.h, 0);
        if (rc == CSC_OK) rc = decode_plane(&cx, &in, fi, is_i, qp, cur[1], prv[1], cx.w / 2, cx.h / 2, 1);
        if (rc == CSC_OK) rc = decode_plane(&cx, &in, fi, is_i, qp, cur[2], prv[2], cx.w / 2, cx.h / 2, 1);
        post_filters(&p, cur, cx.w, cx.h, qp, cx.snap);
        for (int c = 0; c < 3; c++) {
            memcpy(decout + fi * fsz + (c ? cx.ysz + (c - 1) * cx.csz : 0), cur[c], c ? cx.csz : cx.ysz);
            uint8_t* t = prv[c];
            prv[c] = cur[c];
            cur[c] = t;
        }
    }
    if (rc != CSC_OK) goto done;
    out->closed_loop = in.p == cx.main.n && !memcmp(decout, encout, (size_t)frames * fsz);
    if ((flags & CSC_CLOSED_LOOP) && !out->closed_loop) {
        rc = CSC_E_DESYNC;
        goto done;
    }
    const uint8_t* metric_src = decout;
    float acc = 0.0f, acc_u = 0.0f, acc_v = 0.0f, acc_ms = 0.0f;
    for (uint32_t fi = 0; fi < frames; fi++) {
        const uint8_t* d = metric_src + fi * fsz;
        const uint8_t* sy = yuv + (size_t)fi * cx.ysz;
        const uint8_t* su = yuv + (size_t)frames * cx.ysz + (size_t)fi * cx.csz;
        const uint8_t* sv = su + (size_t)frames * cx.csz;
        acc += csc_psnr_plane(sy, d, cx.ysz);
        acc_u += csc_psnr_plane(su, d + cx.ysz, cx.csz);
        acc_v += csc_psnr_plane(sv, d + cx.ysz + cx.csz, cx.csz);
        if (flags & CSC_MSSSIM) {
            float m = 0.0f;
            rc = csc_plane_msssim(sy, d, w, h, &m);
            if (rc != CSC_OK) goto done;
            acc_ms += m;
        }
    }
    if (recon) {
        /* global-planar like the clip: all Y frames, then all U, then all V */
        for (uint32_t fi = 0; fi < frames; fi++) {
            const uint8_t* d = decout + fi * fsz;
            memcpy(recon + (size_t)fi * cx.ysz, d, cx.ysz);
            memcpy(recon + (size_t)frames * cx.ysz + (size_t)fi * cx.csz, d + cx.ysz, cx.csz);
            memcpy(recon + (size_t)frames * (cx.ysz + cx.csz) + (size_t)fi * cx.csz, d + cx.ysz + cx.csz, cx.csz);
        }
    }
    const float nf = (float)frames;
    out->bytes = cx.main.n;
    out->psnr_y = acc / nf;
    out->psnr_u = acc_u / nf;
    out->psnr_v = acc_v / nf;
    out->msssim_y = acc_ms / nf;
    out->dct_calls = enc_dct;
    out->idct_calls = cx.idct_calls;
    for (int k = 0; k < 3; k++) out->sb_unified[k] = cx.sb_dec[k];
    out->enc_mops = (float)(enc_dct + cx.idct_calls) * 64.0f / 1e6f;
    out->dec_mops = (float)cx.idct_calls * 64.0f / 1e6f;
    out->enc_ms = enc_ms;
done:
    free(planes);
    free(stream);
    free(encout);
    free(decout);
    return rc;
}
