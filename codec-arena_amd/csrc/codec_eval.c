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
 *   deblocking ....................... src/codec.zig:1258-1310
 *   frame / plane encode ............. src/codec.zig:1505-2237 (8x8 legacy syntax)
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

#include "cscodec.h"

/* ------------------------------------------------------------------ tables */

static float DCT8[64]; /* t[u*8 + x] = c(u)/2 cos((2x+1) u pi / 16) */
static float GAUSS[11];
static pthread_once_t tables_once = PTHREAD_ONCE_INIT;

static const uint8_t ZZ[64] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
    41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
    30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
};

static inline int zz_band(int k) { return k < 8 ? 0 : (k < 16 ? 1 : 2); }

static void init_tables(void) {
    const float pi = 3.14159265358979f;
    for (int u = 0; u < 8; u++) {
        const float cu = u == 0 ? sqrtf(0.5f) : 1.0f;
        for (int x = 0; x < 8; x++) {
            const float ang = pi * (2.0f * (float)x + 1.0f) * (float)u / 16.0f;
            DCT8[u * 8 + x] = 0.5f * cu * (float)cos((double)ang);
        }
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
}

int csc_genome_supported(const cs_genome* g) {
    if (!g) return 0;
    return clamp01(g->g[16]) < 0.25f && clamp01(g->g[19]) <= 0.5f && clamp01(g->g[20]) <= 0.5f &&
           !(g->g[21] > 0.5f) && !(g->g[22] > 0.5f);
}

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
        q[b] = fmaxf(1.0f, qp * s);
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
    CX_LV_B0 = 34, CX_LV_SIGN = 46,
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

static void bw_subpel(blockw* w, uint32_t code) {
    if (w->ac) rc_ubits(&w->e, CX_SP_B0, 4, code);
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
static void bw_levels(blockw* w, const int32_t lv[64], int last) {
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

static uint32_t br_subpel(blockr* r) {
    if (r->ac) return rc_dubits(&r->d, CX_SP_B0, 4);
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
} codec;

static inline float frame_gop(uint32_t fi, uint32_t key_int) { return (float)(fi % key_int) / (float)key_int; }

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

    uint8_t blk[64], pred[64], res[64], rec[64], hp[64];
    int32_t lv[64];
    float deq[64];
    for (int by = 0; by < nby; by++) {
        const float by_n = ((float)by + 0.5f) / (float)nby;
        for (int bx = 0; bx < nbx; bx++) {
            const int sx = bx * 8, sy = by * 8;
            for (int r = 0; r < 8; r++) memcpy(blk + r * 8, src + (size_t)(sy + r) * pw + sx, 8);
            const int force_intra = is_i || (do_refresh && (uint32_t)by == rrow);
            regctx rc = {0.0f, by_n, cx->cur_gop, force_intra};
            qblock qb;
            int last = 0;
            if (force_intra) {
                if (!is_i) bw_marker(&w, 2);
                band_steps(p, qp, iq_scale, &rc, qb.q);
                qb.dz0 = p->deadzone * reg_mod(p->gain[0], rc.mv_n);
                quantize_block(p, &qb, blk, lv);
                cx->dct_calls++;
                for (int k = 0; k < 64; k++)
                    if (lv[k]) last = k;
                bw_levels(&w, lv, last);
                dequant(qb.q, lv, deq);
                idct8(deq, enc + (size_t)sy * pw + sx, (size_t)pw);
                cx->idct_calls++;
                continue;
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
            uint32_t subpel = 4;
            const int ox = sx + bdx, oy = sy + bdy;
            if (p->halfpel) {
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
                continue;
            }
            bw_marker(&w, subpel != 4 ? 2 : 0);
            bw_mv(&w, bdx, bdy);
            if (subpel != 4) bw_subpel(&w, subpel);
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
            uint8_t* dst = enc + (size_t)sy * pw + sx;
            for (int r = 0; r < 8; r++)
                for (int c = 0; c < 8; c++)
                    dst[(size_t)r * pw + c] = (uint8_t)clampi((int)pred[r * 8 + c] + (int)rec[r * 8 + c] - 128, 0, 255);
        }
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
    encode_plane(cx, fi, is_i, qp, su, cx->enc[1], cx->prev[1], cx->w / 2, cx->h / 2, 1);
    encode_plane(cx, fi, is_i, qp, sv, cx->enc[2], cx->prev[2], cx->w / 2, cx->h / 2, 1);
    deblock_frame(p, cx->enc, cx->w, cx->h, qp);
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
    blockr r = {p->entropy, in, {0}};
    if (r.ac) rc_dec_init(&r.d, in);
    uint8_t pred[64], rec[64];
    float coef[64];
    for (int by = 0; by < nby; by++) {
        const float by_n = ((float)by + 0.5f) / (float)nby;
        for (int bx = 0; bx < nbx; bx++) {
            const int sx = bx * 8, sy = by * 8;
            const int force_intra = is_i || (do_refresh && (uint32_t)by == rrow);
            regctx rc = {0.0f, by_n, cx->cur_gop, force_intra};
            float q[3];
            if (force_intra) {
                if (!is_i) (void)br_marker(&r);
                band_steps(p, qp, iq_scale, &rc, q);
                if (br_levels(&r, q, coef)) return CSC_E_CORRUPT;
                idct8(coef, cur + (size_t)sy * pw + sx, (size_t)pw);
                cx->idct_calls++;
                continue;
            }
            const int flags = br_marker(&r);
            if (flags == 1) continue; /* skip: cur already holds the reference copy */
            int32_t dx, dy;
            if (br_mv(&r, &dx, &dy)) return CSC_E_CORRUPT;
            rc.mv_n = fminf(1.0f, ((float)abs(dx) + (float)abs(dy)) / 8.0f);
            const int ox = sx + dx, oy = sy + dy;
            if (flags == 2) {
                const uint32_t code = br_subpel(&r);
                if (code >= 9) return CSC_E_CORRUPT;
                interp_block(prev, pw, ph, ox, oy, (int)(code / 3) - 1, (int)(code % 3) - 1, pred);
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

int csc_eval_clip(const uint8_t* yuv, uint32_t w, uint32_t h, uint32_t frames, const cs_genome* g,
                  uint32_t flags, csc_clip_stats* out) {
    if (!yuv || !g || !out || w < 16 || h < 16 || (w % 16) || (h % 16) || !frames) return CSC_E_ARGS;
    if (!csc_genome_supported(g)) return CSC_E_UNSUPPORTED;
    pthread_once(&tables_once, init_tables);
    pheno p;
    phenotype(g, &p);
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
    uint8_t* planes = (uint8_t*)malloc(4 * fsz);
    uint8_t* stream = (uint8_t*)malloc(CLIP_STREAM_CAP + PLANE_SCRATCH_CAP);
    uint8_t* encout = (uint8_t*)malloc((size_t)frames * fsz); /* encoder recon, frame-major */
    uint8_t* decout = (uint8_t*)malloc((size_t)frames * fsz);
    int rc = CSC_OK;
    if (!planes || !stream || !encout || !decout) {
        rc = CSC_E_ARGS;
        goto done;
    }
    memset(planes, 128, 4 * fsz);
    cx.enc[0] = planes;
    cx.enc[1] = planes + cx.ysz;
    cx.enc[2] = planes + cx.ysz + cx.csz;
    cx.prev[0] = planes + fsz;
    cx.prev[1] = cx.prev[0] + cx.ysz;
    cx.prev[2] = cx.prev[1] + cx.csz;
    cx.main = (bytes_out){stream, 0, CLIP_STREAM_CAP, 0};
    cx.scratch = (bytes_out){stream + CLIP_STREAM_CAP, 0, PLANE_SCRATCH_CAP, 0};
    for (uint32_t fi = 0; fi < frames; fi++) {
        const size_t fb_start = cx.main.n;
        uint8_t* rec[3] = {encout + fi * fsz, encout + fi * fsz + cx.ysz, encout + fi * fsz + cx.ysz + cx.csz};
        encode_frame(&cx, yuv, fi, rec);
        cx.last_frame_bytes = cx.main.n - fb_start;
    }
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
        rc = decode_plane(&cx, &in, fi, is_i, qp, cur[0], prv[0], cx.w, cx.h, 0);
        if (rc == CSC_OK) rc = decode_plane(&cx, &in, fi, is_i, qp, cur[1], prv[1], cx.w / 2, cx.h / 2, 1);
        if (rc == CSC_OK) rc = decode_plane(&cx, &in, fi, is_i, qp, cur[2], prv[2], cx.w / 2, cx.h / 2, 1);
        deblock_frame(&p, cur, cx.w, cx.h, qp);
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
    const float nf = (float)frames;
    out->bytes = cx.main.n;
    out->psnr_y = acc / nf;
    out->psnr_u = acc_u / nf;
    out->psnr_v = acc_v / nf;
    out->msssim_y = acc_ms / nf;
    out->dct_calls = enc_dct;
    out->idct_calls = cx.idct_calls;
    out->enc_mops = (float)(enc_dct + cx.idct_calls) * 64.0f / 1e6f;
    out->dec_mops = (float)cx.idct_calls * 64.0f / 1e6f;
done:
    free(planes);
    free(stream);
    free(encout);
    free(decout);
    return rc;
}
