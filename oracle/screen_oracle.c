/*
 * screen_oracle.c -- CPU restatement of the codec-arena screen.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP funnel and
 * the timed CPU baseline ("port" of src/screen.zig).  It is never linked into
 * the product library.  Build: oracle/Makefile (-O3 -ffp-contract=off: Zig's
 * default strict float mode performs no multiply-add contraction).
 *
 * Reference lines followed (codec-arena snapshot):
 *   splitmix64 Rng.next / Rng.unit        src/screen.zig:195-213
 *   stream partition + seeds               src/screen.zig:267-287
 *   candidate synthesis (worker)           src/screen.zig:225-255
 *   Model.fromCompact                      src/screen.zig:36-95
 *   Model.score (ordered @reduce)          src/screen.zig:97-112
 *   loadModel                              src/screen.zig:115-127
 *   Heap.offer                             src/screen.zig:151-193
 *   merge + stable sort + output lines     src/screen.zig:306-342
 *   parseLine (incl. the v6 token bug)     src/genome.zig:116-188
 *   clamp01                                src/genome.zig:292-294
 *   elites file loading                    src/main.zig:201-214
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <errno.h>
#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GAMMA 0x9E3779B97F4A7C15ULL

/* ---------------------------------------------------------------- RNG --- */

/* src/screen.zig:202-208 (inlined into the worker loop, as in the
 * reference's ReleaseFast build; or_splitmix_next exports it) */
static inline uint64_t sm_next(uint64_t *state) {
    *state += GAMMA;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t or_splitmix_next(uint64_t *state) { return sm_next(state); }

void or_splitmix_fill(uint64_t seed, uint64_t n, uint64_t *out) {
    uint64_t s = seed;
    for (uint64_t i = 0; i < n; i++) out[i] = sm_next(&s);
}

/* src/screen.zig:210-212: top 24 bits times 2^-24, exact in f32 */
static inline float unit_of(uint64_t *s) {
    return (float)(sm_next(s) >> 40) * (1.0f / 16777216.0f);
}

/* src/screen.zig:284 */
uint64_t or_stream_seed(uint64_t seed, uint64_t t) { return seed ^ (GAMMA * (t + 1)); }

/* ------------------------------------------------------- integer parse --- */

/* Zig std.fmt.parseInt(unsigned, tok, 10): optional sign, '_' between digits,
 * '-' only for the value zero, overflow is an error. */
static int parse_u64(const char *s, size_t n, uint64_t max, uint64_t *out) {
    size_t i = 0;
    int neg = 0;
    if (n == 0) return -1;
    if (s[0] == '+' || s[0] == '-') {
        neg = s[0] == '-';
        i = 1;
    }
    if (i >= n) return -1;
    if (s[i] == '_' || s[n - 1] == '_') return -1;
    uint64_t x = 0;
    for (; i < n; i++) {
        char c = s[i];
        if (c == '_') continue;
        if (c < '0' || c > '9') return -1;
        uint64_t d = (uint64_t)(c - '0');
        if (x > (max - d) / 10) return -1;
        x = x * 10 + d;
    }
    if (neg && x != 0) return -1;
    *out = x;
    return 0;
}

/* Zig std.fmt.parseFloat(f32, tok): correctly rounded; '_' allowed between
 * digits; inf/nan words.  glibc strtof is correctly rounded. */
static int parse_f32(const char *s, size_t n, float *out) {
    char buf[128];
    size_t k = 0;
    if (n == 0 || n >= sizeof buf) return -1;
    if (s[0] == '_' || s[n - 1] == '_') return -1;
    for (size_t i = 0; i < n; i++) {
        if (s[i] == '_') continue;
        if (s[i] == '(' || s[i] == ')') return -1;
        buf[k++] = s[i];
    }
    buf[k] = 0;
    char *end = NULL;
    errno = 0;
    float v = strtof(buf, &end);
    if (end != buf + k) return -1;
    *out = v;
    return 0;
}

/* ------------------------------------------------------------ genome --- */

static inline int is_tok_sep(char c) { return c == ' ' || c == '\t' || c == '\r'; }

/* src/genome.zig:116-188.  Returns 1 = genome, 0 = skipped (null), <0 = error. */
int or_parse_line(const char *line, size_t len, or_genome *out) {
    const char *tok[34];
    size_t tl[34];
    int n = 0;
    size_t i = 0;
    while (i < len) {
        while (i < len && is_tok_sep(line[i])) i++;
        if (i >= len) break;
        size_t b = i;
        while (i < len && !is_tok_sep(line[i])) i++;
        if (n >= 34) return OR_E_BADLINE;
        tok[n] = line + b;
        tl[n] = i - b;
        n++;
    }
    if (!(n == 21 || n == 27 || n == 29 || n == 32 || n == 33 || n == 34)) return 0;
    if (tl[0] > 0 && tok[0][0] == '#') return 0;
    or_genome g;
    memset(&g, 0, sizeof g);
    uint64_t v;
    if (parse_u64(tok[0], tl[0], UINT64_MAX, &v)) return OR_E_PARSE;
    g.id = v;
    if (parse_u64(tok[1], tl[1], UINT32_MAX, &v)) return OR_E_PARSE;
    g.gen = (uint32_t)v;
    if (parse_u64(tok[2], tl[2], UINT64_MAX, &v)) return OR_E_PARSE;
    g.pa = v;
    if (parse_u64(tok[3], tl[3], UINT64_MAX, &v)) return OR_E_PARSE;
    g.pb = v;
    if (parse_u64(tok[4], tl[4], UINT32_MAX, &v)) return OR_E_PARSE;
    g.tribe = (uint32_t)v;
    for (int k = 0; k < 16; k++)
        if (parse_f32(tok[5 + k], tl[5 + k], &g.g[k])) return OR_E_PARSE;
    /* genes 16..22: which token counts fill them (genome.zig:138-163).
     * A 34-token line fills only g[22]; g[16..21] stay 0. */
    if (n == 29 || n == 32 || n == 33) {
        if (parse_f32(tok[21], tl[21], &g.g[16])) return OR_E_PARSE;
        if (parse_f32(tok[22], tl[22], &g.g[17])) return OR_E_PARSE;
    }
    if (n == 32 || n == 33) {
        for (int k = 0; k < 3; k++)
            if (parse_f32(tok[23 + k], tl[23 + k], &g.g[18 + k])) return OR_E_PARSE;
    }
    if (n == 33)
        if (parse_f32(tok[26], tl[26], &g.g[21])) return OR_E_PARSE;
    if (n == 34)
        if (parse_f32(tok[27], tl[27], &g.g[22])) return OR_E_PARSE;
    int rbase = -1;
    if (n == 27) rbase = 21;
    if (n == 29) rbase = 23;
    if (n == 32) rbase = 26;
    if (n == 33) rbase = 27;
    if (n == 34) rbase = 28;
    for (int k = 0; k < OR_NR; k++) {
        if (rbase < 0) {
            g.reg[k] = 0.5f;
        } else if (parse_f32(tok[rbase + k], tl[rbase + k], &g.reg[k])) {
            return OR_E_PARSE;
        }
    }
    *out = g;
    return 1;
}

/* ------------------------------------------------------------- model --- */

/* src/screen.zig:36-95 */
int or_model_from_compact(const float *w, int n, or_model *m) {
    if (!(n == 153 || n == 159 || n == 161 || n == 164 || n == 165 || n == 166)) return OR_E_BADMODEL;
    memset(m, 0, sizeof *m);
    m->w0 = w[0];
    for (int i = 0; i < 16; i++) m->lin[i] = w[1 + i];
    int qstart = 30;
    switch (n) {
    case 166:
        for (int i = 0; i < 7; i++) m->lin[16 + i] = w[17 + i];
        for (int i = 0; i < OR_NR; i++) m->lin[OR_NG + i] = w[24 + i];
        qstart = 30;
        break;
    case 165:
        for (int i = 0; i < 6; i++) m->lin[16 + i] = w[17 + i];
        for (int i = 0; i < OR_NR; i++) m->lin[OR_NG + i] = w[23 + i];
        qstart = 29;
        break;
    case 164:
        for (int i = 0; i < 5; i++) m->lin[16 + i] = w[17 + i];
        for (int i = 0; i < OR_NR; i++) m->lin[OR_NG + i] = w[22 + i];
        qstart = 28;
        break;
    case 161:
        m->lin[16] = w[17];
        m->lin[17] = w[18];
        for (int i = 0; i < OR_NR; i++) m->lin[OR_NG + i] = w[19 + i];
        qstart = 25;
        break;
    case 159:
        /* the six v2 regulator weights land on flat slots 16..21 */
        for (int i = 0; i < OR_NR; i++) m->lin[16 + i] = w[17 + i];
        qstart = 23;
        break;
    default: /* 153 */
        qstart = 17;
        break;
    }
    int idx = qstart;
    for (int i = 0; i < 16; i++)
        for (int j = 0; j <= i; j++) m->quad[i * 16 + j] = w[idx++];
    return OR_OK;
}

static char *read_file_limited(const char *path, size_t limit, size_t *len_out, int *err) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        *err = OR_E_IO;
        return NULL;
    }
    size_t cap = 1 << 16, len = 0;
    char *buf = malloc(cap + 1);
    for (;;) {
        if (len == cap) {
            cap *= 2;
            buf = realloc(buf, cap + 1);
        }
        size_t r = fread(buf + len, 1, cap - len, f);
        len += r;
        if (len > limit) {
            fclose(f);
            free(buf);
            *err = OR_E_IO;
            return NULL;
        }
        if (r == 0) break;
    }
    fclose(f);
    buf[len] = 0;
    *len_out = len;
    return buf;
}

/* src/screen.zig:115-127 */
int or_load_model(const char *path, or_model *m) {
    size_t len;
    int err = 0;
    char *txt = read_file_limited(path, 1u << 20, &len, &err);
    if (!txt) return err;
    const char *p = txt, *e = txt + len;
    const char *tb[200];
    size_t tlen[200];
    int nt = 0;
    while (p < e && nt < 200) {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++;
        if (p >= e) break;
        const char *b = p;
        while (p < e && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++;
        tb[nt] = b;
        tlen[nt] = (size_t)(p - b);
        nt++;
    }
    int rc = OR_E_BADMODEL;
    uint64_t n;
    float w[166];
    if (nt < 1) goto done;
    if (parse_u64(tb[0], tlen[0], UINT64_MAX, &n)) {
        rc = OR_E_PARSE;
        goto done;
    }
    if (!(n == 153 || n == 159 || n == 161 || n == 164 || n == 165 || n == 166)) goto done;
    for (uint64_t i = 0; i < n; i++) {
        if ((int)(1 + i) >= nt) goto done;
        if (parse_f32(tb[1 + i], tlen[1 + i], &w[i])) {
            rc = OR_E_PARSE;
            goto done;
        }
    }
    rc = or_model_from_compact(w, (int)n, m);
done:
    free(txt);
    return rc;
}

/* src/main.zig:201-214: lines split on '\n', trimmed, '#' and bad lines skipped */
int or_load_elites(const char *path, or_genome *out, int cap, int *n_out) {
    size_t len;
    int err = 0;
    char *txt = read_file_limited(path, 1u << 22, &len, &err);
    if (!txt) return err;
    int n = 0;
    size_t i = 0;
    while (i < len) {
        size_t b = i;
        while (i < len && txt[i] != '\n') i++;
        size_t e = i;
        if (i < len) i++;
        while (b < e && is_tok_sep(txt[b])) b++;
        while (e > b && is_tok_sep(txt[e - 1])) e--;
        if (e == b || txt[b] == '#') continue;
        or_genome g;
        if (or_parse_line(txt + b, e - b, &g) == 1) {
            if (n < cap) out[n] = g;
            n++;
        }
    }
    free(txt);
    *n_out = n;
    return n > cap ? OR_E_CAP : OR_OK;
}

/* --------------------------------------------------------------- score --- */

/* src/screen.zig:97-112, no contraction.  Lane j:
 *   acc_j = w0 + lin_j*g_j ; sum_j = ((0 + q[0][j]*(g_j*g_0)) + ...) ;
 *   score = ((lane_0 + lane_1) + ... + lane_15) + (0 + lin_16*g_16 + ... + lin_28*g_28) */
static inline float score_scalar(const or_model *m, const float *g) {
    float lane[16];
    for (int j = 0; j < 16; j++) {
        float acc = m->w0 + m->lin[j] * g[j];
        float sum = 0.0f;
        for (int i = 0; i < 16; i++) sum = sum + m->quad[i * 16 + j] * (g[j] * g[i]);
        lane[j] = acc + sum;
    }
    float s = -0.0f;
    for (int j = 0; j < 16; j++) s = s + lane[j];
    float r = 0.0f;
    for (int i = 16; i < OR_NTOT; i++) r = r + m->lin[i] * g[i];
    return s + r;
}

/* The same function in the form src/screen.zig:97-112 writes it: 16-wide
 * vectors over the lane index j (GCC vector extensions for @Vector(16, f32);
 * each lane's operations and their order are those of or_score, with no
 * contraction), then the ordered lane sum and the linear tail. */
typedef float v16f __attribute__((vector_size(64)));

#define SPLAT16(x) ((v16f){(x), (x), (x), (x), (x), (x), (x), (x), (x), (x), (x), (x), (x), (x), (x), (x)})

static inline float score_simd(const or_model *m, const float *g) {
    v16f gv, lv, wv;
    memcpy(&gv, g, 64);
    memcpy(&lv, m->lin, 64);
    v16f acc = SPLAT16(m->w0);
    acc = acc + lv * gv;
    v16f sum = SPLAT16(0.0f);
    for (int i = 0; i < 16; i++) {
        const v16f row = gv * SPLAT16(g[i]);
        memcpy(&wv, m->quad + i * 16, 64);
        sum = sum + wv * row;
    }
    const v16f lane = acc + sum;
    float s = -0.0f;
    for (int j = 0; j < 16; j++) s = s + lane[j];
    float r = 0.0f;
    for (int i = 16; i < OR_NTOT; i++) r = r + m->lin[i] * g[i];
    return s + r;
}

/* src/genome.zig:292-294: @max(0.0, @min(1.0, x)), i.e. LLVM maxnum/minnum
 * with constant non-NaN operands: a NaN gives the constant (clamp01(NaN) = 1)
 * and an equal pair gives the constant, so clamp01(-0) = +0 (LLVM allows
 * either zero there; the sign of that zero is not pinned by the reference). */
float or_score(const or_model *m, const float *g) { return score_scalar(m, g); }
float or_score_simd(const or_model *m, const float *g) { return score_simd(m, g); }

static inline float clamp01(float x) {
    /* minss(a, b) = a < b ? a : b and maxss(a, b) = a > b ? a : b (the
     * second operand when unordered or equal): branch-free on x86 */
    const __m128 t = _mm_min_ss(_mm_set_ss(x), _mm_set_ss(1.0f));
    return _mm_cvtss_f32(_mm_max_ss(t, _mm_set_ss(0.0f)));
}

float or_clamp01(float x) { return clamp01(x); }

/* One candidate of src/screen.zig:232-249 from draw counter `start`.
 * Returns the genome, the number of draws consumed and the elite indices. */
int or_synth(const or_genome *elites, int ne, uint64_t stream_seed, uint64_t start, float sigma,
             float jump_p, float imm_p, float *g, uint32_t *len_out, uint32_t *ia_out,
             uint32_t *ib_out) {
    uint64_t s = stream_seed + start * GAMMA;
    uint32_t used = 3;
    uint32_t ia = (uint32_t)((sm_next(&s) >> 32) % (uint64_t)ne);
    uint32_t ib = (uint32_t)((sm_next(&s) >> 32) % (uint64_t)ne);
    const or_genome *ea = &elites[ia], *eb = &elites[ib];
    if (unit_of(&s) < imm_p) {
        for (int k = 0; k < OR_NTOT; k++) g[k] = unit_of(&s);
        used += OR_NTOT;
    } else {
        for (int k = 0; k < OR_NTOT; k++) {
            float pa = k < OR_NG ? ea->g[k] : ea->reg[k - OR_NG];
            float pb = k < OR_NG ? eb->g[k] : eb->reg[k - OR_NG];
            float v = (sm_next(&s) & 1) == 0 ? pa : pb;
            float u = unit_of(&s);
            v = v + ((u - 0.5f) * 2.0f) * sigma;
            g[k] = clamp01(v);
        }
        used += 2 * OR_NTOT;
    }
    used += 1;
    if (unit_of(&s) < jump_p) {
        uint32_t idx = (uint32_t)((sm_next(&s) >> 32) % (uint64_t)OR_NTOT);
        g[idx] = unit_of(&s);
        used += 2;
    }
    *len_out = used;
    if (ia_out) *ia_out = ia;
    if (ib_out) *ib_out = ib;
    return 0;
}

/* ---------------------------------------------------------------- heap --- */

typedef struct {
    uint64_t k, n;
    or_slot *slots;
} heap_t;

/* src/screen.zig:166-192; returns 1 when the slot entered the heap */
static int heap_offer(heap_t *h, const or_slot *s) {
    if (h->n < h->k) {
        h->slots[h->n] = *s;
        uint64_t i = h->n;
        h->n++;
        while (i > 0) {
            uint64_t parent = (i - 1) / 2;
            if (h->slots[parent].score <= h->slots[i].score) break;
            or_slot t = h->slots[parent];
            h->slots[parent] = h->slots[i];
            h->slots[i] = t;
            i = parent;
        }
        return 1;
    }
    if (s->score <= h->slots[0].score) return 0;
    h->slots[0] = *s;
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = 2 * i + 2, m = i;
        if (l < h->k && h->slots[l].score < h->slots[m].score) m = l;
        if (r < h->k && h->slots[r].score < h->slots[m].score) m = r;
        if (m == i) break;
        or_slot t = h->slots[m];
        h->slots[m] = h->slots[i];
        h->slots[i] = t;
        i = m;
    }
    return 1;
}

/* -------------------------------------------------------------- worker --- */

/* One job per stream (or chunk).  The jobs sit side by side in one array, so
 * each is aligned to a cache line and the per-candidate count lives in a
 * register: a line shared by two threads' jobs would otherwise bounce between
 * their cores on every candidate, and how much that costs depends on where
 * the scheduler puts the threads. */
typedef struct __attribute__((aligned(64))) {
    const or_model *m;
    const or_genome *elites;
    int ne;
    const or_cfg *cfg;
    uint64_t my_lambda;
    uint64_t seed;
    uint32_t stream;
    int simd; /* score with or_score_simd */
    heap_t heap;
    uint64_t screens;
    /* chunked form: candidates [i0, i0 + my_lambda) of the stream, starting
     * at draw counter start0; every slot the (chunk-local) heap takes is
     * also appended to emit */
    uint64_t i0, start0;
    or_slot *emit;
    uint64_t n_emit, cap_emit;
} job_t;

/* zeroed jobs on cache-line boundaries (sizeof(job_t) is a multiple of 64) */
static job_t *jobs_alloc(uint64_t n) {
    job_t *j = aligned_alloc(64, (n ? n : 1) * sizeof(job_t));
    if (j) memset(j, 0, (n ? n : 1) * sizeof(job_t));
    return j;
}

/* src/screen.zig:225-255; the draw counter is tracked so that each kept slot
 * records where its candidate started. */
static void run_job(job_t *j) {
    uint64_t s = j->seed + j->start0 * GAMMA;
    const int ne = j->ne;
    const float sigma = j->cfg->sigma, jump_p = j->cfg->jump_p, imm_p = j->cfg->imm_p;
    or_slot slot;
    uint64_t start = j->start0, screens = 0;
    for (uint64_t i = j->i0; i < j->i0 + j->my_lambda; i++) {
        const or_genome *ea = &j->elites[(sm_next(&s) >> 32) % (uint64_t)ne];
        const or_genome *eb = &j->elites[(sm_next(&s) >> 32) % (uint64_t)ne];
        float *g = slot.g;
        uint64_t used = 2;
        if (unit_of(&s) < imm_p) {
            for (int k = 0; k < OR_NTOT; k++) g[k] = unit_of(&s);
            used += 1 + OR_NTOT;
        } else {
            /* bit 0 -> parent a, bit 1 -> parent b (a select, not a branch) */
            const or_genome *par[2] = {ea, eb};
            for (int k = 0; k < OR_NTOT; k++) {
                const or_genome *e = par[sm_next(&s) & 1];
                float v = k < OR_NG ? e->g[k] : e->reg[k - OR_NG];
                float u = unit_of(&s);
                v = v + ((u - 0.5f) * 2.0f) * sigma;
                g[k] = clamp01(v);
            }
            used += 1 + 2 * OR_NTOT;
        }
        used += 1;
        if (unit_of(&s) < jump_p) {
            uint32_t idx = (uint32_t)((sm_next(&s) >> 32) % (uint64_t)OR_NTOT);
            g[idx] = unit_of(&s);
            used += 2;
        }
        slot.score = j->simd ? score_simd(j->m, g) : score_scalar(j->m, g);
        slot.pa = ea->id;
        slot.pb = eb->id;
        slot.tribe = ea->tribe;
        slot.stream = j->stream;
        slot.index = i;
        slot.start = start;
        start += used;
        screens++;
        if (heap_offer(&j->heap, &slot) && j->cap_emit) {
            if (j->n_emit == j->cap_emit) {
                j->cap_emit *= 2;
                j->emit = realloc(j->emit, j->cap_emit * sizeof(or_slot));
            }
            j->emit[j->n_emit++] = slot;
        }
    }
    j->screens += screens;
}

typedef struct {
    job_t *jobs;
    uint64_t njobs;
    uint64_t next;
    pthread_mutex_t mu;
} pool_t;

static void *pool_main(void *arg) {
    pool_t *p = (pool_t *)arg;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        uint64_t t = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (t >= p->njobs) break;
        run_job(&p->jobs[t]);
    }
    return NULL;
}

/* stable merge sort, descending by score (std.mem.sort with a.score > b.score) */
static void merge_desc(or_slot *a, or_slot *tmp, uint64_t n) {
    if (n < 2) return;
    uint64_t h = n / 2;
    merge_desc(a, tmp, h);
    merge_desc(a + h, tmp, n - h);
    uint64_t i = 0, j = h, k = 0;
    while (i < h && j < n) {
        if (a[j].score > a[i].score) tmp[k++] = a[j++];
        else tmp[k++] = a[i++];
    }
    while (i < h) tmp[k++] = a[i++];
    while (j < n) tmp[k++] = a[j++];
    memcpy(a, tmp, n * sizeof(or_slot));
}

/* src/screen.zig:257-320.  `workers` OS threads execute the T stream jobs;
 * the result depends only on T (cfg->threads), as in the reference. */
static int screen_jobs(const or_model *m, const or_genome *elites, int ne, const or_cfg *cfg, int workers,
                       int simd, or_slot *out, uint64_t cap, uint64_t *n_out, uint64_t *screens) {
    if (ne < 2) return OR_E_NEEDTWO;
    if (cfg->keep == 0) return OR_E_KEEP0;
    uint64_t T = cfg->threads < cfg->lambda ? cfg->threads : cfg->lambda;
    if (T < 1) T = 1;
    job_t *jobs = jobs_alloc(T);
    uint64_t per = cfg->lambda / T, issued = 0;
    for (uint64_t t = 0; t < T; t++) {
        uint64_t my = (t == T - 1) ? cfg->lambda - issued : per;
        issued += per;
        jobs[t].m = m;
        jobs[t].elites = elites;
        jobs[t].ne = ne;
        jobs[t].cfg = cfg;
        jobs[t].my_lambda = my;
        jobs[t].seed = or_stream_seed(cfg->seed, t);
        jobs[t].stream = (uint32_t)t;
        jobs[t].simd = simd;
        jobs[t].heap.k = cfg->keep;
        jobs[t].heap.n = 0;
        jobs[t].heap.slots = malloc(cfg->keep * sizeof(or_slot));
    }
    if (workers < 1) workers = 1;
    if ((uint64_t)workers > T) workers = (int)T;
    pool_t pool = {jobs, T, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t *th = calloc((size_t)workers, sizeof(pthread_t));
    for (int w = 1; w < workers; w++) pthread_create(&th[w], NULL, pool_main, &pool);
    pool_main(&pool);
    for (int w = 1; w < workers; w++) pthread_join(th[w], NULL);
    free(th);

    uint64_t total = 0, n = 0;
    for (uint64_t t = 0; t < T; t++) {
        total += jobs[t].screens;
        n += jobs[t].heap.n;
    }
    or_slot *all = malloc((n ? n : 1) * sizeof(or_slot));
    or_slot *tmp = malloc((n ? n : 1) * sizeof(or_slot));
    uint64_t k = 0;
    for (uint64_t t = 0; t < T; t++) {
        memcpy(all + k, jobs[t].heap.slots, jobs[t].heap.n * sizeof(or_slot));
        k += jobs[t].heap.n;
        free(jobs[t].heap.slots);
    }
    free(jobs);
    merge_desc(all, tmp, n);
    uint64_t out_n = n < cfg->keep ? n : cfg->keep;
    int rc = OR_OK;
    if (out_n > cap) {
        rc = OR_E_CAP;
        out_n = cap;
    }
    memcpy(out, all, out_n * sizeof(or_slot));
    free(all);
    free(tmp);
    *n_out = out_n;
    *screens = total;
    return rc;
}

int or_screen(const or_model *m, const or_genome *elites, int ne, const or_cfg *cfg, int workers,
              or_slot *out, uint64_t cap, uint64_t *n_out, uint64_t *screens) {
    return screen_jobs(m, elites, ne, cfg, workers, 0, out, cap, n_out, screens);
}

/* or_screen with the 16-wide score form (the reference's SIMD structure). */
int or_screen_simd(const or_model *m, const or_genome *elites, int ne, const or_cfg *cfg, int workers,
                   or_slot *out, uint64_t cap, uint64_t *n_out, uint64_t *screens) {
    return screen_jobs(m, elites, ne, cfg, workers, 1, out, cap, n_out, screens);
}

/* Draw counter after `count` candidates that start at draw counter c: each
 * candidate's length needs only its immigrant test (draw c+3) and its jump
 * test (c+33 or c+62), read in the counter form out(n) = mix(seed + n g). */
static uint64_t skip_candidates(uint64_t seed, uint64_t c, uint64_t count, float imm_p, float jump_p) {
    for (uint64_t i = 0; i < count; i++) {
        uint64_t st = seed + (c + 2) * GAMMA;
        const int imm = unit_of(&st) < imm_p;
        st = seed + (c + (imm ? 32 : 61)) * GAMMA;
        const int jump = unit_of(&st) < jump_p;
        c += (imm ? 33 : 62) + (jump ? 2 : 0);
    }
    return c;
}

/* or_screen in chunks of every stream, for many host threads (the survivors
 * are identical).  Stream t's candidates are cut into `chunks` consecutive
 * chunks; each chunk runs the worker loop with a heap of its own (same
 * capacity and rule) and keeps every slot that heap took.  A candidate that
 * heap rejected has `keep` earlier candidates of its stream scoring at least
 * as much, so the stream's real heap rejects it as well; replaying the kept
 * slots of the chunks in stream order through the real heap therefore gives
 * the real heap exactly, ties and slot order included. */
int or_screen_chunked(const or_model *m, const or_genome *elites, int ne, const or_cfg *cfg, int workers,
                      int chunks, or_slot *out, uint64_t cap, uint64_t *n_out, uint64_t *screens) {
    if (ne < 2) return OR_E_NEEDTWO;
    if (cfg->keep == 0) return OR_E_KEEP0;
    if (chunks < 1) chunks = 1;
    uint64_t T = cfg->threads < cfg->lambda ? cfg->threads : cfg->lambda;
    if (T < 1) T = 1;
    const uint64_t per = cfg->lambda / T;
    const uint64_t nj = T * (uint64_t)chunks;
    job_t *jobs = jobs_alloc(nj);
    uint64_t issued = 0;
    for (uint64_t t = 0; t < T; t++) {
        const uint64_t my = (t == T - 1) ? cfg->lambda - issued : per;
        issued += per;
        const uint64_t seed = or_stream_seed(cfg->seed, t);
        uint64_t i0 = 0, c = 0;
        for (int q = 0; q < chunks; q++) {
            job_t *j = &jobs[t * (uint64_t)chunks + (uint64_t)q];
            const uint64_t n = my / (uint64_t)chunks + ((uint64_t)q < my % (uint64_t)chunks ? 1 : 0);
            j->m = m;
            j->elites = elites;
            j->ne = ne;
            j->cfg = cfg;
            j->my_lambda = n;
            j->seed = seed;
            j->stream = (uint32_t)t;
            j->simd = 1;
            j->heap.k = cfg->keep;
            j->heap.slots = malloc(cfg->keep * sizeof(or_slot));
            j->i0 = i0;
            j->start0 = c;
            j->cap_emit = 1024;
            j->emit = malloc(j->cap_emit * sizeof(or_slot));
            i0 += n;
            c = skip_candidates(seed, c, n, cfg->imm_p, cfg->jump_p);
        }
    }
    if (workers < 1) workers = 1;
    if ((uint64_t)workers > nj) workers = (int)nj;
    pool_t pool = {jobs, nj, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t *th = calloc((size_t)workers, sizeof(pthread_t));
    for (int w = 1; w < workers; w++) pthread_create(&th[w], NULL, pool_main, &pool);
    pool_main(&pool);
    for (int w = 1; w < workers; w++) pthread_join(th[w], NULL);
    free(th);

    /* replay per stream, then the merge of or_screen */
    uint64_t total = 0, n = 0;
    heap_t *real = calloc(T, sizeof(heap_t));
    for (uint64_t t = 0; t < T; t++) {
        real[t].k = cfg->keep;
        real[t].slots = malloc(cfg->keep * sizeof(or_slot));
        for (int q = 0; q < chunks; q++) {
            job_t *j = &jobs[t * (uint64_t)chunks + (uint64_t)q];
            total += j->screens;
            for (uint64_t e = 0; e < j->n_emit; e++) heap_offer(&real[t], &j->emit[e]);
            free(j->emit);
            free(j->heap.slots);
        }
        n += real[t].n;
    }
    free(jobs);
    or_slot *all = malloc((n ? n : 1) * sizeof(or_slot));
    or_slot *tmp = malloc((n ? n : 1) * sizeof(or_slot));
    uint64_t k = 0;
    for (uint64_t t = 0; t < T; t++) {
        memcpy(all + k, real[t].slots, real[t].n * sizeof(or_slot));
        k += real[t].n;
        free(real[t].slots);
    }
    free(real);
    merge_desc(all, tmp, n);
    uint64_t out_n = n < cfg->keep ? n : cfg->keep;
    int rc = OR_OK;
    if (out_n > cap) {
        rc = OR_E_CAP;
        out_n = cap;
    }
    memcpy(out, all, out_n * sizeof(or_slot));
    free(all);
    free(tmp);
    *n_out = out_n;
    *screens = total;
    return rc;
}

/* ------------------------------------------------------------ output --- */

/* Shortest decimal digits M*10^X that read back (strtof) as |v| in f32,
 * the digit string Zig's float formatter starts from. */
static void shortest_f32(float v, uint64_t *M, int *X) {
    if (v == 0.0f) {
        *M = 0;
        *X = 0;
        return;
    }
    char buf[64];
    for (int p = 1; p <= 9; p++) {
        snprintf(buf, sizeof buf, "%.*e", p - 1, (double)v);
        /* buf = d[.ddd]e[+-]xx */
        uint64_t d = 0;
        char *q = buf;
        for (; *q && *q != 'e'; q++)
            if (*q >= '0' && *q <= '9') d = d * 10 + (uint64_t)(*q - '0');
        int e = atoi(q + 1) - (p - 1);
        uint64_t cand[3] = {d, d + 1, d - 1};
        int best = -1;
        long double bestdiff = 0;
        for (int c = 0; c < 3; c++) {
            if (cand[c] == 0) continue;
            char t[64];
            snprintf(t, sizeof t, "%llue%d", (unsigned long long)cand[c], e);
            if (strtof(t, NULL) == v) {
                long double diff = fabsl(strtold(t, NULL) - (long double)v);
                if (best < 0 || diff < bestdiff) {
                    best = c;
                    bestdiff = diff;
                }
            }
            if (c == 0 && best == 0) break;
        }
        if (best >= 0) {
            uint64_t mm = cand[best];
            int xx = e;
            while (mm != 0 && mm % 10 == 0) {
                mm /= 10;
                xx++;
            }
            *M = mm;
            *X = xx;
            return;
        }
    }
    /* unreachable for finite f32: 9 digits always round-trip */
    *M = 0;
    *X = 0;
}

static int ndigits(uint64_t x) {
    int n = 1;
    while (x >= 10) {
        x /= 10;
        n++;
    }
    return n;
}

/* Zig std 0.17 `{d:.4}` on an f32: shortest digits, then round half-up at the
 * digit after the 4th decimal (format_float round(.decimal)), then print the
 * remaining digits with the decimal point and zero padding to 4 places. */
int or_fmt_d4(float v, char *buf, size_t cap) {
    if (isnan(v)) return snprintf(buf, cap, "nan");
    if (isinf(v)) return snprintf(buf, cap, v < 0 ? "-inf" : "inf");
    int neg = signbit(v) != 0;
    uint64_t M;
    int X;
    shortest_f32(fabsf(v), &M, &X);
    const int prec = 4;
    int olen = ndigits(M);
    int round_digit = 0;
    if (X > 0) {
        round_digit = (olen - 1) + prec + X;
    } else {
        int mer = -X;
        if (prec + olen > mer) round_digit = prec + olen - mer;
    }
    if (round_digit < olen) {
        for (int k = round_digit + 1; k < olen; k++) {
            M /= 10;
            X++;
        }
        if (M % 10 >= 5) {
            M = M / 10 + 1;
        } else {
            M /= 10;
        }
        X++;
    }
    char digits[32];
    snprintf(digits, sizeof digits, "%llu", (unsigned long long)M);
    olen = (int)strlen(digits);
    char out[96];
    int o = 0;
    if (neg) out[o++] = '-';
    int dp = X + olen;
    if (dp <= 0) {
        out[o++] = '0';
        out[o++] = '.';
        for (int k = 0; k < -dp; k++) out[o++] = '0';
        memcpy(out + o, digits, (size_t)olen);
        o += olen;
        int written = -dp + olen;
        for (int k = written; k < prec; k++) out[o++] = '0';
    } else if (dp >= olen) {
        memcpy(out + o, digits, (size_t)olen);
        o += olen;
        for (int k = 0; k < dp - olen; k++) out[o++] = '0';
        out[o++] = '.';
        for (int k = 0; k < prec; k++) out[o++] = '0';
    } else {
        memcpy(out + o, digits, (size_t)dp);
        o += dp;
        out[o++] = '.';
        memcpy(out + o, digits + dp, (size_t)(olen - dp));
        o += olen - dp;
        for (int k = olen - dp; k < prec; k++) out[o++] = '0';
    }
    out[o] = 0;
    return snprintf(buf, cap, "%s", out);
}

/* src/screen.zig:326-337: "0 0 pa pb tribe" + g[0..23] at {d:.4} */
int or_write_promoted(const char *path, const or_slot *s, uint64_t n) {
    FILE *f = fopen(path, "wb");
    if (!f) return OR_E_IO;
    char g[64];
    for (uint64_t i = 0; i < n; i++) {
        fprintf(f, "0 0 %llu %llu %u", (unsigned long long)s[i].pa, (unsigned long long)s[i].pb,
                s[i].tribe);
        for (int k = 0; k < 24; k++) {
            or_fmt_d4(s[i].g[k], g, sizeof g);
            fprintf(f, " %s", g);
        }
        fputc('\n', f);
    }
    fclose(f);
    return OR_OK;
}

/* src/screen.zig:338-341 */
int or_write_count(const char *path, uint64_t total) {
    FILE *f = fopen(path, "wb");
    if (!f) return OR_E_IO;
    fprintf(f, "%llu\n", (unsigned long long)total);
    fclose(f);
    return OR_OK;
}
