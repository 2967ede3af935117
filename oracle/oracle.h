/*
 * oracle.h -- CPU restatement of the codec-arena screen (TEST INFRASTRUCTURE).
 *
 * This library is the parity checker for the HIP screening funnel.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only to check or time against; the product path never links it.
 *
 * Every function restates the reference algorithm and cites the line it
 * follows (paths relative to the codec-arena snapshot):
 *   src/screen.zig   -- Model.fromCompact / score, loadModel, Heap, Rng, worker, run
 *   src/genome.zig   -- Genome, parseLine, clamp01
 *   src/main.zig     -- screenCmd elites loading
 * Zig std 0.17.0 (not vendored in the reference) supplies parseInt/parseFloat,
 * std.mem.sort (stable block sort) and the {d:.4} float formatter; their
 * published behaviour is restated where it touches the output.
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NG 23
#define OR_NR 6
#define OR_NTOT 29

typedef struct {
    uint64_t id;
    uint32_t gen;
    uint32_t tribe;
    uint64_t pa;
    uint64_t pb;
    float g[OR_NG];
    float reg[OR_NR];
} or_genome;

typedef struct {
    float w0;
    float lin[OR_NTOT];
    float quad[256];
} or_model;

typedef struct {
    uint64_t lambda;
    uint64_t keep;
    uint64_t seed;
    uint64_t threads;
    float sigma;
    float jump_p;
    float imm_p;
    float _pad;
} or_cfg;

/* One heap slot of the reference plus where the candidate came from. */
typedef struct {
    float score;
    float g[OR_NTOT];
    uint64_t pa;
    uint64_t pb;
    uint32_t tribe;
    uint32_t stream;
    uint64_t index;  /* candidate index within its stream */
    uint64_t start;  /* draw counter before the candidate's first draw */
} or_slot;

/* error codes (mirroring the Zig error set) */
#define OR_OK 0
#define OR_E_BADMODEL (-1)
#define OR_E_NEEDTWO (-2)
#define OR_E_BADARGS (-3)
#define OR_E_IO (-4)
#define OR_E_BADLINE (-5)
#define OR_E_PARSE (-6)
#define OR_E_KEEP0 (-7)
#define OR_E_CAP (-8)

uint64_t or_splitmix_next(uint64_t *state);
void or_splitmix_fill(uint64_t seed, uint64_t n, uint64_t *out);
uint64_t or_stream_seed(uint64_t seed, uint64_t t);

int or_parse_line(const char *line, size_t len, or_genome *out);
int or_model_from_compact(const float *w, int n, or_model *out);
int or_load_model(const char *path, or_model *out);
int or_load_elites(const char *path, or_genome *out, int cap, int *n_out);
float or_score(const or_model *m, const float *g);
float or_clamp01(float x);
float or_score_simd(const or_model *m, const float *g);

int or_synth(const or_genome *elites, int ne, uint64_t stream_seed, uint64_t start,
             float sigma, float jump_p, float imm_p, float *g_out, uint32_t *len_out,
             uint32_t *ia_out, uint32_t *ib_out);

int or_screen(const or_model *m, const or_genome *elites, int ne, const or_cfg *cfg,
              int workers, or_slot *out, uint64_t cap, uint64_t *n_out, uint64_t *screens);

int or_screen_simd(const or_model *m, const or_genome *elites, int ne, const or_cfg *cfg,
                   int workers, or_slot *out, uint64_t cap, uint64_t *n_out, uint64_t *screens);

int or_screen_chunked(const or_model *m, const or_genome *elites, int ne, const or_cfg *cfg, int workers,
                      int chunks, or_slot *out, uint64_t cap, uint64_t *n_out, uint64_t *screens);

int or_fmt_d4(float v, char *buf, size_t cap);
int or_write_promoted(const char *path, const or_slot *s, uint64_t n);
int or_write_count(const char *path, uint64_t total);

#ifdef __cplusplus
}
#endif
#endif
