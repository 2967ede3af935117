/*
 * csfunnel.h -- C ABI of the MI355X (gfx950) screening funnel for codec-arena.
 *
 * The reference has no FFI on this path: brain/evolve.py:431-442 runs the Zig
 * CLI `arena screen ...` as a subprocess and reads two files back.  Each entry
 * point below replaces one piece of that CLI; the python package
 * (codec-arena_amd/csfunnel) binds them with ctypes and provides the
 * flag-compatible `arena screen` command.
 *
 * Conventions: every function returns an int status (CS_OK = 0, negative =
 * error; cs_last_error() gives a message, per thread).  All buffers are
 * caller-owned host memory; device memory is owned by a cs_ctx.  Calls are
 * synchronous.  A cs_ctx is not re-entrant: use one per host thread.
 */
#ifndef CSFUNNEL_H
#define CSFUNNEL_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_NG 23   /* effector genes   (src/genome.zig:14) */
#define CS_NR 6    /* regulator genes  (src/genome.zig:15) */
#define CS_NTOT 29 /* flat gene vector (src/genome.zig:16) */

#define CS_OK 0
#define CS_E_BADMODEL (-1)  /* error.BadModel      src/screen.zig:37,120,123 */
#define CS_E_NEEDTWO (-2)   /* error.NeedTwoElites src/main.zig:214         */
#define CS_E_BADARGS (-3)   /* error.BadArgs       src/main.zig:165         */
#define CS_E_IO (-4)        /* file errors / size limits                     */
#define CS_E_BADLINE (-5)   /* error.BadLine       src/genome.zig:121       */
#define CS_E_PARSE (-6)     /* parseInt / parseFloat errors                  */
#define CS_E_KEEP0 (-7)     /* keep == 0: undefined in the reference (R3)    */
#define CS_E_CAP (-8)       /* caller buffer too small                       */
#define CS_E_HIP (-9)       /* HIP runtime error                             */
#define CS_E_OOM (-10)      /* device allocation failed                      */
#define CS_E_INTERNAL (-11) /* internal consistency check failed             */
#define CS_E_NODEV (-12)    /* no usable gfx950 device                       */

/* Genome: src/genome.zig:78-86 (regulators default to 0.5). */
typedef struct cs_genome {
    uint64_t id;
    uint32_t gen;
    uint32_t tribe;
    uint64_t pa;
    uint64_t pb;
    float g[CS_NG];
    float reg[CS_NR];
} cs_genome;

/* Flat model: src/screen.zig:23-26 (w_quad lower triangle, strict upper = 0). */
typedef struct cs_model {
    float w0;
    float lin[CS_NTOT];
    float quad[256];
} cs_model;

/* src/screen.zig:129-137.  `threads` is semantic: it fixes the RNG stream
 * partition (src/screen.zig:267-287), not host or device concurrency. */
typedef struct cs_screen_cfg {
    uint64_t lambda;
    uint64_t keep;
    uint64_t seed;
    uint64_t threads;
    float sigma;
    float jump_p;
    float imm_p;
    float _pad;
} cs_screen_cfg;

/* One survivor in reference output order: the heap Slot of
 * src/screen.zig:139-149 plus where its candidate came from. */
typedef struct cs_survivor {
    float score;
    float g[CS_NTOT];
    uint64_t pa;
    uint64_t pb;
    uint32_t tribe;
    uint32_t stream; /* RNG stream (reference worker thread) */
    uint64_t index;  /* candidate index within the stream */
    uint64_t start;  /* stream draw counter before the candidate's first draw */
} cs_survivor;

/* Per-call device statistics of the last cs_screen (milliseconds are wall
 * time of the device stages measured with HIP events). */
typedef struct cs_screen_stats {
    uint64_t candidates;   /* counted by the scoring kernel */
    uint64_t positions;    /* RNG draws covered by the offset maps */
    uint64_t blocks;       /* 65536-draw blocks mapped */
    uint64_t records;      /* candidates that passed the running bound */
    uint32_t levels;       /* scoring launches (bound refresh points) */
    uint32_t retries;      /* overflowed launches that were split */
    double ms_map;         /* offset-map kernel */
    double ms_scan;        /* map composition (reduce + down sweeps) */
    double ms_score;       /* synthesis + score + filter kernels */
    double ms_replay;      /* host heap replay + copies */
    double ms_total;
    uint32_t offset_path;  /* 1 = speculative segment walk, 0 = transfer maps */
    uint32_t walk_fixes;   /* walk segments re-walked after verification */
} cs_screen_stats;

typedef struct cs_ctx cs_ctx;

int cs_version(void);
const char *cs_last_error(void);

/* Zig std.fmt.parseInt(u64, tok, 10) / parseFloat(f32, tok) as used by the
 * screen CLI (src/main.zig:178-198) and the parsers below.  0 ok, <0 error. */
int cs_parse_u64(const char *tok, size_t len, uint64_t *out);
int cs_parse_f32(const char *tok, size_t len, float *out);

/* src/genome.zig:116-188 (including the v6 parse of tokens 21..26 as 0).
 * Returns 1 = parsed, 0 = skipped line (wrong token count or '#'), <0 error. */
int cs_parse_pop_line(const char *line, size_t len, cs_genome *out);

/* src/screen.zig:36-95 (Model.fromCompact), n in {153,159,161,164,165,166}. */
int cs_model_from_compact(const float *w, int n, cs_model *out);

/* src/screen.zig:115-127 (loadModel, file <= 1 MiB). */
int cs_load_model(const char *path, cs_model *out);

/* src/main.zig:201-213 (elites file <= 4 MiB, bad lines skipped).
 * *n_out receives the number of genomes; CS_E_CAP if it exceeds cap. */
int cs_load_elites(const char *path, cs_genome *out, int cap, int *n_out);

/* Device context on HIP device `device`. */
int cs_ctx_create(int device, cs_ctx **out);
/* Fails with CS_E_BADARGS while MLP handles created on ctx are alive. */
int cs_ctx_destroy(cs_ctx *ctx);

/* src/screen.zig:257-320 (run): synthesize cfg->lambda candidates in
 * cfg->threads splitmix64 streams on the GPU, score them with the quadratic
 * surface and return the top cfg->keep in the reference's output order.
 * *screens receives the number of candidates scored (== lambda). */
int cs_screen(cs_ctx *ctx, const cs_model *model, const cs_genome *elites, int ne,
              const cs_screen_cfg *cfg, cs_survivor *out, uint64_t cap, uint64_t *n_out,
              uint64_t *screens);

/* No counterpart in the reference.  cs_screen for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same survivors, but every device
 * buffer comes from a static 64 MiB scratch array in the library's code
 * object (one per GPU, held until the process ends), so the call makes no
 * hipMalloc / hipFree.  Record buffers hold at most 2^16 records (larger
 * launches are split and redone).  Calls beyond the scratch size fail with
 * CS_E_CAP before any scoring launch: the speculative-walk path needs about
 * 7.5 bytes per 1000 candidates plus 20 MB (2.65e9 candidates: 40 MB, the
 * limit is about 6e9); the transfer-map path (rates below 0.004, or forced)
 * also 512 bytes per 65536-draw block, so about 9e7 candidates at most.
 * Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU race.  Host
 * bookkeeping (heaps, record replay) still uses the C++ heap. */
int cs_screen_noalloc(cs_ctx *ctx, const cs_model *model, const cs_genome *elites, int ne,
                      const cs_screen_cfg *cfg, cs_survivor *out, uint64_t cap, uint64_t *n_out, uint64_t *screens);

int cs_get_stats(cs_ctx *ctx, cs_screen_stats *out);

/* Testing hook, no counterpart in the reference: records per record buffer
 * of the scoring launches (default 2^20; cap in [2048, 2^20]).  A launch that
 * emits more is split by streams, then by blocks, and redone. */
int cs_set_record_cap(cs_ctx *ctx, uint32_t cap);

/* How candidate offsets are recovered: 0 = automatic (speculative walk when
 * both event rates are >= 0.004, verified, else transfer maps), 1 = always
 * transfer maps.  Both give identical survivors; exposed for tests and
 * measurement. */
int cs_set_offset_mode(cs_ctx *ctx, int mode);

/* ---- Learned surrogate (no counterpart in the reference) --------------
 * A 16 -> 256 -> 256 -> 1 GELU MLP over the 16 legacy effector genes, trained
 * on the device with BF16 MFMA forward/backward kernels (FP32 accumulation,
 * FP32 master weights, AdamW as torch.optim.AdamW) and used as an alternative
 * screen scorer in BF16 or FP8 (E4M3: a per-tensor scale for the layer-2
 * activations, per-block E8M0 scales for W2). Gene and target
 * pointers are DEVICE memory: genes [n][ldg] f32 (the first 16 columns are
 * the inputs), targets [n] f32. */
typedef struct cs_mlp cs_mlp;
#define CS_MLP_PARAMS 70401 /* W1[16][256], b1[256], W2[256][256], b2, w3, b3 */

int cs_mlp_create(cs_ctx *ctx, uint64_t seed, cs_mlp **out);
int cs_mlp_destroy(cs_mlp *mlp);
int cs_mlp_get_params(cs_mlp *mlp, float *host_params);
int cs_mlp_set_params(cs_mlp *mlp, const float *host_params);
/* FP8 screen: H1 is multiplied by h_scale before E4M3 rounding; h_scale must
 * be a power of two in [2^-15, 2^15] (else CS_E_BADARGS).  W2 needs no scale
 * from the caller: it is rounded in blocks of 32 k values, each with its own
 * E8M0 scale (the layer-2 MFMA's block scales). */
int cs_mlp_set_fp8_scale(cs_mlp *mlp, float h_scale);
/* out[n] (device) = MLP(genes); *h1max (optional) = max |H1| for FP8 calibration */
int cs_mlp_forward(cs_mlp *mlp, const float *genes, int ldg, uint64_t n, float *out, float *h1max);
/* loss = mean (MLP(genes) - y)^2 and its gradient (host_grads, optional) */
int cs_mlp_grad(cs_mlp *mlp, const float *genes, int ldg, const float *y, int n, float *loss, float *host_grads);
/* one AdamW step (beta 0.9 / 0.999, eps 1e-8) on a batch; *loss = batch MSE before the step */
int cs_mlp_train_step(cs_mlp *mlp, const float *genes, int ldg, const float *y, int n, float lr, float wd,
                      float *loss);

#define CS_SCORER_QUAD 0     /* src/screen.zig:97-112, exact */
#define CS_SCORER_MLP_BF16 1 /* learned surrogate, BF16 MFMA */
#define CS_SCORER_MLP_FP8 2  /* learned surrogate, FP8 MFMA for layer 2 */
#define CS_SCORER_MLP_F64 3  /* learned surrogate in FP64 (FP64 MFMA layer 2): the
                                reference the reduced-precision screens' top-k
                                recall is measured against */

typedef struct cs_screen_opts {
    int scorer;        /* CS_SCORER_* */
    int _pad;
    cs_mlp *mlp;       /* required by the MLP scorers */
    float *dump_genes;  /* optional DEVICE [lambda][29]: every candidate's genes, stream order */
    float *dump_scores; /* DEVICE [lambda], required with dump_genes: every candidate's score */
    double *scores64;   /* optional HOST [cap]: each survivor's ranking key (the FP64
                           score for CS_SCORER_MLP_F64, the f32 score otherwise) */
} cs_screen_opts;

/* cs_screen with a choice of scorer and optional dumps; the synthesis,
 * stream partition, heaps and output order are those of cs_screen.  model may
 * be NULL for the MLP scorers. */
int cs_screen_ex(cs_ctx *ctx, const cs_model *model, const cs_genome *elites, int ne, const cs_screen_cfg *cfg,
                 const cs_screen_opts *opts, cs_survivor *out, uint64_t cap, uint64_t *n_out, uint64_t *screens);

/* ---- FP64 ridge refit (brain/evolve.py:211-268) on the device ----------
 * Pointers are DEVICE memory (e.g. torch tensors); ng = 16 selects the
 * 153-feature layout [1, g0..g15, g_i g_j (j<=i)], ng = 29 the 166-feature
 * layout of brain/evolve.py:211-220.  Calls synchronize the context stream
 * before returning; the caller must have finished writing the inputs. */

/* X'X (p x p, row-major), X'y (p) and y'y of the feature matrix X of
 * genes [n][ng] (the products of brain/evolve.py:264-265). */
int cs_ridge_gram_f64(cs_ctx *ctx, const double *genes, int ng, const double *y, uint64_t n, double *gram,
                      double *xty, double *yy);

/* w = solve(gram + alpha I, xty + alpha prior) by Cholesky
 * (brain/evolve.py:266); prior may be NULL (zero). */
int cs_ridge_solve_f64(cs_ctx *ctx, const double *gram, const double *xty, int p, double alpha,
                       const double *prior, double *w);

/* No counterpart in the reference: the mean ridge loss
 *   L(w) = (|X w - y|^2 + alpha |w - prior|^2) / n
 * whose minimizer is refit_surrogate's solution, and dL/dw, in one pass over
 * the data.  loss: device double[1]; grad: device double[p]. */
int cs_ridge_loss_grad_f64(cs_ctx *ctx, const double *genes, int ng, const double *y, uint64_t n,
                           const double *w, const double *prior, double alpha, double *loss, double *grad);

/* Zig `{d:.4}` of an f32 (src/screen.zig:327). Returns the length written. */
int cs_format_gene(float v, char *buf, size_t cap);

/* src/screen.zig:322-337: "0 0 pa pb tribe" + 24 genes at {d:.4} per line. */
int cs_write_promoted(const char *path, const cs_survivor *s, uint64_t n);

/* No counterpart in the reference (its writer above drops genes 24..28, so
 * brain/evolve.py:449, which keeps only 5 + 29 = 34-token lines, promotes
 * none of them): the same lines with all 29 genes at {d:.4}, the v6 token
 * count that src/genome.zig:116-188 and brain/evolve.py:445-454 read. */
int cs_write_promoted_v6(const char *path, const cs_survivor *s, uint64_t n);

/* src/screen.zig:338-341: "<total>\n". */
int cs_write_count(const char *path, uint64_t total);

/* No counterpart in the reference: a measurement probe for the rooflines of
 * the screen kernels.  Times one instruction class (8 per iteration: 0
 * v_add_u32, 1 v_mul_lo_u32, 2 v_mul_hi_u32, 3 v_mad_u64_u32, 4 v_mul_f32,
 * 5 v_lshl_add_u64, 6 v_alignbit_b32, 7 v_pk_fma_f16, 8 v_mfma_f32_16x16x32_bf16,
 * 9 the same MFMAs with 32 v_pk_fma_f16 beside them, 10 v_mfma_f32_16x16x128_f8f6f4
 * on E4M3 operands, 11 v_lshrrev_b64) at 8 waves per SIMD on
 * `device`; returns the time per iteration per SIMD (ns), the s_memtime
 * counts per iteration per SIMD (median over waves) and the SIMD count. */
int cs_valu_rate_probe(int device, int op, uint32_t iters, double *ns_per_iter, double *cycles_per_iter,
                       uint32_t *simds);

#ifdef __cplusplus
}
#endif
#endif
