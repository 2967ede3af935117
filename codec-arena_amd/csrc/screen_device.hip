// screen_device.hip -- gfx950 screening funnel: candidate synthesis, quadratic
// scoring and exact top-K for the codec-arena `arena screen` call.
//
// Reference semantics (codec-arena snapshot):
//   src/screen.zig:195-213  splitmix64 stream, unit = (out >> 40) * 2^-24
//   src/screen.zig:225-255  worker: per candidate 2 parent draws, immigrant
//                           test, 29 x (crossover bit, mutation unit) or
//                           29 immigrant units, jump test (+ 2 jump draws)
//   src/screen.zig:97-112   Model.score, strict f32, ordered lane sum
//   src/screen.zig:151-193  per-stream bounded min-heap
//   src/screen.zig:257-320  stream partition, merge, stable sort
//
// Device pipeline for one call (all streams at once):
//   1. k_map      every 65536-draw block of every stream: immigrant/jump event
//                 bits of every draw (LDS bitmaps), then the transfer map
//                 phase -> (exit phase, candidates) for all 64 entry phases.
//   2. k_reduce / k_down   compose the maps in groups of 64 up to one map per
//                 stream, then push the true entry (phase, first candidate
//                 index) back down to every block.
//   3. k_score    per block (one wavefront): walk the true candidate chain with
//                 ballot speculation, synthesize each candidate in registers,
//                 score it and emit only candidates above the stream's running
//                 bound.  Launched in geometrically growing block ranges; the
//                 host replays the emitted candidates through the reference
//                 heap between launches, so the bound is the heap root of the
//                 prefix already processed and the final heaps are exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device_ctx.h"
#include "mlp_common.h"

namespace csf {

constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;
constexpr uint64_t MIX1 = 0xBF58476D1CE4E5B9ull;
constexpr uint64_t MIX2 = 0x94D049BB133111EBull;

constexpr int BLOCK_LOG = 16;
constexpr uint64_t BLOCK = 1ull << BLOCK_LOG;  // draws per block
// A scoring run covers up to two blocks.  Queue codes of its candidates:
// start - run start (< 2 BLOCK + 62: 18 bits), index - first index (at most
// 2 BLOCK / 33 < 4096 candidates: 12 bits), immigrant flag (bit 30), jump
// flag (bit 31).
constexpr int RUN_POS_BITS = 18;
constexpr int RUN_IDX_BITS = 12;
static_assert(2 * BLOCK + 62 <= (1ull << RUN_POS_BITS) && 2 * BLOCK / 33 < (1ull << RUN_IDX_BITS) &&
                  RUN_POS_BITS + RUN_IDX_BITS == 30,
              "queue code layout");
constexpr int SUBSEG = 4096;                   // draws per walked sub-segment
constexpr int NSUB = (int)(BLOCK / SUBSEG);    // 16
constexpr int BITWORDS = (int)(BLOCK / 64) + 1;  // + 64 draws of look-ahead

struct StreamInfo {
    uint64_t seed;       // seed ^ (gamma * (t + 1))
    uint64_t my_lambda;  // candidates of this stream
    uint64_t blk_first;  // index of the stream's first block in block arrays
    uint64_t nblocks;    // blocks mapped for this stream
    uint64_t cand_base;  // candidates of the streams before this one
};

struct LaunchSeg {  // one stream's block range in a scoring launch
    uint64_t flat_first;
    uint64_t b0;
    float bound;  // NaN: accept every candidate (heap not full yet)
    uint32_t _pad;
    double bound64;  // the same bound for the FP64 scorer
};

struct DevModel {
    float w0;
    float lin[CS_NTOT];
    float tri[136];  // lower triangle, row-major: (i, j<=i) at i*(i+1)/2 + j
};
typedef const __attribute__((address_space(4))) DevModel *ConstModel;

// The quadratic model as the LDS-elites scorer keeps it in LDS, in the order
// the score reads it: w0, then for each lane j = 0..15 the run lin[j],
// w(j,j), w(j+1,j), ..., w(15,j) (17 - j values), then lin[16..28].
constexpr int LDS_MODEL_FLOATS = 1 + 152 + 13;
constexpr int LDS_MODEL_BYTES = (LDS_MODEL_FLOATS * 4 + 15) / 16 * 16;
__host__ __device__ constexpr int lds_model_run(int j) { return 1 + 17 * j - j * (j - 1) / 2; }
// DevModel index of LDS model slot i
__device__ inline int lds_model_src(int i) {
    if (i == 0) return 0;  // w0
    if (i >= 1 + 152) return 1 + 16 + (i - 153);  // lin[16..28]
    int j = 0;
    while (j < 15 && lds_model_run(j + 1) <= i) ++j;
    const int r = i - lds_model_run(j);  // 0: lin[j]; r >= 1: w(j + r - 1, j)
    if (r == 0) return 1 + j;
    const int row = j + r - 1;
    return 1 + CS_NTOT + row * (row + 1) / 2 + j;
}

// ------------------------------------------------------------ splitmix64

// state after n increments is seed + n*gamma; draw n is mix(that).  The
// mixer is written on 32-bit halves so that each 64x64 product keeps only the
// words needed.  The state arrives as a register pair (one v_lshl_add_u64
// from the generation's base), so its shift is one v_lshrrev_b64 (one
// instruction where alignbit + shift are two); the product's halves come
// from separate instructions, so its shift stays alignbit + shift (a 64-bit
// shift would first need moves to pair them).
struct Mid {
    uint32_t lo, hi;  // u = x ^ (x >> 27), x = (s ^ (s >> 30)) * MIX1
};
__device__ __forceinline__ Mid mix_mid(uint64_t s) {
    const uint64_t t = s ^ (s >> 30);
    const uint32_t tlo = (uint32_t)t, thi = (uint32_t)(t >> 32);
    const uint64_t p = (uint64_t)tlo * (uint32_t)MIX1;
    const uint32_t xlo = (uint32_t)p;
    const uint32_t xhi = (uint32_t)(p >> 32) + tlo * (uint32_t)(MIX1 >> 32) + thi * (uint32_t)MIX1;
    return Mid{xlo ^ __builtin_amdgcn_alignbit(xhi, xlo, 27), xhi ^ (xhi >> 27)};
}
// high word of z = u * MIX2
__device__ __forceinline__ uint32_t mix_zhi(Mid u) {
    return __umulhi(u.lo, (uint32_t)MIX2) + u.lo * (uint32_t)(MIX2 >> 32) + u.hi * (uint32_t)MIX2;
}
// (out >> 32): parent / jump indices
__device__ __forceinline__ uint32_t draw_hi(uint64_t s) {
    const uint32_t h = mix_zhi(mix_mid(s));
    return h ^ (h >> 31);
}
// (out >> 40): the 24-bit unit numerator
__device__ __forceinline__ uint32_t draw_u24(uint64_t s) { return mix_zhi(mix_mid(s)) >> 8; }
// (out & 1): crossover bit, = bit0(z) ^ bit31(z) with z = u * MIX2.  MIX2 is
// odd, so bit0(z) = bit0(u.lo), and adding u.lo 2^31 (mod 2^32) to the low
// word of z flips its bit 31 exactly when bit0(u.lo) = 1: the bit is bit 31
// of u.lo * (MIX2 + 2^31) mod 2^32, i.e. the sign of one 32-bit product.
__device__ __forceinline__ int32_t draw_bit_signed(uint64_t s) {
    return (int32_t)(mix_mid(s).lo * (uint32_t)(MIX2 + 0x80000000ull));
}
__device__ __forceinline__ uint32_t draw_bit(uint64_t s) { return (uint32_t)draw_bit_signed(s) >> 31; }
// (out >> 40) - 2^23 as a signed integer: the unit numerator m centred.
// (zhi + 2^31) >> 8, arithmetic, is m - 2^23; the 2^31 rides in the 64-bit
// addend of the product that forms the high word.
__device__ __forceinline__ int32_t draw_centered(uint64_t s) {
    const Mid u = mix_mid(s);
    const uint64_t p = (uint64_t)u.lo * (uint32_t)MIX2 + 0x8000000000000000ull;
    const uint32_t zb = (uint32_t)(p >> 32) + u.lo * (uint32_t)(MIX2 >> 32) + u.hi * (uint32_t)MIX2;
    return (int32_t)zb >> 8;
}
__device__ __forceinline__ float unit_f(uint32_t m) { return (float)m * 0x1p-24f; }

// a % d for 32-bit a by Lemire, Kaser & Kurz, "Faster remainder by direct
// computation" (2019): M = floor((2^64-1)/d) + 1.
__device__ __forceinline__ uint32_t fastmod(uint32_t a, uint64_t M, uint32_t d) {
    return (uint32_t)__umul64hi(M * (uint64_t)a, (uint64_t)d);
}

// base + the number of bits of m below this lane (two v_mbcnt)
__device__ __forceinline__ uint32_t below(uint64_t m, uint32_t base) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}

__device__ __forceinline__ int find_seg(const uint64_t *first, int T, uint64_t f) {
    int lo = 0, hi = T;  // first[lo] <= f < first[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (first[mid] <= f) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t bit_at(const uint64_t *w, uint32_t p) {
    return (uint32_t)(w[p >> 6] >> (p & 63)) & 1u;
}

// ---------------------------------------------------------------- k_map

// One workgroup (4 waves) per block.  Bits: word i, bit l <-> draw
// block_start + 64 i + l.  Candidate starts c (the draw counter before the
// candidate) in [block_start, block_start + BLOCK) need the immigrant bit at
// c+3 and the jump bit at c+33 (immigrant) or c+62.
__global__ __launch_bounds__(256) void k_map(const StreamInfo *__restrict__ streams,
                                             const uint64_t *__restrict__ blk_first, int T,
                                             uint64_t nblk, uint32_t immT, uint32_t jumpT,
                                             uint64_t *__restrict__ maps) {
    __shared__ uint64_t imm_bits[BITWORDS];
    __shared__ uint64_t jmp_bits[BITWORDS];
    __shared__ uint16_t sub_map[NSUB][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint64_t f = blockIdx.x; f < nblk; f += gridDim.x) {
        const int t = find_seg(blk_first, T, f);
        const uint64_t b = f - blk_first[t];
        const uint64_t seed = streams[t].seed;
        const uint64_t p0 = b << BLOCK_LOG;
        // 1. event bits: wave w covers words [256 w, 256 w + 256); wave 3 also word 1024
        uint64_t s = seed + (p0 + (uint64_t)wave * 16384u + (uint64_t)lane) * GAMMA;
        const uint64_t step = 64ull * GAMMA;
#pragma unroll 4
        for (int it = 0; it < 256; ++it) {
            const uint32_t m = draw_u24(s);
            const uint64_t bi = __ballot(m < immT);
            const uint64_t bj = __ballot(m < jumpT);
            if (lane == 0) {
                imm_bits[wave * 256 + it] = bi;
                jmp_bits[wave * 256 + it] = bj;
            }
            s += step;
        }
        if (wave == 3) {
            const uint32_t m = draw_u24(s);
            const uint64_t bi = __ballot(m < immT);
            const uint64_t bj = __ballot(m < jumpT);
            if (lane == 0) {
                imm_bits[1024] = bi;
                jmp_bits[1024] = bj;
            }
        }
        __syncthreads();
        // 2. walk: lane = entry phase, 4 interleaved sub-segments per wave
        uint32_t c[4], e[4], n[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t base = (uint32_t)(wave * 4 + q) * SUBSEG;
            c[q] = base + lane;
            e[q] = base + SUBSEG;
            n[q] = 0;
        }
        while (c[0] < e[0] || c[1] < e[1] || c[2] < e[2] || c[3] < e[3]) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (c[q] < e[q]) {
                    const uint32_t im = bit_at(imm_bits, c[q] + 3);
                    const uint32_t j33 = bit_at(jmp_bits, c[q] + 33);
                    const uint32_t j62 = bit_at(jmp_bits, c[q] + 62);
                    const uint32_t L = im ? (j33 ? 35u : 33u) : (j62 ? 64u : 62u);
                    c[q] += L;
                    n[q] += 1;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) sub_map[wave * 4 + q][lane] = (uint16_t)((n[q] << 6) | (c[q] - e[q]));
        __syncthreads();
        // 3. compose the 16 sub-maps into the block map
        if (wave == 0) {
            uint32_t ph = (uint32_t)lane;
            uint64_t cnt = 0;
#pragma unroll
            for (int k = 0; k < NSUB; ++k) {
                const uint32_t v = sub_map[k][ph];
                ph = v & 63u;
                cnt += v >> 6;
            }
            maps[f * 64 + lane] = (cnt << 8) | ph;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------- composition

// Group g of stream t covers child maps [64 g, 64 g + 64) of that stream.
// Map word: (count << 8) | exit phase.
__global__ __launch_bounds__(64) void k_reduce(const uint64_t *__restrict__ in,
                                               const uint64_t *__restrict__ in_first,
                                               const uint64_t *__restrict__ out_first, int T,
                                               uint64_t ngroups, uint64_t *__restrict__ out) {
    const int lane = threadIdx.x;
    for (uint64_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const int t = find_seg(out_first, T, gi);
        const uint64_t g = gi - out_first[t];
        const uint64_t nin = in_first[t + 1] - in_first[t];
        const uint64_t c0 = g * 64;
        const int nc = (int)min<uint64_t>(64, nin - c0);
        const uint64_t *base = in + (in_first[t] + c0) * 64;
        uint64_t m[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) m[k] = k < nc ? base[k * 64 + lane] : 0;
        uint32_t ph = (uint32_t)lane;
        uint64_t cnt = 0;
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            if (k < nc) {
                const uint64_t v = __shfl(m[k], (int)ph);
                ph = (uint32_t)(v & 255u);
                cnt += v >> 8;
            }
        }
        out[gi * 64 + lane] = (cnt << 8) | ph;
    }
}

// Entries: (first candidate index << 8) | entry phase.
__global__ __launch_bounds__(64) void k_down(const uint64_t *__restrict__ in,
                                             const uint64_t *__restrict__ in_first,
                                             const uint64_t *__restrict__ out_first, int T,
                                             uint64_t ngroups,
                                             const uint64_t *__restrict__ parent_entry,
                                             uint64_t *__restrict__ child_entry) {
    const int lane = threadIdx.x;
    for (uint64_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const int t = find_seg(out_first, T, gi);
        const uint64_t g = gi - out_first[t];
        const uint64_t nin = in_first[t + 1] - in_first[t];
        const uint64_t c0 = g * 64;
        const int nc = (int)min<uint64_t>(64, nin - c0);
        const uint64_t *base = in + (in_first[t] + c0) * 64;
        uint64_t m[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) m[k] = k < nc ? base[k * 64 + lane] : 0;
        const uint64_t pe = parent_entry[gi];
        uint32_t ph = (uint32_t)(pe & 255u);
        uint64_t idx = pe >> 8;
        uint64_t mine = 0;
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            if (k < nc) {
                if (lane == k) mine = (idx << 8) | ph;
                const uint64_t v = __shfl(m[k], (int)ph);
                ph = (uint32_t)(v & 255u);
                idx += v >> 8;
            }
        }
        if (lane < nc) child_entry[in_first[t] + c0 + lane] = mine;
    }
}

// --------------------------------------------------------------- k_walk

// Segments of 2^24 draws (256 blocks) by default: the walk is close to the
// VALU issue limit, and on the production shard 2^24 beat 2^25 (9.2 vs
// 10.7 ms: twice the waves, 1.125x the work) and 2^23 (10.4 ms: 1.2x more
// burn-in work).  The true candidate chain of a stream
// is found segment by segment in parallel: each wavefront starts a chain at a
// guessed start BURN draws before its segment and follows it with ballot
// speculation; two chains that ever share a start coincide from then on, and
// chains started at different points merge after ~3e5 draws on average at
// the reference's rates.  The guess is verified on the host: segment s is on
// the true chain iff its first start >= segment start (x_s) equals the first
// start >= that point found by segment s-1 (y_{s-1}); segment 0 starts at the
// true stream start.  Unverified segments are re-walked from y_{s-1}.
// Segment length: seg_blocks blocks.  By default the generation's blocks are
// cut into as many equal segments as the device keeps k_walk waves resident
// at once (eight per SIMD), so that every segment runs in one round with no
// tail of late waves, but never into segments shorter than MIN_SEG_BLOCKS
// blocks (2^24 draws), below which the burn-in dominates.  For measurement,
// CSFUNNEL_WALK_SEG_LOG in [18, 30] at context creation fixes 2^seg_log
// draws per segment instead.
constexpr uint64_t MIN_SEG_BLOCKS = 256;
constexpr uint64_t MAX_SEG_BLOCKS = 1ull << (30 - BLOCK_LOG);  // 2^30 draws
constexpr uint64_t WALK_WAVES_PER_SIMD = 8;
constexpr uint64_t BURN = 1ull << 22;  // ~14x the mean merge distance

struct SegOut {
    uint64_t x;      // first candidate start >= segment start
    uint64_t y;      // first candidate start >= segment end
    uint64_t total;  // candidates starting inside the segment
};

// Follows one stream's chain from `c` (a candidate start) and records, for
// every block boundary B in [seg_start, seg_end], the first start >= B and
// the number of starts in [seg_start, that start).  The chain position c, B
// and the counts are the same in every lane (scalar registers); lane j's
// counter for draw c + 62 j + k is (seed + (c + k) gamma) + 62 j gamma, one
// 64-bit add per lane.  A chain that passes seg_end without resolving it
// (impossible: a round spans < 4032 draws) stores y = ~0 and stops, which the
// host's chain check rejects.
//
// The tests on positions use p = c - seg_start and b = B - seg_start as
// signed 32-bit values (the scalar unit compares 32-bit values only, so
// 64-bit ones go through the VALU): a segment spans at most 2^30 draws (the
// host caps seg_blocks), the walk starts at most BURN draws before it and
// stops within 4096 + 4032 draws after its end.
__device__ void walk_segment(uint64_t seed, uint64_t c, uint64_t seg_start, uint64_t seg_end, uint32_t immT,
                             uint32_t jumpT, uint64_t *__restrict__ entries_seg, SegOut *__restrict__ out,
                             int lane) {
    uint64_t lane_step = (uint64_t)(62u * (uint32_t)lane) * GAMMA;
    asm volatile("" : "+v"(lane_step));  // opaque: keeps (c + 3) gamma + lane_step from being refactored
    const int32_t e = (int32_t)(seg_end - seg_start);
    int32_t p = (int32_t)(int64_t)(c - seg_start);
    int32_t b = 0;    // next block boundary to resolve, B - seg_start
    uint64_t r = 0;   // starts >= seg_start seen so far
    for (;;) {
        const uint64_t x3 = seed + (c + 3) * GAMMA + lane_step;  // draw c + 62 j + 3
        const bool im = draw_u24(x3) < immT;
        // the jump test: draw c + 62 j + 33 (immigrant) or + 62
        const bool jp = draw_u24(x3 + (im ? 30ull * GAMMA : 59ull * GAMMA)) < jumpT;
        const uint32_t L = im ? (jp ? 35u : 33u) : (jp ? 64u : 62u);
        const uint64_t stop = __ballot(im) | __ballot(jp);
        const uint32_t n = stop ? (uint32_t)__ffsll((unsigned long long)stop) : 64u;
        // valid starts this round: c + 62 k, k < n (ascending); the last one
        const int32_t last = p + 62 * (int32_t)(n - 1);
        // lanes before the segment start do not count
        uint32_t ks = 0;
        if (p < 0) ks = last < 0 ? n : (uint32_t)(61 - p) / 62u;
        // a round spans < 4032 draws, so at most one boundary falls inside
        // it; B <= last <=> the first start >= B is one of this round's
        if (b <= last) {
            const uint32_t kb = p >= b ? 0u : (uint32_t)(b - p + 61) / 62u;
            const int32_t first = p + 62 * (int32_t)kb;  // >= b >= 0
            const uint64_t rel = r + (kb - min(kb, ks));
            if (b == e) {
                if (lane == 0) {
                    out->y = seg_start + (uint32_t)first;
                    out->total = rel;
                }
                return;
            }
            if (b == 0 && lane == 0) out->x = seg_start + (uint32_t)first;
            if (lane == 0) entries_seg[(uint32_t)b >> BLOCK_LOG] = (rel << 8) | (uint32_t)(first - b);
            b += (int32_t)BLOCK;
        }
        r += n - ks;
        const uint32_t step = 62u * (n - 1) + (uint32_t)__builtin_amdgcn_readlane(L, (int)(n - 1));
        c += step;
        p += (int32_t)step;
        if (p > e + 4096) {
            if (lane == 0) {
                out->y = ~0ull;
                out->total = 0;
            }
            return;
        }
    }
}

__global__ __launch_bounds__(64) void k_walk(const StreamInfo *__restrict__ streams,
                                             const uint64_t *__restrict__ seg_first, int T, uint64_t nseg,
                                             uint32_t immT, uint32_t jumpT, uint64_t *__restrict__ entries,
                                             SegOut *__restrict__ segout, int64_t fix_flat, uint64_t fix_start,
                                             uint64_t seg_blocks) {
    const int lane = threadIdx.x;
    for (uint64_t f = blockIdx.x; f < nseg; f += gridDim.x) {
        const uint64_t ff = fix_flat >= 0 ? (uint64_t)fix_flat : f;
        const int t = find_seg(seg_first, T, ff);
        const uint64_t s = ff - seg_first[t];
        const StreamInfo st = streams[t];
        const uint64_t b0 = s * seg_blocks;
        const uint64_t seg_start = b0 << BLOCK_LOG;
        const uint64_t seg_end = min(b0 + seg_blocks, st.nblocks) << BLOCK_LOG;
        uint64_t c;
        if (fix_flat >= 0) c = fix_start;
        else c = seg_start > BURN ? seg_start - BURN : 0;  // 0: the true stream start
        walk_segment(st.seed, c, seg_start, seg_end, immT, jumpT, entries + st.blk_first + b0, segout + ff, lane);
        if (fix_flat >= 0) break;
    }
}

// entries[b] += segment base index (per stream prefix of segment totals)
__global__ void k_rebase(const StreamInfo *__restrict__ streams, const uint64_t *__restrict__ blk_first, int T,
                         uint64_t nblk, const uint64_t *__restrict__ seg_first,
                         const uint64_t *__restrict__ seg_base, uint64_t *__restrict__ entries, uint64_t seg_blocks) {
    for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nblk; f += (uint64_t)gridDim.x * blockDim.x) {
        const int t = find_seg(blk_first, T, f);
        const uint64_t b = f - blk_first[t];
        entries[f] += seg_base[seg_first[t] + b / seg_blocks] << 8;
    }
}

// -------------------------------------------------------------- k_score

constexpr int QCAP = 128;  // per-wave ring of queued candidate starts

struct ScoreArgs {
    const StreamInfo *streams;
    const LaunchSeg *segs;
    const uint64_t *seg_first;  // T+1 prefix of blocks in this launch
    int T;
    uint64_t nflat;
    const uint64_t *entries;  // per block: (first index << 8) | phase
    const float *elites;      // [ne][29]
    uint32_t ne;
    uint64_t ne_magic;
    float sigma;
    float sig24;  // sigma * 2^-24 (exact when the fast mutation path is used)
    uint32_t immT;
    uint32_t jumpT;
    const DevModel *model;  // device copy, read with scalar loads
    Record *rec;
    uint32_t rec_cap;
    uint32_t *rec_count;
    unsigned long long *cand_count;
    MlpScreenW mlp;       // learned-surrogate weights (MLP scorers)
    float *dump_genes;    // optional [lambda][29]: every candidate, stream order
    float *dump_scores;   // optional [lambda]
};

// LDS of the MLP scorers (per workgroup): W2 fragments, b1 | b2 | w3, and per
// wave the candidates' genes (bf16) and their scores.
constexpr size_t MLP_W2_BF16_BYTES = 8 * 16 * 64 * 16;
constexpr size_t MLP_W2_FP8_BYTES = 8 * 16 * 64 * 8;
constexpr size_t MLP_W2S_BYTES = 8 * 64 * 4;  // FP8: E8M0 scale bytes of the W2 fragments
constexpr size_t MLP_VEC_BYTES = 2 * MLP_H * 4 + (MLP_H / 2) * 4;  // b1 | b2 | w3 as f16 pairs
constexpr size_t MLP_WAVE_BYTES = 64 * MLP_IN * 2;  // genes (bf16), then the scores over them
// waves per workgroup: the MLP variants run 8 (two per SIMD) around one LDS
// copy of W2; the quadratic scorer runs 4
enum { SC_QUAD = 0, SC_MLP_BF16 = 1, SC_MLP_FP8 = 2, SC_MLP_F64 = 3 };
template <int SC> constexpr int score_waves() { return SC == SC_QUAD || SC == SC_MLP_F64 ? 4 : 8; }

// LDS of the FP64 scorer (per workgroup of 4 waves): W1 n-major and
// b1 | b2 | w3 in FP64, then per wave the 64 candidates' genes 0..15 and
// their scores in FP64.
constexpr size_t F64_W1_BYTES = (size_t)MLP_H * MLP_IN * 8;
constexpr size_t F64_VEC_BYTES = 3 * (size_t)MLP_H * 8;
constexpr size_t F64_WAVE_BYTES = 64 * (size_t)MLP_IN * 8 + 64 * 8;

struct MlpLds {
    const void *w2;        // W2 fragments (bf16 or fp8)
    const uint32_t *w2s;   // fp8: E8M0 block-scale bytes of the W2 fragments
    const bf16x4 *w1f;     // W1^T fragments, global (fp8 screen: times h_scale)
    const float *b1, *b2;  // biases (fp8 screen: b1 h_scale; both screens: b2 GELU_Q)
    const uint32_t *w3h;   // w3 as f16 pairs, [n-tile pair][lane group][4]
    __bf16 *xs;
    float *sc;
    const double *w1d, *vd;  // FP64 scorer
    double *xd, *scd;
};

typedef double dbl4 __attribute__((ext_vector_type(4)));

// GELU, tanh form, in FP64 (the host reference's definition)
__device__ __forceinline__ double gelu64(double z) {
    return 0.5 * z * (1.0 + tanh(0.7978845608028654 * (z + 0.044715 * z * z * z)));
}

// The MLP in FP64 for the wave's 64 candidates (genes in xd), one 16-candidate
// tile at a time.  Layer 2 runs on v_mfma_f64_16x16x4f64 as Z2^T = W2^T H1^T:
// lane l supplies B = H1^T[n][c] with n = 4 ns + l / 16, c = l % 16, which it
// computes itself (layer 1 and GELU on the VALU), and A = W2[n][16 mt + l % 16]
// from the packed fragments; D[m][c] lands in register r of lane l with
// m = 16 mt + 4 r + l / 16, c = l % 16, so each lane finishes layer 3 for 64
// hidden units of its candidate and two cross-lane adds complete the sum.
__device__ double mlp_wave_f64(const MlpScreenW &w, const MlpLds &ml, int lane) {
    const int lr = lane & 15, lq = lane >> 4;
    const double *b1 = ml.vd, *b2 = ml.vd + MLP_H, *w3 = ml.vd + 2 * MLP_H;
    for (int ct = 0; ct < 4; ++ct) {
        const int c = ct * 16 + lr;
        double x[MLP_IN];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) x[i] = ml.xd[c * MLP_IN + i];
        dbl4 acc[16];
#pragma unroll
        for (int mt = 0; mt < 16; ++mt) acc[mt] = dbl4{0.0, 0.0, 0.0, 0.0};
        for (int ns = 0; ns < MLP_H / 4; ++ns) {
            const int n = 4 * ns + lq;
            const double *wr = ml.w1d + n * MLP_IN;
            double z = 0.0;
#pragma unroll
            for (int i = 0; i < MLP_IN; ++i) z = __builtin_fma(x[i], wr[i], z);
            const double hv = gelu64(z + b1[n]);
            const double *a2 = w.w2d + (size_t)ns * 16 * 64 + lane;
#pragma unroll
            for (int mt = 0; mt < 16; ++mt) acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[mt * 64], hv, acc[mt], 0, 0, 0);
        }
        double part = 0.0;
#pragma unroll
        for (int mt = 0; mt < 16; ++mt) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = 16 * mt + 4 * r + lq;
                part = __builtin_fma(w3[m], gelu64(acc[mt][r] + b2[m]), part);
            }
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (lq == 0) ml.scd[c] = part + w.b3d;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return ml.scd[lane];
}

// FP64 copies of the MLP for mlp_wave_f64, from the FP32 master parameters:
// W1 n-major, W2 as A fragments, b1 | b2 | w3.
__global__ void k_pack_f64(const float *__restrict__ p, double *__restrict__ w1d, double *__restrict__ w2d,
                           double *__restrict__ vd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < MLP_H * MLP_H) {
        const int l = i & 63, mt = (i >> 6) & 15, ns = i >> 10;
        w2d[i] = (double)p[P_W2 + (4 * ns + (l >> 4)) * MLP_H + 16 * mt + (l & 15)];
    }
    if (i < MLP_H * MLP_IN) {
        const int n = i / MLP_IN, k = i % MLP_IN;
        w1d[i] = (double)p[P_W1 + k * MLP_H + n];
    }
    if (i < 3 * MLP_H) vd[i] = (double)p[(i < MLP_H ? P_B1 : i < 2 * MLP_H ? P_B2 - MLP_H : P_W3 - 2 * MLP_H) + i];
}

// One wavefront scores its 64 candidates (genes already in xs) with the MLP.
// Layer 1 computes H1^T = W1^T X^T on v_mfma_f32_16x16x16_bf16, so each
// lane's accumulators are, for candidate l%16 of a group of 16, the hidden
// units 16 mt + 4 (l/16) + r; packed two tiles at a time they are exactly the
// H1^T fragments of layer 2 in the permuted K order of w2_k_of().  Layer 2
// runs on v_mfma_f32_16x16x32_{bf16,fp8} over pairs of n-tiles, so that a
// lane holds z2 of eight hidden units of one candidate: GELU runs on them as
// four f16 pairs and layer 3 is four v_dot2c_f32_f16 with the matching w3
// pairs.
template <int SC>
__device__ __forceinline__ float mlp_wave(const MlpScreenW &w, const MlpLds &ml, int lane) {
    const int q = lane >> 4, j = lane & 15;
    // Wave priorities: two waves share each SIMD, and a batch is VALU-only
    // synthesis followed by layer 1 (VALU-heavy, few MFMAs) and layer 2
    // (MFMA-bound).  Raising the priority through the network, highest in
    // layer 2, lets a wave in its MFMA phase issue each MFMA as soon as the
    // matrix core frees while the other wave's synthesis fills the VALU
    // slots in between (measured per 2.65e9 generation: BF16 483 -> 449 ms,
    // FP8 391 -> 353 ms; the reverse order, or raising layer 2 alone, gains
    // less).
    __builtin_amdgcn_s_setprio(1);
    typedef typename std::conditional<SC == SC_MLP_FP8, uint64_t, bf16x8>::type Frag;
    // BF16: W1 fragments are fetched afresh for every batch (an opaque copy
    // of the pointer): held across batches they would be 32 more live
    // registers in the scoring loop and get spilled to scratch, whose reloads
    // then stall every layer-1 MFMA.  The FP8 variant (fewer registers) keeps
    // them.
    uint64_t w1p = (uint64_t)ml.w1f;
    if constexpr (SC == SC_MLP_BF16) asm volatile("" : "+s"(w1p));
    const bf16x4 *w1f = (const bf16x4 *)w1p;
    bf16x4 w1v[16];
#pragma unroll
    for (int mt = 0; mt < 16; ++mt) w1v[mt] = w1f[mt * 64 + lane];
    const h8 q1v = (h8)(_Float16)w.q1;
    Frag af[4][8];
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
        const s16x4 bx = *reinterpret_cast<const s16x4 *>(ml.xs + (c4 * 16 + j) * MLP_IN + 4 * q);
#pragma unroll
        for (int m4 = 0; m4 < 16; m4 += 4) {
            // four m-tiles, then their eight GELU pairs as two 8-wide vectors
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[t] = *reinterpret_cast<const f32x4 *>(ml.b1 + 16 * (m4 + t) + 4 * q);
                const bf16x4 wa = w1v[m4 + t];
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(*reinterpret_cast<const s16x4 *>(&wa), bx, acc[t],
                                                                   0, 0, 0);
            }
            h8 xa, xb;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const h2 x01 = pk_rtz(acc[t][0], acc[t][1]), x23 = pk_rtz(acc[t][2], acc[t][3]);
                xa[2 * t] = x01[0];
                xa[2 * t + 1] = x01[1];
                xb[2 * t] = x23[0];
                xb[2 * t + 1] = x23[1];
            }
            h8 ga, gb;
            if constexpr (SC == SC_MLP_FP8) {
                ga = xa * gelu_sig_v<h8>(xa, q1v);
                gb = xb * gelu_sig_v<h8>(xb, q1v);
            } else {  // xa, xb are v = x q; H1 carries the factor q on to layer 2
                ga = xa * gelu_sig_scaled<h8>(xa);
                gb = xb * gelu_sig_scaled<h8>(xb);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
            const int mt = m4 + t;
            const h2 g01 = h2{ga[2 * t], ga[2 * t + 1]}, g23 = h2{gb[2 * t], gb[2 * t + 1]};
            if constexpr (SC == SC_MLP_FP8) {
                typedef short s16x2 __attribute__((ext_vector_type(2)));
                s16x2 pk = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(s16x2{0, 0}, g01, 1.0f, false);
                pk = __builtin_amdgcn_cvt_scalef32_pk_fp8_f16(pk, g23, 1.0f, true);
                const uint64_t u = __builtin_bit_cast(uint32_t, pk);
                if (mt & 1) af[c4][mt >> 1] |= u << 32;
                else af[c4][mt >> 1] = u;
            } else {
                const bf16x2 b01 = __builtin_convertvector(__builtin_convertvector(g01, f32x2), bf16x2);
                const bf16x2 b23 = __builtin_convertvector(__builtin_convertvector(g23, f32x2), bf16x2);
                af[c4][mt >> 1][(mt & 1) * 4 + 0] = b01[0];
                af[c4][mt >> 1][(mt & 1) * 4 + 1] = b01[1];
                af[c4][mt >> 1][(mt & 1) * 4 + 2] = b23[0];
                af[c4][mt >> 1][(mt & 1) * 4 + 3] = b23[1];
            }
            }
        }
    }
    // Layer 2 runs transposed, Z2^T = W2^T H1^T: the W2 fragments are the A
    // operands and the H1 fragments the B operands (both fragment layouts
    // are lane -> row/column l%16, k-group l/16), so a lane's accumulators are
    // four consecutive hidden units of one candidate and its layer-3 partial
    // sum is a single register per group of 16 candidates.
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // The n-tile pairs are software-pipelined over two accumulator sets: the
    // MFMAs of pair np + 1 sit in the same block as the GELU / layer-3
    // epilogue of pair np, so the VALU work of one hides under the matrix
    // work of the other inside one wave.
    auto mfma_pair = [&](int np, f32x4 (&ca)[4], f32x4 (&cb)[4]) {
        const f32x4 ba = *reinterpret_cast<const f32x4 *>(ml.b2 + 32 * np + 4 * q);
        const f32x4 bb = *reinterpret_cast<const f32x4 *>(ml.b2 + 32 * np + 16 + 4 * q);
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) {
            ca[c4] = ba;
            cb[c4] = bb;
        }
        if constexpr (SC == SC_MLP_FP8) {
            // CDNA4's v_mfma_scale_f32_16x16x128_f8f6f4 (E4M3 operands): four
            // k-blocks per instruction, W2 staged as 32-byte fragments
            // [k-quad][n-tile][lane] and H1 as the four k-blocks' 8-byte
            // fragments side by side, in the same k order.  The scale
            // operands undo W2's per-block E8M0 scales (byte 2 m + (n-tile &
            // 1) of this lane's word for pair np, laid out by k_refresh) and
            // H1's h_scale, so the accumulators hold v = q z2.
            const int sw = (int)ml.w2s[np * 64 + lane];
            auto kquad = [&](auto mc) {  // mc: the k-quad, a compile-time constant (it selects scale bytes)
                constexpr int m = decltype(mc)::value;
                const i32x8 wa = reinterpret_cast<const i32x8 *>(ml.w2)[(m * 16 + 2 * np) * 64 + lane];
                const i32x8 wb = reinterpret_cast<const i32x8 *>(ml.w2)[(m * 16 + 2 * np + 1) * 64 + lane];
#pragma unroll
                for (int c4 = 0; c4 < 4; ++c4) {
                    const i32x8 hx = __builtin_bit_cast(
                        i32x8, u64x4{(uint64_t)af[c4][4 * m], (uint64_t)af[c4][4 * m + 1], (uint64_t)af[c4][4 * m + 2],
                                     (uint64_t)af[c4][4 * m + 3]});
                    ca[c4] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(wa, hx, ca[c4], 0, 0, 2 * m, sw, 0,
                                                                              w.h_e8m0);
                    cb[c4] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(wb, hx, cb[c4], 0, 0, 2 * m + 1, sw,
                                                                              0, w.h_e8m0);
                }
            };
            kquad(std::integral_constant<int, 0>{});
            kquad(std::integral_constant<int, 1>{});
        } else {
        // W2 fragments stream through per k-block: eight independent MFMAs each
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) {
            const Frag wa = reinterpret_cast<const Frag *>(ml.w2)[(kb * 16 + 2 * np) * 64 + lane];
            const Frag wb = reinterpret_cast<const Frag *>(ml.w2)[(kb * 16 + 2 * np + 1) * 64 + lane];
#pragma unroll
            for (int c4 = 0; c4 < 4; ++c4) {
                ca[c4] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, af[c4][kb], ca[c4], 0, 0, 0);
                cb[c4] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb, af[c4][kb], cb[c4], 0, 0, 0);
            }
        }
        }
    };
    auto epilogue = [&](int np, const f32x4 (&ca)[4], const f32x4 (&cb)[4]) {
        // w3 of this lane's hidden units: {4q, 4q+1}, {4q+2, 4q+3} of both n-tiles
        const uint4 w3v = *reinterpret_cast<const uint4 *>(ml.w3h + (np * 4 + q) * 4);
        // GELU of each candidate's eight hidden units as one 8-wide vector
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) {
            float z[8];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                z[r] = ca[c4][r];
                z[4 + r] = cb[c4][r];
            }
            h8 x;
#pragma unroll
            for (int p2 = 0; p2 < 4; ++p2) {
                const h2 xp = pk_rtz(z[2 * p2], z[2 * p2 + 1]);
                x[2 * p2] = xp[0];
                x[2 * p2 + 1] = xp[1];
            }
            const h8 g = x * gelu_sig_scaled<h8>(x);  // x = q z2 (q folded into W2 and b2)
            float acc3 = part[c4];
            acc3 = __builtin_amdgcn_fdot2(h2{g[0], g[1]}, __builtin_bit_cast(h2, w3v.x), acc3, false);
            acc3 = __builtin_amdgcn_fdot2(h2{g[2], g[3]}, __builtin_bit_cast(h2, w3v.y), acc3, false);
            acc3 = __builtin_amdgcn_fdot2(h2{g[4], g[5]}, __builtin_bit_cast(h2, w3v.z), acc3, false);
            part[c4] = __builtin_amdgcn_fdot2(h2{g[6], g[7]}, __builtin_bit_cast(h2, w3v.w), acc3, false);
        }
    };
    __builtin_amdgcn_s_setprio(2);
    f32x4 xa[4], xb[4], ya[4], yb[4];
    mfma_pair(0, xa, xb);
    for (int np = 0; np < 6; np += 2) {
        mfma_pair(np + 1, ya, yb);
        epilogue(np, xa, xb);
        mfma_pair(np + 2, xa, xb);
        epilogue(np + 1, ya, yb);
    }
    mfma_pair(7, ya, yb);
    epilogue(6, xa, xb);
    epilogue(7, ya, yb);
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
        float v = part[c4];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (q == 0) ml.sc[c4 * 16 + j] = v + w.b3;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float res = ml.sc[lane];
    __builtin_amdgcn_s_setprio(0);
    return res;
}

// Genes [K0, K1) of one candidate.  s0 is the stream state of the first gene
// draw; gene k uses draws s0 + 2k (crossover bit) and s0 + 2k + 1 (mutation
// unit), or s0 + k (immigrant unit).
typedef const __attribute__((address_space(3))) float *LdsRow;

// Gene k of parent a or b, picked by the sign of the crossover draw.  With
// the elites in LDS the two rows' addresses are made opaque, so the pick is
// one select of a 32-bit LDS address and the gene offset folds into the
// ds_read (the compiler otherwise selects the elite index and multiplies it
// by the row length for every gene).
template <bool LDS>
struct ParentRows {
    LdsRow la, lb;
    const float *ga, *gb;
    __device__ __forceinline__ ParentRows(const float *el, uint32_t ia, uint32_t ib) {
        if constexpr (LDS) {
            uint32_t oa = (uint32_t)(uintptr_t)(LdsRow)el + ia * (uint32_t)(CS_NTOT * 4);
            uint32_t ob = (uint32_t)(uintptr_t)(LdsRow)el + ib * (uint32_t)(CS_NTOT * 4);
            asm volatile("" : "+v"(oa), "+v"(ob));
            la = (LdsRow)(uintptr_t)oa;
            lb = (LdsRow)(uintptr_t)ob;
        } else {
            ga = el + ia * CS_NTOT;
            gb = el + ib * CS_NTOT;
        }
    }
    __device__ __forceinline__ float pick(int32_t bs, int k) const {
        if constexpr (LDS) return (bs < 0 ? lb : la)[k];
        else return (bs < 0 ? gb : ga)[k];
    }
};

template <bool IMM, bool FAST_MUT, int K0, int K1, bool LDS = false>
__device__ __forceinline__ void synth_genes(float *g, uint64_t s0, const float *el, uint32_t ia, uint32_t ib,
                                            float sigma, float sig24) {
    if (IMM) {
#pragma unroll
        for (int k = K0; k < K1; ++k) g[k] = unit_f(draw_u24(s0 + (uint64_t)k * GAMMA));
    } else {
        const ParentRows<LDS> rows(el, ia, ib);
#pragma unroll
        for (int k = K0; k < K1; ++k) {
            const int32_t bs = draw_bit_signed(s0 + (uint64_t)(2 * k) * GAMMA);
            const float par = rows.pick(bs, k);
            if (FAST_MUT) {
                // ((u - 0.5) * 2) * sigma with u = m 2^-24 has the real value
                // (m - 2^23) * (sigma 2^-23), which is rounded once when
                // sigma 2^-23 is exact; the clamp of a finite sum to [0, 1]
                // is one median-of-three.
                const int32_t c = draw_centered(s0 + (uint64_t)(2 * k + 1) * GAMMA);
                const float d = (float)c * (sig24 * 2.0f);
                g[k] = __builtin_amdgcn_fmed3f(par + d, 0.0f, 1.0f);
            } else {
                const uint32_t mu = draw_u24(s0 + (uint64_t)(2 * k + 1) * GAMMA);
                float d = (unit_f(mu) - 0.5f) * 2.0f;
                d = d * sigma;
                const float v = par + d;
                g[k] = fmaxf(0.0f, fminf(1.0f, v));
            }
        }
    }
}

// Records of the lanes of an MLP batch that pass the bound: the genome is
// synthesized again in full (only genes 0..15 were needed for the score).
// Out of line: batches emit rarely once the bounds are set, and keeping this
// code out of the scoring loop keeps the loop's instruction footprint small.
template <bool FAST_MUT>
__device__ __attribute__((noinline)) void emit_mlp_records(Record *rec, uint32_t *rec_count, uint32_t rec_cap,
                                                           const float *el, float sigma, float sig24, bool imm,
                                                           uint64_t s0, uint32_t ia, uint32_t ib, uint32_t gi,
                                                           float jval, float score, double score64, bool pass,
                                                           uint64_t pm, int lane, uint32_t stream, uint64_t cidx,
                                                           uint64_t c) {
    float g[CS_NTOT];
    if (imm) synth_genes<true, FAST_MUT, 0, CS_NTOT>(g, s0, el, ia, ib, sigma, sig24);
    else synth_genes<false, FAST_MUT, 0, CS_NTOT>(g, s0, el, ia, ib, sigma, sig24);
#pragma unroll
    for (int k = 0; k < CS_NTOT; ++k) g[k] = (k == (int)gi) ? jval : g[k];
    uint32_t basei = 0;
    if (lane == 0) basei = atomicAdd(rec_count, (uint32_t)__popcll(pm));
    basei = __shfl(basei, 0);
    if (pass) {
        const uint32_t slot = basei + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
        if (slot < rec_cap) {
            Record *o = rec + slot;
            o->score = score;
            o->stream = stream;
            o->index = cidx;
            o->start = c;
            o->ia = ia;
            o->ib = ib;
#pragma unroll
            for (int k = 0; k < CS_NTOT; ++k) o->g[k] = g[k];
            o->_pad = 0;
            o->score64 = score64;
        }
    }
}

// IMM: 0 normal candidates, 1 immigrants, 2 either, by the uniform imm_rt
// (one inlined copy of the scorer for both kinds: the MLP variants).
template <int IMM, bool LDS_ELITES, bool FAST_MUT, int SC, bool DUMP>
__device__ __forceinline__ void score_batch(const ScoreArgs &a, const MlpLds &ml, const float *mdl,
                                            const float *el, const uint32_t *q, uint32_t head,
                                            uint32_t count, uint64_t bstart, uint64_t first,
                                            uint64_t seed, uint32_t stream, float bound, double bound64,
                                            int lane, bool imm_rt = false) {
    const bool imm = IMM == 2 ? imm_rt : IMM == 1;
    const bool act = (uint32_t)lane < count;
    const uint32_t code = act ? q[(head + (uint32_t)lane) & (QCAP - 1)] : 0u;
    const uint64_t c = bstart + (code & ((1u << RUN_POS_BITS) - 1u));
    const uint64_t cidx = first + ((code >> RUN_POS_BITS) & ((1u << RUN_IDX_BITS) - 1u));
    const bool jp = (code >> 31) & 1u;
    uint64_t s = seed + (c + 1) * GAMMA;
    const uint32_t ia = fastmod(draw_hi(s), a.ne_magic, a.ne);
    s += GAMMA;
    const uint32_t ib = fastmod(draw_hi(s), a.ne_magic, a.ne);
    // the immigrant test draw (one step on) is already known; genes start two steps on
    const uint64_t s0 = s + 2 * GAMMA;
    // The MLP reads genes 0..15 only: the whole genome is synthesized again
    // after scoring, and only in batches that emit (or dump) a candidate, so
    // no gene stays live across the network.
    constexpr int KH = SC == SC_QUAD ? CS_NTOT : MLP_IN;
    float g[CS_NTOT];
    if (imm) synth_genes<true, FAST_MUT, 0, KH, LDS_ELITES>(g, s0, el, ia, ib, a.sigma, a.sig24);
    else synth_genes<false, FAST_MUT, 0, KH, LDS_ELITES>(g, s0, el, ia, ib, a.sigma, a.sig24);
    // jump test draw at s0 + ngene draws (already known), then index and value
    const uint64_t sj = s0 + (uint64_t)(imm ? CS_NTOT : 2 * CS_NTOT) * GAMMA;
    uint32_t gi = CS_NTOT;
    float jval = 0.0f;
    // jumps are rare (jump_p per candidate): the gene replacement runs only
    // in batches where some lane jumped (a wave-uniform branch)
    if (__ballot(jp)) {
        if (jp) {
            gi = draw_hi(sj + GAMMA) % (uint32_t)CS_NTOT;
            jval = unit_f(draw_u24(sj + 2 * GAMMA));
        }
#pragma unroll
        for (int k = 0; k < KH; ++k) g[k] = (k == (int)gi) ? jval : g[k];
    }
    float score;
    double s64 = 0.0;
    if constexpr (SC == SC_MLP_F64) {
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) ml.xd[lane * MLP_IN + k] = (double)g[k];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        s64 = mlp_wave_f64(a.mlp, ml, lane);
        score = (float)s64;
    } else if constexpr (SC != SC_QUAD) {
        bf16x8 x0, x1;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            x0[k] = (__bf16)g[k];
            x1[k] = (__bf16)g[8 + k];
        }
        *reinterpret_cast<bf16x8 *>(ml.xs + lane * MLP_IN) = x0;
        *reinterpret_cast<bf16x8 *>(ml.xs + lane * MLP_IN + 8) = x1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        score = mlp_wave<SC>(a.mlp, ml, lane);
    } else {
    // Model.score: lanes j = 0..15 of acc + sum, ordered sum, then the
    // linear tail over genes 16..28 (src/screen.zig:97-112).  The weights are
    // fetched per batch through an opaque uniform pointer so they stream
    // through scalar loads instead of being pinned (and spilled) for the
    // whole kernel.
    //
    // The reference starts each lane's sum and the tail at +0; here they
    // start at their first product, which saves 17 adds.  Every partial
    // result is then the reference's or both are zeros (adding a zero to a
    // nonzero value returns that value exactly, so the two can differ only
    // in the sign of a zero result), and the reference's score is never -0
    // (a round-to-nearest sum is -0 only when both addends are, and its
    // sums start at +0), so the final + 0.0f returns the reference's bits.
    float S = 0.0f, r;
    if constexpr (LDS_ELITES) {
        // broadcast LDS reads of this lane run; the address is re-opaqued
        // per run so that only one run's weights are live at a time
        float S2 = 0.0f;
        uint32_t ma = (uint32_t)(uintptr_t)mdl;
        asm volatile("" : "+v"(ma));
        const LdsRow w = (LdsRow)(uintptr_t)ma;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float acc = w[0] + w[lds_model_run(j)] * g[j];
            float sum = w[lds_model_run(j) + 1] * (g[j] * g[j]);
#pragma unroll
            for (int i = j + 1; i < 16; ++i) sum = sum + w[lds_model_run(j) + 1 + (i - j)] * (g[j] * g[i]);
            const float ln = acc + sum;
            S2 = (j == 0) ? ln : S2 + ln;
        }
        r = w[153] * g[16];
#pragma unroll
        for (int i = 17; i < CS_NTOT; ++i) r = r + w[153 + (i - 16)] * g[i];
        S = S2;
    } else {
    uint64_t mp = (uint64_t)a.model;
    asm volatile("" : "+s"(mp));
    ConstModel md = (ConstModel)mp;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float acc = md->w0 + md->lin[j] * g[j];
        float sum = md->tri[j * (j + 1) / 2 + j] * (g[j] * g[j]);
#pragma unroll
        for (int i = j + 1; i < 16; ++i) sum = sum + md->tri[i * (i + 1) / 2 + j] * (g[j] * g[i]);
        const float ln = acc + sum;
        S = (j == 0) ? ln : S + ln;
    }
    r = md->lin[16] * g[16];
#pragma unroll
    for (int i = 17; i < CS_NTOT; ++i) r = r + md->lin[i] * g[i];
    }
    score = (S + r) + 0.0f;
    }
    if constexpr (SC != SC_MLP_F64) s64 = (double)score;
    const bool pass = act && (SC == SC_MLP_F64 ? !(s64 <= bound64) : !(score <= bound));
    const uint64_t pm = __ballot(pass);
    if constexpr (KH < CS_NTOT && !DUMP) {
        if (pm)
            emit_mlp_records<FAST_MUT>(a.rec, a.rec_count, a.rec_cap, el, a.sigma, a.sig24, imm, s0, ia, ib, gi, jval,
                                       score, s64, pass, pm, lane, stream, cidx, c);
        return;
    }
    if constexpr (KH < CS_NTOT) {
        if (DUMP || pm) {
            if (imm) synth_genes<true, FAST_MUT, 0, CS_NTOT>(g, s0, el, ia, ib, a.sigma, a.sig24);
            else synth_genes<false, FAST_MUT, 0, CS_NTOT>(g, s0, el, ia, ib, a.sigma, a.sig24);
#pragma unroll
            for (int k = 0; k < CS_NTOT; ++k) g[k] = (k == (int)gi) ? jval : g[k];
        }
    }
    if (DUMP && act) {
        const uint64_t o = a.streams[stream].cand_base + cidx;
#pragma unroll
        for (int k = 0; k < CS_NTOT; ++k) a.dump_genes[o * CS_NTOT + k] = g[k];
        a.dump_scores[o] = score;
    }
    if (pm) {
        uint32_t basei = 0;
        if (lane == 0) basei = atomicAdd(a.rec_count, (uint32_t)__popcll(pm));
        basei = __shfl(basei, 0);
        if (pass) {
            const uint32_t slot = basei + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
            if (slot < a.rec_cap) {
                Record *o = a.rec + slot;
                o->score = score;
                o->stream = stream;
                o->index = cidx;
                o->start = c;
                o->ia = ia;
                o->ib = ib;
#pragma unroll
                for (int k = 0; k < CS_NTOT; ++k) o->g[k] = g[k];
                o->_pad = 0;
                o->score64 = s64;
            }
        }
    }
}

template <bool LDS_ELITES, bool FAST_MUT, int SC, bool DUMP = false>
__global__ __launch_bounds__(64 * score_waves<SC>()) void k_score(ScoreArgs a) {
    constexpr int NW = score_waves<SC>();
    extern __shared__ __attribute__((aligned(16))) uint32_t smem[];
    uint32_t *qbase = smem;  // NW waves x 2 rings x QCAP
    // quadratic scorer with elites in LDS: the model (lane-run order) first
    constexpr bool LDS_MODEL = LDS_ELITES && SC == SC_QUAD;
    float *mdl_lds = reinterpret_cast<float *>(smem + NW * 2 * QCAP);
    float *el_lds = mdl_lds + (LDS_MODEL ? LDS_MODEL_BYTES / 4 : 0);
    // readfirstlane: the wave index is uniform, and saying so keeps the run
    // bookkeeping derived from it (block, stream, chain position, ring
    // tails) in scalar registers instead of on the VALU
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const float *el = a.elites;
    const float *mdl = mdl_lds;
    if (LDS_ELITES) {
        for (uint32_t i = threadIdx.x; i < a.ne * CS_NTOT; i += blockDim.x) el_lds[i] = a.elites[i];
        if (LDS_MODEL) {
            const float *src = reinterpret_cast<const float *>(a.model);
            for (uint32_t i = threadIdx.x; i < (uint32_t)LDS_MODEL_FLOATS; i += blockDim.x)
                mdl_lds[i] = src[lds_model_src((int)i)];
        }
        __syncthreads();
        el = el_lds;
    }
    MlpLds ml{};
    if constexpr (SC == SC_MLP_BF16 || SC == SC_MLP_FP8) {
        // after the rings: W2 fragments (FP8: then their scale bytes), b1 |
        // b2 | w3 pairs, then per wave the genes (the scores reuse that
        // buffer once layer 1 has read it).  The FP8 screen carries H1 scaled
        // by the power of two h_scale (exact on W1 and b1); its layer-2 MFMAs
        // undo that and W2's block scales.
        char *base = reinterpret_cast<char *>(smem + NW * 2 * QCAP);
        const size_t w2b = SC == SC_MLP_FP8 ? MLP_W2_FP8_BYTES + MLP_W2S_BYTES : MLP_W2_BF16_BYTES;
        const uint4 *src = SC == SC_MLP_FP8 ? reinterpret_cast<const uint4 *>(a.mlp.w2f8)
                                            : reinterpret_cast<const uint4 *>(a.mlp.w2f);
        if constexpr (SC == SC_MLP_FP8) {
            // [k-block][n-tile][lane] x 8 bytes -> [k-quad][n-tile][lane] x 32 bytes
            const uint64_t *s8 = reinterpret_cast<const uint64_t *>(src);
            uint64_t *d8 = reinterpret_cast<uint64_t *>(base);
            for (uint32_t i = threadIdx.x; i < 8u * 16u * 64u; i += blockDim.x) {
                const uint32_t kb = i >> 10, nt = (i >> 6) & 15u, ln = i & 63u;
                d8[(((kb >> 2) * 16u + nt) * 64u + ln) * 4u + (kb & 3u)] = s8[i];
            }
            uint32_t *ds = reinterpret_cast<uint32_t *>(base + MLP_W2_FP8_BYTES);
            for (uint32_t i = threadIdx.x; i < 8u * 64u; i += blockDim.x) ds[i] = a.mlp.w2s8[i];
        } else {
            uint4 *dst = reinterpret_cast<uint4 *>(base);
            for (uint32_t i = threadIdx.x; i < w2b / 16; i += blockDim.x) dst[i] = src[i];
        }
        // The GELU input scale q: BF16 folds it into W1 and b1, so that H1
        // carries it into layer 2; FP8 into W2 (its layer-1 GELU takes
        // q1 = q / h_scale); both fold it into b2 and 1/q into w3, so the
        // layer-2 accumulators hold v = z2 q directly
        const float hs = SC == SC_MLP_FP8 ? a.mlp.h_scale : GELU_Q;
        const float zs = GELU_Q;
        float *vec = reinterpret_cast<float *>(base + w2b);
        uint32_t *w3h = reinterpret_cast<uint32_t *>(vec + 2 * MLP_H);
        for (uint32_t i = threadIdx.x; i < MLP_H; i += blockDim.x) {
            vec[i] = a.mlp.b1[i] * hs;
            vec[MLP_H + i] = a.mlp.b2[i] * zs;
        }
        for (uint32_t i = threadIdx.x; i < MLP_H / 2; i += blockDim.x) {
            // [np][q][k]: the pairs {4q, 4q+1}, {4q+2, 4q+3} of n-tile 2np (k = 0, 1) and 2np+1 (k = 2, 3)
            const uint32_t p = i >> 4, qq = (i >> 2) & 3u, k = i & 3u;
            const uint32_t h = 32 * p + (k >= 2 ? 16 : 0) + 4 * qq + 2 * (k & 1u);
            const float w3s = 1.0f / GELU_Q;  // both screens feed layer 3 with q GELU(z2)
            const h2 pr{(_Float16)(a.mlp.w3[h] * w3s), (_Float16)(a.mlp.w3[h + 1] * w3s)};
            w3h[i] = __builtin_bit_cast(uint32_t, pr);
        }
        char *wv = base + w2b + MLP_VEC_BYTES + (size_t)wave * MLP_WAVE_BYTES;
        ml.w2 = base;
        ml.w2s = reinterpret_cast<const uint32_t *>(base + MLP_W2_FP8_BYTES);
        ml.w1f = SC == SC_MLP_FP8 ? a.mlp.w1fh : a.mlp.w1f;
        ml.b1 = vec;
        ml.b2 = vec + MLP_H;
        ml.w3h = w3h;
        ml.xs = reinterpret_cast<__bf16 *>(wv);
        ml.sc = reinterpret_cast<float *>(wv);
        __syncthreads();
    } else if constexpr (SC == SC_MLP_F64) {
        // after the rings: W1 n-major and b1 | b2 | w3 in FP64, then per wave
        // the genes and the scores
        double *w1 = reinterpret_cast<double *>(smem + NW * 2 * QCAP);
        double *vd = w1 + MLP_H * MLP_IN;
        for (uint32_t i = threadIdx.x; i < (uint32_t)(MLP_H * MLP_IN); i += blockDim.x) w1[i] = a.mlp.w1d[i];
        for (uint32_t i = threadIdx.x; i < (uint32_t)(3 * MLP_H); i += blockDim.x) vd[i] = a.mlp.vd[i];
        double *wv = vd + 3 * MLP_H + (size_t)wave * (F64_WAVE_BYTES / 8);
        ml.w1d = w1;
        ml.vd = vd;
        ml.xd = wv;
        ml.scd = wv + 64 * MLP_IN;
        __syncthreads();
    }
    uint32_t *qN = qbase + wave * 2 * QCAP;
    uint32_t *qI = qN + QCAP;
    unsigned long long screened = 0;
    // Each wave takes pairs of flat blocks; two blocks of one stream are
    // walked as one run (the chain crosses the block boundary by itself), so
    // the partly filled batches at the end of a run come once per two blocks.
    // The host keeps nflat below 2^31, so flat block indices are 32-bit
    // (fewer scalar registers to hold across the scoring loop).
    const uint32_t nflat = (uint32_t)a.nflat;
    for (uint32_t fp = blockIdx.x * NW + wave; 2 * fp < nflat; fp += gridDim.x * NW) {
    const uint32_t fend = min(2 * fp + 2, nflat);
    for (uint32_t fb = 2 * fp, nb = 1; fb < fend; fb += nb) {
        const int t = find_seg(a.seg_first, a.T, fb);
        nb = (fb + 1 < fend && fb + 1 < a.seg_first[t + 1]) ? 2 : 1;
        const LaunchSeg sg = a.segs[t];
        const StreamInfo st = a.streams[t];
        const uint64_t b = sg.b0 + (fb - sg.flat_first);
        const uint64_t ent = a.entries[st.blk_first + b];
        const uint64_t first = ent >> 8;
        if (first >= st.my_lambda) continue;
        const uint64_t bstart = b << BLOCK_LOG;
        // The chain position p and the candidate count i are kept relative
        // to the run's start (bstart, first): both fit 32 bits (a run spans
        // at most two blocks), so the loop's tests are scalar 32-bit
        // compares.  iend caps i at the stream's last candidate.
        const uint32_t pend = nb * BLOCK;
        const uint64_t lam_run = st.my_lambda - first;  // > 0
        const uint32_t iend = __builtin_amdgcn_readfirstlane((lam_run >> 32) ? 0xffffffffu : (uint32_t)lam_run);
        uint32_t p = (uint32_t)(ent & 255u), i = 0;
        const uint64_t seed3 = st.seed + (bstart + 3) * GAMMA;
        uint32_t headN = 0, tailN = 0, headI = 0, tailI = 0;
        uint64_t lane_step = (uint64_t)(62u * (uint32_t)lane) * GAMMA;
        asm volatile("" : "+v"(lane_step));  // opaque: keeps p gamma + lane_step from being refactored
        // one step of the chain walk: up to 64 candidate starts into the rings
        auto enqueue = [&]() {
            // speculate: lane j starts at c + 62 j, c = bstart + p (no
            // immigrant, no jump before it).  p, i and n are the same in
            // every lane (scalar registers); lane j's counter for draw
            // c + 62 j + 3 is (seed + (bstart + 3) gamma) + p gamma + 62 j
            // gamma, one 64-bit multiply-add per lane.
            const uint64_t x3 = seed3 + (uint64_t)p * GAMMA + lane_step;
            const bool im = draw_u24(x3) < a.immT;
            // the jump test: draw c + 62 j + 33 (immigrant) or + 62
            const bool jp = draw_u24(x3 + (im ? 30ull * GAMMA : 59ull * GAMMA)) < a.jumpT;
            const uint32_t L = im ? (jp ? 35u : 33u) : (jp ? 64u : 62u);
            // the masks are formed on the scalar unit from the two ballots
            const uint64_t bim = __ballot(im), stop = bim | __ballot(jp);
            uint32_t n = stop ? (uint32_t)__ffsll((unsigned long long)stop) : 64u;
            const uint32_t room = (pend - p + 61) / 62u;  // p < pend
            // readfirstlane (free on a scalar value) keeps these two scalar
            // mins from being fused into a VALU min3
            n = __builtin_amdgcn_readfirstlane(min(n, room));
            n = min(n, iend - i);
            const bool take = (uint32_t)lane < n;
            const uint32_t code = (p + 62u * (uint32_t)lane) | ((i + (uint32_t)lane) << RUN_POS_BITS) |
                                  ((uint32_t)im << 30) | ((uint32_t)jp << 31);
            const uint64_t mtake = n >= 64 ? ~0ull : (1ull << n) - 1ull;
            const uint64_t mN = mtake & ~bim, mI = mtake & bim;
            // ring slot: the tail plus the mask's bits below this lane (mbcnt adds the tail)
            if (take && !im) qN[below(mN, tailN) & (QCAP - 1)] = code;
            if (take && im) qI[below(mI, tailI) & (QCAP - 1)] = code;
            tailN += (uint32_t)__popcll(mN);
            tailI += (uint32_t)__popcll(mI);
            p += 62u * (n - 1) + (uint32_t)__builtin_amdgcn_readlane(L, (int)(n - 1));
            i += n;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        if constexpr (SC == SC_QUAD) {
            while (p < pend && i < iend) {
                enqueue();
                while (tailN - headN >= 64) {
                    score_batch<0, LDS_ELITES, FAST_MUT, SC, DUMP>(a, ml, mdl, el, qN, headN, 64, bstart, first, st.seed,
                                                                   (uint32_t)t, sg.bound, sg.bound64, lane);
                    headN += 64;
                    screened += 64;
                }
                if (tailI - headI >= 64) {
                    score_batch<1, LDS_ELITES, FAST_MUT, SC, DUMP>(a, ml, mdl, el, qI, headI, 64, bstart, first, st.seed,
                                                                   (uint32_t)t, sg.bound, sg.bound64, lane);
                    headI += 64;
                    screened += 64;
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (tailN > headN) {
                score_batch<0, LDS_ELITES, FAST_MUT, SC, DUMP>(a, ml, mdl, el, qN, headN, tailN - headN, bstart, first,
                                                               st.seed, (uint32_t)t, sg.bound, sg.bound64, lane);
                screened += tailN - headN;
            }
            if (tailI > headI) {
                score_batch<1, LDS_ELITES, FAST_MUT, SC, DUMP>(a, ml, mdl, el, qI, headI, tailI - headI, bstart, first,
                                                               st.seed, (uint32_t)t, sg.bound, sg.bound64, lane);
                screened += tailI - headI;
            }
        } else {
            // One call site for every batch (full or final, either kind):
            // the rings are refilled only when both hold fewer than 64, so
            // they never exceed 127 < QCAP entries.
            bool more = p < pend && i < iend;
            for (;;) {
                if (more && tailN - headN < 64 && tailI - headI < 64) {
                    enqueue();
                    more = p < pend && i < iend;
                    continue;
                }
                bool im;
                uint32_t cnt;
                if (tailN - headN >= 64) { im = false; cnt = 64; }
                else if (tailI - headI >= 64) { im = true; cnt = 64; }
                else if (tailN > headN) { im = false; cnt = tailN - headN; }
                else if (tailI > headI) { im = true; cnt = tailI - headI; }
                else break;
                score_batch<2, LDS_ELITES, FAST_MUT, SC, DUMP>(a, ml, mdl, el, im ? qI : qN, im ? headI : headN, cnt,
                                                               bstart, first, st.seed, (uint32_t)t, sg.bound, sg.bound64, lane,
                                                               im);
                if (im) headI += cnt;
                else headN += cnt;
                screened += cnt;
                __builtin_amdgcn_wave_barrier();
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    }
    if (lane == 0 && screened) atomicAdd(a.cand_count, screened);
}

// ------------------------------------------------------------------ host

static uint32_t event_threshold(float p) {
    // number of 24-bit numerators m with (float)m * 2^-24 < p
    if (!(p > 0.0f)) return 0;
    double x = std::ceil((double)p * 16777216.0);
    if (x > 16777216.0) x = 16777216.0;
    return (uint32_t)x;
}

static double ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

static float nan_f() { return std::nanf(""); }

// ---------------------------------------------- allocation-free scratch

// Device memory of cs_screen_noalloc: every buffer of such a call is carved
// out of this array, which the code object brings with it, so the call never
// allocates.  Each GPU holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (kept per device in scratch_address).  Every call on one GPU
// shares the one array: calls that overlap on a GPU race.
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
constexpr uint32_t NOALLOC_REC_CAP = 1u << 16;  // records per record buffer (2 x 9.96 MB)
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES] = {};

// The device address of this GPU's scratch array; `device` must be current.
static int scratch_address(int device, void **out) {
    if (device < 0 || device >= MAX_DEVICES) return fail(CS_E_NODEV, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(CS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return CS_OK;
}

struct HostState {
    cs_ctx *ctx;
    ScreenBufs *B = nullptr;    // device buffers of this call
    char *arena = nullptr;      // non-null: carve B out of this scratch memory
    size_t arena_used = 0, arena_cap = 0;
    uint32_t rec_cap = 0;       // records per record buffer
    int T;
    std::vector<StreamInfo> si;
    std::vector<StreamHeap> heaps;
    const cs_genome *elites;
    ScoreArgs sa;
    bool lds_elites;
    bool fast_mut;
    int sc = SC_QUAD;
    size_t smem;
    uint64_t records = 0, cand = 0;
    uint32_t walk_fixes = 0;
    bool walked = false;
    uint32_t launches = 0, retries = 0;
    double ms_score = 0, ms_replay = 0;
};

// Device memory for buffer b of this call: the context's buffer, grown with
// hipMalloc, or (allocation-free calls) the next 256-byte aligned piece of the
// scratch array.  A scratch piece already carved and large enough is reused.
static int acquire(HostState &h, DevBuf &b, size_t bytes) {
    if (!h.arena) return ensure(b, bytes);
    if (b.p && b.bytes >= bytes) return CS_OK;
    const size_t off = (h.arena_used + 255) & ~(size_t)255;
    if (off > h.arena_cap || bytes > h.arena_cap - off)
        return fail(CS_E_CAP, "cs_screen_noalloc: this call needs more than the " +
                                  std::to_string(h.arena_cap >> 20) + " MiB scratch array (at least " +
                                  std::to_string((off + bytes + (1 << 20) - 1) >> 20) + " MiB)");
    b.p = h.arena + off;
    b.bytes = bytes;
    h.arena_used = off + bytes;
    return CS_OK;
}

// A scoring launch uses one of two buffer sets (launch segments, record
// buffer, counters, start/end events), so that the next launch can run while
// the host replays this one's records.
static LaunchSeg *buf_segs(HostState &h, int buf) { return (LaunchSeg *)h.B->segs.p + (size_t)buf * h.T; }
static uint64_t *buf_seg_first(HostState &h, int buf) { return (uint64_t *)h.B->seg_first.p + (size_t)buf * (h.T + 1); }
static Record *buf_rec(HostState &h, int buf) { return (Record *)h.B->rec.p + (size_t)buf * h.rec_cap; }
static char *buf_counters(HostState &h, int buf) { return (char *)h.B->counters.p + 32 * buf; }
static hipEvent_t buf_ev_start(cs_ctx *ctx, int buf) { return ctx->ev[buf ? 6 : 4]; }
static hipEvent_t buf_ev_end(cs_ctx *ctx, int buf) { return ctx->ev[buf ? 7 : 5]; }

// Launches the scoring of the block ranges [lo[t], hi[t]) of every stream
// with buffer set `buf`; each stream's bound is its heap's root as replayed
// so far (a bound from fewer candidates is lower, so the launch emits a
// superset of the records the heap can take).  *launched = false for empty
// ranges.
static int launch_ranges(HostState &h, const std::vector<uint64_t> &lo, const std::vector<uint64_t> &hi, int buf,
                         bool *launched) {
    cs_ctx *ctx = h.ctx;
    *launched = false;
    const int T = h.T;
    std::vector<LaunchSeg> segs(T);
    std::vector<uint64_t> first(T + 1, 0);
    for (int t = 0; t < T; ++t) {
        uint64_t n = hi[t] > lo[t] ? hi[t] - lo[t] : 0;
        first[t + 1] = first[t] + n;
        segs[t].flat_first = first[t];
        segs[t].b0 = lo[t];
        StreamHeap &hp = h.heaps[t];
        segs[t].bound = hp.full() ? (float)hp.root() : nan_f();  // exact: f32 keys
        segs[t].bound64 = hp.full() ? hp.root() : (double)nan_f();
        segs[t]._pad = 0;
    }
    const uint64_t nflat = first[T];
    if (nflat == 0) return CS_OK;
    if (nflat >= (1ull << 31))  // k_score indexes flat blocks with 32 bits
        return fail(CS_E_CAP, "a scoring launch of 2^31 or more blocks (lambda too large for one call)");
    int rc;
    if ((rc = acquire(h, h.B->segs, sizeof(LaunchSeg) * T * 2))) return rc;
    if ((rc = acquire(h, h.B->seg_first, sizeof(uint64_t) * (T + 1) * 2))) return rc;
    HIPCHK(hipMemcpyAsync(buf_segs(h, buf), segs.data(), sizeof(LaunchSeg) * T, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(buf_seg_first(h, buf), first.data(), sizeof(uint64_t) * (T + 1),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(buf_counters(h, buf), 0, 16, ctx->stream));
    ScoreArgs a = h.sa;
    a.segs = buf_segs(h, buf);
    a.seg_first = buf_seg_first(h, buf);
    a.nflat = nflat;
    a.rec = buf_rec(h, buf);
    a.rec_cap = h.rec_cap;
    a.rec_count = (uint32_t *)buf_counters(h, buf);
    a.cand_count = (unsigned long long *)(buf_counters(h, buf) + 8);
    const unsigned nw = h.sc == SC_QUAD ? score_waves<SC_QUAD>()
                        : h.sc == SC_MLP_F64 ? score_waves<SC_MLP_F64>() : score_waves<SC_MLP_BF16>();
    uint64_t grid = (nflat + nw - 1) / nw;
    if (grid > (1u << 20)) grid = 1u << 20;
    HIPCHK(hipEventRecord(buf_ev_start(ctx, buf), ctx->stream));
    if (h.sc != SC_QUAD) {
        // persistent: the workgroup loads the MLP weights into LDS once
        const uint64_t per_cu = std::max<uint64_t>(1, (160u * 1024u) / h.smem);
        grid = std::min<uint64_t>(grid, (uint64_t)ctx->n_cu * per_cu);
    }
    auto go = [&](auto kern) {
        if (h.smem > 64 * 1024)
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)h.smem);
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * nw), h.smem, ctx->stream, a);
    };
    if (a.dump_genes) {  // evaluation runs: every candidate's genes and score
        if (h.sc == SC_MLP_BF16) go(k_score<false, false, SC_MLP_BF16, true>);
        else if (h.sc == SC_MLP_FP8) go(k_score<false, false, SC_MLP_FP8, true>);
        else if (h.sc == SC_MLP_F64) go(k_score<false, false, SC_MLP_F64, true>);
        else go(k_score<false, false, SC_QUAD, true>);
    } else if (h.sc == SC_MLP_F64) {
        if (h.fast_mut) go(k_score<false, true, SC_MLP_F64>);
        else go(k_score<false, false, SC_MLP_F64>);
    } else if (h.sc == SC_MLP_BF16) {
        if (h.fast_mut) go(k_score<false, true, SC_MLP_BF16>);
        else go(k_score<false, false, SC_MLP_BF16>);
    } else if (h.sc == SC_MLP_FP8) {
        if (h.fast_mut) go(k_score<false, true, SC_MLP_FP8>);
        else go(k_score<false, false, SC_MLP_FP8>);
    } else if (h.lds_elites && h.fast_mut) {
        go(k_score<true, true, SC_QUAD>);
    } else if (h.lds_elites) {
        go(k_score<true, false, SC_QUAD>);
    } else if (h.fast_mut) {
        go(k_score<false, true, SC_QUAD>);
    } else {
        go(k_score<false, false, SC_QUAD>);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(buf_ev_end(ctx, buf), ctx->stream));
    *launched = true;
    return CS_OK;
}

static int score_ranges(HostState &h, const std::vector<uint64_t> &lo, const std::vector<uint64_t> &hi, int buf);

// Waits for the launch of buffer set `buf` (the next launch may still run on
// the stream), then replays its records through the stream heaps.  On
// record-buffer overflow the ranges are split in two and scored again in
// order with the same buffer set, before any later launch is replayed.
static int finish_ranges(HostState &h, const std::vector<uint64_t> &lo, const std::vector<uint64_t> &hi, int buf) {
    cs_ctx *ctx = h.ctx;
    const int T = h.T;
    int rc;
    HIPCHK(hipEventSynchronize(buf_ev_end(ctx, buf)));
    h.ms_score += ms_between(buf_ev_start(ctx, buf), buf_ev_end(ctx, buf));
    h.launches++;
    auto t0 = std::chrono::steady_clock::now();
    // the stream is non-blocking: these copies do not wait for a later launch
    unsigned char cbuf[16];
    HIPCHK(hipMemcpy(cbuf, buf_counters(h, buf), 16, hipMemcpyDeviceToHost));
    uint32_t cnt;
    unsigned long long cand;
    std::memcpy(&cnt, cbuf, 4);
    std::memcpy(&cand, cbuf + 8, 8);
    if (cnt > h.rec_cap) {
        // overflow: split and redo in order (streams are independent)
        h.retries++;
        std::vector<uint64_t> lo1 = lo, hi1 = hi, lo2 = lo, hi2 = hi;
        int active = 0;
        for (int t = 0; t < T; ++t) active += hi[t] > lo[t];
        if (active > 1) {
            int seen = 0;
            for (int t = 0; t < T; ++t) {
                if (hi[t] <= lo[t]) continue;
                if (seen < active / 2) lo2[t] = hi2[t] = 0;
                else lo1[t] = hi1[t] = 0;
                ++seen;
            }
        } else {
            for (int t = 0; t < T; ++t) {
                if (hi[t] <= lo[t]) continue;
                uint64_t mid = lo[t] + (hi[t] - lo[t]) / 2;
                if (mid == lo[t]) return fail(CS_E_INTERNAL, "record buffer too small for one block");
                hi1[t] = mid;
                lo2[t] = mid;
            }
        }
        h.ms_replay += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = score_ranges(h, lo1, hi1, buf))) return rc;
        return score_ranges(h, lo2, hi2, buf);
    }
    h.records += cnt;
    h.cand += cand;
    if (cnt) {
        if (ctx->host_rec.size() < cnt) ctx->host_rec.resize(cnt);
        HIPCHK(hipMemcpy(ctx->host_rec.data(), buf_rec(h, buf), sizeof(Record) * cnt, hipMemcpyDeviceToHost));
        // replay order (stream, candidate index): sort 16-byte keys, not the records
        std::vector<std::pair<uint64_t, uint64_t>> order(cnt);
        for (uint32_t i = 0; i < cnt; ++i)
            order[i] = {((uint64_t)ctx->host_rec[i].stream << 32) | i, ctx->host_rec[i].index};
        std::sort(order.begin(), order.end(), [](const std::pair<uint64_t, uint64_t> &x,
                                                 const std::pair<uint64_t, uint64_t> &y) {
            return (x.first >> 32) != (y.first >> 32) ? (x.first >> 32) < (y.first >> 32) : x.second < y.second;
        });
        for (uint32_t k = 0; k < cnt; ++k) {
            const Record &r = ctx->host_rec[(uint32_t)order[k].first];
            cs_survivor s;
            // the FP64 scorer ranks by its FP64 score (kept as the key)
            const double key = h.sc == SC_MLP_F64 ? r.score64 : (double)r.score;
            s.score = h.sc == SC_MLP_F64 ? (float)r.score64 : r.score;
            std::memcpy(s.g, r.g, sizeof s.g);
            s.pa = h.elites[r.ia].id;
            s.pb = h.elites[r.ib].id;
            s.tribe = h.elites[r.ia].tribe;
            s.stream = r.stream;
            s.index = r.index;
            s.start = r.start;
            h.heaps[r.stream].offer(s, key);
        }
    }
    h.ms_replay += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return CS_OK;
}

// One launch and its replay, in order.
static int score_ranges(HostState &h, const std::vector<uint64_t> &lo, const std::vector<uint64_t> &hi, int buf) {
    bool launched;
    int rc;
    if ((rc = launch_ranges(h, lo, hi, buf, &launched))) return rc;
    return launched ? finish_ranges(h, lo, hi, buf) : CS_OK;
}

// Exact offsets from transfer maps of every block (any rates).
static int offsets_by_maps(cs_ctx *ctx, HostState &h, uint64_t nblk, const std::vector<uint64_t> &blk_first,
                           uint32_t immT, uint32_t jumpT) {
    const int T = h.T;
    int rc;
    if ((rc = acquire(h, h.B->maps, sizeof(uint64_t) * 64 * nblk))) return rc;
    // ---- 1. block maps
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    {
        uint64_t grid = std::min<uint64_t>(nblk, 1u << 22);
        hipLaunchKernelGGL(k_map, dim3((unsigned)grid), dim3(256), 0, ctx->stream,
                           (const StreamInfo *)h.B->streams.p, (const uint64_t *)h.B->blk_first.p, T, nblk, immT,
                           jumpT, (uint64_t *)h.B->maps.p);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    // ---- 2. compose: levels of maps until one map per stream
    std::vector<std::vector<uint64_t>> lvl_first;  // per level, T+1 prefix
    lvl_first.push_back(blk_first);
    std::vector<uint64_t> lvl_off;  // element offset of each upper level in upper buffers
    uint64_t upper_total = 0;
    for (;;) {
        const auto &cur = lvl_first.back();
        bool done = true;
        for (int t = 0; t < T; ++t)
            if (cur[t + 1] - cur[t] > 1) done = false;
        if (done) break;
        std::vector<uint64_t> nx(T + 1, 0);
        for (int t = 0; t < T; ++t) nx[t + 1] = nx[t] + (cur[t + 1] - cur[t] + 63) / 64;
        lvl_off.push_back(upper_total);
        upper_total += nx[T];
        lvl_first.push_back(nx);
    }
    const int L = (int)lvl_first.size() - 1;
    if ((rc = acquire(h, h.B->upper_maps, sizeof(uint64_t) * 64 * (upper_total + 1)))) return rc;
    if ((rc = acquire(h, h.B->upper_entries, sizeof(uint64_t) * (upper_total + 1)))) return rc;
    if ((rc = acquire(h, h.B->level_first, sizeof(uint64_t) * (T + 1) * (L + 1)))) return rc;
    // host staging vectors stay alive until the stream synchronizes below
    std::vector<uint64_t> flat_first;
    for (auto &v : lvl_first) flat_first.insert(flat_first.end(), v.begin(), v.end());
    HIPCHK(hipMemcpyAsync(h.B->level_first.p, flat_first.data(), flat_first.size() * 8, hipMemcpyHostToDevice,
                          ctx->stream));
    auto lvl_maps = [&](int l) -> uint64_t * {
        return l == 0 ? (uint64_t *)h.B->maps.p : (uint64_t *)h.B->upper_maps.p + lvl_off[l - 1] * 64;
    };
    auto lvl_entries = [&](int l) -> uint64_t * {
        return l == 0 ? (uint64_t *)h.B->entries.p : (uint64_t *)h.B->upper_entries.p + lvl_off[l - 1];
    };
    auto lvl_firstp = [&](int l) -> const uint64_t * { return (const uint64_t *)h.B->level_first.p + (size_t)l * (T + 1); };
    for (int l = 0; l < L; ++l) {
        const uint64_t ng = lvl_first[l + 1][T];
        unsigned grid = (unsigned)std::min<uint64_t>(ng, 1u << 20);
        hipLaunchKernelGGL(k_reduce, dim3(grid), dim3(64), 0, ctx->stream, lvl_maps(l), lvl_firstp(l),
                           lvl_firstp(l + 1), T, ng, lvl_maps(l + 1));
        HIPCHK(hipGetLastError());
    }
    // coverage check: the top map at phase 0 counts every candidate whose
    // start lies in the mapped blocks
    std::vector<uint64_t> top(T), top_entries(T, 0);
    {
        uint64_t *tm = lvl_maps(L);
        for (int t = 0; t < T; ++t)
            HIPCHK(hipMemcpyAsync(&top[t], tm + lvl_first[L][t] * 64, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(lvl_entries(L), top_entries.data(), 8 * T, hipMemcpyHostToDevice, ctx->stream));
    }
    for (int l = L - 1; l >= 0; --l) {
        const uint64_t ng = lvl_first[l + 1][T];
        unsigned grid = (unsigned)std::min<uint64_t>(ng, 1u << 20);
        hipLaunchKernelGGL(k_down, dim3(grid), dim3(64), 0, ctx->stream, lvl_maps(l), lvl_firstp(l),
                           lvl_firstp(l + 1), T, ng, lvl_entries(l + 1), lvl_entries(l));
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int t = 0; t < T; ++t)
        if ((top[t] >> 8) < h.si[t].my_lambda)
            return fail(CS_E_INTERNAL, "block budget did not cover the stream's candidates");
    return CS_OK;
}

// Offsets from the speculative segment walk; *ok = false when too many
// segments failed verification (the caller then uses the maps).
static int offsets_by_walk(cs_ctx *ctx, HostState &h, uint64_t nblk, const std::vector<uint64_t> &blk_first,
                           uint32_t immT, uint32_t jumpT, bool *ok) {
    const int T = h.T;
    int rc;
    *ok = false;
    // ceil(nb / seg_blocks) <= nb / seg_blocks + 1 per stream, so seg_blocks
    // >= nblk / (resident - T) keeps the segment count within the resident
    // waves
    const uint64_t resident = WALK_WAVES_PER_SIMD * 4 * (uint64_t)ctx->n_cu;
    uint64_t seg_blocks = MIN_SEG_BLOCKS;
    if (ctx->walk_seg_log) seg_blocks = 1ull << (ctx->walk_seg_log - BLOCK_LOG);
    else if (resident > (uint64_t)T)
        seg_blocks = std::max(MIN_SEG_BLOCKS, (nblk + resident - T - 1) / (resident - T));
    // walk_segment tests positions as 32-bit offsets within a segment
    seg_blocks = std::min(seg_blocks, MAX_SEG_BLOCKS);
    std::vector<uint64_t> seg_first(T + 1, 0);
    for (int t = 0; t < T; ++t) {
        const uint64_t nb = h.si[t].nblocks;
        seg_first[t + 1] = seg_first[t] + (nb + seg_blocks - 1) / seg_blocks;
    }
    const uint64_t nseg = seg_first[T];
    if ((rc = acquire(h, h.B->seg_first_buf, sizeof(uint64_t) * (T + 1)))) return rc;
    if ((rc = acquire(h, h.B->segout, sizeof(SegOut) * nseg))) return rc;
    if ((rc = acquire(h, h.B->seg_base, sizeof(uint64_t) * nseg))) return rc;
    HIPCHK(hipMemcpyAsync(h.B->seg_first_buf.p, seg_first.data(), 8 * (T + 1), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    {
        unsigned grid = (unsigned)std::min<uint64_t>(nseg, 1u << 20);
        hipLaunchKernelGGL(k_walk, dim3(grid), dim3(64), 0, ctx->stream, (const StreamInfo *)h.B->streams.p,
                           (const uint64_t *)h.B->seg_first_buf.p, T, nseg, immT, jumpT, (uint64_t *)h.B->entries.p,
                           (SegOut *)h.B->segout.p, (int64_t)-1, (uint64_t)0, seg_blocks);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    std::vector<SegOut> so(nseg);
    HIPCHK(hipMemcpyAsync(so.data(), h.B->segout.p, sizeof(SegOut) * nseg, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (uint64_t f = 0; f < nseg; ++f)
        if (so[f].y == ~0ull) return fail(CS_E_INTERNAL, "segment walk passed its end");
    // verify the chain segment by segment; re-walk segments that started on
    // another chain from the verified exit of their predecessor
    uint32_t fixes = 0;
    for (int t = 0; t < T; ++t) {
        for (uint64_t f = seg_first[t] + 1; f < seg_first[t + 1]; ++f) {
            if (so[f].x == so[f - 1].y) continue;
            if (++fixes > 32) return CS_OK;  // *ok stays false
            hipLaunchKernelGGL(k_walk, dim3(1), dim3(64), 0, ctx->stream, (const StreamInfo *)h.B->streams.p,
                               (const uint64_t *)h.B->seg_first_buf.p, T, (uint64_t)1, immT, jumpT,
                               (uint64_t *)h.B->entries.p, (SegOut *)h.B->segout.p, (int64_t)f, so[f - 1].y,
                               seg_blocks);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(&so[f], (SegOut *)h.B->segout.p + f, sizeof(SegOut), hipMemcpyDeviceToHost,
                                  ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            if (so[f].x != so[f - 1].y || so[f].y == ~0ull)
                return fail(CS_E_INTERNAL, "segment fixup did not start on the chain");
        }
    }
    std::vector<uint64_t> base(nseg, 0);
    for (int t = 0; t < T; ++t) {
        uint64_t acc = 0;
        for (uint64_t f = seg_first[t]; f < seg_first[t + 1]; ++f) {
            base[f] = acc;
            acc += so[f].total;
        }
        if (acc < h.si[t].my_lambda) return fail(CS_E_INTERNAL, "block budget did not cover the stream's candidates");
    }
    HIPCHK(hipMemcpyAsync(h.B->seg_base.p, base.data(), 8 * nseg, hipMemcpyHostToDevice, ctx->stream));
    {
        unsigned grid = (unsigned)std::min<uint64_t>((nblk + 255) / 256, 1u << 16);
        hipLaunchKernelGGL(k_rebase, dim3(grid), dim3(256), 0, ctx->stream, (const StreamInfo *)h.B->streams.p,
                           (const uint64_t *)h.B->blk_first.p, T, nblk, (const uint64_t *)h.B->seg_first_buf.p,
                           (const uint64_t *)h.B->seg_base.p, (uint64_t *)h.B->entries.p, seg_blocks);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    h.walk_fixes = fixes;
    *ok = true;
    return CS_OK;
}

// noalloc: take every device buffer from the context device's scratch array
// (no hipMalloc / hipFree), with record buffers of at most NOALLOC_REC_CAP.
static int screen_impl(cs_ctx *ctx, const cs_model *model, const cs_genome *elites, int ne,
                       const cs_screen_cfg *cfg, const cs_screen_opts *opts, cs_survivor *out, uint64_t cap,
                       uint64_t *n_out, uint64_t *screens, bool noalloc = false) {
    auto wall0 = std::chrono::steady_clock::now();
    ctx->stats = cs_screen_stats{};
    if (ne < 2) return fail(CS_E_NEEDTWO, "NeedTwoElites");
    if (cfg->keep == 0) return fail(CS_E_KEEP0, "keep must be >= 1 (keep=0 is undefined in the reference)");
    const int sc = opts ? opts->scorer : SC_QUAD;
    if (sc < SC_QUAD || sc > SC_MLP_F64) return fail(CS_E_BADARGS, "unknown scorer");
    if (sc != SC_QUAD && (!opts->mlp || opts->mlp->ctx != ctx))
        return fail(CS_E_BADARGS, "the MLP scorers need an MLP created on this context");
    if (sc == SC_QUAD && !model) return fail(CS_E_BADARGS, "the quadratic scorer needs a model");
    // finite scores: every term is bounded by |w| since genes lie in [0, 1]
    if (sc == SC_QUAD) {
        const float *w = &model->w0;
        for (size_t i = 0; i < sizeof(cs_model) / 4; ++i)
            if (!(std::fabs(w[i]) <= 1e30f))
                return fail(CS_E_BADMODEL, "model weights must be finite and |w| <= 1e30 on the device path");
    }
    const uint64_t lambda = cfg->lambda;
    uint64_t T64 = std::min<uint64_t>(cfg->threads, lambda);
    if (T64 < 1) T64 = 1;
    if (T64 > (1u << 24)) return fail(CS_E_BADARGS, "threads must be <= 2^24");
    const int T = (int)T64;
    HostState h;
    ScreenBufs carved;  // noalloc: pieces of the scratch array
    h.ctx = ctx;
    h.B = noalloc ? &carved : &ctx->sb;
    if (noalloc) {
        if (!ctx->scratch) return fail(CS_E_INTERNAL, "scratch array address unknown for this device");
        h.arena = (char *)ctx->scratch;
        h.arena_cap = SCRATCH_BYTES;
    }
    h.rec_cap = noalloc ? std::min<uint32_t>(ctx->rec_cap, NOALLOC_REC_CAP) : ctx->rec_cap;
    h.T = T;
    h.elites = elites;
    h.si.resize(T);
    h.heaps.resize(T);
    const uint32_t immT = event_threshold(cfg->imm_p), jumpT = event_threshold(cfg->jump_p);
    // draws per candidate: mean and spread, for the block budget
    const double pi = immT / 16777216.0, pj = jumpT / 16777216.0;
    const double meanL = pi * (33 + 2 * pj) + (1 - pi) * (62 + 2 * pj);
    const double varL = pi * (1 - pi) * 29.0 * 29.0 + 4.0 * pj * (1 - pj) + 1.0;
    const uint64_t per = lambda / T64;
    uint64_t issued = 0, nblk = 0;
    std::vector<uint64_t> blk_first(T + 1, 0);
    for (int t = 0; t < T; ++t) {
        const uint64_t my = (t == T - 1) ? lambda - issued : per;
        issued += per;
        h.si[t].seed = cfg->seed ^ (GAMMA * (uint64_t)(t + 1));
        h.si[t].my_lambda = my;
        h.si[t].cand_base = per * (uint64_t)t;
        h.heaps[t].reset(cfg->keep);
        const double need = (double)my * meanL + 16.0 * std::sqrt((double)my * varL + 1.0) + 128.0;
        uint64_t nb = (uint64_t)std::ceil(need / (double)BLOCK) + 1;
        h.si[t].blk_first = nblk;
        h.si[t].nblocks = nb;
        blk_first[t] = nblk;
        nblk += nb;
    }
    blk_first[T] = nblk;
    int rc;
    // The quadratic screen launches level k + 1 before it replays level k's
    // records (two record buffers), see the scoring loop below.
    const bool overlap = sc == SC_QUAD && (opts ? opts->dump_genes : nullptr) == nullptr;
    // ---- device buffers
    if ((rc = acquire(h, h.B->streams, sizeof(StreamInfo) * T))) return rc;
    if ((rc = acquire(h, h.B->blk_first, sizeof(uint64_t) * (T + 1)))) return rc;
    if ((rc = acquire(h, h.B->entries, sizeof(uint64_t) * nblk))) return rc;
    if ((rc = acquire(h, h.B->elites, sizeof(float) * CS_NTOT * (size_t)ne))) return rc;
    if ((rc = acquire(h, h.B->rec, sizeof(Record) * h.rec_cap * (overlap ? 2 : 1)))) return rc;
    if ((rc = acquire(h, h.B->counters, 64))) return rc;
    std::vector<float> eg((size_t)ne * CS_NTOT);
    for (int e = 0; e < ne; ++e) {
        for (int k = 0; k < CS_NG; ++k) eg[(size_t)e * CS_NTOT + k] = elites[e].g[k];
        for (int k = 0; k < CS_NR; ++k) eg[(size_t)e * CS_NTOT + CS_NG + k] = elites[e].reg[k];
    }
    HIPCHK(hipMemcpyAsync(h.B->streams.p, h.si.data(), sizeof(StreamInfo) * T, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(h.B->blk_first.p, blk_first.data(), sizeof(uint64_t) * (T + 1), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(h.B->elites.p, eg.data(), eg.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    // ---- 1-2. candidate offsets: speculative walk when the event rates make
    // chains merge quickly, the transfer maps otherwise (or on failure)
    const uint32_t walk_min = (uint32_t)(0.004 * 16777216.0);
    bool walked = false;
    if (immT >= walk_min && jumpT >= walk_min && immT < (1u << 24) && jumpT < (1u << 24) && !ctx->force_maps) {
        if ((rc = offsets_by_walk(ctx, h, nblk, blk_first, immT, jumpT, &walked))) return rc;
    }
    if (!walked) {
        if ((rc = offsets_by_maps(ctx, h, nblk, blk_first, immT, jumpT))) return rc;
    }
    h.walked = walked;
    // ---- 3. scoring launches over geometrically growing block ranges
    DevModel md{};
    if (sc == SC_QUAD) {
        md.w0 = model->w0;
        for (int k = 0; k < CS_NTOT; ++k) md.lin[k] = model->lin[k];
        for (int i = 0; i < 16; ++i)
            for (int j = 0; j <= i; ++j) md.tri[i * (i + 1) / 2 + j] = model->quad[i * 16 + j];
    }
    if ((rc = acquire(h, h.B->model, sizeof(DevModel)))) return rc;
    HIPCHK(hipMemcpyAsync(h.B->model.p, &md, sizeof(DevModel), hipMemcpyHostToDevice, ctx->stream));
    h.sc = sc;
    h.sa.mlp = MlpScreenW{};
    h.sa.dump_genes = opts ? opts->dump_genes : nullptr;
    h.sa.dump_scores = opts ? opts->dump_scores : nullptr;
    if ((h.sa.dump_genes == nullptr) != (h.sa.dump_scores == nullptr))
        return fail(CS_E_BADARGS, "dump_genes and dump_scores go together");
    if (sc != SC_QUAD) {
        cs_mlp *ml = opts->mlp;
        const float *pp = (const float *)ml->params.p;
        float b3 = 0.0f;
        HIPCHK(hipMemcpyAsync(&b3, pp + P_B3, 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        MlpScreenW &mw = h.sa.mlp;
        mw.w1f = (const bf16x4 *)ml->w1f.p;
        mw.w1fh = (const bf16x4 *)ml->w1fh.p;
        mw.w2f = (const bf16x8 *)ml->w2f.p;
        mw.w2f8 = (const uint64_t *)ml->w2f8.p;
        mw.w2s8 = (const uint32_t *)ml->w2s8.p;
        mw.b1 = pp + P_B1;
        mw.b2 = pp + P_B2;
        mw.w3 = pp + P_W3;
        mw.b3 = b3;
        mw.h_scale = ml->h_scale;
        {
            int eh = 0;
            (void)std::frexp(ml->h_scale, &eh);  // h_scale = 2^(eh - 1)
            mw.h_e8m0 = 127 - (eh - 1);
        }
        mw.q1 = sc == SC_MLP_FP8 ? GELU_Q / ml->h_scale : GELU_Q;
        if (sc == SC_MLP_F64) {
            // FP64 copies of the FP32 master parameters, repacked per call
            const size_t nd = (size_t)MLP_H * MLP_H + (size_t)MLP_H * MLP_IN + 3 * MLP_H;
            if ((rc = ensure(ml->f64pack, nd * 8))) return rc;
            double *w2d = (double *)ml->f64pack.p, *w1d = w2d + MLP_H * MLP_H, *vd = w1d + MLP_H * MLP_IN;
            hipLaunchKernelGGL(k_pack_f64, dim3(MLP_H * MLP_H / 256), dim3(256), 0, ctx->stream, pp, w1d, w2d, vd);
            HIPCHK(hipGetLastError());
            mw.w1d = w1d;
            mw.w2d = w2d;
            mw.vd = vd;
            mw.b3d = (double)b3;
        }
    }
    ScoreArgs &a = h.sa;
    a.streams = (const StreamInfo *)h.B->streams.p;
    a.T = T;
    a.entries = (const uint64_t *)h.B->entries.p;
    a.elites = (const float *)h.B->elites.p;
    a.ne = (uint32_t)ne;
    a.ne_magic = UINT64_MAX / (uint64_t)ne + 1;
    a.sigma = cfg->sigma;
    a.sig24 = cfg->sigma * 0x1p-24f;
    {
        // fast mutation needs sigma * 2^-24 exact and finite parents/sigma
        bool ok = std::isfinite(cfg->sigma) && (double)a.sig24 == (double)cfg->sigma * 0x1p-24;
        for (int e = 0; e < ne && ok; ++e) {
            for (int k = 0; k < CS_NG; ++k) ok = ok && std::isfinite(elites[e].g[k]);
            for (int k = 0; k < CS_NR; ++k) ok = ok && std::isfinite(elites[e].reg[k]);
        }
        h.fast_mut = ok;
    }
    a.immT = immT;
    a.jumpT = jumpT;
    a.model = (const DevModel *)h.B->model.p;
    a.rec = (Record *)h.B->rec.p;
    a.rec_cap = h.rec_cap;
    a.rec_count = (uint32_t *)h.B->counters.p;
    a.cand_count = (unsigned long long *)((char *)h.B->counters.p + 8);
    const size_t el_bytes = (size_t)ne * CS_NTOT * 4;
    h.lds_elites = sc == SC_QUAD && el_bytes <= 48 * 1024 && h.sa.dump_genes == nullptr;
    if (sc == SC_QUAD) {
        h.smem = score_waves<SC_QUAD>() * 2 * QCAP * 4 + (h.lds_elites ? LDS_MODEL_BYTES + el_bytes : 0);
    } else if (sc == SC_MLP_F64) {
        const size_t nw = score_waves<SC_MLP_F64>();
        h.smem = nw * 2 * QCAP * 4 + F64_W1_BYTES + F64_VEC_BYTES + nw * F64_WAVE_BYTES;
    } else {
        const size_t nw = score_waves<SC_MLP_BF16>();
        h.smem = nw * 2 * QCAP * 4 + (sc == SC_MLP_FP8 ? MLP_W2_FP8_BYTES + MLP_W2S_BYTES : MLP_W2_BF16_BYTES) + MLP_VEC_BYTES +
                 nw * MLP_WAVE_BYTES;
    }
    {
        // The quadratic screen launches level k + 1 before it replays level
        // k's records, so the replay runs beside the next launch; level k + 1
        // then filters with the bound of the levels before k, which is lower
        // and so still passes every candidate the heaps can take.
        std::vector<uint64_t> lo(T), hi(T), plo, phi;
        bool pending = false;
        int buf = 0, pbuf = 0;
        uint64_t maxnb = 0;
        for (int t = 0; t < T; ++t) maxnb = std::max(maxnb, h.si[t].nblocks);
        uint64_t a0 = 0, a1 = 1;
        while (a0 < maxnb) {
            for (int t = 0; t < T; ++t) {
                lo[t] = std::min(a0, h.si[t].nblocks);
                hi[t] = std::min(a1, h.si[t].nblocks);
            }
            if (overlap) {
                bool launched;
                if ((rc = launch_ranges(h, lo, hi, buf, &launched))) return rc;
                if (pending && (rc = finish_ranges(h, plo, phi, pbuf))) return rc;
                pending = launched;
                if (launched) {
                    plo = lo;
                    phi = hi;
                    pbuf = buf;
                    buf ^= 1;
                }
            } else if ((rc = score_ranges(h, lo, hi, 0))) {
                return rc;
            }
            a0 = a1;
            a1 *= 4;
        }
        if (pending && (rc = finish_ranges(h, plo, phi, pbuf))) return rc;
    }
    // ---- 4. merge heaps t-major, stable sort by score descending
    auto tm0 = std::chrono::steady_clock::now();
    std::vector<HeapItem> all;
    for (int t = 0; t < T; ++t) all.insert(all.end(), h.heaps[t].slots.begin(), h.heaps[t].slots.end());
    std::stable_sort(all.begin(), all.end(), [](const HeapItem &x, const HeapItem &y) { return x.key > y.key; });
    uint64_t out_n = std::min<uint64_t>(all.size(), cfg->keep);
    if (out_n > cap) return fail(CS_E_CAP, "survivor buffer too small");
    for (uint64_t i = 0; i < out_n; ++i) out[i] = all[i].s;
    if (opts && opts->scores64)
        for (uint64_t i = 0; i < out_n; ++i) opts->scores64[i] = all[i].key;
    *n_out = out_n;
    *screens = lambda;
    h.ms_replay += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tm0).count();
    cs_screen_stats &S = ctx->stats;
    S.candidates = h.cand;
    S.positions = nblk * BLOCK;
    S.blocks = nblk;
    S.records = h.records;
    S.levels = h.launches;
    S.retries = h.retries;
    S.ms_map = ms_between(ctx->ev[0], ctx->ev[1]);
    S.ms_scan = ms_between(ctx->ev[1], ctx->ev[2]);
    S.ms_score = h.ms_score;
    S.ms_replay = h.ms_replay;
    S.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    S.offset_path = h.walked ? 1u : 0u;
    S.walk_fixes = h.walk_fixes;
    return CS_OK;
}

}  // namespace csf

using namespace csf;

extern "C" {

int cs_ctx_create(int device, cs_ctx **out) {
    if (!out) return fail(CS_E_BADARGS, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
        return fail(CS_E_NODEV, "no HIP device " + std::to_string(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(CS_E_NODEV, std::string("device is ") + prop.gcnArchName + ", libcsfunnel is compiled for gfx950 only");
    HIPCHK(hipSetDevice(device));
    cs_ctx *c = new cs_ctx();
    c->device = device;
    c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    c->walk_seg_log = 0;  // balanced segments
    if (const char *e = std::getenv("CSFUNNEL_WALK_SEG_LOG")) {
        const int v = std::atoi(e);
        if (v >= 18 && v <= 30) c->walk_seg_log = v;
    }
    if (int rc = scratch_address(device, &c->scratch)) {
        delete c;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(CS_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    for (auto &ev : c->ev) {
        if ((e = hipEventCreate(&ev)) != hipSuccess) {
            for (auto &ev2 : c->ev)
                if (ev2) (void)hipEventDestroy(ev2);
            (void)hipStreamDestroy(c->stream);
            delete c;
            return fail(CS_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    *out = c;
    return CS_OK;
}

int cs_ctx_destroy(cs_ctx *c) {
    if (!c) return CS_OK;
    if (c->n_mlp > 0) return fail(CS_E_BADARGS, "context still has live MLP handles; destroy them first");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    csf::ScreenBufs &b = c->sb;
    csf::DevBuf *bufs[] = {&b.streams, &b.blk_first, &b.maps, &b.entries, &b.upper_maps, &b.upper_entries,
                           &b.level_first, &b.elites, &b.segs, &b.seg_first, &b.rec, &b.counters, &b.model,
                           &b.seg_first_buf, &b.segout, &b.seg_base, &c->gram_part, &c->ridge_part,
                           &c->ridge_scalar};
    for (auto *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return CS_OK;
}

int cs_screen(cs_ctx *ctx, const cs_model *model, const cs_genome *elites, int ne, const cs_screen_cfg *cfg,
              cs_survivor *out, uint64_t cap, uint64_t *n_out, uint64_t *screens) {
    if (!ctx || !model || !cfg || !n_out || !screens || (ne > 0 && !elites) || (cap > 0 && !out))
        return fail(CS_E_BADARGS, "null argument");
    clear_error();
    HIPCHK(hipSetDevice(ctx->device));
    return screen_impl(ctx, model, elites, ne, cfg, nullptr, out, cap, n_out, screens);
}

int cs_screen_ex(cs_ctx *ctx, const cs_model *model, const cs_genome *elites, int ne, const cs_screen_cfg *cfg,
                 const cs_screen_opts *opts, cs_survivor *out, uint64_t cap, uint64_t *n_out, uint64_t *screens) {
    if (!ctx || !cfg || !opts || !n_out || !screens || (ne > 0 && !elites) || (cap > 0 && !out))
        return fail(CS_E_BADARGS, "null argument");
    clear_error();
    HIPCHK(hipSetDevice(ctx->device));
    return screen_impl(ctx, model, elites, ne, cfg, opts, out, cap, n_out, screens);
}

int cs_screen_noalloc(cs_ctx *ctx, const cs_model *model, const cs_genome *elites, int ne,
                      const cs_screen_cfg *cfg, cs_survivor *out, uint64_t cap, uint64_t *n_out, uint64_t *screens) {
    if (!ctx || !model || !cfg || !n_out || !screens || (ne > 0 && !elites) || (cap > 0 && !out))
        return fail(CS_E_BADARGS, "null argument");
    clear_error();
    HIPCHK(hipSetDevice(ctx->device));
    return screen_impl(ctx, model, elites, ne, cfg, nullptr, out, cap, n_out, screens, true);
}

int cs_set_record_cap(cs_ctx *ctx, uint32_t cap) {
    if (!ctx || cap < 2048 || cap > (1u << 20))
        return fail(CS_E_BADARGS, "record capacity must lie in [2048, 2^20]");
    ctx->rec_cap = cap;
    return CS_OK;
}

int cs_set_offset_mode(cs_ctx *ctx, int mode) {
    if (!ctx || mode < 0 || mode > 1) return fail(CS_E_BADARGS, "offset mode must be 0 (auto) or 1 (maps)");
    ctx->force_maps = mode == 1;
    return CS_OK;
}

int cs_get_stats(cs_ctx *ctx, cs_screen_stats *out) {
    if (!ctx || !out) return fail(CS_E_BADARGS, "null argument");
    *out = ctx->stats;
    return CS_OK;
}

}  // extern "C"
