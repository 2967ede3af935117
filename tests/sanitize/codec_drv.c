/* Host-side memory and UB check of the codec label source: built with
 * -fsanitize=address,undefined by tests/test_host_sanitizers.py.  Encodes and
 * decodes random clips (noise, ramps, near-flat) of several sizes, 1-4
 * frames, with random and 0/1 genomes (every gene-gated tool, the 16x16
 * superblock layer included), every flag combination and quantizer
 * overrides.  Exit status 3: a range-coded stream whose decoder output left
 * the encoder recon, or a corrupt or oversized stream (none may occur; the
 * byte-varint stream's drift is the reference's own and is allowed). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cscodec.h"

static uint64_t st = 88172645463325252ull;
static uint64_t xr(void) {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}
static float uf(void) { return (float)(xr() >> 40) * 0x1p-24f; }

int main(int argc, char **argv) {
    const int iters = argc > 1 ? atoi(argv[1]) : 600;
    const uint32_t sizes[][2] = {{16, 16}, {32, 16}, {16, 64}, {48, 32}, {64, 48}, {80, 64}, {24, 40}};
    int ok = 0, unsupported = 0, args = 0, drift = 0, bad = 0;
    for (int it = 0; it < iters; it++) {
        const uint32_t w = sizes[it % 7][0], h = sizes[it % 7][1], frames = 1 + (uint32_t)(it % 4);
        const size_t n = (size_t)w * h * 3 / 2 * frames;
        uint8_t *yuv = malloc(n), *rec = malloc(n);
        const int kind = it % 3;
        for (size_t i = 0; i < n; i++)
            yuv[i] = kind == 0 ? (uint8_t)xr() : kind == 1 ? (uint8_t)((i * 7 + (size_t)it) & 255) : (uint8_t)(128 + (xr() & 7));
        cs_genome g;
        memset(&g, 0, sizeof g);
        for (int k = 0; k < CS_NG; k++) g.g[k] = (it % 5 == 0) ? ((xr() & 1) ? 1.0f : 0.0f) : uf();
        for (int k = 0; k < CS_NR; k++) g.reg[k] = uf();
        csc_clip_stats s;
        const uint32_t flags = (uint32_t)(it % 8);
        const float q = (it % 6 == 0) ? 0.5f + uf() * 3.0f : 1.0f;
        const int rc = csc_eval_clip_ex(yuv, w, h, frames, &g, flags, q, (it & 1) ? rec : NULL, &s);
        if (rc == CSC_OK) ok++;
        else if (rc == CSC_E_UNSUPPORTED) unsupported++;
        else if (rc == CSC_E_ARGS) args++;
        else if (rc == CSC_E_DESYNC && !(g.g[18] > 0.5f)) drift++;
        else {
            printf("it %d %ux%ux%u flags %u rc %d g18 %.3f\n", it, w, h, frames, flags, rc, g.g[18]);
            bad = 1;
        }
        free(yuv);
        free(rec);
    }
    printf("ok %d unsupported %d args %d varint drift %d\n", ok, unsupported, args, drift);
    return bad ? 3 : 0;
}
