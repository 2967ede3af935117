/* Host-side memory and UB check of the CPU oracle (test infrastructure):
 * built with -fsanitize=address,undefined by tests/test_host_sanitizers.py.
 * Runs the three screen forms (scalar, 16-wide, chunked) over small sizes,
 * thread counts and keeps, requires them to agree, and formats every
 * survivor's score.  argv: model file, elites file. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    static or_model m;
    static or_genome el[4096];
    int ne = 0;
    if (or_load_model(argv[1], &m) || or_load_elites(argv[2], el, 4096, &ne)) return 2;
    const uint64_t lams[] = {1, 7, 1000, 30000};
    const uint64_t keeps[] = {1, 5, 256};
    int bad = 0;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 3; b++)
            for (uint64_t T = 1; T <= 8; T *= 2) {
                or_cfg cfg;
                memset(&cfg, 0, sizeof cfg);
                cfg.lambda = lams[a];
                cfg.keep = keeps[b];
                cfg.seed = 11 + (uint64_t)a;
                cfg.threads = T;
                cfg.sigma = 0.12f;
                cfg.jump_p = 0.02f;
                cfg.imm_p = 0.01f;
                const uint64_t cap = keeps[b] * T + 16;
                or_slot *o1 = calloc(cap, sizeof *o1), *o2 = calloc(cap, sizeof *o2), *o3 = calloc(cap, sizeof *o3);
                uint64_t n1 = 0, n2 = 0, n3 = 0, s1 = 0, s2 = 0, s3 = 0;
                const int r1 = or_screen(&m, el, ne, &cfg, 4, o1, cap, &n1, &s1);
                const int r2 = or_screen_simd(&m, el, ne, &cfg, 3, o2, cap, &n2, &s2);
                const int r3 = or_screen_chunked(&m, el, ne, &cfg, 5, 3, o3, cap, &n3, &s3);
                if (r1 || r2 || r3 || n1 != n2 || n1 != n3 || s1 != cfg.lambda || s2 != s1 || s3 != s1 ||
                    memcmp(o1, o2, n1 * sizeof *o1) || memcmp(o1, o3, n1 * sizeof *o1)) {
                    printf("mismatch lambda=%llu keep=%llu T=%llu\n", (unsigned long long)lams[a],
                           (unsigned long long)keeps[b], (unsigned long long)T);
                    bad = 1;
                }
                char buf[64];
                for (uint64_t i = 0; i < n1; i++) or_fmt_d4(o1[i].score, buf, sizeof buf);
                free(o1);
                free(o2);
                free(o3);
            }
    /* short buffers and special values */
    const float vals[] = {0.0f, -0.0f, 1e-5f, 0.00005f, 0.99995f, 12345.678f, -3.25f, 1e30f};
    char small[8];
    for (int i = 0; i < 8; i++) (void)or_fmt_d4(vals[i], small, sizeof small);
    puts(bad ? "FAIL" : "ok");
    return bad;
}
