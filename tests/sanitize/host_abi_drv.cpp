// Host-side memory and UB check of the C ABI's host code (funnel_host.cpp:
// token and line parsing, model and elite loading, the gene printer and the
// survivor / count writers): built with g++ -fsanitize=address,undefined by
// tests/test_host_sanitizers.py.  argv: model file, elites file, output dir.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "csfunnel.h"

int main(int argc, char **argv) {
    if (argc != 4) return 2;
    cs_genome g;
    const char *lines[] = {
        "1 0 0 0 7 0.1 0.9 0.2 0.3 0.4 0.5 0.6 0.7 0.8 0.9 0.05 0.5 0.5 0.5 0.11 0.22",
        "", "#x", "1 2 3", "-1 0 0 0 7 0.1", "   \t  ",
        "1_0 0 0 0 7 0.1 0.9 0.2 0.3 0.4 0.5 0.6 0.7 0.8 0.9 0.05 0.5 0.5 0.5 0.11 0.22 x",
        "1 0 0 0 7 0.1 0.9 0.2 0.3 0.4 0.5 0.6 0.7 0.8 0.9 0.05 0.5 0.5 0.5 0.11 0.22 0.9 0.1 0.5 0.5 0.5 0.7 1 1 1 1 1 1 1 1"};
    for (const char *l : lines) (void)cs_parse_pop_line(l, strlen(l), &g);
    float f;
    uint64_t u;
    const char *toks[] = {"0", "18446744073709551615", "18446744073709551616", "1e39", "-0", "nan", "0x10",
                          "1_000", "", "+", "-", "1e", ".", "1.5e-45", "123456789012345678901234567890"};
    for (const char *t : toks) {
        (void)cs_parse_u64(t, strlen(t), &u);
        (void)cs_parse_f32(t, strlen(t), &f);
    }
    char buf[64], small[4];
    const float vals[] = {0.0f, -0.0f, 1e-5f, 0.00005f, 0.99995f, 12345.678f, 1e30f, 3.4e38f, -3.4e38f};
    for (float v : vals) {
        (void)cs_format_gene(v, buf, sizeof buf);
        (void)cs_format_gene(v, small, sizeof small);
    }
    float w[170];
    for (int i = 0; i < 170; i++) w[i] = (float)i * 0.01f;
    cs_model m;
    for (int n = 0; n <= 170; n++) (void)cs_model_from_compact(w, n, &m);
    if (cs_load_model(argv[1], &m) != CS_OK) return 2;
    static cs_genome el[1000];
    int ne = 0;
    if (cs_load_elites(argv[2], el, 1000, &ne) != CS_OK || ne < 2) return 2;
    (void)cs_load_elites(argv[2], el, 3, &ne);  // buffer too small: an error, no overflow
    (void)cs_load_model("/nonexistent/model", &m);
    cs_survivor s[40];
    memset(s, 0, sizeof s);
    for (int i = 0; i < 40; i++) {
        s[i].score = (float)i * -1.25f;
        for (int k = 0; k < CS_NTOT; k++) s[i].g[k] = (float)((i * 31 + k) % 97) / 97.0f;
        s[i].pa = (uint64_t)i;
        s[i].pb = (uint64_t)i * 7;
        s[i].tribe = (uint32_t)(i % 12);
    }
    const std::string dir = argv[3];
    if (cs_write_promoted((dir + "/p.txt").c_str(), s, 40) != CS_OK) return 4;
    if (cs_write_promoted_v6((dir + "/p6.txt").c_str(), s, 40) != CS_OK) return 4;
    if (cs_write_count((dir + "/c.txt").c_str(), 2650000000ull) != CS_OK) return 4;
    (void)cs_write_promoted("/nonexistent/dir/p.txt", s, 40);
    puts("ok");
    return 0;
}
