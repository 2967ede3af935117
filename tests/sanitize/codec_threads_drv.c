/* Thread-safety check of the codec label source ("no shared mutable state",
 * include/cscodec.h): six threads encode and decode their own random clips
 * at once; built with -fsanitize=thread by tests/test_host_sanitizers.py. */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cscodec.h"

static uint64_t next(uint64_t *st) {
    *st ^= *st << 13;
    *st ^= *st >> 7;
    *st ^= *st << 17;
    return *st;
}

static void *work(void *arg) {
    uint64_t st = 1234567 + (uint64_t)(uintptr_t)arg * 7919;
    for (int it = 0; it < 24; it++) {
        const uint32_t w = 32, h = 32, frames = 2;
        const size_t n = (size_t)w * h * 3 / 2 * frames;
        uint8_t *yuv = malloc(n);
        for (size_t i = 0; i < n; i++) yuv[i] = (uint8_t)next(&st);
        cs_genome g;
        memset(&g, 0, sizeof g);
        for (int k = 0; k < CS_NG; k++) g.g[k] = (float)(next(&st) >> 40) * 0x1p-24f;
        for (int k = 0; k < CS_NR; k++) g.reg[k] = 0.5f;
        csc_clip_stats s;
        (void)csc_eval_clip(yuv, w, h, frames, &g, (uint32_t)(it % 8), &s);
        free(yuv);
    }
    return NULL;
}

int main(void) {
    pthread_t t[6];
    for (uintptr_t i = 0; i < 6; i++)
        if (pthread_create(&t[i], NULL, work, (void *)i)) return 2;
    for (int i = 0; i < 6; i++) pthread_join(t[i], NULL);
    puts("ok");
    return 0;
}
