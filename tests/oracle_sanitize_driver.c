/* Drives every function of the CPU oracle (oracle/ibp_oracle.c) on random
 * tables, built together with it under AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/test_oracle.py: any out-of-bounds access,
 * use after free, shift overflow or other undefined behaviour aborts the run.
 * It also checks the round trips the oracle promises (fetch of the packed
 * slots, decompress of the condensed stream, per-cluster forms), so a run
 * that completes has exercised every path with meaningful data.
 *
 * Usage: oracle_sanitize_driver <seed> <cases>; exit status 0 = all passed. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ibp_oracle.h"

static uint64_t rng_state;

static uint64_t rnd(void)
{
    /* splitmix64 */
    uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static int fail(const char *what, int w, int64_t n, int64_t v)
{
    fprintf(stderr, "oracle sanitize driver: %s (w=%d n=%lld v=%lld)\n", what, w, (long long)n, (long long)v);
    return 1;
}

/* One table of n rows of v elements of w bytes: each row is a noisy copy of one
 * of a few base rows (so some columns are invariant), about a third of the
 * rows pure noise (raw rows). */
static void fill_table(unsigned char *x, int w, int64_t n, int64_t v)
{
    unsigned char *base = (unsigned char *)malloc((size_t)(4 * v * w));
    for (int64_t i = 0; i < 4 * v * w; ++i) base[i] = (unsigned char)rnd();
    for (int64_t r = 0; r < n; ++r) {
        const int kind = (int)(rnd() % 3);
        const unsigned char *b = base + (rnd() % 4) * v * w;
        for (int64_t k = 0; k < v * w; ++k) {
            unsigned char byte = kind == 0 ? (unsigned char)rnd() : b[k];
            if (kind == 1 && (k % w) == 0) byte ^= (unsigned char)(rnd() & 3);
            x[r * v * w + k] = byte;
        }
    }
    free(base);
}

static int one_case(int w, int64_t n, int64_t v)
{
    const size_t row = (size_t)(v * w), bytes = (size_t)n * row;
    unsigned char *x = (unsigned char *)malloc(bytes + 1);
    fill_table(x, w, n, v);
    unsigned char *mask = (unsigned char *)calloc(row + 1, 1), *bitval = (unsigned char *)calloc(row + 1, 1);
    float chosen = 0.0f;
    const float thr = (rnd() % 2) ? -1.0f : 0.75f;
    oracle_preprocess(w, x, n, v, thr, mask, bitval, &chosen);

    /* sizes, slots, bitmask, fetch of every row and of a random index */
    int64_t *sizes = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    uint64_t ctr = 0;
    oracle_get_compress_size(w, x, n, v, mask, bitval, NULL, n, sizes, &ctr);
    unsigned char *slots = (unsigned char *)malloc(bytes + 1);
    int32_t *bitmask = (int32_t *)calloc((size_t)((n + 31) / 32 + 1), sizeof(int32_t));
    oracle_compress(w, x, slots, n, v, mask, bitval, NULL, n, bitmask);
    unsigned char *out = (unsigned char *)malloc(bytes + 1);
    oracle_decompress_fetch(w, slots, n, v, mask, bitval, bitmask, NULL, n, out);
    if (memcmp(out, x, bytes) != 0) return fail("fetch of the slots differs from the rows", w, n, v);
    const int64_t m = n > 0 ? 1 + (int64_t)(rnd() % (uint64_t)(2 * n)) : 0;
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    for (int64_t i = 0; i < m; ++i) idx[i] = (int64_t)(rnd() % (uint64_t)n);
    unsigned char *outi = (unsigned char *)malloc((size_t)m * row + 1);
    oracle_decompress_fetch(w, slots, n, v, mask, bitval, bitmask, idx, m, outi);
    for (int64_t i = 0; i < m; ++i)
        if (memcmp(outi + i * row, x + idx[i] * row, row) != 0) return fail("indexed fetch differs", w, n, v);

    /* in place, and the multi-threaded form */
    unsigned char *inplace = (unsigned char *)malloc(bytes + 1);
    memcpy(inplace, x, bytes);
    int32_t *bm2 = (int32_t *)calloc((size_t)((n + 31) / 32 + 1), sizeof(int32_t));
    oracle_compress(w, inplace, inplace, n, v, mask, bitval, NULL, n, bm2);
    if (memcmp(inplace, slots, bytes) != 0) return fail("in-place pack differs", w, n, v);
    oracle_compress_mt(w, x, inplace, n, v, mask, bitval, NULL, NULL, n, bm2, 3);
    if (memcmp(inplace, slots, bytes) != 0) return fail("multi-threaded pack differs", w, n, v);

    /* condensed stream with an index, then its decompress */
    int64_t *offs = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m + 1));
    const int64_t total = oracle_compress_condensed(w, x, n, v, mask, bitval, NULL, idx, m, NULL, offs);
    unsigned char *stream = (unsigned char *)malloc((size_t)total + 1);
    oracle_compress_condensed(w, x, n, v, mask, bitval, NULL, idx, m, stream, offs);
    oracle_decompress_condensed(w, stream, offs, m, v, mask, bitval, NULL, NULL, m, outi);
    for (int64_t i = 0; i < m; ++i)
        if (memcmp(outi + i * row, x + idx[i] * row, row) != 0) return fail("condensed round trip differs", w, n, v);

    /* k-means and the per-cluster forms */
    if ((w == 4 || w == 8) && n > 0) {
        const int K = 1 + (int)(rnd() % 5);
        unsigned char *km = (unsigned char *)calloc((size_t)K * row + 1, 1);
        unsigned char *kb = (unsigned char *)calloc((size_t)K * row + 1, 1);
        int32_t *cl = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
        oracle_kmeans(w, x, n, v, K, (int64_t)(rnd() % 100), km, kb, cl);
        for (int64_t r = 0; r < n; ++r)
            if (cl[r] < 0 || cl[r] >= K) return fail("cluster id out of range", w, n, v);
        oracle_get_compress_size_clustered(w, x, n, v, km, kb, cl, NULL, n, sizes, NULL);
        oracle_compress_clustered(w, x, slots, n, v, km, kb, cl, NULL, n, bitmask);
        oracle_decompress_fetch_clustered(w, slots, n, v, km, kb, cl, bitmask, NULL, n, out);
        if (memcmp(out, x, bytes) != 0) return fail("per-cluster round trip differs", w, n, v);
        free(km);
        free(kb);
        free(cl);
    }
    free(x);
    free(mask);
    free(bitval);
    free(sizes);
    free(slots);
    free(bitmask);
    free(out);
    free(idx);
    free(outi);
    free(inplace);
    free(bm2);
    free(offs);
    free(stream);
    return 0;
}

int main(int argc, char **argv)
{
    rng_state = argc > 1 ? strtoull(argv[1], NULL, 10) : 1;
    const int cases = argc > 2 ? atoi(argv[2]) : 40;
    static const int widths[4] = {1, 2, 4, 8};
    static const int64_t lengths[] = {1, 2, 3, 7, 8, 9, 33, 50, 64, 65, 130, 300};
    for (int c = 0; c < cases; ++c) {
        const int w = widths[c % 4];
        const int64_t v = lengths[rnd() % (sizeof(lengths) / sizeof(lengths[0]))];
        const int64_t n = (int64_t)(rnd() % 300);
        if (one_case(w, n, v)) return 1;
    }
    int steps = 0;
    const float t = oracle_sweep_last(&steps);
    if (steps != 7 || t != 1.0f) return fail("sweep", 0, 0, 0);
    printf("oracle sanitize driver: %d cases passed\n", cases);
    return 0;
}
