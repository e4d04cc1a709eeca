/*
 * ibp_oracle.c -- CPU restatement of the reference IBP algorithm (the checker).
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never does.
 *
 * Reference: AKKamath/InvariantBitPacking (CUDA).  The reference cannot be
 * built in this environment (CUDA-only sources), so this file restates its
 * arithmetic in plain C, function by function, with file:line citations.
 *
 * Full-width shifts.  The reference writes `(1ll << n) - 1` with n == B and
 * shifts by CLZ(~mask) == B (ibp_compress_dev.cuh:47,53-55,80;
 * ibp_decompress_dev.cuh:315,327-328).  PTX `shl` clamps counts >= width to a
 * zero result, so those expressions mean "all ones" and "zero".  C (and x86,
 * and gfx950's v_lshlrev_b64) take the count modulo the width instead, so every
 * shift here goes through shl_b()/shr_b()/lowmask(), which implement the
 * clamping semantics explicitly.
 *
 * Parity status of each function is recorded in the test suite that pins it:
 *   pack/unpack/size/gather  -- pinned (hand-derived vectors, bit-list brute
 *                               force, closed-form size, round trips)
 *   preprocess               -- pinned (numpy brute force, constant-column and
 *                               nothing-saved special cases, float sweep probe)
 *   random_shuffle           -- pinned against libstdc++ std::random_shuffle
 *   kmeans                   -- pinned on separable data and invariants; tie
 *                               order and the 8-byte create_mask_many counters
 *                               are defined here (reference behaviour is a race
 *                               / out-of-bounds shared memory): parity unpinned
 *                               for those two aspects.
 */
#include "ibp_oracle.h"

#include <limits.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ---- element access and clamped bit helpers ---------------------------- */

static inline uint64_t ones_b(int B) { return B >= 64 ? ~0ull : ((1ull << B) - 1ull); }

/* (1 << n) - 1 with n in [0, 64]: PTX-clamped meaning of the reference's
 * `(1ll << n) - 1ll`. */
static inline uint64_t lowmask(int n) { return n >= 64 ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1ull)); }

/* B-bit logical shifts that yield 0 for counts >= B (PTX shl/shr clamping). */
static inline uint64_t shl_b(uint64_t v, int s, int B) { return s >= B ? 0ull : ((v << s) & ones_b(B)); }
static inline uint64_t shr_b(uint64_t v, int s, int B) { return s >= B ? 0ull : ((v & ones_b(B)) >> s); }

static inline uint64_t ld(const void *base, int64_t i, int w)
{
    uint64_t v = 0;
    memcpy(&v, (const uint8_t *)base + i * w, (size_t)w); /* little-endian host */
    return v;
}

static inline void st(void *base, int64_t i, int w, uint64_t v)
{
    memcpy((uint8_t *)base + i * w, &v, (size_t)w);
}

/* POPC / CLZ of a B-bit value (include/ibp_helpers.cuh:13-47; CLZ(0) == B). */
static inline int popc_b(uint64_t v, int B) { return __builtin_popcountll(v & ones_b(B)); }
static inline int clz_b(uint64_t v, int B)
{
    v &= ones_b(B);
    if (v == 0) return B;
    return __builtin_clzll(v) - (64 - B);
}

static inline int64_t align_up(int64_t a, int64_t w) { return (a + w - 1) / w * w; }

/* ---- preprocess ---------------------------------------------------------- */

void oracle_set_threads(int threads)
{
    if (threads > 0) omp_set_num_threads(threads);
}

void oracle_count_bits(int w, const void *data, int64_t num_vecs, int64_t vec_size, int32_t *cnt)
{
    const int B = 8 * w;
    memset(cnt, 0, sizeof(int32_t) * (size_t)(vec_size * B));
    /* ibp_preproc_kernels.cuh:19-28: every row, every column, every bit.  Row
     * blocks are counted per thread and the integer counts added: the sum
     * does not depend on the order. */
#pragma omp parallel
    {
        int32_t *part = (int32_t *)calloc((size_t)(vec_size * B), sizeof(int32_t));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < num_vecs; ++i)
            for (int64_t j = 0; j < vec_size; ++j) {
                uint64_t val = ld(data, i * vec_size + j, w);
                for (int bit = 0; bit < B; ++bit)
                    if (val & (1ull << bit)) part[j * B + bit] += 1;
            }
#pragma omp critical
        for (int64_t k = 0; k < vec_size * B; ++k) cnt[k] += part[k];
        free(part);
    }
}

void oracle_create_mask(int w, const int32_t *cnt, int64_t num_vecs, int64_t vec_size,
                        float threshold, void *mask, void *bitval)
{
    const int B = 8 * w;
    /* ibp_preproc_kernels.cuh:41: `bit_count > threshold * num_vecs` -- float
     * times unsigned long long converts the count of rows to float and
     * multiplies in float; the int32 count is compared as float. */
    volatile float hi = threshold * (float)(uint64_t)num_vecs;
    /* :44: `bit_count < (1.0 - threshold) * num_vecs` -- double arithmetic. */
    volatile double lo = (1.0 - (double)threshold) * (double)(uint64_t)num_vecs;
    for (int64_t i = 0; i < vec_size; ++i) {
        uint64_t val = 0, masker = 0;
        for (int j = 0; j < B; ++j) {
            int32_t c = cnt[i * B + j];
            if ((float)c > hi) {
                val |= (1ull << j);
                masker |= (1ull << j);
            } else if ((double)c < lo) {
                masker |= (1ull << j);
            }
        }
        st(bitval, i, w, val);
        st(mask, i, w, masker);
    }
}

uint64_t oracle_check_feats(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                            const void *mask, const void *bitval)
{
    const int B = 8 * w;
    uint64_t saved = 0;
    /* ibp_preproc_kernels.cuh:60-75: per row, bits of flagged elements; only
     * the excess over vec_size metadata bits counts. */
#pragma omp parallel for schedule(static) reduction(+ : saved)
    for (int64_t i = 0; i < num_vecs; ++i) {
        uint64_t ctr = 0;
        for (int64_t j = 0; j < vec_size; ++j) {
            uint64_t v = ld(data, i * vec_size + j, w);
            uint64_t m = ld(mask, j, w), b = ld(bitval, j, w);
            if ((v & m) == b) ctr += (uint64_t)popc_b(m, B);
        }
        if (ctr > (uint64_t)vec_size) saved += ctr - (uint64_t)vec_size;
    }
    return saved;
}

float oracle_sweep_last(int *steps)
{
    /* ibp_preproc_host.cuh:257: float loop variable, += double 0.05. */
    int n = 0;
    float last = 0.f;
    for (volatile float t = 0.7f; t <= 1.0f; t = (float)((double)t + 0.05)) {
        last = t;
        ++n;
    }
    if (steps) *steps = n;
    return last;
}

int oracle_preprocess(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                      float threshold, void *mask, void *bitval, float *chosen_threshold)
{
    const int B = 8 * w;
    int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(vec_size * B));
    void *d_mask = calloc((size_t)vec_size, (size_t)w);
    void *d_bitval = calloc((size_t)vec_size, (size_t)w);
    oracle_count_bits(w, data, num_vecs, vec_size, cnt);

    /* ibp_preproc_host.cuh:61-62: float min/max thresholds. */
    float min_thresh = threshold < 0 ? 0.7f : threshold;
    float max_thresh = threshold < 0 ? 1.0f : threshold;
    double max_saved = 0; /* :64 */
    int chosen = -1, step = 0;
    for (volatile float t = min_thresh; t <= max_thresh; t = (float)((double)t + 0.05), ++step) {
        oracle_create_mask(w, cnt, num_vecs, vec_size, t, d_mask, d_bitval);
        uint64_t saved = oracle_check_feats(w, data, num_vecs, vec_size, d_mask, d_bitval);
        /* :92-99: first strict maximum, compared as double. */
        if ((double)saved > max_saved) {
            memcpy(mask, d_mask, (size_t)(vec_size * w));
            memcpy(bitval, d_bitval, (size_t)(vec_size * w));
            max_saved = (double)saved;
            chosen = step;
            if (chosen_threshold) *chosen_threshold = t;
        }
    }
    free(cnt);
    free(d_mask);
    free(d_bitval);
    return chosen;
}

/* ---- compressed size ----------------------------------------------------- */

int64_t oracle_row_size(int w, const void *row, int64_t vec_size, const void *mask, const void *bitval)
{
    const int B = 8 * w;
    const int64_t vec_bytes = vec_size * w;
    uint64_t saved_bits = 0;
    /* ibp_misc_dev.cuh:27-31 */
    for (int64_t j = 0; j < vec_size; ++j) {
        uint64_t x = ld(row, j, w), m = ld(mask, j, w), b = ld(bitval, j, w);
        if ((x & m) == b) saved_bits += (uint64_t)popc_b(m, B);
    }
    /* :36-42 */
    int64_t metadata_size = (vec_size + 7) / 8;
    int64_t data_size = vec_bytes - (int64_t)(saved_bits / 8);
    metadata_size = align_up(metadata_size, w);
    data_size = align_up(data_size, w);
    int64_t comp = metadata_size + data_size;
    return comp < vec_bytes ? comp : vec_bytes;
}

/* Mask row of source row r: clusters[r] (NULL: row 0, which is what the
 * reference applies to every row, src/compress.cu:221-232). */
static inline int64_t mask_row(const int32_t *clusters, int64_t r) { return clusters ? clusters[r] : 0; }

void oracle_get_compress_size_clustered(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                                        const void *masks, const void *bitvals, const int32_t *clusters,
                                        const int64_t *index, int64_t n, int64_t *sizes, uint64_t *ctr)
{
    (void)num_vecs;
    const int64_t vb = vec_size * w;
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = index ? index[i] : i; /* ibp_misc_kernels.cuh:18-24 */
        const int64_t c = mask_row(clusters, r);
        sizes[i] = oracle_row_size(w, (const uint8_t *)data + r * vb, vec_size, (const uint8_t *)masks + c * vb,
                                   (const uint8_t *)bitvals + c * vb);
        if (ctr) *ctr += (uint64_t)sizes[i]; /* :27-30, never reset */
    }
}

void oracle_get_compress_size(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                              const void *mask, const void *bitval, const int64_t *index,
                              int64_t n, int64_t *sizes, uint64_t *ctr)
{
    oracle_get_compress_size_clustered(w, data, num_vecs, vec_size, mask, bitval, NULL, index, n, sizes, ctr);
}

/* ---- slot codec ---------------------------------------------------------- */

void oracle_pack_row(int w, const void *row, int64_t vec_size, const void *mask,
                     const void *bitval, void *slot)
{
    const int B = 8 * w;
    /* ibp_compress_dev.cuh:13-15: metadata (one flag bit per element) rounded
     * up to the element size. */
    const int64_t bitmask_offset = align_up((vec_size + 7) / 8, w);
    memset(slot, 0, (size_t)(vec_size * w)); /* zeroed workspace, ibp_compress_host.cuh:19 */
    int64_t bitshift = 0;
    for (int64_t j = 0; j < vec_size; ++j) {
        const uint64_t val0 = ld(row, j, w), m = ld(mask, j, w), b = ld(bitval, j, w);
        uint64_t compressed_val = 0;
        int insert_size = B - popc_b(m, B); /* :35 */
        if ((val0 & m) == b) {
            /* :39-58: gather the non-mask bits of val, MSB first, into the top
             * of compressed_val, one run of zeros of the mask at a time. */
            int num_bits = insert_size;
            int temp_insert_size = clz_b(m, B);
            int num_inserted = 0;
            uint64_t local_mask = m, val = val0;
            while (num_bits > 0) {
                num_inserted += temp_insert_size;
                compressed_val |= val & shl_b(lowmask(temp_insert_size), B - num_inserted, B);
                num_bits -= temp_insert_size;
                local_mask = shl_b(local_mask, temp_insert_size, B);
                int shift = clz_b(~local_mask, B);
                local_mask = shl_b(local_mask, shift, B);
                val = shl_b(val, shift, B);
                temp_insert_size = clz_b(local_mask, B);
                if (temp_insert_size > num_bits) temp_insert_size = num_bits;
            }
            /* :60-63: header flag, T-word j / B, bit j % B (LSB first). */
            st(slot, j / B, w, ld(slot, j / B, w) | (1ull << (j % B)));
        } else {
            /* :64-73: element stored whole; header bit stays clear. */
            insert_size = B;
            compressed_val = val0;
        }
        /* :75-92: append insert_size bits MSB first at the bit cursor. */
        int fin_ins_bits = 0;
        while (insert_size > 0) {
            int64_t elem_offset = (bitmask_offset + bitshift / 8) / w;
            int bit_offset = (int)((bitmask_offset * 8 + bitshift) % B);
            int insert_bits = insert_size < B - bit_offset ? insert_size : B - bit_offset;
            uint64_t insert = shr_b(shl_b(compressed_val, fin_ins_bits, B) &
                                        shl_b(lowmask(insert_bits), B - insert_bits, B),
                                    bit_offset, B);
            st(slot, elem_offset, w, ld(slot, elem_offset, w) | insert);
            fin_ins_bits += insert_bits;
            insert_size -= insert_bits;
            bitshift += insert_bits;
        }
    }
}

void oracle_unpack_row(int w, const void *slot, int64_t vec_size, const void *mask,
                       const void *bitval, void *out)
{
    const int B = 8 * w;
    const int64_t bitmask_offset = align_up((vec_size + 7) / 8, w); /* ibp_decompress_dev.cuh:137-139 */
    const uint8_t *working_data = (const uint8_t *)slot + bitmask_offset;
    int64_t bitshift = 0;
    for (int64_t i = 0; i < vec_size; ++i) {
        /* :202-215: header flag and element length. */
        const int compressed_feat = (int)((ld(slot, i / B, w) >> (i % B)) & 1ull);
        const uint64_t m = ld(mask, i, w);
        int num_bits = B - (compressed_feat ? popc_b(m, B) : 0);
        uint64_t local_mask = m;
        uint64_t temp_dest;
        int temp_read_size;
        /* :294-304 */
        if (compressed_feat) {
            temp_dest = ld(bitval, i, w);
            temp_read_size = clz_b(local_mask, B);
            if (temp_read_size > num_bits) temp_read_size = num_bits;
        } else {
            temp_dest = 0;
            temp_read_size = num_bits;
        }
        /* :308-334: read runs from the bit stream and drop them into the zero
         * positions of the mask, MSB first. */
        int fin_read_bits = 0;
        while (num_bits > 0) {
            int bit_offset = (int)(bitshift % B);
            int read_bits = temp_read_size < B - bit_offset ? temp_read_size : B - bit_offset;
            uint64_t word = ld(working_data, bitshift / B, w);
            temp_dest |= shr_b(shl_b(word, bit_offset, B) & shl_b(lowmask(read_bits), B - read_bits, B),
                               fin_read_bits, B);
            fin_read_bits += read_bits;
            num_bits -= read_bits;
            bitshift += read_bits;
            if (compressed_feat) {
                local_mask = shl_b(local_mask, read_bits, B);
                int shift = clz_b(~local_mask, B);
                local_mask = shl_b(local_mask, shift, B);
                fin_read_bits += shift;
                temp_read_size = clz_b(local_mask, B);
                if (temp_read_size > num_bits) temp_read_size = num_bits;
            } else {
                temp_read_size = num_bits;
            }
        }
        st(out, i, w, temp_dest);
    }
}

void oracle_compress_clustered(int w, const void *input, void *output, int64_t num_vecs,
                               int64_t vec_size, const void *masks, const void *bitvals,
                               const int32_t *clusters, const int64_t *index, int64_t n, int32_t *bitmask)
{
    (void)num_vecs;
    const int64_t vec_bytes = vec_size * w;
    uint8_t *slot = (uint8_t *)malloc((size_t)vec_bytes);
    memset(bitmask, 0, sizeof(int32_t) * (size_t)((n + 31) / 32)); /* ibp_compress_host.cuh:11-13 */
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = index ? index[i] : i; /* ibp_compress_kernel.cuh:22-25 */
        const uint8_t *src = (const uint8_t *)input + r * vec_bytes;
        uint8_t *dst = (uint8_t *)output + i * vec_bytes;
        const int64_t c = mask_row(clusters, r);
        const uint8_t *m = (const uint8_t *)masks + c * vec_bytes, *b = (const uint8_t *)bitvals + c * vec_bytes;
        /* :28-31: pack only if strictly smaller (int32 closed form, same value). */
        int64_t compressed = oracle_row_size(w, src, vec_size, m, b);
        if (compressed < vec_bytes) {
            oracle_pack_row(w, src, vec_size, m, b, slot);
            memcpy(dst, slot, (size_t)vec_bytes); /* :36-39 */
            bitmask[i / 32] |= (int32_t)(1u << (i % 32)); /* :40-46 */
        } else if (dst != src) {
            memmove(dst, src, (size_t)vec_bytes); /* :48-49 */
        }
    }
    free(slot);
}

void oracle_compress_mt(int w, const void *input, void *output, int64_t num_vecs, int64_t vec_size,
                        const void *masks, const void *bitvals, const int32_t *clusters, const int64_t *index,
                        int64_t n, int32_t *bitmask, int threads)
{
    const int64_t vec_bytes = vec_size * w;
    const int64_t groups = (n + 31) / 32;
    /* 32 rows per task: a task owns its bitmask word. */
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int64_t g = 0; g < groups; ++g) {
        const int64_t lo = g * 32, cnt = (lo + 32 < n ? lo + 32 : n) - lo;
        const uint8_t *in = (const uint8_t *)input + (index ? 0 : lo * vec_bytes);
        const int32_t *cl = clusters && !index ? clusters + lo : clusters;
        oracle_compress_clustered(w, in, (uint8_t *)output + lo * vec_bytes, num_vecs, vec_size, masks, bitvals,
                                  cl, index ? index + lo : NULL, cnt, bitmask + g);
    }
}

void oracle_compress(int w, const void *input, void *output, int64_t num_vecs, int64_t vec_size,
                     const void *mask, const void *bitval, const int64_t *index, int64_t n,
                     int32_t *bitmask)
{
    oracle_compress_clustered(w, input, output, num_vecs, vec_size, mask, bitval, NULL, index, n, bitmask);
}

void oracle_decompress_fetch_clustered(int w, const void *comp, int64_t num_vecs, int64_t vec_size,
                                       const void *masks, const void *bitvals, const int32_t *clusters,
                                       const int32_t *bitmask, const int64_t *index, int64_t n, void *out)
{
    (void)num_vecs;
    const int64_t vec_bytes = vec_size * w;
    for (int64_t k = 0; k < n; ++k) {
        int64_t r = index ? index[k] : k;
        const uint8_t *src = (const uint8_t *)comp + r * vec_bytes;
        uint8_t *dst = (uint8_t *)out + k * vec_bytes;
        const int64_t c = mask_row(clusters, r);
        /* ibp_decompress_kernel.cuh:95-102 */
        if ((uint32_t)bitmask[r / 32] & (1u << (r % 32)))
            oracle_unpack_row(w, src, vec_size, (const uint8_t *)masks + c * vec_bytes,
                              (const uint8_t *)bitvals + c * vec_bytes, dst);
        else
            memcpy(dst, src, (size_t)vec_bytes);
    }
}

void oracle_decompress_fetch(int w, const void *comp, int64_t num_vecs, int64_t vec_size,
                             const void *mask, const void *bitval, const int32_t *bitmask,
                             const int64_t *index, int64_t n, void *out)
{
    oracle_decompress_fetch_clustered(w, comp, num_vecs, vec_size, mask, bitval, NULL, bitmask, index, n, out);
}

/* ---- condensed stream -------------------------------------------------------- */

int64_t oracle_compress_condensed(int w, const void *input, int64_t num_vecs, int64_t vec_size,
                                  const void *masks, const void *bitvals, const int32_t *clusters,
                                  const int64_t *index, int64_t n, void *out, int64_t *offsets)
{
    (void)num_vecs;
    const int64_t vec_bytes = vec_size * w;
    uint8_t *slot = (uint8_t *)malloc((size_t)vec_bytes);
    int64_t start = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = index ? index[i] : i; /* ibp_compress_kernel.cuh:76-80 */
        const uint8_t *src = (const uint8_t *)input + r * vec_bytes;
        const int64_t c = mask_row(clusters, r);
        const uint8_t *m = (const uint8_t *)masks + c * vec_bytes, *b = (const uint8_t *)bitvals + c * vec_bytes;
        /* src/compress.cu:276-285: sizes, then their inclusive scan. */
        const int64_t size = oracle_row_size(w, src, vec_size, m, b);
        /* ibp_compress_kernel.cuh:66-73,82-92: row i occupies [offsets[i-1],
         * offsets[i]); a row whose size is V*w is copied raw, any other is the
         * first `size` bytes of its slot (the rest of the slot is zero). */
        if (out) {
            if (size != vec_bytes) {
                oracle_pack_row(w, src, vec_size, m, b, slot);
                memcpy((uint8_t *)out + start, slot, (size_t)size);
            } else {
                memcpy((uint8_t *)out + start, src, (size_t)vec_bytes);
            }
        }
        start += size;
        offsets[i] = start;
    }
    free(slot);
    return start;
}

void oracle_decompress_condensed(int w, const void *comp, const int64_t *offsets, int64_t num_rows,
                                 int64_t vec_size, const void *masks, const void *bitvals,
                                 const int32_t *clusters, const int64_t *index, int64_t n, void *out)
{
    (void)num_rows;
    const int64_t vec_bytes = vec_size * w;
    uint8_t *slot = (uint8_t *)malloc((size_t)vec_bytes);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t r = index ? index[k] : k;
        const int64_t lo = r == 0 ? 0 : offsets[r - 1], len = offsets[r] - lo;
        uint8_t *dst = (uint8_t *)out + k * vec_bytes;
        const int64_t c = mask_row(clusters, r);
        if (len == vec_bytes) {
            memcpy(dst, (const uint8_t *)comp + lo, (size_t)vec_bytes);
        } else {
            /* the stored prefix of the slot; its zero tail was not stored */
            memset(slot, 0, (size_t)vec_bytes);
            memcpy(slot, (const uint8_t *)comp + lo, (size_t)len);
            oracle_unpack_row(w, slot, vec_size, (const uint8_t *)masks + c * vec_bytes,
                              (const uint8_t *)bitvals + c * vec_bytes, dst);
        }
    }
    free(slot);
}

/* ---- k-means --------------------------------------------------------------- */

void oracle_random_shuffle(int32_t *a, int64_t n)
{
    /* libstdc++ bits/stl_algo.h random_shuffle(first, last): for i in
     * [first+1, last): j = first + rand() % ((i - first) + 1); swap if i != j. */
    if (n <= 0) return;
    for (int64_t i = 1; i < n; ++i) {
        int64_t j = (int64_t)(rand() % (int)(i + 1));
        if (i != j) {
            int32_t t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
}

/* ibp_preproc_kmeans.cuh:106-150 (cluster_vecs) with a defined tie rule. */
static void km_assign(int w, const void *data, int64_t N, int64_t V, int K,
                      const uint64_t *centroids, int32_t *cluster)
{
    const int B = 8 * w;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        int best = 0;
        int64_t best_d = INT64_MAX;
        for (int c = 0; c < K; ++c) {
            int64_t d = 0;
            for (int64_t k = 0; k < V; ++k)
                d += popc_b(ld(data, i * V + k, w) ^ centroids[(int64_t)c * V + k], B);
            if (d <= best_d) { /* '<=': the highest index among equals wins */
                best_d = d;
                best = c;
            }
        }
        cluster[i] = best;
    }
}

/* ibp_preproc_kmeans.cuh:196-217 (compute_new_centroids) counts. */
static void km_count(int w, const void *data, int64_t N, int64_t V, int K, const int32_t *cluster,
                     int32_t *count, int32_t *bits_set)
{
    const int B = 8 * w;
    memset(count, 0, sizeof(int32_t) * (size_t)K);
    memset(bits_set, 0, sizeof(int32_t) * (size_t)((int64_t)K * V * B));
    for (int64_t i = 0; i < N; ++i) count[cluster[i]] += 1;
    /* Columns are split over threads: each counter has one writer. */
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < V; ++k)
        for (int64_t i = 0; i < N; ++i) {
            const int c = cluster[i];
            uint64_t x = ld(data, i * V + k, w);
            for (int bit = 0; bit < B; ++bit)
                if (x & (1ull << bit)) bits_set[((int64_t)c * V + k) * B + bit] += 1;
        }
}

void oracle_kmeans(int w, const void *data, int64_t N, int64_t V, int K, int64_t seed,
                   void *masks, void *bitvals, int32_t *clusters)
{
    const int B = 8 * w;
    uint64_t *centroids = (uint64_t *)calloc((size_t)((int64_t)K * V), sizeof(uint64_t));
    int32_t *count = (int32_t *)calloc((size_t)K, sizeof(int32_t));
    int32_t *bits_set = (int32_t *)calloc((size_t)((int64_t)K * V * B), sizeof(int32_t));
    memset(clusters, 0, sizeof(int32_t) * (size_t)N); /* ibp_preproc_host.cuh:147 */

    /* :150-151: centroid 0 is row 0.  :176-188: centroids 1..min(K,N)-1 are
     * rows indices[1..] of a random_shuffle of 0..N-1.  Centroids K > N are
     * left unset by the reference (raw device memory); they are zero here. */
    if (N > 0)
        for (int64_t k = 0; k < V; ++k) centroids[k] = ld(data, k, w);
    if (seed >= 0) srand((unsigned)seed);
    {
        int32_t *indices = (int32_t *)malloc(sizeof(int32_t) * (size_t)(N > 0 ? N : 1));
        for (int64_t i = 0; i < N; ++i) indices[i] = (int32_t)i;
        oracle_random_shuffle(indices, N);
        for (int64_t c = 1; c < (K < N ? K : N); ++c)
            for (int64_t k = 0; k < V; ++k)
                centroids[c * V + k] = ld(data, (int64_t)indices[c] * V + k, w);
        free(indices);
    }

    /* :191-233: ten Lloyd iterations. */
    for (int it = 0; it < 10; ++it) {
        km_assign(w, data, N, V, K, centroids, clusters);
        km_count(w, data, N, V, K, clusters, count, bits_set);
        /* ibp_preproc_kmeans.cuh:219-236 (construct_bitmasks): one 32-lane warp
         * per centroid; lane l visits columns l, l+32, ... and its accumulator
         * is never cleared between columns, so column j receives the OR of the
         * majority bits of columns l, l+32, ..., j. */
        for (int c = 0; c < K; ++c)
            for (int lane = 0; lane < 32; ++lane) {
                uint64_t bitmask = 0;
                for (int64_t id = lane; id < V; id += 32) {
                    for (int bit = 0; bit < B; ++bit)
                        if (bits_set[((int64_t)c * V + id) * B + bit] > count[c] / 2)
                            bitmask |= (1ull << bit);
                    centroids[(int64_t)c * V + id] = bitmask;
                }
            }
    }

    /* ibp_preproc_host.cuh:235,257-268: create_mask_many runs at 0.9 and then
     * at every step of the 0.7..1.0 sweep, each time overwriting all K rows;
     * the result is the one of the last sweep value (1.0f).
     * ibp_preproc_kmeans.cuh:273-293 with one float counter per bit. */
    const float t = oracle_sweep_last(NULL);
    km_count(w, data, N, V, K, clusters, count, bits_set);
    for (int c = 0; c < K; ++c) {
        volatile float hi = t * (float)count[c];
        volatile double lo = (1.0 - (double)t) * (double)count[c];
        for (int64_t id = 0; id < V; ++id) {
            uint64_t val = 0, masker = 0;
            for (int bit = 0; bit < B; ++bit) {
                float bs = (float)bits_set[((int64_t)c * V + id) * B + bit];
                if (bs >= hi) {
                    val |= (1ull << bit);
                    masker |= (1ull << bit);
                } else if ((double)bs <= lo) {
                    masker |= (1ull << bit);
                }
            }
            st(masks, (int64_t)c * V + id, w, masker);
            st(bitvals, (int64_t)c * V + id, w, val);
        }
    }
    free(centroids);
    free(count);
    free(bits_set);
}
