/*
 * ibp_oracle.h -- CPU restatement of the Invariant Bit Packing (IBP) algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The
 * product (libibp) never links, loads or calls it.
 *
 * Data are passed as raw bytes.  Elements are unsigned little-endian integers
 * of `w` bytes (w in {1,2,4,8}; the slot codec takes w in {4,8}).  B = 8*w.
 * Every function cites the reference file:line it restates (paths relative to
 * the AKKamath/InvariantBitPacking tree).
 */
#ifndef IBP_ORACLE_H
#define IBP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/preproc/ibp_preproc_kernels.cuh:8-31 (count_bit_kernel):
 * cnt[j*B + k] = number of rows whose column j has bit k set. */
/* Threads of the OpenMP loops below (<= 0 leaves the OpenMP default). */
void oracle_set_threads(int threads);

void oracle_count_bits(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                       int32_t *cnt);

/* include/preproc/ibp_preproc_kernels.cuh:34-51 (create_mask) */
void oracle_create_mask(int w, const int32_t *cnt, int64_t num_vecs, int64_t vec_size,
                        float threshold, void *mask, void *bitval);

/* include/preproc/ibp_preproc_kernels.cuh:54-77 (check_feats) */
uint64_t oracle_check_feats(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                            const void *mask, const void *bitval);

/* include/preproc/ibp_preproc_host.cuh:24-115 (preproc_data) with
 * src/preprocess.cu:17-76.  threshold < 0 selects the 0.7..1.0 sweep.
 * mask/bitval must be zero-initialised by the caller (torch::zeros,
 * src/preprocess.cu:46-47).  Returns the index of the chosen threshold in the
 * sweep, or -1 when no threshold saved any bit.  chosen_threshold (optional)
 * receives it. */
int oracle_preprocess(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                      float threshold, void *mask, void *bitval, float *chosen_threshold);

/* include/misc/ibp_misc_dev.cuh:20-46 (check_compress_size_blk): bytes of the
 * slot for one row, min(V*w, align_w(ceil(V/8)) + align_w(V*w - floor(S/8))). */
int64_t oracle_row_size(int w, const void *row, int64_t vec_size, const void *mask,
                        const void *bitval);

/* src/compress.cu:25-174 + include/misc/ibp_misc_kernels.cuh:9-32.
 * index may be NULL (then n rows 0..n-1).  ctr may be NULL; otherwise the sum
 * of sizes is ADDED to *ctr (never reset, ibp_misc_kernels.cuh:27-30). */
void oracle_get_compress_size(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                              const void *mask, const void *bitval,
                              const int64_t *index, int64_t n, int64_t *sizes,
                              uint64_t *ctr);

/* include/compress/ibp_compress_dev.cuh:8-97 (compress_and_write): writes the
 * slot of one row into `slot` (V*w bytes, fully overwritten: header, payload,
 * zero tail). */
void oracle_pack_row(int w, const void *row, int64_t vec_size, const void *mask,
                     const void *bitval, void *slot);

/* include/compress/ibp_compress_kernel.cuh:10-56 + ibp_compress_host.cuh:8-27:
 * out slot i <- row index[i] (or row i), packed when its size is < V*w, raw
 * otherwise.  bitmask has ceil(n/32) words and is fully rewritten.
 * `input` and `output` may alias only when index is NULL. */
void oracle_compress(int w, const void *input, void *output, int64_t num_vecs,
                     int64_t vec_size, const void *mask, const void *bitval,
                     const int64_t *index, int64_t n, int32_t *bitmask);

/* include/decompress/ibp_decompress_dev.cuh:279-336 (decode arithmetic of
 * decompress_fetch_cpu): inverse of oracle_pack_row. */
void oracle_unpack_row(int w, const void *slot, int64_t vec_size, const void *mask,
                       const void *bitval, void *out);

/* src/decompress.cu:10-156 + include/decompress/ibp_decompress_kernel.cuh:63-103:
 * out[k] = unpack(slot[index[k]]) (packed rows) or the raw slot.  index may be
 * NULL (then n = num_vecs rows in order). */
void oracle_decompress_fetch(int w, const void *comp, int64_t num_vecs, int64_t vec_size,
                             const void *mask, const void *bitval, const int32_t *bitmask,
                             const int64_t *index, int64_t n, void *out);

/* Per-cluster forms of the three functions above.  masks/bitvals are [K x V]
 * (the preprocess_kmeans output); clusters (NULL: row 0 for every row, which
 * is what the reference applies, src/compress.cu:221-232 and
 * src/decompress.cu:130-134) names the mask row of each row: of each SOURCE
 * row (dataset order) for size and compress, of each STORED row for fetch. */
void oracle_get_compress_size_clustered(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                                        const void *masks, const void *bitvals, const int32_t *clusters,
                                        const int64_t *index, int64_t n, int64_t *sizes, uint64_t *ctr);
void oracle_compress_clustered(int w, const void *input, void *output, int64_t num_vecs,
                               int64_t vec_size, const void *masks, const void *bitvals,
                               const int32_t *clusters, const int64_t *index, int64_t n, int32_t *bitmask);
void oracle_decompress_fetch_clustered(int w, const void *comp, int64_t num_vecs, int64_t vec_size,
                                       const void *masks, const void *bitvals, const int32_t *clusters,
                                       const int32_t *bitmask, const int64_t *index, int64_t n, void *out);

/* oracle_compress_clustered on `threads` OpenMP threads, 32 rows per task
 * (same output).  Used to time the host-side compression path. */
void oracle_compress_mt(int w, const void *input, void *output, int64_t num_vecs, int64_t vec_size,
                        const void *masks, const void *bitvals, const int32_t *clusters, const int64_t *index,
                        int64_t n, int32_t *bitmask, int threads);

/* src/compress.cu:238-313 + include/compress/ibp_compress_kernel.cuh:59-94
 * (compress_condensed): rows index[i] (or i) stored back to back, row i in
 * bytes [offsets[i-1], offsets[i]) (offsets[-1] = 0) -- the inclusive scan of
 * the row sizes.  A row of size V*w is stored raw, any other as the first
 * `size` bytes of its slot.  out (NULL: offsets only) must hold the returned
 * total.  clusters as above (per source row). */
int64_t oracle_compress_condensed(int w, const void *input, int64_t num_vecs, int64_t vec_size,
                                  const void *masks, const void *bitvals, const int32_t *clusters,
                                  const int64_t *index, int64_t n, void *out, int64_t *offsets);

/* Inverse of oracle_compress_condensed (the reference has none: its
 * `decompress` binding is commented out, src/ibp_api.cu:134): out[k] = stored
 * row index[k] (or k), unpacked when its stored length differs from V*w.
 * clusters per stored row. */
void oracle_decompress_condensed(int w, const void *comp, const int64_t *offsets, int64_t num_rows,
                                 int64_t vec_size, const void *masks, const void *bitvals,
                                 const int32_t *clusters, const int64_t *index, int64_t n, void *out);

/* libstdc++ std::random_shuffle(first, last) over 0..n-1 driven by rand()
 * (include/preproc/ibp_preproc_host.cuh:176-180). */
void oracle_random_shuffle(int32_t *indices, int64_t n);

/* include/preproc/ibp_preproc_host.cuh:119-284 + ibp_preproc_kmeans.cuh:106-297.
 * If seed >= 0 the libc generator is re-seeded with srand(seed) first; otherwise
 * the process-wide rand() state continues, as in the reference.  masks/bitvals
 * are [K x V]; clusters is [N].  Tie rule for equal Hamming distances: the
 * highest centroid index wins (the reference leaves it to an atomicExch race). */
void oracle_kmeans(int w, const void *data, int64_t num_vecs, int64_t vec_size,
                   int num_clusters, int64_t seed, void *masks, void *bitvals,
                   int32_t *clusters);

/* The last value of `for(float t = 0.7; t <= 1.0; t += 0.05)`
 * (ibp_preproc_host.cuh:257) and the number of steps it takes. */
float oracle_sweep_last(int *steps);

#ifdef __cplusplus
}
#endif
#endif /* IBP_ORACLE_H */
