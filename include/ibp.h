/*
 * ibp.h -- C ABI of libibp, the MI355X-native Invariant Bit Packing library.
 *
 * Plain pointers and sizes only (no torch types).  Every entry point returns an
 * ibp_status; on failure ibp_last_error() (thread-local) holds the message.
 * Work is enqueued on the caller's HIP stream (`stream` is a hipStream_t; NULL
 * is the null stream).  Unless stated otherwise a call does not synchronise.
 *
 * Element width `elem_bytes` is 1, 2, 4 or 8 bytes.  The reference's codec
 * takes 4 or 8 (src/compress.cu:191-192); 1- and 2-byte elements use the
 * same slot format with 8- and 16-bit words.
 * Index arrays hold `index_bytes` = 4 (int32) or 8 (int64) signed entries.
 * Data pointers may be device memory or pinned (hipHostMalloc) host memory.
 *
 * Each declaration names the reference interface it replaces
 * (AKKamath/InvariantBitPacking, file:line).
 */
#ifndef IBP_H
#define IBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ibp_status {
    IBP_OK = 0,
    IBP_ERR_INVALID = 1,     /* bad argument (the reference's TORCH_CHECK failures) */
    IBP_ERR_HIP = 2,         /* HIP runtime error (the reference exits the process) */
    IBP_ERR_NOMEM = 3,       /* allocation failed / admission impossible */
    IBP_ERR_UNSUPPORTED = 4, /* valid request outside what this build implements */
    IBP_ERR_BUSY = 5         /* serving: request not finished yet */
} ibp_status;

/* Where a data pointer lives; IBP_LOC_AUTO asks the HIP runtime. */
typedef enum ibp_loc { IBP_LOC_AUTO = 0, IBP_LOC_DEVICE = 1, IBP_LOC_HOST = 2 } ibp_loc;

/* ---- library state --------------------------------------------------------- */

/* Thread-local description of the last failure (empty string if none). */
const char *ibp_last_error(void);

/* Library version string. */
const char *ibp_version(void);

/* Replaces print_debug_msg (src/ibp_api.cu:9-12; ibp/__init__.py:8-9). */
void ibp_set_debug(int flag);

/* Number of rows that device kernels skipped because their index was out of
 * range since the last reset (the reference reads out of bounds instead).  The
 * counter lives in device memory; reading it synchronises the device. */
int ibp_index_errors(uint64_t *count, int reset);

/* ---- preprocess -------------------------------------------------------------- */

/* Replaces preprocess (src/preprocess.cu:17-76) -> ibp::preproc_data
 * (include/preproc/ibp_preproc_host.cuh:24-115).  data [num_vecs x vec_size];
 * threshold < 0 runs the 0.7..1.0 sweep.  mask/bitval [vec_size] (device) are
 * written only when some threshold saves bits: the caller zero-initialises them
 * (the reference returns torch::zeros in that case).  chosen_step (device
 * int32, optional) receives the sweep index picked or -1.  workspace: device
 * memory of ibp_preprocess_workspace_size() bytes, or NULL to let the library
 * allocate (stream-ordered).  Asynchronous. */
size_t ibp_preprocess_workspace_size(int elem_bytes, int64_t vec_size, float threshold);
int ibp_preprocess(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                   float threshold, void *mask, void *bitval, int32_t *chosen_step,
                   void *workspace, size_t workspace_bytes, void *stream);

/* Replaces preprocess_kmeans (src/preprocess.cu:88-139) -> ibp::preproc_kmeans
 * (include/preproc/ibp_preproc_host.cuh:119-284).  masks/bitvals [K x V]
 * (device), clusters [num_vecs] int32 (device or pinned host).  seed >= 0
 * draws the random_shuffle initialisation from the sequence rand() returns
 * after srand(seed), computed privately (the process-wide generator is not
 * touched); seed < 0 continues the process-wide rand() state like the
 * reference.  threshold is accepted and ignored, as in the reference.
 * Synchronises the stream before returning. */
int ibp_preprocess_kmeans(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                          int num_clusters, float threshold, int64_t seed, void *masks,
                          void *bitvals, int32_t *clusters, void *stream);

/* ---- row layouts -------------------------------------------------------------- */

/* Optional description of how compressed rows are stored and coded, taken by
 * the *_ex entry points (NULL = the reference's slot layout with mask row 0).
 *  - offsets != NULL: CONDENSED stream (compress_condensed, src/compress.cu:
 *    238-313): stored row s is bytes [offsets[s-1], offsets[s]) of the stream
 *    (offsets[-1] = 0), i.e. offsets is the inclusive scan of the row sizes.
 *    A row whose length is V*w is raw; any other is its slot cut at its size.
 *    stream_bytes bounds every read of the stream.  No bitmask is needed.
 *  - clusters != NULL: mask/bitval are [num_clusters x V] (preprocess_kmeans
 *    output) and row r is coded with row clusters[r].  For size and compress
 *    clusters is indexed by SOURCE row (the dataset's row, as preprocess_kmeans
 *    returns it); for fetch by STORED row.  The reference applies mask row 0 to
 *    every row (src/compress.cu:221-232, src/decompress.cu:130-134).
 * Rows whose byte range or cluster id is invalid are fetched as zero rows and
 * counted like out-of-range indices (ibp_index_errors). */
typedef struct ibp_layout {
    const int64_t *offsets;
    int64_t stream_bytes;
    const int32_t *clusters;
    int64_t num_clusters;
} ibp_layout;

/* ---- compressed size and pack ----------------------------------------------- */

/* Replaces get_compress_size (src/compress.cu:25-174) ->
 * check_compress_size_kernel (include/misc/ibp_misc_kernels.cuh:9-32).
 * sizes [n] int64 (device or pinned host).  index NULL means rows 0..n-1.
 * total (device uint64, optional) is ADDED to, never reset.  Asynchronous. */
int ibp_get_compress_size(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                          const void *mask, const void *bitval, const void *index,
                          int index_bytes, int64_t n, int64_t *sizes, uint64_t *total,
                          void *stream);
/* ibp_get_compress_size with per-row mask rows (layout->clusters; offsets
 * must be NULL).  layout NULL is ibp_get_compress_size. */
int ibp_get_compress_size_ex(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                             const void *mask, const void *bitval, const void *index,
                             int index_bytes, int64_t n, int64_t *sizes, uint64_t *total,
                             const ibp_layout *layout, void *stream);

/* First half of compress_condensed (src/compress.cu:238-287): the row sizes
 * of get_compress_size and their inclusive scan, into offsets [n] (device or
 * pinned host).  offsets[n-1] is the stream's byte length; the caller reads it
 * (after synchronising) to size the stream.  layout: clusters only.
 * Asynchronous. */
int ibp_condensed_offsets(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                          const void *mask, const void *bitval, const void *index, int index_bytes,
                          int64_t n, int64_t *offsets, const ibp_layout *layout, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * ibp_condensed_offsets with no allocation of any kind.  The scan's tile sums
 * live in a fixed device array owned by the library (512 KiB per device, held
 * until the process ends), so n is limited to 268,435,456 rows (larger n is
 * IBP_ERR_INVALID) and offsets must be device memory.  Every call on a device
 * shares that array: calls that overlap on one device (e.g. on two streams)
 * race and must be ordered by the caller.  Asynchronous. */
int ibp_condensed_offsets_noalloc(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                                  const void *mask, const void *bitval, const void *index, int index_bytes,
                                  int64_t n, int64_t *offsets, const ibp_layout *layout, void *stream);

/* Replaces compress_inplace (src/compress.cu:176-236) -> ibp::compress_inplace
 * (include/compress/ibp_compress_host.cuh:8-27).  Slot i of `output` receives
 * row index[i] (or row i) of `input`, packed when that makes it smaller.
 * output == input gives the in-place form; with an index the two must not
 * alias (the reference races there).  bitmask [ceil(n/32)] int32 (device) is
 * fully written, no pre-zeroing needed.  Asynchronous. */
int ibp_compress(int elem_bytes, const void *input, void *output, int64_t num_vecs,
                 int64_t vec_size, const void *mask, const void *bitval, const void *index,
                 int index_bytes, int64_t n, int32_t *bitmask, void *stream);
/* ibp_compress with a layout.  layout->offsets (from ibp_condensed_offsets)
 * makes it the second half of compress_condensed (src/compress.cu:288-312 ->
 * compress_condensed_kernel, include/compress/ibp_compress_kernel.cuh:59-94):
 * row index[i] (or i) is written to bytes [offsets[i-1], offsets[i]) of
 * `output` (the stream, layout->stream_bytes long, must not overlap input);
 * bitmask may then be NULL, otherwise it is written as for slots.
 * layout->clusters selects each source row's mask row. */
int ibp_compress_ex(int elem_bytes, const void *input, void *output, int64_t num_vecs,
                    int64_t vec_size, const void *mask, const void *bitval, const void *index,
                    int index_bytes, int64_t n, int32_t *bitmask, const ibp_layout *layout,
                    void *stream);

/* ---- fetch (decompress) ------------------------------------------------------- */

/* Replaces decompress_fetch (src/decompress.cu:10-156) -> ibp::decompress_fetch
 * (include/decompress/ibp_decompress_host.cuh:9-168).
 * out[k] = unpack(comp[index[k]]) for k < n (index NULL: n rows in order).
 * comp [num_vecs x vec_size] may be pinned host memory (read over PCIe inside
 * the kernel) or device memory.  comp_len: expected compressed row length in
 * elements (<= 0 means vec_size), a performance hint only.  impl: -1 auto,
 * 0/1/5 unpack (identical output), 2 raw gather without unpack, 3/4 the
 * reference's transfer-only baselines (wrong output by design).  nblks/nthreads
 * <= 0 pick defaults.  comp_loc (ibp_loc) says where comp lives; AUTO asks the
 * runtime.  Host-resident rows of <= 64 elements default to half a block per
 * CU: enough to saturate PCIe, leaving CUs to other streams' kernels.  Out-of-range indices produce zero rows and are counted
 * (ibp_index_errors).  No allocation, no host synchronisation: safe under
 * stream capture. */
int ibp_decompress_fetch(int elem_bytes, void *out, const void *comp, int64_t num_vecs,
                         int64_t vec_size, const void *mask, const void *bitval,
                         const int32_t *bitmask, const void *index, int index_bytes, int64_t n,
                         int comp_len, int nblks, int nthreads, int impl, int comp_loc,
                         void *stream);
/* ibp_decompress_fetch with a layout.  With layout->offsets, comp is a
 * condensed stream of num_vecs stored rows and bitmask may be NULL: this is
 * the stream's decompress, which the reference declares but never built
 * (src/ibp_api.cu:134).  With layout->clusters, stored row s is unpacked with
 * mask row clusters[s].  impl 2/3/4 take neither. */
int ibp_decompress_fetch_ex(int elem_bytes, void *out, const void *comp, int64_t num_vecs,
                            int64_t vec_size, const void *mask, const void *bitval,
                            const int32_t *bitmask, const void *index, int index_bytes, int64_t n,
                            int comp_len, int nblks, int nthreads, int impl, int comp_loc,
                            const ibp_layout *layout, void *stream);

/* Launch shape of ibp_decompress_fetch for rows of <= 64 elements: rows kept
 * in flight per wave (2, 4, 8 or 16) and resident blocks per CU (1..32).
 * Process-wide; values are unaffected. */
int ibp_set_fetch_tuning(int rows_in_flight, int blocks_per_cu);

/* ---- batched fetch (no reference counterpart: many tables in one launch) ---- */

/* One table fetch of a batch: fields as in ibp_decompress_fetch. */
typedef struct ibp_fetch_task {
    const void *comp;
    const void *mask;
    const void *bitval;
    const int32_t *bitmask;
    const void *index; /* NULL: rows 0..n-1 */
    void *out;
    int64_t num_vecs;
    int64_t vec_size;
    int64_t n;
    int32_t elem_bytes;
    int32_t index_bytes;
    int32_t comp_len; /* hint, <= 0 means vec_size */
    int32_t comp_loc; /* ibp_loc */
    /* Condensed stream (as ibp_layout.offsets / stream_bytes): the inclusive
     * end byte offset of each of the num_vecs stored rows in comp, or NULL for
     * the slot layout; bitmask may then be NULL.  A plan with condensed tasks
     * holds only condensed tasks, all with rows of <= 64 elements of one
     * element width and one index width. */
    const int64_t *offsets;
    int64_t stream_bytes;
} ibp_fetch_task;

typedef struct ibp_plan ibp_plan;

/* Builds a plan for a fixed set of tasks (pointers are captured: refill the
 * index/output buffers in place between launches).  The plan owns a device work
 * queue; all tasks run in ONE kernel launch: segmented (each task its own
 * blocks) when every task has rows of <= 64 elements of one element width and
 * one index width (always for condensed tasks), else persistent over a work
 * queue. */
int ibp_plan_create(const ibp_fetch_task *tasks, int num_tasks, ibp_plan **plan);
int ibp_plan_destroy(ibp_plan *plan);
/* Launch on a stream (graph-capturable: no allocation, no memset node). */
int ibp_plan_launch(ibp_plan *plan, void *stream);
/* Times candidate launch configurations on `stream` (runs the plan several
 * times) and keeps the fastest; returns its mean time in *best_ms. */
int ibp_plan_autotune(ibp_plan *plan, void *stream, float *best_ms);
/* Captures the launch into a HIP graph owned by the plan; replays with
 * ibp_plan_graph_launch. */
int ibp_plan_capture_graph(ibp_plan *plan, void *stream);
int ibp_plan_graph_launch(ibp_plan *plan, void *stream);
/* Describes the chosen configuration (rows per work item, grid, ...). */
int ibp_plan_describe(ibp_plan *plan, char *buf, size_t len);
/* The launch configuration a plan uses (after ibp_plan_autotune: the best
 * one found) as 7 integers: rows per work item of tables with rows of <= 64
 * elements and of wider ones, rows in flight per wave (2, 4 or 8), blocks,
 * work items claimed per queue atomic, dynamic queue (0/1), segmented launch
 * where the tasks allow it (0/1).  ibp_plan_set_config applies such a
 * configuration (for instance one an earlier autotune stored) without
 * timing anything; a value out of range is IBP_ERR_INVALID.  A captured
 * graph is dropped by set_config and must be captured again. */
int ibp_plan_get_config(ibp_plan *plan, int32_t cfg[7]);
int ibp_plan_set_config(ibp_plan *plan, const int32_t cfg[7]);

/* ---- serving (no reference counterpart: continuous batching of requests) ---- */

typedef struct ibp_server ibp_server;

typedef struct ibp_server_config {
    size_t hbm_budget_bytes; /* output arena size (admission by free HBM) */
    int num_streams;         /* concurrent batches on the device */
    int max_batch_requests;  /* requests merged into one launch */
    int64_t max_batch_rows;  /* rows merged into one launch */
} ibp_server_config;

int ibp_server_create(const ibp_server_config *cfg, ibp_server **srv);
/* Stops the server: launches in flight are completed, pending requests are
 * dropped, all memory is freed.  No other call on the same server may be in
 * progress or follow.  The other ibp_server_* calls are thread-safe with
 * respect to each other (one submitting thread and one waiting thread is the
 * usual pattern); a request must be waited for and released exactly once. */
int ibp_server_destroy(ibp_server *srv);
/* Registers a compressed tensor set; returns its id in *set_id. */
int ibp_server_register(ibp_server *srv, int elem_bytes, const void *comp, int64_t num_vecs,
                        int64_t vec_size, const void *mask, const void *bitval,
                        const int32_t *bitmask, int comp_len, int comp_loc, int *set_id);
/* Submits a request: fetch rows `index` (int64, host memory, copied at submit)
 * of set `set_id`.  Returns immediately with *req_id; the request is admitted
 * when its output fits the arena.  Requests are admitted in submission order. */
int ibp_server_submit(ibp_server *srv, int set_id, const int64_t *index, int64_t n,
                      int64_t *req_id);
/* ibp_server_submit with the request's arrival time (ibp_server_clock_us
 * microseconds, not in the future; <= 0 means now): latencies and the
 * server's p50/p99 count from it, e.g. from a request's scheduled arrival. */
int ibp_server_submit_at(ibp_server *srv, int set_id, const int64_t *index, int64_t n,
                         double arrival_us, int64_t *req_id);
/* The server's clock: host steady-clock microseconds. */
double ibp_server_clock_us(void);
/* Waits for a request (timeout_ms < 0: forever).  On success *out points to
 * the [n x vec_size] result in the device arena, valid until
 * ibp_server_release(req_id); *latency_us is arrival -> completion. */
int ibp_server_wait(ibp_server *srv, int64_t req_id, int timeout_ms, void **out,
                    double *latency_us);
/* Returns a finished request's output region to the arena at once: the
 * caller must have no work in flight that still reads it. */
int ibp_server_release(ibp_server *srv, int64_t req_id);
/* Returns the region once the work queued on `stream` (hipStream_t) before
 * this call has finished: an event is recorded on `stream`, and later requests
 * are placed in the region only after it completes.  The request id is
 * invalid on return. */
int ibp_server_release_after(ibp_server *srv, int64_t req_id, void *stream);
/* Statistics since creation: completed requests, launches, rows, p50/p99
 * latency (us), peak arena use (bytes), admissions that had to wait. */
typedef struct ibp_server_stats {
    int64_t completed;
    int64_t launches;
    int64_t rows;
    double p50_us;
    double p99_us;
    double mean_batch_requests;
    size_t peak_arena_bytes;
    int64_t waited_admissions;
} ibp_server_stats;
int ibp_server_stats_get(ibp_server *srv, ibp_server_stats *st);
/* Diagnostics: host steady-clock microseconds at which a request arrived,
 * was admitted, entered the launch path, had its launch issued and was seen
 * complete (times[0..4]; 0 = not reached).  Valid until ibp_server_release. */
int ibp_server_request_times(ibp_server *srv, int64_t req_id, double *times);
/* Test hook: pretend the arena has only `bytes` free (0 disables). */
int ibp_server_set_budget_override(ibp_server *srv, size_t bytes);

/* ---- file loading ------------------------------------------------------------ */

/* The copy step of read_shared (src/ibp_api.cu:52-122): reads the whole file
 * at `path` into dst (host memory, `capacity` bytes; bytes past the file's end
 * are left as they are) with `threads` reader threads (<= 0: 16).  A file larger
 * than capacity is IBP_ERR_INVALID (the reference copies past the tensor).
 * *bytes_read receives the file size.  Synchronous, no GPU work. */
int ibp_read_file(const char *path, void *dst, size_t capacity, int threads, size_t *bytes_read);

/* ---- self-test of the element-level bit arithmetic on the host ------------ */
/* Runs the codec's element functions (compiled for host) against a brute-force
 * bit-by-bit model on `trials` random cases; returns IBP_OK or IBP_ERR_INVALID
 * with the first mismatch in ibp_last_error(). */
int ibp_selftest_bitops(int trials, uint64_t seed);
/* Runs the server's admission bookkeeping (output arena, pinned index arena,
 * in-order admission) on the host with a mock launcher: a hand-built case in
 * which the queue head can only be admitted by taking back the pinned index
 * space of later requests, then `trials` random submit / launch / complete /
 * release sequences on small arenas, each of which must drain with every
 * request's indices intact.  No GPU work. */
int ibp_selftest_serve_admission(int trials, uint64_t seed);
/* Compares the private copy of glibc's rand() sequence that preprocess_kmeans
 * uses for seed >= 0 with rand() itself for n draws after srand(seed) (this
 * re-seeds the process-wide generator).  No GPU work. */
int ibp_selftest_rand(uint32_t seed, int64_t n);
/* Runs the server's split copy of request indices into pinned memory (the
 * submitting thread and helper threads claiming 64 KiB parts) on host
 * buffers: `trials` copies of random sizes from each of three submitting
 * threads at once, with idle gaps that put the helpers to sleep; every copy
 * must match its source and leave the bytes around it untouched.  No GPU
 * work. */
int ibp_selftest_index_copy(int trials, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* IBP_H */
