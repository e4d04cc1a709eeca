// ibp_batch.hip -- batched fetch: many tables of uneven row counts and widths
// fetched and unpacked by ONE kernel launch.
//
// No reference counterpart (the reference issues one decompress_fetch per
// table, e.g. tests/dlrm_comp.py:79-81); values are those of
// ibp_decompress_fetch on each table.
//
// Two launch forms.  Segmented: when every task has rows of <= 64 elements of
// one element width and one index width (always for condensed streams), each
// task owns a range of blocks that runs the plain fetch's loop over its rows.
// Persistent: otherwise, waves take work items = (task, range of rows), by
// static striding or by claiming them with one atomic on a queue head; the
// head is reset by a small kernel launched before on the same stream, so a
// launch is at most two kernel nodes and can be captured into a HIP graph (no
// memset node).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ibp_batch.h"
#include "ibp_internal.h"

namespace ibp {

__global__ void batch_reset_kernel(unsigned long long *head)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *head = 0;
}

// Rows [b, e) of task t.  KINDS (kTasks* flags) says which kinds of task the
// launch holds: a kernel compiled for one kind needs far fewer registers than
// one holding every path, so more waves fit.  A small-row task's column table
// is built per item (two loads per lane); keeping it across items would hold
// registers for the whole loop.
template <int ROWS, int KINDS>
__device__ __forceinline__ void run_item(const TaskDev &t, int64_t b, int64_t e)
{
    const FetchParams &p = t.p;
    const bool i32 = p.index != nullptr && t.index_bytes == 4;
    const int kind = !t.small ? kTasksLarge : t.elem_bytes == 8 ? kTasksSmall8 : kTasksSmall4;
    if constexpr ((KINDS & kTasksLarge) != 0) {
        if (KINDS == kTasksLarge || kind == kTasksLarge) {
            if (t.elem_bytes == 8) {
                if (i32) fetch_large_range<uint64_t, int32_t>(p, b, e, 1);
                else fetch_large_range<uint64_t, int64_t>(p, b, e, 1);
            } else {
                if (i32) fetch_large_range<uint32_t, int32_t>(p, b, e, 1);
                else fetch_large_range<uint32_t, int64_t>(p, b, e, 1);
            }
            return;
        }
    }
    if constexpr ((KINDS & kTasksSmall8) != 0) {
        if (KINDS == kTasksSmall8 || kind == kTasksSmall8) {
            const SmallCols<uint64_t> c = load_small_cols<uint64_t>(p.mask, p.bitval, (int)p.vec_size);
            if (i32) fetch_small_range<uint64_t, int32_t, ROWS>(p, c, b, e, ROWS);
            else fetch_small_range<uint64_t, int64_t, ROWS>(p, c, b, e, ROWS);
            return;
        }
    }
    if constexpr ((KINDS & kTasksSmall4) != 0) {
        const SmallCols<uint32_t> c = load_small_cols<uint32_t>(p.mask, p.bitval, (int)p.vec_size);
        if (i32) fetch_small_range<uint32_t, int32_t, ROWS>(p, c, b, e, ROWS);
        else fetch_small_range<uint32_t, int64_t, ROWS>(p, c, b, e, ROWS);
    }
}

// Tasks whose first work item is staged in LDS for the item -> task lookup.
constexpr int kLdsTasks = 1024;

// Persistent kernel over the work items of all tasks.  Static mode: wave w
// takes items w, w + W, ... (no atomics; items are equal-sized except each
// task's last).  Dynamic mode: waves claim `claim` consecutive items per
// atomic on the queue head until it is drained (uneven item costs).  The
// item -> task lookup is a binary search over first-item offsets staged in
// LDS, skipped while consecutive items stay in the current task.
template <int ROWS, int KINDS>
__global__ __launch_bounds__(256) void batch_kernel(const TaskDev *__restrict__ tasks, int ntasks, int64_t total_items,
                                                    unsigned long long *head, int claim, int dynamic)
{
    __shared__ int64_t first_item[kLdsTasks];
    const int lane = lane_id();
    const bool staged = ntasks <= kLdsTasks;
    if (staged)
        for (int k = threadIdx.x; k < ntasks; k += blockDim.x) first_item[k] = tasks[k].item_begin;
    __syncthreads();
    int cur = -1;
    int64_t cur_lo = 0, cur_hi = 0; // items [cur_lo, cur_hi) belong to task `cur`
    TaskDev t;
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    int64_t next = wave; // static mode: next item of this wave
    for (;;) {
        int64_t first, last;
        if (dynamic) {
            unsigned long long claimed = 0;
            if (lane == 0) claimed = atomicAdd(head, (unsigned long long)claim);
            first = readfirstlane((int64_t)claimed);
            last = first + claim;
        } else {
            first = next;
            last = next + 1;
            next += nwaves;
        }
        if (first >= total_items) break; // every wave reaches this exit
        if (last > total_items) last = total_items;
        for (int64_t item = first; item < last; ++item) {
            if (item < cur_lo || item >= cur_hi) {
                int lo = 0, hi = ntasks - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) / 2;
                    const int64_t fm = staged ? first_item[mid] : tasks[mid].item_begin;
                    if (fm <= item) lo = mid;
                    else hi = mid - 1;
                }
                // The search ran on LDS values: make the task index visibly
                // uniform so the descriptor is a scalar load held in SGPRs.
                cur = readfirstlane(lo);
                t = tasks[cur];
                cur_lo = t.item_begin;
                cur_hi = cur + 1 < ntasks ? (staged ? first_item[cur + 1] : tasks[cur + 1].item_begin) : total_items;
                cur_hi = readfirstlane(cur_hi);
            }
            const int64_t b = (item - t.item_begin) * t.rows_per_item;
            int64_t e = b + t.rows_per_item;
            if (e > t.p.n) e = t.p.n;
            run_item<ROWS, KINDS>(t, b, e);
        }
    }
}

// Segmented launch: task t owns blocks [block_begin, block_begin + blocks)
// and runs the plain fetch's loop over its rows on that sub-grid.  The block ->
// task search runs on blockIdx, so it and the task's parameters are scalar.
template <typename T, typename IdxT, int ROWS>
__global__ __launch_bounds__(256) void batch_segmented_kernel(const TaskDev *__restrict__ tasks, int ntasks)
{
    const int blk = (int)blockIdx.x;
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (tasks[mid].block_begin <= blk) lo = mid;
        else hi = mid - 1;
    }
    const TaskDev &t = tasks[lo];
    const FetchParams &p = t.p;
    const SmallCols<T> col = load_small_cols<T>(p.mask, p.bitval, (int)p.vec_size);
    const int64_t waves_per_block = blockDim.x / kWave;
    const int64_t wave = (int64_t)(blk - t.block_begin) * waves_per_block + wave_in_block_id();
    const int64_t nwaves = (int64_t)t.blocks * waves_per_block;
    fetch_small_range<T, IdxT, ROWS>(p, col, wave * ROWS, p.n, nwaves * ROWS);
}

// Segmented launch of condensed tasks: as batch_segmented_kernel, with the
// condensed fetch's loop (two offsets per row, reads bounded by the row's
// range and the stream).
template <typename T, typename IdxT>
__global__ __launch_bounds__(256) void batch_segmented_condensed_kernel(const TaskDev *__restrict__ tasks, int ntasks)
{
    constexpr int ROWS = 8;
    const int blk = (int)blockIdx.x;
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (tasks[mid].block_begin <= blk) lo = mid;
        else hi = mid - 1;
    }
    const TaskDev &t = tasks[lo];
    const FetchParams &p = t.p;
    const SmallCols<T> col = load_small_cols<T>(p.mask, p.bitval, (int)p.vec_size);
    const int64_t waves_per_block = blockDim.x / kWave;
    const int64_t wave = (int64_t)(blk - t.block_begin) * waves_per_block + wave_in_block_id();
    const int64_t nwaves = (int64_t)t.blocks * waves_per_block;
    const bool nt = rows_fill_lines(p.vec_size, (int)sizeof(T));
    with_runs<T>(col, nt, [&](auto runs, auto ntc) {
        fetch_small_range_meta<T, IdxT, ROWS, kLayoutCondensed, decltype(ntc)::value, decltype(runs)::value>(
            p, col, wave * ROWS, p.n, nwaves * ROWS);
    });
}

BatchConfig adapt_config(const BatchConfig &base, int64_t total_rows, int num_cus)
{
    BatchConfig c = base;
    const int64_t waves = base.blocks > 0 ? (int64_t)base.blocks * 4 : (int64_t)num_cus * 32;
    int64_t rpi = total_rows / waves;
    rpi = (rpi + c.rows_in_flight - 1) / c.rows_in_flight * c.rows_in_flight;
    if (rpi < c.rows_in_flight) rpi = c.rows_in_flight;
    if (rpi > base.rows_per_item_small) rpi = base.rows_per_item_small;
    c.rows_per_item_small = (int)rpi;
    if (total_rows < waves * 4) c.claim = 1;
    return c;
}

// Rows in flight of a segmented launch (4 or 8).
static int seg_rows(const BatchConfig &cfg) { return cfg.rows_in_flight >= 8 ? 8 : 4; }

int64_t fill_tasks(const ibp_fetch_task *tasks, int n, const BatchConfig &cfg, int max_blocks, TaskDev *out,
                   BatchShape *shape)
{
    int64_t items = 0, rows = 0;
    int kinds = 0, idx_mask = 0;
    for (int i = 0; i < n; ++i) {
        const ibp_fetch_task &t = tasks[i];
        TaskDev d;
        memset(&d, 0, sizeof(d));
        d.p.out = t.out;
        d.p.comp = t.comp;
        d.p.mask = t.mask;
        d.p.bitval = t.bitval;
        d.p.bitmask = (const uint32_t *)t.bitmask;
        d.p.index = t.index;
        d.p.num_vecs = t.num_vecs;
        d.p.vec_size = t.vec_size;
        d.p.n = t.n;
        const bool host = t.comp_loc == IBP_LOC_HOST || (t.comp_loc == IBP_LOC_AUTO && pointer_is_host(t.comp));
        d.p.first_words = first_words_for(t.vec_size, t.elem_bytes, t.comp_len, host);
        d.p.err = (t.index || t.offsets) ? index_error_counter() : nullptr;
        d.p.offsets = t.offsets;
        d.p.stream_bytes = t.stream_bytes;
        d.elem_bytes = t.elem_bytes;
        d.index_bytes = t.index_bytes;
        d.small = t.vec_size <= kWave;
        kinds |= !d.small ? kTasksLarge : t.elem_bytes == 8 ? kTasksSmall8 : kTasksSmall4;
        idx_mask |= (t.index != nullptr && t.index_bytes == 4) ? 1 : 2;
        d.rows_per_item = d.small ? cfg.rows_per_item_small : cfg.rows_per_item_large;
        d.item_begin = items;
        items += (t.n + d.rows_per_item - 1) / d.rows_per_item;
        rows += t.n;
        out[i] = d;
    }
    BatchShape sh;
    sh.kinds = kinds ? kinds : kTasksAll;
    // Condensed tasks (checked at plan creation: all small, one width, one
    // index width) have only the segmented form.
    const bool condensed = n > 0 && tasks[0].offsets != nullptr;
    if (condensed) sh.layout = kLayoutCondensed;
    if ((cfg.segmented || condensed) && (kinds == kTasksSmall8 || kinds == kTasksSmall4) &&
        (idx_mask == 1 || idx_mask == 2) && rows > 0) {
        // Blocks per task: one group of rows per wave, as the plain fetch
        // launches, shrunk in proportion to the task's rows when the total
        // exceeds max_blocks (each task with rows keeps at least one block).
        const int64_t per_block = (int64_t)seg_rows(cfg) * 4;
        int64_t natural = 0;
        for (int i = 0; i < n; ++i) natural += (out[i].p.n + per_block - 1) / per_block;
        const int64_t cap = max_blocks > 0 ? max_blocks : (int64_t)1 << 22;
        int64_t begin = 0;
        for (int i = 0; i < n; ++i) {
            int64_t b = (out[i].p.n + per_block - 1) / per_block;
            if (natural > cap && b > 0) {
                b = (int64_t)((double)cap * (double)out[i].p.n / (double)rows);
                if (b < 1) b = 1;
            }
            out[i].block_begin = (int32_t)begin;
            out[i].blocks = (int32_t)b;
            begin += b;
        }
        sh.seg_blocks = (int)begin;
        sh.elem_bytes = kinds == kTasksSmall8 ? 8 : 4;
        sh.index_bytes = idx_mask == 1 ? 4 : 8;
    }
    if (shape) *shape = sh;
    return items;
}

template <typename T, typename IdxT>
void launch_segmented(int rows, int blocks, hipStream_t s, const TaskDev *dev_tasks, int ntasks, int layout)
{
    if (layout == kLayoutCondensed) {
        hipLaunchKernelGGL((batch_segmented_condensed_kernel<T, IdxT>), dim3(blocks), dim3(256), 0, s, dev_tasks,
                           ntasks);
        return;
    }
    if (rows == 8)
        hipLaunchKernelGGL((batch_segmented_kernel<T, IdxT, 8>), dim3(blocks), dim3(256), 0, s, dev_tasks, ntasks);
    else
        hipLaunchKernelGGL((batch_segmented_kernel<T, IdxT, 4>), dim3(blocks), dim3(256), 0, s, dev_tasks, ntasks);
}

template <int ROWS>
void launch_rows(int classes, dim3 grid, dim3 block, hipStream_t s, const TaskDev *dev_tasks, int ntasks,
                 int64_t total_items, unsigned long long *head, const BatchConfig &cfg)
{
    // One kind of task (the usual plan): the kernel compiled for that kind.
    switch (classes) {
    case kTasksSmall8:
        hipLaunchKernelGGL((batch_kernel<ROWS, kTasksSmall8>), grid, block, 0, s, dev_tasks, ntasks, total_items, head,
                           cfg.claim, cfg.dynamic);
        break;
    case kTasksSmall4:
        hipLaunchKernelGGL((batch_kernel<ROWS, kTasksSmall4>), grid, block, 0, s, dev_tasks, ntasks, total_items, head,
                           cfg.claim, cfg.dynamic);
        break;
    case kTasksLarge:
        hipLaunchKernelGGL((batch_kernel<ROWS, kTasksLarge>), grid, block, 0, s, dev_tasks, ntasks, total_items, head,
                           cfg.claim, cfg.dynamic);
        break;
    default:
        hipLaunchKernelGGL((batch_kernel<ROWS, kTasksAll>), grid, block, 0, s, dev_tasks, ntasks, total_items, head,
                           cfg.claim, cfg.dynamic);
        break;
    }
}

int reset_queue_head(unsigned long long *head, hipStream_t s)
{
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, head);
    IBP_LAUNCH_CHECK();
    return IBP_OK;
}

int launch_batch(const TaskDev *dev_tasks, int ntasks, int64_t total_items, unsigned long long *head,
                 const BatchConfig &cfg, hipStream_t s, const BatchShape &shape)
{
    if (shape.seg_blocks > 0) {
        const int rows = seg_rows(cfg), b = shape.seg_blocks;
        if (shape.elem_bytes == 8) {
            if (shape.index_bytes == 4) launch_segmented<uint64_t, int32_t>(rows, b, s, dev_tasks, ntasks, shape.layout);
            else launch_segmented<uint64_t, int64_t>(rows, b, s, dev_tasks, ntasks, shape.layout);
        } else {
            if (shape.index_bytes == 4) launch_segmented<uint32_t, int32_t>(rows, b, s, dev_tasks, ntasks, shape.layout);
            else launch_segmented<uint32_t, int64_t>(rows, b, s, dev_tasks, ntasks, shape.layout);
        }
        IBP_LAUNCH_CHECK();
        return IBP_OK;
    }
    const int classes = shape.kinds;
    // No more blocks than there is work for (claims of `claim` items in dynamic
    // mode, one item per wave in static mode; four waves per block).
    const int64_t per_wave = cfg.dynamic ? cfg.claim : 1;
    int64_t blocks = (total_items + per_wave * 4 - 1) / (per_wave * 4);
    if (blocks > cfg.blocks) blocks = cfg.blocks;
    if (blocks < 1) blocks = 1;
    const dim3 grid((unsigned)blocks), block(256);
    if (cfg.dynamic) hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, head);
    switch (cfg.rows_in_flight) {
    case 2: launch_rows<2>(classes, grid, block, s, dev_tasks, ntasks, total_items, head, cfg); break;
    case 8: launch_rows<8>(classes, grid, block, s, dev_tasks, ntasks, total_items, head, cfg); break;
    default: launch_rows<4>(classes, grid, block, s, dev_tasks, ntasks, total_items, head, cfg); break;
    }
    IBP_LAUNCH_CHECK();
    return IBP_OK;
}

} // namespace ibp

using namespace ibp;

struct ibp_plan {
    std::vector<ibp_fetch_task> tasks;
    std::vector<TaskDev> host_tasks;
    TaskDev *dev_tasks = nullptr;
    unsigned long long *head = nullptr;
    int64_t total_items = 0;
    BatchShape shape;
    BatchConfig cfg{128, 2, 8, 0, 4, 0, 1};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    int device = -1;
};

namespace {

int rebuild(ibp_plan *pl)
{
    pl->host_tasks.resize(pl->tasks.size());
    pl->total_items = fill_tasks(pl->tasks.data(), (int)pl->tasks.size(), pl->cfg, pl->cfg.blocks * 2,
                                 pl->host_tasks.data(), &pl->shape);
    IBP_HIP(hipMemcpy(pl->dev_tasks, pl->host_tasks.data(), sizeof(TaskDev) * pl->host_tasks.size(),
                      hipMemcpyHostToDevice));
    if (pl->graph_exec) {
        (void)hipGraphExecDestroy(pl->graph_exec);
        pl->graph_exec = nullptr;
    }
    if (pl->graph) {
        (void)hipGraphDestroy(pl->graph);
        pl->graph = nullptr;
    }
    return IBP_OK;
}

int launch(ibp_plan *pl, hipStream_t s)
{
    return launch_batch(pl->dev_tasks, (int)pl->host_tasks.size(), pl->total_items, pl->head, pl->cfg, s,
                        pl->shape);
}

} // namespace

extern "C" int ibp_plan_create(const ibp_fetch_task *tasks, int num_tasks, ibp_plan **plan)
{
    IBP_CHECK(plan != nullptr, "null plan pointer");
    IBP_CHECK(num_tasks >= 1 && tasks != nullptr, "a plan needs at least one task");
    for (int i = 0; i < num_tasks; ++i) {
        const ibp_fetch_task &t = tasks[i];
        IBP_CHECK(t.elem_bytes == 4 || t.elem_bytes == 8, "task %d: element size must be 4 or 8", i);
        IBP_CHECK(t.vec_size > 0 && t.n >= 0 && t.num_vecs >= 0, "task %d: invalid shape", i);
        IBP_CHECK(t.index != nullptr || t.n <= t.num_vecs, "task %d: n exceeds num_vecs without an index", i);
        IBP_CHECK(t.index == nullptr || t.index_bytes == 4 || t.index_bytes == 8, "task %d: index must be int32/int64", i);
        IBP_CHECK(t.n == 0 || (t.comp && t.mask && t.bitval && (t.bitmask || t.offsets) && t.out),
                  "task %d: null pointer", i);
        if (tasks[0].offsets != nullptr || t.offsets != nullptr) {
            const ibp_fetch_task &t0 = tasks[0];
            IBP_CHECK(t.offsets != nullptr && t0.offsets != nullptr,
                      "task %d: a plan's tasks are all condensed or all in the slot layout", i);
            IBP_CHECK(t.vec_size <= 64 && t.elem_bytes == t0.elem_bytes &&
                          (t.index ? t.index_bytes : 8) == (t0.index ? t0.index_bytes : 8),
                      "task %d: condensed tasks need rows of <= 64 elements of one element width and one index width",
                      i);
            IBP_CHECK(t.stream_bytes >= 0, "task %d: stream_bytes must be >= 0", i);
        }
    }
    ibp_plan *pl = new ibp_plan();
    pl->tasks.assign(tasks, tasks + num_tasks);
    (void)hipGetDevice(&pl->device);
    const DeviceInfo &dev = device_info();
    pl->cfg.blocks = dev.num_cus * 8;
    if (hipMalloc((void **)&pl->dev_tasks, sizeof(TaskDev) * (size_t)num_tasks) != hipSuccess ||
        hipMalloc((void **)&pl->head, sizeof(unsigned long long)) != hipSuccess) {
        ibp_plan_destroy(pl);
        return set_error(IBP_ERR_NOMEM, "plan allocation failed");
    }
    const int st = rebuild(pl);
    if (st != IBP_OK) {
        ibp_plan_destroy(pl);
        return st;
    }
    *plan = pl;
    return IBP_OK;
}

extern "C" int ibp_plan_destroy(ibp_plan *pl)
{
    if (!pl) return IBP_OK;
    if (pl->graph_exec) (void)hipGraphExecDestroy(pl->graph_exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->capture_stream) (void)hipStreamDestroy(pl->capture_stream);
    if (pl->dev_tasks) (void)hipFree(pl->dev_tasks);
    if (pl->head) (void)hipFree(pl->head);
    delete pl;
    return IBP_OK;
}

extern "C" int ibp_plan_launch(ibp_plan *pl, void *stream)
{
    IBP_TRACE("ibp_plan_launch");
    IBP_CHECK(pl != nullptr, "null plan");
    return launch(pl, as_stream(stream));
}

extern "C" int ibp_plan_autotune(ibp_plan *pl, void *stream, float *best_ms)
{
    IBP_TRACE("ibp_plan_autotune");
    IBP_CHECK(pl != nullptr, "null plan");
    hipStream_t s = as_stream(stream);
    const DeviceInfo &dev = device_info();
    std::vector<BatchConfig> cands;
    for (int rif : {4, 8})
        for (int mult : {4, 8, 16, 32}) cands.push_back(BatchConfig{128, 2, rif, dev.num_cus * mult, 1, 0, 1});
    for (int rpi : {32, 64, 128, 256})
        for (int rif : {4, 8})
            for (int mult : {8, 16}) {
                cands.push_back(BatchConfig{rpi, rpi >= 128 ? 4 : 2, rif, dev.num_cus * mult, 1, 0, 0});
                for (int claim : {1, 8})
                    cands.push_back(BatchConfig{rpi, rpi >= 128 ? 4 : 2, rif, dev.num_cus * mult, claim, 1, 0});
            }
    struct Events { // destroyed on every return, the error ones included
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    IBP_HIP(hipEventCreate(&ev.e0));
    IBP_HIP(hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    float best = 1e30f;
    BatchConfig best_cfg = pl->cfg;
    for (const BatchConfig &c : cands) {
        pl->cfg = c;
        int st = rebuild(pl);
        if (st != IBP_OK) return st;
        st = launch(pl, s); // warm-up
        if (st != IBP_OK) return st;
        IBP_HIP(hipEventRecord(e0, s));
        const int reps = 5;
        for (int r = 0; r < reps; ++r)
            if ((st = launch(pl, s)) != IBP_OK) return st;
        IBP_HIP(hipEventRecord(e1, s));
        IBP_HIP(hipEventSynchronize(e1));
        float ms = 0;
        IBP_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= reps;
        if (ms < best) {
            best = ms;
            best_cfg = c;
        }
    }
    pl->cfg = best_cfg;
    const int st = rebuild(pl);
    if (st != IBP_OK) return st;
    if (best_ms) *best_ms = best;
    dprintf("[ibp] autotune: rows/item %d/%d, rows in flight %d, blocks %d, segmented %d: %.4f ms\n",
            best_cfg.rows_per_item_small, best_cfg.rows_per_item_large, best_cfg.rows_in_flight, best_cfg.blocks,
            pl->shape.seg_blocks, best);
    return IBP_OK;
}

extern "C" int ibp_plan_capture_graph(ibp_plan *pl, void *stream)
{
    IBP_CHECK(pl != nullptr, "null plan");
    (void)stream;
    if (pl->graph_exec) return IBP_OK;
    if (!pl->capture_stream) IBP_HIP(hipStreamCreateWithFlags(&pl->capture_stream, hipStreamNonBlocking));
    IBP_HIP(hipStreamBeginCapture(pl->capture_stream, hipStreamCaptureModeThreadLocal));
    const int st = launch(pl, pl->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(pl->capture_stream, &g);
    if (st != IBP_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (st != IBP_OK) return st;
        return set_error(IBP_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    pl->graph = g;
    IBP_HIP(hipGraphInstantiate(&pl->graph_exec, pl->graph, nullptr, nullptr, 0));
    return IBP_OK;
}

extern "C" int ibp_plan_graph_launch(ibp_plan *pl, void *stream)
{
    IBP_TRACE("ibp_plan_graph_launch");
    IBP_CHECK(pl != nullptr, "null plan");
    if (!pl->graph_exec) {
        const int st = ibp_plan_capture_graph(pl, stream);
        if (st != IBP_OK) return st;
    }
    IBP_HIP(hipGraphLaunch(pl->graph_exec, as_stream(stream)));
    return IBP_OK;
}

extern "C" int ibp_plan_get_config(ibp_plan *pl, int32_t cfg[7])
{
    IBP_CHECK(pl != nullptr && cfg != nullptr, "null argument");
    const BatchConfig &c = pl->cfg;
    const int32_t v[7] = {c.rows_per_item_small, c.rows_per_item_large, c.rows_in_flight, c.blocks,
                          c.claim, c.dynamic, c.segmented};
    for (int k = 0; k < 7; ++k) cfg[k] = v[k];
    return IBP_OK;
}

extern "C" int ibp_plan_set_config(ibp_plan *pl, const int32_t cfg[7])
{
    IBP_CHECK(pl != nullptr && cfg != nullptr, "null argument");
    const DeviceInfo &dev = device_info();
    IBP_CHECK(cfg[0] >= 1 && cfg[0] <= (1 << 20) && cfg[1] >= 1 && cfg[1] <= (1 << 20),
              "rows per work item must be in [1, 2^20]");
    IBP_CHECK(cfg[2] == 2 || cfg[2] == 4 || cfg[2] == 8, "rows in flight must be 2, 4 or 8");
    IBP_CHECK(cfg[3] >= 1 && cfg[3] <= dev.num_cus * 64, "blocks must be in [1, %d]", dev.num_cus * 64);
    IBP_CHECK(cfg[4] >= 1 && cfg[4] <= 1024, "claim must be in [1, 1024]");
    IBP_CHECK((cfg[5] == 0 || cfg[5] == 1) && (cfg[6] == 0 || cfg[6] == 1), "dynamic and segmented are 0 or 1");
    pl->cfg = BatchConfig{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6]};
    return rebuild(pl);
}

extern "C" int ibp_plan_describe(ibp_plan *pl, char *buf, size_t len)
{
    IBP_CHECK(pl != nullptr && buf != nullptr, "null argument");
    snprintf(buf, len,
             "{\"tasks\": %zu, \"items\": %lld, \"rows_per_item_small\": %d, \"rows_per_item_large\": %d, "
             "\"rows_in_flight\": %d, \"blocks\": %d, \"claim\": %d, \"dynamic\": %d, \"segmented_blocks\": %d, "
             "\"graph\": %s}",
             pl->tasks.size(), (long long)pl->total_items, pl->cfg.rows_per_item_small, pl->cfg.rows_per_item_large,
             pl->cfg.rows_in_flight, pl->cfg.blocks, pl->cfg.claim, pl->cfg.dynamic, pl->shape.seg_blocks,
             pl->graph_exec ? "true" : "false");
    return IBP_OK;
}
