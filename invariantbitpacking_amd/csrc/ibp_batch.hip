// ibp_batch.hip -- batched fetch: many tables of uneven row counts and widths
// fetched and unpacked by ONE persistent kernel from a device-side work queue.
//
// No reference counterpart (the reference issues one decompress_fetch per
// table, e.g. tests/dlrm_comp.py:79-81); values are those of
// ibp_decompress_fetch on each table.
//
// Work item = (task, range of rows).  Waves claim items with one atomic on the
// queue head; the head is reset by a small kernel launched before the
// persistent kernel on the same stream, so a launch is two kernel nodes and can
// be captured into a HIP graph (no memset node).
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

// Rows [b, e) of task t.  Column tables of small-row tasks are cached across
// consecutive items of the same task (cached_task).
template <int ROWS>
__device__ __forceinline__ void run_item(const TaskDev &t, int task, int64_t b, int64_t e, int &cached_task,
                                         SmallCols<uint64_t> &c64, SmallCols<uint32_t> &c32)
{
    const FetchParams &p = t.p;
    const bool i32 = p.index != nullptr && t.index_bytes == 4;
    if (t.small) {
        if (t.elem_bytes == 8) {
            if (cached_task != task) c64 = load_small_cols<uint64_t>(p.mask, p.bitval, (int)p.vec_size);
            if (i32) fetch_small_range<uint64_t, int32_t, ROWS>(p, c64, b, e, ROWS);
            else fetch_small_range<uint64_t, int64_t, ROWS>(p, c64, b, e, ROWS);
        } else {
            if (cached_task != task) c32 = load_small_cols<uint32_t>(p.mask, p.bitval, (int)p.vec_size);
            if (i32) fetch_small_range<uint32_t, int32_t, ROWS>(p, c32, b, e, ROWS);
            else fetch_small_range<uint32_t, int64_t, ROWS>(p, c32, b, e, ROWS);
        }
        cached_task = task;
    } else {
        if (t.elem_bytes == 8) {
            if (i32) fetch_large_range<uint64_t, int32_t>(p, b, e, 1);
            else fetch_large_range<uint64_t, int64_t>(p, b, e, 1);
        } else {
            if (i32) fetch_large_range<uint32_t, int32_t>(p, b, e, 1);
            else fetch_large_range<uint32_t, int64_t>(p, b, e, 1);
        }
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
template <int ROWS>
__global__ __launch_bounds__(256) void batch_kernel(const TaskDev *tasks, int ntasks, int64_t total_items,
                                                    unsigned long long *head, int claim, int dynamic)
{
    __shared__ int64_t first_item[kLdsTasks];
    const int lane = lane_id();
    const bool staged = ntasks <= kLdsTasks;
    if (staged)
        for (int k = threadIdx.x; k < ntasks; k += blockDim.x) first_item[k] = tasks[k].item_begin;
    __syncthreads();
    int cached_task = -1;
    int cur = -1;
    int64_t cur_lo = 0, cur_hi = 0; // items [cur_lo, cur_hi) belong to task `cur`
    TaskDev t;
    SmallCols<uint64_t> c64;
    SmallCols<uint32_t> c32;
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
                cur = lo;
                t = tasks[lo];
                cur_lo = t.item_begin;
                cur_hi = lo + 1 < ntasks ? (staged ? first_item[lo + 1] : tasks[lo + 1].item_begin) : total_items;
            }
            const int64_t b = (item - t.item_begin) * t.rows_per_item;
            int64_t e = b + t.rows_per_item;
            if (e > t.p.n) e = t.p.n;
            run_item<ROWS>(t, cur, b, e, cached_task, c64, c32);
        }
    }
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

int64_t fill_tasks(const ibp_fetch_task *tasks, int n, const BatchConfig &cfg, TaskDev *out)
{
    int64_t items = 0;
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
        d.p.first_words = first_words_for(t.vec_size, t.elem_bytes, t.comp_len);
        d.p.err = t.index ? index_error_counter() : nullptr;
        d.elem_bytes = t.elem_bytes;
        d.index_bytes = t.index_bytes;
        d.small = t.vec_size <= kWave;
        d.rows_per_item = d.small ? cfg.rows_per_item_small : cfg.rows_per_item_large;
        d.item_begin = items;
        items += (t.n + d.rows_per_item - 1) / d.rows_per_item;
        out[i] = d;
    }
    return items;
}

int launch_batch(const TaskDev *dev_tasks, int ntasks, int64_t total_items, unsigned long long *head,
                 const BatchConfig &cfg, hipStream_t s)
{
    // No more blocks than there is work for (claims of `claim` items in dynamic
    // mode, one item per wave in static mode; four waves per block).
    const int64_t per_wave = cfg.dynamic ? cfg.claim : 1;
    int64_t blocks = (total_items + per_wave * 4 - 1) / (per_wave * 4);
    if (blocks > cfg.blocks) blocks = cfg.blocks;
    if (blocks < 1) blocks = 1;
    const dim3 grid((unsigned)blocks), block(256);
    if (cfg.dynamic) hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, head);
    switch (cfg.rows_in_flight) {
    case 2: hipLaunchKernelGGL((batch_kernel<2>), grid, block, 0, s, dev_tasks, ntasks, total_items, head, cfg.claim, cfg.dynamic); break;
    case 8: hipLaunchKernelGGL((batch_kernel<8>), grid, block, 0, s, dev_tasks, ntasks, total_items, head, cfg.claim, cfg.dynamic); break;
    default: hipLaunchKernelGGL((batch_kernel<4>), grid, block, 0, s, dev_tasks, ntasks, total_items, head, cfg.claim, cfg.dynamic); break;
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
    BatchConfig cfg{128, 2, 8, 0, 4, 0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    int device = -1;
};

namespace {

int rebuild(ibp_plan *pl)
{
    pl->host_tasks.resize(pl->tasks.size());
    pl->total_items = fill_tasks(pl->tasks.data(), (int)pl->tasks.size(), pl->cfg, pl->host_tasks.data());
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
    return launch_batch(pl->dev_tasks, (int)pl->host_tasks.size(), pl->total_items, pl->head, pl->cfg, s);
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
        IBP_CHECK(t.n == 0 || (t.comp && t.mask && t.bitval && t.bitmask && t.out), "task %d: null pointer", i);
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
    IBP_CHECK(pl != nullptr, "null plan");
    return launch(pl, as_stream(stream));
}

extern "C" int ibp_plan_autotune(ibp_plan *pl, void *stream, float *best_ms)
{
    IBP_CHECK(pl != nullptr, "null plan");
    hipStream_t s = as_stream(stream);
    const DeviceInfo &dev = device_info();
    std::vector<BatchConfig> cands;
    for (int rpi : {32, 64, 128, 256})
        for (int rif : {4, 8})
            for (int mult : {8, 16}) {
                cands.push_back(BatchConfig{rpi, rpi >= 128 ? 4 : 2, rif, dev.num_cus * mult, 1, 0});
                for (int claim : {1, 8})
                    cands.push_back(BatchConfig{rpi, rpi >= 128 ? 4 : 2, rif, dev.num_cus * mult, claim, 1});
            }
    hipEvent_t e0, e1;
    IBP_HIP(hipEventCreate(&e0));
    IBP_HIP(hipEventCreate(&e1));
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
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pl->cfg = best_cfg;
    const int st = rebuild(pl);
    if (st != IBP_OK) return st;
    if (best_ms) *best_ms = best;
    dprintf("[ibp] autotune: rows/item %d/%d, rows in flight %d, blocks %d: %.4f ms\n", best_cfg.rows_per_item_small,
            best_cfg.rows_per_item_large, best_cfg.rows_in_flight, best_cfg.blocks, best);
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
    if (st != IBP_OK) return st;
    if (e != hipSuccess) return set_error(IBP_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    pl->graph = g;
    IBP_HIP(hipGraphInstantiate(&pl->graph_exec, pl->graph, nullptr, nullptr, 0));
    return IBP_OK;
}

extern "C" int ibp_plan_graph_launch(ibp_plan *pl, void *stream)
{
    IBP_CHECK(pl != nullptr, "null plan");
    if (!pl->graph_exec) {
        const int st = ibp_plan_capture_graph(pl, stream);
        if (st != IBP_OK) return st;
    }
    IBP_HIP(hipGraphLaunch(pl->graph_exec, as_stream(stream)));
    return IBP_OK;
}

extern "C" int ibp_plan_describe(ibp_plan *pl, char *buf, size_t len)
{
    IBP_CHECK(pl != nullptr && buf != nullptr, "null argument");
    snprintf(buf, len,
             "{\"tasks\": %zu, \"items\": %lld, \"rows_per_item_small\": %d, \"rows_per_item_large\": %d, "
             "\"rows_in_flight\": %d, \"blocks\": %d, \"claim\": %d, \"dynamic\": %d, \"graph\": %s}",
             pl->tasks.size(), (long long)pl->total_items, pl->cfg.rows_per_item_small, pl->cfg.rows_per_item_large,
             pl->cfg.rows_in_flight, pl->cfg.blocks, pl->cfg.claim, pl->cfg.dynamic, pl->graph_exec ? "true" : "false");
    return IBP_OK;
}
