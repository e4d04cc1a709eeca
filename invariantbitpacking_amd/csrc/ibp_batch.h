// ibp_batch.h -- internal interface of the batched fetch kernels,
// shared by the batch plan (ibp_batch.hip) and the server (ibp_serve.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ibp.h"
#include "ibp_fetch_dev.h"

namespace ibp {

// Kinds of task in a launch (bit flags): rows of <= 64 elements of 8 or of 4
// bytes, and wider rows.
enum : int { kTasksSmall8 = 1, kTasksSmall4 = 2, kTasksLarge = 4, kTasksAll = 7 };

struct TaskDev {
    FetchParams p;
    int64_t item_begin;
    int32_t elem_bytes;
    int32_t index_bytes;
    int32_t rows_per_item;
    int32_t small;
    int32_t block_begin; // segmented launch: the task's first block
    int32_t blocks;      // segmented launch: the task's blocks
};

struct BatchConfig {
    int rows_per_item_small; // rows of a work item, tables with V <= 64
    int rows_per_item_large; // rows of a work item, wider tables
    int rows_in_flight;      // 2, 4 or 8 (small rows)
    int blocks;              // persistent grid (256 threads per block)
    int claim;               // work items claimed per queue atomic (dynamic mode)
    int dynamic;             // 1: atomic work queue, 0: static item striding
    int segmented;           // 1: a segmented launch where the tasks allow it
};

// What fill_tasks found: the kinds of task present and, when the launch is
// segmented (every task has rows of <= 64 elements of one element width and
// one index width), the grid.  A segmented launch gives each task its own
// range of blocks, in which the task runs exactly the plain fetch's loop; it
// needs no work queue and the registers of the plain kernel.
struct BatchShape {
    int kinds = kTasksAll;
    int seg_blocks = 0;  // 0: persistent launch over work items
    int elem_bytes = 8;  // segmented launch
    int index_bytes = 8; // segmented launch (8 also for tasks without index)
    int layout = kLayoutSlots; // kLayoutCondensed: condensed tasks (always segmented)
};

// A launch shape for `total_rows` rows: work items small enough that every
// resident wave gets work (latency), no smaller than one group of
// rows_in_flight rows and no larger than base.rows_per_item_small.
BatchConfig adapt_config(const BatchConfig &base, int64_t total_rows, int num_cus);

// Fills `out[0..n)` from the public task descriptions; returns the number of
// work items and fills *shape.  `max_blocks` caps a segmented grid.
int64_t fill_tasks(const ibp_fetch_task *tasks, int n, const BatchConfig &cfg, int max_blocks, TaskDev *out,
                   BatchShape *shape);

// The plain fetch's launcher (ibp_fetch.hip): one table, parameters by value.
int fetch_launch(const FetchParams &p, int elem_bytes, int index_bytes, int impl, int comp_len, int nblks,
                 int nthreads, hipStream_t stream, int layout, bool host_resident);

// Launches the one-thread kernel that zeroes the work-queue head on `s` (also
// used to make the runtime set up a stream's hardware queue ahead of time).
int reset_queue_head(unsigned long long *head, hipStream_t s);

// Launches the fetch of the tasks on `s`: the segmented kernel when
// shape.seg_blocks > 0, else the persistent kernel compiled for shape.kinds
// (dynamic mode: after a one-thread kernel that resets the queue head).
int launch_batch(const TaskDev *dev_tasks, int ntasks, int64_t total_items, unsigned long long *head,
                 const BatchConfig &cfg, hipStream_t s, const BatchShape &shape);

} // namespace ibp
