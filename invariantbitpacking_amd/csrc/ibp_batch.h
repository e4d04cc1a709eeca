// ibp_batch.h -- internal interface of the persistent batched fetch kernel,
// shared by the batch plan (ibp_batch.hip) and the server (ibp_serve.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ibp.h"
#include "ibp_fetch_dev.h"

namespace ibp {

struct TaskDev {
    FetchParams p;
    int64_t item_begin;
    int32_t elem_bytes;
    int32_t index_bytes;
    int32_t rows_per_item;
    int32_t small;
};

struct BatchConfig {
    int rows_per_item_small; // rows of a work item, tables with V <= 64
    int rows_per_item_large; // rows of a work item, wider tables
    int rows_in_flight;      // 2, 4 or 8 (small rows)
    int blocks;              // persistent grid (256 threads per block)
    int claim;               // work items claimed per queue atomic (dynamic mode)
    int dynamic;             // 1: atomic work queue, 0: static item striding
};

// A launch shape for `total_rows` rows: work items small enough that every
// resident wave gets work (latency), no smaller than one group of
// rows_in_flight rows and no larger than base.rows_per_item_small.
BatchConfig adapt_config(const BatchConfig &base, int64_t total_rows, int num_cus);

// Fills `out[0..n)` from the public task descriptions; returns the number of
// work items.
int64_t fill_tasks(const ibp_fetch_task *tasks, int n, const BatchConfig &cfg, TaskDev *out);

// Launches the persistent kernel on `s` (dynamic mode: after a one-thread
// kernel that resets the queue head).
int launch_batch(const TaskDev *dev_tasks, int ntasks, int64_t total_items, unsigned long long *head,
                 const BatchConfig &cfg, hipStream_t s);

} // namespace ibp
