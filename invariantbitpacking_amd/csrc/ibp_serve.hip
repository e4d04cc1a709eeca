// ibp_serve.hip -- continuous-batching fetch server.
//
// No reference counterpart (the reference has no serving loop; its callers
// issue decompress_fetch per request).  A request = (registered tensor set,
// host index array).  Requests are admitted when their output fits the
// device arena (admission by free HBM), merged with the other admitted
// requests into ONE batched launch (ibp_batch.hip), and launched on one of
// `num_streams` streams as soon as a stream is free, so requests that arrive
// while others run are batched and overlap on the device.  Each request's
// output region and index range are its own, so a result never depends on
// which other requests shared its launch.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ibp_batch.h"
#include "ibp_index_copy.h"
#include "ibp_internal.h"

using namespace ibp;

// 1: launches whose requests all have rows of <= 64 elements of one width run
// segmented (each request its own blocks, the plain fetch's loop).
#ifndef IBP_SERVE_SEGMENTED
#define IBP_SERVE_SEGMENTED 1
#endif

namespace {

using Clock = std::chrono::steady_clock;

double now_us()
{
    return std::chrono::duration<double, std::micro>(Clock::now().time_since_epoch()).count();
}

struct Set {
    int elem_bytes;
    const void *comp;
    int64_t num_vecs, vec_size;
    const void *mask, *bitval;
    const int32_t *bitmask;
    int comp_len, comp_loc;
};

enum class St { Pending, Admitted, Running, Done };

struct Req {
    int set = 0;
    std::vector<int64_t> index; // only while the request holds no pinned index space
    int64_t n = 0;
    size_t out_off = 0, out_bytes = 0;
    size_t idx_off = 0;
    bool idx_pinned = false;
    double t_arrival = 0, t_admit = 0, t_launch = 0, t_issued = 0, t_done = 0;
    St st = St::Pending;
    bool waited = false;
};

// First-fit allocator over [0, size) with coalescing; 256-byte granules.
struct Arena {
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_list; // offset -> length
    static size_t granules(size_t bytes)
    {
        bytes = (bytes + 255) / 256 * 256;
        return bytes == 0 ? 256 : bytes;
    }
    void init(size_t s)
    {
        size = s;
        used = peak = 0;
        free_list.clear();
        free_list[0] = s;
    }
    bool alloc(size_t bytes, size_t cap, size_t *off)
    {
        bytes = granules(bytes);
        if (used + bytes > cap) return false;
        for (auto it = free_list.begin(); it != free_list.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t roff = it->first + bytes;
                free_list.erase(it);
                if (rest) free_list[roff] = rest;
                used += bytes;
                peak = std::max(peak, used);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = granules(bytes);
        used -= bytes;
        auto it = free_list.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_list.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_list.erase(next);
        }
        if (it != free_list.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_list.erase(it);
            }
        }
    }
};

// Request bookkeeping and admission, free of HIP calls (the host self-test
// drives it with a mock launcher).  Requests are admitted in arrival order
// when their output fits the output arena and their indices are in the pinned
// index arena, where the kernel reads them in place.  A request takes pinned
// index space at submit when some is free; one that finds none keeps its
// indices in host memory and is copied in when it reaches the head of the
// queue.  The head can then always be admitted once the running requests
// finish: if the pinned arena is too full for it, the pending requests behind
// it give their pinned space back (their indices are copied out again), since
// they cannot be admitted before it anyway.
struct Admission {
    Arena out, pin;
    char *pin_mem = nullptr;
    size_t budget_override = 0;
    std::unordered_map<int64_t, Req> reqs;
    std::deque<int64_t> pending, admitted;
    int64_t waited = 0, reclaimed = 0;

    size_t out_cap() const { return budget_override ? std::min(budget_override, out.size) : out.size; }

    // Pinned index space for the head request `r`, reclaiming the space of the
    // later pending requests (newest first) while it does not fit.
    bool pin_head(Req &r, size_t *off)
    {
        const size_t bytes = sizeof(int64_t) * (size_t)r.n;
        if (pin.alloc(bytes, pin.size, off)) return true;
        for (size_t q = pending.size(); q-- > 1;) {
            Req &o = reqs[pending[q]];
            if (!o.idx_pinned) continue;
            const int64_t *src = (const int64_t *)(pin_mem + o.idx_off);
            o.index.assign(src, src + o.n);
            pin.release(o.idx_off, sizeof(int64_t) * (size_t)o.n);
            o.idx_pinned = false;
            ++reclaimed;
            if (pin.alloc(bytes, pin.size, off)) return true;
        }
        return false;
    }

    // Admits pending requests in order while they fit.
    void admit(double now)
    {
        while (!pending.empty()) {
            Req &r = reqs[pending.front()];
            size_t off = 0, ioff = r.idx_off;
            bool ok = out.alloc(r.out_bytes, out_cap(), &off);
            if (ok && !r.idx_pinned && !pin_head(r, &ioff)) {
                out.release(off, r.out_bytes);
                ok = false;
            }
            if (!ok) {
                if (!r.waited) {
                    r.waited = true;
                    ++waited;
                }
                return; // FIFO: later requests wait behind the head
            }
            if (!r.idx_pinned) {
                memcpy(pin_mem + ioff, r.index.data(), sizeof(int64_t) * (size_t)r.n);
                std::vector<int64_t>().swap(r.index);
                r.idx_off = ioff;
                r.idx_pinned = true;
            }
            r.out_off = off;
            r.st = St::Admitted;
            r.t_admit = now;
            admitted.push_back(pending.front());
            pending.pop_front();
        }
    }

    // A launched request's kernel has finished: its pinned indices are free.
    void complete(Req &r, double now)
    {
        pin.release(r.idx_off, sizeof(int64_t) * (size_t)r.n);
        r.idx_pinned = false;
        r.st = St::Done;
        r.t_done = now;
    }
};

// An output region released with a stream: freed once the work the caller
// queued on that stream before the release has finished.
struct DeferredFree {
    size_t off, bytes;
    hipEvent_t ev;
};

// One launch slot: a stream and the batch's task table, staged in pinned host
// memory and copied to the device once per launch.  Request indices stay in
// the server's pinned index arena and are read by the kernel in place.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    TaskDev *host_tasks = nullptr, *dev_tasks = nullptr;
    unsigned long long *head = nullptr;
    std::vector<int64_t> reqs;
    bool busy = false;
};

} // namespace

struct ibp_server {
    ibp_server_config cfg{};
    int device = 0;
    char *arena_mem = nullptr;
    char *pin_mem = nullptr; // pinned host memory of request indices
    Admission adm;
    BatchConfig bcfg{256, 4, 8, 0, 1, 0, IBP_SERVE_SEGMENTED};
    std::vector<Set> sets;
    std::vector<Slot> slots;
    std::vector<DeferredFree> deferred;
    std::vector<hipEvent_t> event_pool;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    IndexCopier copier;
    bool stop = false;
    int64_t next_id = 1;
    int64_t completed = 0, launches = 0, rows = 0, launched_reqs = 0;
    std::vector<double> latencies;
    std::string worker_error;

    void fail_locked(const char *what, hipError_t e)
    {
        if (worker_error.empty()) worker_error = std::string(what) + ": " + hipGetErrorString(e);
        cv_done.notify_all();
    }

    bool launch_locked(Slot &sl)
    {
        const double t_launch = now_us();
        std::vector<ibp_fetch_task> tasks;
        int64_t nrows = 0;
        sl.reqs.clear();
        while (!adm.admitted.empty() && (int)sl.reqs.size() < cfg.max_batch_requests) {
            Req &r = adm.reqs[adm.admitted.front()];
            if (!sl.reqs.empty() && nrows + r.n > cfg.max_batch_rows) break;
            const Set &s = sets[r.set];
            ibp_fetch_task t;
            t.comp = s.comp;
            t.mask = s.mask;
            t.bitval = s.bitval;
            t.bitmask = s.bitmask;
            t.index = pin_mem + r.idx_off;
            t.out = arena_mem + r.out_off;
            t.num_vecs = s.num_vecs;
            t.vec_size = s.vec_size;
            t.n = r.n;
            t.elem_bytes = s.elem_bytes;
            t.index_bytes = 8;
            t.comp_len = s.comp_len;
            t.comp_loc = s.comp_loc;
            t.offsets = nullptr;
            t.stream_bytes = 0;
            tasks.push_back(t);
            nrows += r.n;
            r.st = St::Running;
            sl.reqs.push_back(adm.admitted.front());
            adm.admitted.pop_front();
        }
        if (sl.reqs.empty()) return false;
        const BatchConfig cfg_now = adapt_config(bcfg, nrows, device_info().num_cus);
        BatchShape shape;
        const int64_t items =
            fill_tasks(tasks.data(), (int)tasks.size(), cfg_now, device_info().num_cus * 16, sl.host_tasks, &shape);
        int st;
        if (tasks.size() == 1) {
            // One request: the plain fetch, its parameters passed by value
            // (no task-table copy before the kernel).
            const ibp_fetch_task &t = tasks[0];
            st = fetch_launch(sl.host_tasks[0].p, t.elem_bytes, t.index_bytes, -1,
                              t.comp_len > 0 ? t.comp_len : (int)t.vec_size, 0, 0, sl.stream, kLayoutSlots,
                              t.comp_loc == IBP_LOC_HOST);
        } else {
            const hipError_t e = hipMemcpyAsync(sl.dev_tasks, sl.host_tasks, sizeof(TaskDev) * tasks.size(),
                                                hipMemcpyHostToDevice, sl.stream);
            st = e == hipSuccess ? launch_batch(sl.dev_tasks, (int)tasks.size(), items, sl.head, cfg_now, sl.stream, shape)
                                 : IBP_ERR_HIP;
        }
        hipError_t rec = hipSuccess;
        if (st == IBP_OK) rec = hipEventRecord(sl.done, sl.stream);
        if (st != IBP_OK) {
            if (worker_error.empty()) worker_error = std::string("launch failed: ") + ibp_last_error();
            cv_done.notify_all();
        } else if (rec != hipSuccess) {
            fail_locked("hipEventRecord", rec);
        }
        const double t_issued = now_us();
        for (int64_t id : sl.reqs) {
            adm.reqs[id].t_launch = t_launch;
            adm.reqs[id].t_issued = t_issued;
        }
        sl.busy = true;
        ++launches;
        launched_reqs += (int64_t)sl.reqs.size();
        rows += nrows;
        return true;
    }

    // Completed launches: their requests are done.  A failed launch or a
    // faulted kernel stops the server: every wait then reports the error.
    void poll_locked()
    {
        for (Slot &sl : slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) fail_locked("fetch kernel", q);
            const double t = now_us();
            for (int64_t id : sl.reqs) {
                Req &r = adm.reqs[id];
                adm.complete(r, t);
                latencies.push_back(r.t_done - r.t_arrival);
                ++completed;
            }
            sl.reqs.clear();
            sl.busy = false;
            cv_done.notify_all();
        }
    }

    // Output regions released with a stream are returned to the arena once
    // the caller's work queued before the release has finished.
    void poll_deferred_locked()
    {
        for (size_t i = 0; i < deferred.size();) {
            const hipError_t q = hipEventQuery(deferred[i].ev);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            if (q != hipSuccess) fail_locked("caller stream", q);
            adm.out.release(deferred[i].off, deferred[i].bytes);
            event_pool.push_back(deferred[i].ev);
            deferred[i] = deferred.back();
            deferred.pop_back();
        }
    }

    void run()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            poll_locked();
            poll_deferred_locked();
            adm.admit(now_us());
            for (Slot &sl : slots)
                if (!sl.busy && !adm.admitted.empty() && worker_error.empty()) launch_locked(sl);
            bool busy = false;
            for (Slot &sl : slots) busy = busy || sl.busy;
            // Idle unless a launch is in flight, or a pending request waits
            // for a region whose release is still in flight on a stream.
            const bool deferred_wait = !deferred.empty() && !adm.pending.empty();
            if (!busy && !deferred_wait && (adm.admitted.empty() || !worker_error.empty())) {
                cv_work.wait(lk);
            } else {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }
};

extern "C" int ibp_server_create(const ibp_server_config *cfg, ibp_server **out)
{
    IBP_CHECK(cfg && out, "null argument");
    IBP_CHECK(cfg->hbm_budget_bytes >= 4096, "hbm_budget_bytes too small");
    IBP_CHECK(cfg->num_streams >= 1 && cfg->num_streams <= 16, "num_streams must be in [1, 16]");
    IBP_CHECK(cfg->max_batch_requests >= 1 && cfg->max_batch_rows >= 1, "batch limits must be positive");
    ibp_server *srv = new ibp_server();
    srv->cfg = *cfg;
    (void)hipGetDevice(&srv->device);
    srv->bcfg.blocks = device_info().num_cus * 8;
    auto fail = [&](const char *what, hipError_t e) {
        ibp_server_destroy(srv);
        return set_error(IBP_ERR_NOMEM, "server %s failed: %s", what, hipGetErrorString(e));
    };
    hipError_t e = hipMalloc((void **)&srv->arena_mem, cfg->hbm_budget_bytes);
    if (e != hipSuccess) return fail("arena allocation", e);
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail("setup", e);
    srv->adm.out.init(cfg->hbm_budget_bytes);
    // Room for four requests of max_batch_rows indices (at least 64 MiB):
    // any single request fits once the running ones have finished.
    size_t pin_bytes = sizeof(int64_t) * (size_t)cfg->max_batch_rows * 4;
    if (pin_bytes < ((size_t)64 << 20)) pin_bytes = (size_t)64 << 20;
    if ((e = hipHostMalloc((void **)&srv->pin_mem, pin_bytes, 0)) != hipSuccess) return fail("pinned index arena", e);
    srv->adm.pin.init(pin_bytes);
    srv->adm.pin_mem = srv->pin_mem;
    srv->slots.resize((size_t)cfg->num_streams);
    for (Slot &sl : srv->slots) {
        // Streams keep the default priority.
        if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
        if ((e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess) return fail("event", e);
        const size_t stage = sizeof(TaskDev) * (size_t)cfg->max_batch_requests;
        if ((e = hipMalloc((void **)&sl.dev_tasks, stage)) != hipSuccess) return fail("task table", e);
        if ((e = hipHostMalloc((void **)&sl.host_tasks, stage, 0)) != hipSuccess) return fail("task staging", e);
        if ((e = hipMalloc((void **)&sl.head, sizeof(unsigned long long))) != hipSuccess) return fail("queue", e);
        // The runtime sets up a stream's hardware queue at its first launch
        // (tens of ms): do that here, not in the first request that lands on it.
        const int st = reset_queue_head(sl.head, sl.stream);
        if (st != IBP_OK || (e = hipStreamSynchronize(sl.stream)) != hipSuccess) {
            const std::string why = st != IBP_OK ? ibp_last_error() : hipGetErrorString(e);
            ibp_server_destroy(srv);
            return set_error(IBP_ERR_HIP, "server stream warm-up failed: %s", why.c_str());
        }
    }
    srv->worker = std::thread([srv] { srv->run(); });
    *out = srv;
    return IBP_OK;
}

extern "C" int ibp_server_destroy(ibp_server *srv)
{
    if (!srv) return IBP_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    for (Slot &sl : srv->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.dev_tasks) (void)hipFree(sl.dev_tasks);
        if (sl.host_tasks) (void)hipHostFree(sl.host_tasks);
        if (sl.head) (void)hipFree(sl.head);
    }
    // Regions released with a stream: the caller's reads of the arena finish first.
    for (DeferredFree &d : srv->deferred) {
        (void)hipEventSynchronize(d.ev);
        (void)hipEventDestroy(d.ev);
    }
    for (hipEvent_t ev : srv->event_pool) (void)hipEventDestroy(ev);
    if (srv->arena_mem) (void)hipFree(srv->arena_mem);
    if (srv->pin_mem) (void)hipHostFree(srv->pin_mem);
    delete srv;
    return IBP_OK;
}

extern "C" int ibp_server_register(ibp_server *srv, int elem_bytes, const void *comp, int64_t num_vecs,
                                   int64_t vec_size, const void *mask, const void *bitval, const int32_t *bitmask,
                                   int comp_len, int comp_loc, int *set_id)
{
    IBP_CHECK(srv && set_id, "null argument");
    IBP_CHECK(elem_bytes == 4 || elem_bytes == 8, "element size must be 4 or 8");
    IBP_CHECK(comp && mask && bitval && bitmask && vec_size > 0 && num_vecs >= 0, "invalid tensor set");
    std::lock_guard<std::mutex> lk(srv->mu);
    // resolved once here: the worker fills task descriptors per launch
    const int loc = comp_loc == IBP_LOC_AUTO ? (pointer_is_host(comp) ? IBP_LOC_HOST : IBP_LOC_DEVICE) : comp_loc;
    srv->sets.push_back(Set{elem_bytes, comp, num_vecs, vec_size, mask, bitval, bitmask, comp_len, loc});
    *set_id = (int)srv->sets.size() - 1;
    return IBP_OK;
}

extern "C" double ibp_server_clock_us(void) { return now_us(); }

extern "C" int ibp_server_submit_at(ibp_server *srv, int set_id, const int64_t *index, int64_t n, double arrival_us,
                                    int64_t *req_id)
{
    IBP_TRACE("ibp_server_submit_at");
    const double t_call = now_us(); // latency includes copying the index
    IBP_CHECK(srv && req_id, "null argument");
    IBP_CHECK(n >= 1 && index != nullptr, "a request needs at least one index");
    IBP_CHECK(arrival_us <= t_call, "arrival time %.1f us is in the future (now %.1f us)", arrival_us, t_call);
    std::unique_lock<std::mutex> lk(srv->mu);
    IBP_CHECK(set_id >= 0 && set_id < (int)srv->sets.size(), "unknown set %d", set_id);
    IBP_CHECK(n <= srv->cfg.max_batch_rows, "request of %lld rows exceeds max_batch_rows (%lld)", (long long)n,
              (long long)srv->cfg.max_batch_rows);
    const Set &s = srv->sets[set_id];
    const size_t bytes = (size_t)n * (size_t)s.vec_size * (size_t)s.elem_bytes;
    if (bytes > srv->adm.out.size)
        return set_error(IBP_ERR_NOMEM, "request output (%zu bytes) can never fit the arena (%zu bytes)", bytes,
                         srv->adm.out.size);
    const int64_t id = srv->next_id++;
    size_t ioff = 0;
    const bool pinned = srv->adm.pin.alloc(sizeof(int64_t) * (size_t)n, srv->adm.pin.size, &ioff);
    lk.unlock();
    // The copy runs outside the lock; the request is not queued until it is done.
    std::vector<int64_t> vec;
    if (pinned) srv->copier.copy(srv->pin_mem + ioff, (const char *)index, sizeof(int64_t) * (size_t)n);
    else vec.assign(index, index + n);
    lk.lock();
    Req &r = srv->adm.reqs[id];
    r.set = set_id;
    r.index.swap(vec);
    r.idx_pinned = pinned;
    r.idx_off = ioff;
    r.n = n;
    r.out_bytes = bytes;
    r.t_arrival = arrival_us > 0 ? arrival_us : t_call;
    srv->adm.pending.push_back(id);
    lk.unlock();
    srv->cv_work.notify_all();
    *req_id = id;
    return IBP_OK;
}

extern "C" int ibp_server_submit(ibp_server *srv, int set_id, const int64_t *index, int64_t n, int64_t *req_id)
{
    return ibp_server_submit_at(srv, set_id, index, n, 0.0, req_id);
}

extern "C" int ibp_server_wait(ibp_server *srv, int64_t req_id, int timeout_ms, void **out, double *latency_us)
{
    IBP_TRACE("ibp_server_wait");
    IBP_CHECK(srv, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    IBP_CHECK(srv->adm.reqs.count(req_id) != 0, "unknown request %lld", (long long)req_id);
    // Looked up again after every wake-up: submits may rehash the table.
    auto ready = [&] {
        auto it = srv->adm.reqs.find(req_id);
        return it == srv->adm.reqs.end() || it->second.st == St::Done || !srv->worker_error.empty();
    };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, ready);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return set_error(IBP_ERR_BUSY, "request %lld not finished", (long long)req_id);
    }
    if (!srv->worker_error.empty()) return set_error(IBP_ERR_HIP, "server worker: %s", srv->worker_error.c_str());
    auto it = srv->adm.reqs.find(req_id);
    IBP_CHECK(it != srv->adm.reqs.end(), "request %lld was released while waited for", (long long)req_id);
    const Req &r = it->second;
    if (out) *out = srv->arena_mem + r.out_off;
    if (latency_us) *latency_us = r.t_done - r.t_arrival;
    return IBP_OK;
}

extern "C" int ibp_server_release_after(ibp_server *srv, int64_t req_id, void *stream)
{
    IBP_CHECK(srv, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->adm.reqs.find(req_id);
        IBP_CHECK(it != srv->adm.reqs.end(), "unknown request %lld", (long long)req_id);
        IBP_CHECK(it->second.st == St::Done, "request %lld is still in flight", (long long)req_id);
        hipEvent_t ev = nullptr;
        if (!srv->event_pool.empty()) {
            ev = srv->event_pool.back();
            srv->event_pool.pop_back();
        } else {
            IBP_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        }
        const hipError_t e = hipEventRecord(ev, as_stream(stream));
        if (e != hipSuccess) {
            srv->event_pool.push_back(ev);
            return set_error(IBP_ERR_HIP, "hipEventRecord on the caller's stream: %s", hipGetErrorString(e));
        }
        srv->deferred.push_back(DeferredFree{it->second.out_off, it->second.out_bytes, ev});
        srv->adm.reqs.erase(it);
    }
    srv->cv_work.notify_all();
    return IBP_OK;
}

extern "C" int ibp_server_release(ibp_server *srv, int64_t req_id)
{
    IBP_CHECK(srv, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->adm.reqs.find(req_id);
        IBP_CHECK(it != srv->adm.reqs.end(), "unknown request %lld", (long long)req_id);
        IBP_CHECK(it->second.st == St::Done, "request %lld is still in flight", (long long)req_id);
        srv->adm.out.release(it->second.out_off, it->second.out_bytes);
        srv->adm.reqs.erase(it);
    }
    srv->cv_work.notify_all();
    return IBP_OK;
}

extern "C" int ibp_server_stats_get(ibp_server *srv, ibp_server_stats *st)
{
    IBP_CHECK(srv && st, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    st->completed = srv->completed;
    st->launches = srv->launches;
    st->rows = srv->rows;
    std::vector<double> l = srv->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) { return l.empty() ? 0.0 : l[(size_t)std::min<double>(l.size() - 1, p * (l.size() - 1) + 0.5)]; };
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
    st->mean_batch_requests = srv->launches ? (double)srv->launched_reqs / (double)srv->launches : 0.0;
    st->peak_arena_bytes = srv->adm.out.peak;
    st->waited_admissions = srv->adm.waited;
    return IBP_OK;
}

extern "C" int ibp_server_set_budget_override(ibp_server *srv, size_t bytes)
{
    IBP_CHECK(srv, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->adm.budget_override = bytes;
    }
    srv->cv_work.notify_all();
    return IBP_OK;
}

extern "C" int ibp_server_request_times(ibp_server *srv, int64_t req_id, double *times)
{
    IBP_CHECK(srv && times, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->adm.reqs.find(req_id);
    IBP_CHECK(it != srv->adm.reqs.end(), "unknown request %lld", (long long)req_id);
    const Req &r = it->second;
    times[0] = r.t_arrival;
    times[1] = r.t_admit;
    times[2] = r.t_launch;
    times[3] = r.t_issued;
    times[4] = r.t_done;
    return IBP_OK;
}

// ---- host self-test of the admission logic ----------------------------------------

namespace {

// Checks that a request's indices reached the pinned arena intact.
bool indices_intact(const Admission &a, int64_t id, const Req &r)
{
    const int64_t *p = (const int64_t *)(a.pin_mem + r.idx_off);
    for (int64_t k = 0; k < r.n; ++k)
        if (p[k] != id * 1000003 + k) return false;
    return true;
}

// Submits like ibp_server_submit: pinned space at once when some is free.
void mock_submit(Admission &a, int64_t id, int64_t n, size_t out_bytes)
{
    std::vector<int64_t> idx((size_t)n);
    for (int64_t k = 0; k < n; ++k) idx[(size_t)k] = id * 1000003 + k;
    Req &r = a.reqs[id];
    size_t off = 0;
    r.n = n;
    r.out_bytes = out_bytes;
    if (a.pin.alloc(sizeof(int64_t) * (size_t)n, a.pin.size, &off)) {
        memcpy(a.pin_mem + off, idx.data(), sizeof(int64_t) * (size_t)n);
        r.idx_pinned = true;
        r.idx_off = off;
    } else {
        r.index.swap(idx);
    }
    a.pending.push_back(id);
}

// Mock launcher: admitted requests start running (their pinned indices are
// checked), up to `max_running` at a time.
int mock_launch(Admission &a, std::vector<int64_t> &running, size_t max_running)
{
    while (!a.admitted.empty() && running.size() < max_running) {
        const int64_t id = a.admitted.front();
        Req &r = a.reqs[id];
        if (!r.idx_pinned || !indices_intact(a, id, r))
            return set_error(IBP_ERR_INVALID, "request %lld launched without its indices", (long long)id);
        r.st = St::Running;
        running.push_back(id);
        a.admitted.pop_front();
    }
    return IBP_OK;
}

int check_arenas_empty(const Admission &a)
{
    if (a.pin.used != 0 || a.out.used != 0)
        return set_error(IBP_ERR_INVALID, "arenas not empty after the drain (pinned %zu, output %zu bytes)",
                         a.pin.used, a.out.used);
    return IBP_OK;
}

// The case a strictly FIFO queue without reclaim deadlocks on: the head has no
// pinned indices, later requests took all the pinned space at submit, and the
// running request's space alone is too little for the head.
int admission_reclaim_case()
{
    std::vector<char> mem(8 * 256);
    Admission a;
    a.pin_mem = mem.data();
    a.pin.init(mem.size());
    a.out.init(1 << 20);
    mock_submit(a, 1, 4 * 32, 4096); // runs, holds 4 of 8 pinned granules
    a.admit(0);
    std::vector<int64_t> running;
    int st = mock_launch(a, running, 1);
    if (st != IBP_OK) return st;
    mock_submit(a, 2, 6 * 32, 4096); // needs 6 granules: none at submit
    mock_submit(a, 3, 2 * 32, 4096); // takes 2
    mock_submit(a, 4, 2 * 32, 4096); // takes 2: pinned space full
    if (a.reqs[2].idx_pinned || !a.reqs[3].idx_pinned || !a.reqs[4].idx_pinned)
        return set_error(IBP_ERR_INVALID, "reclaim case: unexpected submit-time pinning");
    a.complete(a.reqs[1], 0); // frees 4 granules; the head needs 6
    running.clear();
    a.out.release(a.reqs[1].out_off, a.reqs[1].out_bytes);
    a.reqs.erase(1);
    a.admit(0);
    if (a.admitted.empty() || a.admitted.front() != 2)
        return set_error(IBP_ERR_INVALID, "reclaim case: the head request was not admitted");
    while (!a.admitted.empty() || !a.pending.empty()) {
        st = mock_launch(a, running, 1);
        if (st != IBP_OK) return st;
        for (int64_t id : running) {
            a.complete(a.reqs[id], 0);
            a.out.release(a.reqs[id].out_off, a.reqs[id].out_bytes);
            a.reqs.erase(id);
        }
        running.clear();
        a.admit(0);
    }
    return check_arenas_empty(a);
}

// Random submits, launches, completions and releases on small arenas; after
// the last submit everything must drain (no admission deadlock) with every
// request's indices intact.
int admission_random_case(std::mt19937_64 &rng, int steps)
{
    const size_t pin_granules = 4 + rng() % 12;
    std::vector<char> mem(pin_granules * 256);
    Admission a;
    a.pin_mem = mem.data();
    a.pin.init(mem.size());
    a.out.init((8 + rng() % 24) * 256);
    const int64_t max_n = (int64_t)pin_granules * 32; // a request fits the empty pinned arena
    const size_t max_out = a.out.size;
    std::vector<int64_t> running, done;
    int64_t next = 1;
    for (int step = 0; step < steps || !a.pending.empty() || !a.admitted.empty() || !running.empty() || !done.empty();
         ++step) {
        if (step > steps + 100000) return set_error(IBP_ERR_INVALID, "admission did not drain");
        const int op = (int)(rng() % 4);
        if (op == 0 && step < steps) {
            const int64_t n = 1 + (int64_t)(rng() % (uint64_t)max_n);
            mock_submit(a, next++, n, 1 + rng() % max_out);
        } else if (op == 1 && !running.empty()) { // a launch finishes
            const size_t k = rng() % running.size();
            a.complete(a.reqs[running[k]], 0);
            done.push_back(running[k]);
            running[k] = running.back();
            running.pop_back();
        } else if (op == 2 && !done.empty()) { // the caller releases an output
            const size_t k = rng() % done.size();
            a.out.release(a.reqs[done[k]].out_off, a.reqs[done[k]].out_bytes);
            a.reqs.erase(done[k]);
            done[k] = done.back();
            done.pop_back();
        }
        a.admit(0);
        const int st = mock_launch(a, running, 1 + rng() % 4);
        if (st != IBP_OK) return st;
    }
    return check_arenas_empty(a);
}

} // namespace

extern "C" int ibp_selftest_serve_admission(int trials, uint64_t seed)
{
    clear_error();
    int st = admission_reclaim_case();
    std::mt19937_64 rng(seed);
    for (int t = 0; st == IBP_OK && t < trials; ++t) st = admission_random_case(rng, 200);
    return st;
}

// Host self-test of IndexCopier: copies of random sizes (below, at and far
// above the split threshold, with odd tails) from several submitting threads
// at once, in bursts and after idle gaps long enough for the helpers to fall
// asleep; every destination byte is checked, and the bytes around it must be
// untouched.
extern "C" int ibp_selftest_index_copy(int trials, uint64_t seed)
{
    clear_error();
    IndexCopier cp;
    std::atomic<int> bad{0};
    auto worker = [&](uint64_t s) {
        std::mt19937_64 rng(s);
        std::vector<unsigned char> src, dst;
        for (int t = 0; t < trials && bad.load() == 0; ++t) {
            const size_t sizes[] = {1, 4096, IndexCopier::kMinSplit - 8, IndexCopier::kMinSplit,
                                    IndexCopier::kMinSplit + 8, (size_t)2 << 20};
            size_t n = sizes[rng() % 6];
            if (rng() % 2) n = 1 + rng() % ((size_t)3 << 20);
            src.resize(n);
            for (size_t i = 0; i < n; ++i) src[i] = (unsigned char)(rng() >> 56);
            dst.assign(n + 128, 0xA5);
            cp.copy((char *)dst.data() + 64, (const char *)src.data(), n);
            for (size_t i = 0; i < 64; ++i)
                if (dst[i] != 0xA5 || dst[64 + n + i] != 0xA5) {
                    bad.store(1);
                    return;
                }
            if (memcmp(dst.data() + 64, src.data(), n) != 0) {
                bad.store(2);
                return;
            }
            if (rng() % 16 == 0) std::this_thread::sleep_for(std::chrono::microseconds(300 + rng() % 2000));
        }
    };
    std::vector<std::thread> th;
    for (int k = 0; k < 3; ++k) th.emplace_back(worker, seed + (uint64_t)k * 7919);
    for (std::thread &t : th) t.join();
    if (bad.load()) return set_error(IBP_ERR_INVALID, "index copy self-test: %s",
                                     bad.load() == 1 ? "bytes outside the destination changed" : "copy differs");
    return IBP_OK;
}
