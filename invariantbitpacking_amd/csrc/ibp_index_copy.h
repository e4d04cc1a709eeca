// ibp_index_copy.h -- the serving loop's copy of request indices into
// pinned memory, split over threads (host code only, no HIP calls).
#pragma once
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace ibp {

// Copies of request indices into the pinned arena, split over the submitting
// thread and kCopyThreads helpers for large requests (one memcpy of a 250K-
// index request is tens of microseconds of the submitter's time).  A copy is
// cut into 64 KiB parts claimed from one counter that also holds the copy's
// number and part count, so a helper that wakes late can only claim a part
// of the copy in progress, and the submitter never waits for a helper that
// has not woken up: it takes the remaining parts itself and waits only for
// parts already claimed.  Helpers spin for kSpinUs after their last part
// (bursts keep them awake), then sleep on a condition variable.  One copy
// runs at a time.
inline void cpu_relax() { __builtin_ia32_pause(); }

using CopyClock = std::chrono::steady_clock;

class IndexCopier {
  public:
    static constexpr int kCopyThreads = 3;
    static constexpr size_t kPart = (size_t)64 << 10;
    static constexpr size_t kMinSplit = (size_t)128 << 10;
    static constexpr double kSpinUs = 200.0;

    IndexCopier()
    {
        for (int k = 0; k < kCopyThreads; ++k) th_.emplace_back([this] { helper(); });
    }
    ~IndexCopier()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_.store(true);
        }
        cv_.notify_all();
        for (std::thread &t : th_) t.join();
    }

    void copy(char *dst, const char *src, size_t bytes)
    {
        const uint64_t parts = (bytes + kPart - 1) / kPart;
        if (bytes < kMinSplit || parts >= kPartMask - 8) {
            memcpy(dst, src, bytes);
            return;
        }
        std::lock_guard<std::mutex> one(job_mu_);
        const uint64_t job = (++jobs_) & kJobMask;
        done_.store(0, std::memory_order_relaxed);
        dst_.store(dst, std::memory_order_relaxed);
        src_.store(src, std::memory_order_relaxed);
        bytes_.store(bytes, std::memory_order_relaxed);
        // job | parts | next part: the fields above are published with it
        claim_.store((job << 40) | (parts << 20), std::memory_order_release);
        if (sleepers_.load(std::memory_order_acquire) > 0) {
            std::lock_guard<std::mutex> lk(m_);
            cv_.notify_all();
        }
        run_parts();
        while ((uint64_t)done_.load(std::memory_order_acquire) < parts) cpu_relax();
    }

  private:
    static constexpr uint64_t kPartMask = (1u << 20) - 1, kJobMask = (1u << 24) - 1;

    // Claims and copies parts of the copy in progress until none is left.
    // The counter value of a claim carries its copy's part count, so a claim
    // past the last part is recognised without reading any field; a claim
    // below it is a part of the copy in progress, which cannot end (nor the
    // fields change) before that part is done.
    void run_parts()
    {
        for (;;) {
            const uint64_t c = claim_.fetch_add(1, std::memory_order_acq_rel);
            const uint64_t p = c & kPartMask, parts = (c >> 20) & kPartMask;
            if (p >= parts) return;
            const size_t bytes = bytes_.load(std::memory_order_relaxed);
            const size_t lo = (size_t)p * kPart;
            const size_t hi = lo + kPart < bytes ? lo + kPart : bytes;
            memcpy(dst_.load(std::memory_order_relaxed) + lo, src_.load(std::memory_order_relaxed) + lo, hi - lo);
            done_.fetch_add(1, std::memory_order_acq_rel);
        }
    }

    void helper()
    {
        uint64_t seen = claim_.load(std::memory_order_acquire) >> 40;
        CopyClock::time_point idle = CopyClock::now();
        while (!stop_.load(std::memory_order_acquire)) {
            const uint64_t job = claim_.load(std::memory_order_acquire) >> 40;
            if (job != seen) {
                seen = job;
                run_parts();
                idle = CopyClock::now();
                continue;
            }
            if (std::chrono::duration<double, std::micro>(CopyClock::now() - idle).count() < kSpinUs) {
                for (int i = 0; i < 64; ++i) cpu_relax();
                continue;
            }
            std::unique_lock<std::mutex> lk(m_);
            sleepers_.fetch_add(1, std::memory_order_acq_rel);
            // a system_clock deadline: ThreadSanitizer follows this wait (it
            // does not follow the steady_clock one); the predicate is checked
            // on every wake-up, so a clock step only changes a sleep's length
            cv_.wait_until(lk, std::chrono::system_clock::now() + std::chrono::milliseconds(5), [&] {
                return stop_.load(std::memory_order_acquire) ||
                       (claim_.load(std::memory_order_acquire) >> 40) != seen;
            });
            sleepers_.fetch_sub(1, std::memory_order_acq_rel);
            idle = CopyClock::now();
        }
    }

    std::vector<std::thread> th_;
    std::mutex job_mu_, m_;
    std::condition_variable cv_;
    std::atomic<bool> stop_{false};
    std::atomic<uint64_t> claim_{0};
    std::atomic<int> done_{0}, sleepers_{0};
    std::atomic<char *> dst_{nullptr};
    std::atomic<const char *> src_{nullptr};
    std::atomic<size_t> bytes_{0};
    uint64_t jobs_ = 0; // guarded by job_mu_
};

} // namespace ibp
