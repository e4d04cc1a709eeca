// ibp_internal.h -- host-side helpers shared by the libibp translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdarg.h>
#include <stdint.h>

#include <string>

#include "../../include/ibp.h"

namespace ibp {

// Thread-local last error; set_error returns the status for `return set_error(...)`.
int set_error(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
bool debug_enabled();
void dprintf(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Per-device properties cached on first use (no per-call attribute queries).
struct DeviceInfo {
    int device = -1;
    int num_cus = 0;
    int max_lds_per_block = 0;
    int warp_size = 64;
    int num_xcds = 8;
};
const DeviceInfo &device_info();

// Device counter of skipped out-of-range rows (one per device).
unsigned long long *index_error_counter();

// Returns true when `p` is host memory visible to the device (pinned).
bool pointer_is_host(const void *p);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// A named range around one C ABI call, seen by `rocprofv3 --marker-trace`
// beside the kernels the call launches (without a profiler attached the
// markers cost a few tens of nanoseconds).
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

// Words of the speculative first read of a packed row of <= 64 elements.
int first_words_for(int64_t vec_size, int elem_bytes, int comp_len, bool host_resident);

} // namespace ibp

#define IBP_HIP(call)                                                                             \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return ::ibp::set_error(IBP_ERR_HIP, "%s:%d: %s failed: %s", __FILE__, __LINE__, #call, \
                                    hipGetErrorString(_e));                                       \
    } while (0)

#define IBP_CHECK(cond, ...)                                          \
    do {                                                              \
        if (!(cond)) return ::ibp::set_error(IBP_ERR_INVALID, __VA_ARGS__); \
    } while (0)

// After a kernel launch.
#define IBP_LAUNCH_CHECK() IBP_HIP(hipGetLastError())

// The marker range of the entry point it opens.
#define IBP_TRACE(name) ::ibp::TraceRange ibp_trace_range_(name)
