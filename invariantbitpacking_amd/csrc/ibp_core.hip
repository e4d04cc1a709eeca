// ibp_core.hip -- library state: errors, debug flag, device properties, the
// index-error counter and the host self-test of the element bit arithmetic.
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <random>
#include <vector>

#include "ibp_bits.h"
#include "ibp_internal.h"

namespace ibp {

namespace {
thread_local std::string g_last_error;
std::atomic<int> g_debug{0};
std::mutex g_dev_mu;
DeviceInfo g_info[64];
unsigned long long *g_err_counter[64] = {nullptr};
} // namespace

// The index-error counter is a device variable of this code object, so no call
// allocates, clears or synchronises for it (a fetch may be the first call, and
// may be under stream capture).  It is zero when the library is loaded.
static __device__ unsigned long long err_counter;

int set_error(int status, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    if (g_debug.load()) fprintf(stderr, "[ibp] error: %s\n", buf);
    return status;
}

void clear_error() { g_last_error.clear(); }

bool debug_enabled() { return g_debug.load() != 0; }

void dprintf(const char *fmt, ...)
{
    if (!g_debug.load()) return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stdout, fmt, ap);
    va_end(ap);
    fflush(stdout);
}

const DeviceInfo &device_info()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    DeviceInfo &d = g_info[dev];
    if (d.device == dev) return d;
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (d.device != dev) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            d.num_cus = prop.multiProcessorCount;
            d.max_lds_per_block = (int)prop.sharedMemPerBlock;
            d.warp_size = prop.warpSize;
        } else {
            d.num_cus = 256;
            d.max_lds_per_block = 65536;
        }
        d.device = dev;
    }
    return d;
}

unsigned long long *index_error_counter()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (g_err_counter[dev]) return g_err_counter[dev];
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (!g_err_counter[dev]) {
        // Each device holds its own copy of err_counter at its own address.
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(err_counter)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_err_counter[dev] = (unsigned long long *)p;
    }
    return g_err_counter[dev];
}

bool pointer_is_host(const void *p)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return true; // unknown to HIP: treat as host memory
    }
    return attr.type == hipMemoryTypeHost;
}

} // namespace ibp

using namespace ibp;

extern "C" const char *ibp_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *ibp_version(void) { return "ibp-mi355x 0.1.0 (gfx950)"; }

extern "C" void ibp_set_debug(int flag) { g_debug.store(flag ? 1 : 0); }

extern "C" int ibp_index_errors(uint64_t *count, int reset)
{
    unsigned long long *c = index_error_counter();
    IBP_CHECK(c != nullptr, "cannot allocate the index-error counter");
    unsigned long long v = 0;
    IBP_HIP(hipDeviceSynchronize());
    IBP_HIP(hipMemcpy(&v, c, sizeof(v), hipMemcpyDeviceToHost));
    if (reset) IBP_HIP(hipMemset(c, 0, sizeof(v)));
    if (count) *count = v;
    return IBP_OK;
}

// ---- host self-test of the element bit arithmetic --------------------------------

namespace {

template <typename T>
T brute_extract(T x, T free)
{
    const int B = Bits<T>::B;
    T out = 0;
    int pos = B - 1;
    for (int k = B - 1; k >= 0; --k)
        if ((free >> k) & 1) {
            out |= T((x >> k) & 1) << pos;
            --pos;
        }
    return out;
}

template <typename T>
T brute_deposit(T v, T free)
{
    const int B = Bits<T>::B;
    T out = 0;
    int pos = B - 1;
    for (int k = B - 1; k >= 0; --k)
        if ((free >> k) & 1) {
            out |= T((v >> pos) & 1) << k;
            --pos;
        }
    return out;
}

template <typename T>
int selftest_width(int trials, std::mt19937_64 &rng)
{
    const int B = Bits<T>::B;
    for (int t = 0; t < trials; ++t) {
        T free;
        switch (t % 6) {
        case 0: free = T(rng()); break;
        case 1: free = T(~T(0)); break;
        case 2: free = T(0); break;
        case 3: free = shr(himask<T>(int(rng() % (B + 1))), int(rng() % B)); break;
        case 4: free = T(rng()) & T(rng()); break;
        default: free = T(rng()) | T(rng()); break;
        }
        const T x = T(rng());
        const T e = extract_msb<T>(x, free), eb = brute_extract<T>(x, free);
        if (e != eb)
            return set_error(IBP_ERR_INVALID, "extract_msb<%d>(%llx, %llx) = %llx, brute force %llx", B,
                             (unsigned long long)x, (unsigned long long)free, (unsigned long long)e,
                             (unsigned long long)eb);
        const T d = deposit_msb<T>(x, free), db = brute_deposit<T>(x, free);
        if (d != db)
            return set_error(IBP_ERR_INVALID, "deposit_msb<%d>(%llx, %llx) = %llx, brute force %llx", B,
                             (unsigned long long)x, (unsigned long long)free, (unsigned long long)d,
                             (unsigned long long)db);
        // round trip
        if (deposit_msb<T>(e, free) != T(x & free))
            return set_error(IBP_ERR_INVALID, "deposit(extract) round trip failed at width %d", B);
        const int o = int(rng() % B);
        const T w0 = T(rng()), w1 = T(rng());
        T want = 0;
        for (int k = 0; k < B; ++k) {
            const int p = o + k; // stream bit p, MSB-first across w0, w1
            const T word = p < B ? w0 : w1;
            const int bit = (B - 1) - (p % B);
            want |= T((word >> bit) & 1) << (B - 1 - k);
        }
        if (stream_bits<T>(w0, w1, o) != want)
            return set_error(IBP_ERR_INVALID, "stream_bits<%d> mismatch at offset %d", B, o);
    }
    return IBP_OK;
}

} // namespace

extern "C" int ibp_selftest_bitops(int trials, uint64_t seed)
{
    clear_error();
    std::mt19937_64 rng(seed);
    int st = selftest_width<uint8_t>(trials, rng);
    if (st == IBP_OK) st = selftest_width<uint16_t>(trials, rng);
    if (st == IBP_OK) st = selftest_width<uint32_t>(trials, rng);
    if (st == IBP_OK) st = selftest_width<uint64_t>(trials, rng);
    return st;
}
