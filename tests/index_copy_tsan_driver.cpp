// Drives the serving loop's split index copy (invariantbitpacking_amd/csrc/
// ibp_index_copy.h, host code only) from three submitting threads at once,
// with idle gaps that put its helper threads to sleep, built with
// ThreadSanitizer by tests/test_capi.py: a data race aborts the run, and every
// copy is checked byte for byte.
//
// Usage: index_copy_tsan_driver <copies per thread>; exit status 0 = passed.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "ibp_index_copy.h"

int main(int argc, char **argv)
{
    const int copies = argc > 1 ? atoi(argv[1]) : 40;
    ibp::IndexCopier cp;
    std::atomic<int> bad{0};
    auto submitter = [&](uint64_t seed) {
        std::mt19937_64 rng(seed);
        std::vector<unsigned char> src, dst;
        for (int t = 0; t < copies && !bad.load(); ++t) {
            const size_t n = (rng() % 2) ? (size_t)(1 + rng() % ((size_t)3 << 20)) : ((size_t)256 << 10);
            src.resize(n);
            for (size_t i = 0; i < n; ++i) src[i] = (unsigned char)(rng() >> 56);
            dst.assign(n + 128, 0xA5);
            cp.copy((char *)dst.data() + 64, (const char *)src.data(), n);
            if (memcmp(dst.data() + 64, src.data(), n) != 0) bad.store(1);
            for (size_t i = 0; i < 64; ++i)
                if (dst[i] != 0xA5 || dst[64 + n + i] != 0xA5) bad.store(2);
            if (rng() % 8 == 0) std::this_thread::sleep_for(std::chrono::microseconds(500 + rng() % 3000));
        }
    };
    std::vector<std::thread> th;
    for (int k = 0; k < 3; ++k) th.emplace_back(submitter, 5 + (uint64_t)k);
    for (std::thread &t : th) t.join();
    if (bad.load()) {
        fprintf(stderr, "index copy: %s\n", bad.load() == 1 ? "copy differs" : "bytes outside the copy changed");
        return 1;
    }
    printf("index copy: %d copies per thread passed\n", copies);
    return 0;
}
