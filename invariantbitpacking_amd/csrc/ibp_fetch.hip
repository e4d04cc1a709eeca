// ibp_fetch.hip -- the plain per-call decompress_fetch.
//
// Replaces src/decompress.cu:10-156 and include/decompress/ibp_decompress_host.cuh:9-168.
// Differences by design (values are identical):
//  * no index sort and no temporary tensors: each wave gathers its own rows
//    (the reference's CUB sort only orders the work, src/decompress.cu:64-98);
//  * no zero-fill of the output: every element is written;
//  * no per-call device queries: properties are cached per device;
//  * out-of-range indices give zero rows and bump a device counter instead of
//    reading out of bounds.
#include <atomic>

#include "ibp_fetch_launch.h"

namespace ibp {

namespace {

// Launch shape of the plain fetch (ibp_set_fetch_tuning): rows in flight
// per wave for rows of <= 64 elements, and resident blocks per CU.
std::atomic<int> g_rows_in_flight{8};
std::atomic<int> g_blocks_per_cu{16};

} // namespace

int fetch_launch(const FetchParams &p, int elem_bytes, int index_bytes, int impl, int comp_len, int nblks,
                 int nthreads, hipStream_t stream, int layout, bool host_resident)
{
    const DeviceInfo &dev = device_info();
    // Blocks are at most 256 threads (four waves); larger requests are capped.
    const int block = (nthreads > 0 && nthreads <= 256) ? nthreads : 256;
    int rows = g_rows_in_flight.load();
#ifndef IBP_NO_SMALL_CALL_ROWS
    // Small calls (few rows): fewer rows per wave, so that the grid still
    // holds about two waves per SIMD; such a call is bound by the latency of
    // its index -> slot -> store chain, not by bandwidth.
    if (p.vec_size <= kWave && elem_bytes >= 4) {
        const int64_t want = (int64_t)dev.num_cus * 8;
        while (rows > 2 && (p.n + rows - 1) / rows < want) rows /= 2;
    }
#endif
    int64_t waves_needed = p.vec_size <= kWave ? (p.n + rows - 1) / rows : p.n;
    int64_t grid = nblks > 0 ? nblks : (waves_needed * kWave + block - 1) / block;
    int64_t cap = (int64_t)dev.num_cus * g_blocks_per_cu.load();
    // Small rows read over PCIe: half a block per CU keeps enough rows in
    // flight to saturate the link (measured: 32..128 blocks run at the full
    // rate on MI355X) and leaves the other CUs to kernels on other streams.
    if (host_resident && p.vec_size <= kWave) cap = dev.num_cus / 2 > 0 ? dev.num_cus / 2 : 1;
    if (nblks <= 0 && grid > cap) grid = cap;
    if (grid < 1) grid = 1;
    if (debug_enabled()) {
        // What the reference reports about its launch (ibp_decompress_host.cuh:
        // 52-61): which codec runs, where the rows live, the launch shape.
        const char *kind = (impl == 2 || impl == 3 || impl == 4) ? "transfer baseline"
                           : p.vec_size <= kWave ? (layout == kLayoutSlots ? "wave per row group" : "wave per row group (layout)")
                                                 : "wave per row";
        dprintf("[ibp] decompress_fetch: %s, impl %d, %s rows, V=%lld, w=%d, n=%lld, grid %lld x %d, rows in flight %d, "
                "first read %d words, comp_len %d\n",
                kind, impl, host_resident ? "host (PCIe)" : "device", (long long)p.vec_size, elem_bytes,
                (long long)p.n, (long long)grid, block, p.vec_size <= kWave ? rows : 1, p.first_words, comp_len);
    }
    switch (elem_bytes) {
    case 1: return fetch_launch_width<uint8_t>(p, layout, index_bytes, impl, comp_len, rows, (int)grid, block, stream);
    case 2: return fetch_launch_width<uint16_t>(p, layout, index_bytes, impl, comp_len, rows, (int)grid, block, stream);
    case 4: return fetch_launch_width<uint32_t>(p, layout, index_bytes, impl, comp_len, rows, (int)grid, block, stream);
    default: return fetch_launch_width<uint64_t>(p, layout, index_bytes, impl, comp_len, rows, (int)grid, block, stream);
    }
}

// Words of the speculative first read of a packed small row: the expected
// compressed length plus some slack, at least the header.  Rows scatter a
// word or two around the average length; a row longer than the first read
// costs its group a second memory round trip.  From HBM that round trip costs
// more than a few extra bytes (GNN minibatch: 38.2 us with one word of slack,
// 36.0 us with three).  Over PCIe, where reads go out in 128-byte lines and
// bytes are the bound, one word of slack is best (the same rows from pinned
// memory: 1.760 ms with one word, 1.813 ms with three).
#ifndef IBP_FIRST_READ_SLACK
#define IBP_FIRST_READ_SLACK 3
#endif
#ifndef IBP_FIRST_READ_SLACK_HOST
#define IBP_FIRST_READ_SLACK_HOST 1
#endif
int first_words_for(int64_t vec_size, int elem_bytes, int comp_len, bool host_resident)
{
    const int64_t hw = header_bytes(vec_size, elem_bytes) / elem_bytes;
    const int slack = host_resident ? IBP_FIRST_READ_SLACK_HOST : IBP_FIRST_READ_SLACK;
    int64_t e = comp_len > 0 ? (int64_t)comp_len + slack : vec_size;
    if (e < hw + 1) e = hw + 1;
    if (e > vec_size) e = vec_size;
    return (int)e;
}

} // namespace ibp

using namespace ibp;

extern "C" int ibp_decompress_fetch_ex(int elem_bytes, void *out, const void *comp, int64_t num_vecs,
                                       int64_t vec_size, const void *mask, const void *bitval,
                                       const int32_t *bitmask, const void *index, int index_bytes, int64_t n,
                                       int comp_len, int nblks, int nthreads, int impl, int comp_loc,
                                       const ibp_layout *layout, void *stream)
{
    IBP_TRACE("ibp_decompress_fetch_ex");
    IBP_CHECK(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8,
              "Input tensor must be 1, 2, 4 or 8-byte datatype (got %d)", elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0 && n >= 0, "invalid shape num_vecs=%lld vec_size=%lld n=%lld",
              (long long)num_vecs, (long long)vec_size, (long long)n);
    IBP_CHECK(index == nullptr || index_bytes == 4 || index_bytes == 8, "index must be int32 or int64");
    IBP_CHECK(impl >= -1 && impl <= 5, "unknown impl %d", impl);
    IBP_CHECK(nthreads <= 0 || (nthreads % 64 == 0 && nthreads <= 1024), "nthreads must be a multiple of 64 up to 1024");
    const bool condensed = layout && layout->offsets;
    const bool clustered = layout && layout->clusters;
    IBP_CHECK(!clustered || layout->num_clusters >= 1, "num_clusters must be >= 1 with clusters");
    IBP_CHECK(!condensed || layout->stream_bytes >= 0, "stream_bytes must be >= 0");
    IBP_CHECK(!(condensed || clustered) || (impl != 2 && impl != 3 && impl != 4),
              "impl %d (transfer baseline) takes the slot layout without clusters", impl);
    if (n == 0) return IBP_OK;
    IBP_CHECK(out && comp && mask && bitval && (bitmask || condensed), "null pointer argument");
    IBP_CHECK(index != nullptr || n <= num_vecs, "without an index n (%lld) must not exceed num_vecs (%lld)",
              (long long)n, (long long)num_vecs);
    FetchParams p;
    p.out = out;
    p.comp = comp;
    p.mask = mask;
    p.bitval = bitval;
    p.bitmask = (const uint32_t *)bitmask;
    p.index = index;
    p.num_vecs = num_vecs;
    p.vec_size = vec_size;
    p.n = n;
    const bool host = comp_loc == IBP_LOC_HOST || (comp_loc == IBP_LOC_AUTO && pointer_is_host(comp));
    p.first_words = first_words_for(vec_size, elem_bytes, comp_len, host);
    p.pad = 0;
    p.err = (index || condensed || clustered) ? index_error_counter() : nullptr;
    int lay = kLayoutSlots;
    if (condensed) {
        lay |= kLayoutCondensed;
        p.offsets = layout->offsets;
        p.stream_bytes = layout->stream_bytes;
    }
    if (clustered) {
        lay |= kLayoutClustered;
        p.clusters = layout->clusters;
        p.num_clusters = layout->num_clusters;
    }
    const int cl = comp_len > 0 ? comp_len : (int)vec_size;
    return fetch_launch(p, elem_bytes, index_bytes, impl, cl, nblks, nthreads, as_stream(stream), lay, host);
}

extern "C" int ibp_decompress_fetch(int elem_bytes, void *out, const void *comp, int64_t num_vecs,
                                    int64_t vec_size, const void *mask, const void *bitval,
                                    const int32_t *bitmask, const void *index, int index_bytes, int64_t n,
                                    int comp_len, int nblks, int nthreads, int impl, int comp_loc, void *stream)
{
    return ibp_decompress_fetch_ex(elem_bytes, out, comp, num_vecs, vec_size, mask, bitval, bitmask, index,
                                   index_bytes, n, comp_len, nblks, nthreads, impl, comp_loc, nullptr, stream);
}

extern "C" int ibp_set_fetch_tuning(int rows_in_flight, int blocks_per_cu)
{
    IBP_CHECK(rows_in_flight == 2 || rows_in_flight == 4 || rows_in_flight == 8 || rows_in_flight == 16,
              "rows_in_flight must be 2, 4, 8 or 16");
    IBP_CHECK(blocks_per_cu >= 1 && blocks_per_cu <= 32, "blocks_per_cu must be in [1, 32]");
    g_rows_in_flight.store(rows_in_flight);
    g_blocks_per_cu.store(blocks_per_cu);
    return IBP_OK;
}
