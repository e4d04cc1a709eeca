// ibp_fetch_launch.h -- the plain fetch's kernels and their launcher for one
// element width.  Each width is instantiated in its own translation unit
// (ibp_fetch_u64.hip, ibp_fetch_u32.hip, ibp_fetch_narrow.hip) so the many
// specialisations compile in parallel.
#pragma once
#include "ibp_fetch_dev.h"
#include "ibp_internal.h"

namespace ibp {

template <typename T, typename IdxT, int ROWS>
__global__ __launch_bounds__(256) void fetch_small_kernel(FetchParams p)
{
    const SmallCols<T> col = load_small_cols<T>(p.mask, p.bitval, (int)p.vec_size);
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    fetch_small_range<T, IdxT, ROWS>(p, col, wave * ROWS, p.n, nwaves * ROWS);
}

// Rows of <= 64 elements in the condensed and/or clustered layouts.
template <typename T, typename IdxT, int ROWS, int LAYOUT>
__global__ __launch_bounds__(256) void fetch_small_meta_kernel(FetchParams p)
{
    const SmallCols<T> col = load_small_cols<T>(p.mask, p.bitval, (int)p.vec_size);
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const bool nt = rows_fill_lines(p.vec_size, (int)sizeof(T));
    if constexpr ((LAYOUT & kLayoutClustered) != 0) {
        // per-row mask rows: the generic deposit (the run tables are per table)
        if (nt)
            fetch_small_range_meta<T, IdxT, ROWS, LAYOUT, true, -1>(p, col, wave * ROWS, p.n, nwaves * ROWS);
        else
            fetch_small_range_meta<T, IdxT, ROWS, LAYOUT, false, -1>(p, col, wave * ROWS, p.n, nwaves * ROWS);
    } else {
        with_runs<T>(col, nt, [&](auto runs, auto ntc) {
            fetch_small_range_meta<T, IdxT, ROWS, LAYOUT, decltype(ntc)::value, decltype(runs)::value>(
                p, col, wave * ROWS, p.n, nwaves * ROWS);
        });
    }
}

template <typename T, typename IdxT, int LAYOUT>
__global__ __launch_bounds__(256) void fetch_large_kernel(FetchParams p)
{
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    fetch_large_range<T, IdxT, LAYOUT>(p, wave, p.n, nwaves);
}

// Transfer baselines of the reference (include/decompress/ibp_decompress_host.cuh:114-166):
//  mode 2: raw gather of whole slots, no unpack (IndexSelectMultiKernelAligned);
//  mode 3: packed rows read comp_len words and write slot[k % comp_len]
//          (decompress_fake_kernel), raw rows copied;
//  mode 4: packed rows read comp_len words and write nothing, raw rows
//          untouched (decompress_fake_kernel2).
template <typename T, typename IdxT>
__global__ __launch_bounds__(256) void fetch_baseline_kernel(FetchParams p, int mode, int comp_len)
{
    const int lane = lane_id();
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int64_t V = p.vec_size;
    for (int64_t i = wave; i < p.n; i += nwaves) {
        const int64_t s = readfirstlane(load_index<IdxT>(p.index, i));
        T *dst = (T *)p.out + i * V;
        if (s < 0 || s >= p.num_vecs) {
            for (int64_t k = lane; k < V; k += kWave) dst[k] = T(0);
            continue;
        }
        const T *slot = (const T *)p.comp + s * V;
        const bool packed = (p.bitmask[s >> 5] >> (s & 31)) & 1u;
        if (mode == 2 || !packed) {
            if (mode != 4)
                for (int64_t k = lane; k < V; k += kWave) dst[k] = slot[k];
            continue;
        }
        const int64_t cl = comp_len < V ? (comp_len > 0 ? comp_len : 1) : V;
        if (mode == 3 && cl <= kWave) {
            // One read of the first cl words, replicated through the row.
            const T x = lane < cl ? slot[lane] : T(0);
            for (int64_t k0 = 0; k0 < V; k0 += kWave) {
                const int64_t k = k0 + lane;
                const T y = shfl(x, (int)(k % cl));
                if (k < V) dst[k] = y;
            }
        } else if (mode == 3) {
            for (int64_t k = lane; k < V; k += kWave) dst[k] = slot[k % cl];
        } else {
            T acc = T(0);
            for (int64_t k = lane; k < cl; k += kWave) acc ^= slot[k];
            const uint64_t keep = (uint64_t)acc;
            asm volatile("" : : "v"(keep)); // the reads are the point of this mode
        }
    }
}

// Launches the fetch of one table: layout (kLayout* flags), impl (2/3/4 the
// transfer baselines), rows in flight per wave for rows of <= 64 elements
// (2/4/8/16 for 4- and 8-byte elements in the slot layout; 8 otherwise).
template <typename T, typename IdxT, int LAYOUT>
void launch_meta(const FetchParams &p, int grid, int block, hipStream_t stream)
{
    if (p.vec_size <= kWave)
        hipLaunchKernelGGL((fetch_small_meta_kernel<T, IdxT, 8, LAYOUT>), dim3(grid), dim3(block), 0, stream, p);
    else
        hipLaunchKernelGGL((fetch_large_kernel<T, IdxT, LAYOUT>), dim3(grid), dim3(block), 0, stream, p);
}

template <typename T, typename IdxT>
void launch_slots_small(const FetchParams &p, int rows, int grid, int block, hipStream_t stream)
{
    if constexpr (sizeof(T) >= 4) {
        switch (rows) {
        case 16: hipLaunchKernelGGL((fetch_small_kernel<T, IdxT, 16>), dim3(grid), dim3(block), 0, stream, p); return;
        case 4: hipLaunchKernelGGL((fetch_small_kernel<T, IdxT, 4>), dim3(grid), dim3(block), 0, stream, p); return;
        case 2: hipLaunchKernelGGL((fetch_small_kernel<T, IdxT, 2>), dim3(grid), dim3(block), 0, stream, p); return;
        default: break;
        }
    }
    hipLaunchKernelGGL((fetch_small_kernel<T, IdxT, 8>), dim3(grid), dim3(block), 0, stream, p);
}

template <typename T, typename IdxT>
int launch_typed(const FetchParams &p, int layout, int impl, int comp_len, int rows, int grid, int block,
                 hipStream_t stream)
{
    if (layout == kLayoutCondensed) {
        launch_meta<T, IdxT, kLayoutCondensed>(p, grid, block, stream);
    } else if (layout == kLayoutClustered) {
        launch_meta<T, IdxT, kLayoutClustered>(p, grid, block, stream);
    } else if (layout == (kLayoutCondensed | kLayoutClustered)) {
        launch_meta<T, IdxT, kLayoutCondensed | kLayoutClustered>(p, grid, block, stream);
    } else if (impl == 2 || impl == 3 || impl == 4) {
        hipLaunchKernelGGL((fetch_baseline_kernel<T, IdxT>), dim3(grid), dim3(block), 0, stream, p, impl, comp_len);
    } else if (p.vec_size <= kWave) {
        launch_slots_small<T, IdxT>(p, rows, grid, block, stream);
    } else {
        hipLaunchKernelGGL((fetch_large_kernel<T, IdxT, kLayoutSlots>), dim3(grid), dim3(block), 0, stream, p);
    }
    IBP_LAUNCH_CHECK();
    return IBP_OK;
}

template <typename T>
int fetch_launch_width(const FetchParams &p, int layout, int index_bytes, int impl, int comp_len, int rows, int grid,
                       int block, hipStream_t stream)
{
    if (p.index == nullptr || index_bytes == 8)
        return launch_typed<T, int64_t>(p, layout, impl, comp_len, rows, grid, block, stream);
    return launch_typed<T, int32_t>(p, layout, impl, comp_len, rows, grid, block, stream);
}

extern template int fetch_launch_width<uint64_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);
extern template int fetch_launch_width<uint32_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);
extern template int fetch_launch_width<uint16_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);
extern template int fetch_launch_width<uint8_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);

} // namespace ibp
