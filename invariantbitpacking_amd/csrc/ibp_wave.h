// ibp_wave.h -- wave64 cross-lane primitives for gfx950 (CDNA4).
//
// All helpers must be called with the whole wave active (wave-uniform control
// flow): DPP and readlane read other lanes' registers regardless of EXEC.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ibp {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Index of the calling wave within its block and across the grid.  Computed
// through readfirstlane so the compiler keeps them (and every row index
// derived from them) in scalar registers.
__device__ __forceinline__ int wave_in_block_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave)); }
__device__ __forceinline__ int64_t global_wave_id()
{
    return (int64_t)blockIdx.x * (int64_t)(blockDim.x / kWave) + wave_in_block_id();
}

// Inclusive prefix sum over the 64 lanes (int32), DPP row shifts within each
// row of 16 lanes followed by the two row broadcasts of GFX9-family DPP.
__device__ __forceinline__ int wave_incl_scan(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); // row_bcast:31
    return x;
}

__device__ __forceinline__ int readlane(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ __forceinline__ uint32_t readlane(uint32_t x, int lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}
__device__ __forceinline__ uint64_t readlane(uint64_t x, int lane)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t readlane(int64_t x, int lane) { return (int64_t)readlane((uint64_t)x, lane); }
__device__ __forceinline__ uint8_t readlane(uint8_t x, int lane) { return (uint8_t)readlane((uint32_t)x, lane); }
__device__ __forceinline__ uint16_t readlane(uint16_t x, int lane) { return (uint16_t)readlane((uint32_t)x, lane); }

__device__ __forceinline__ int readfirstlane(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t readfirstlane(int64_t x)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Sum over the wave, result in every lane (uniform).
__device__ __forceinline__ int wave_sum(int x) { return readlane(wave_incl_scan(x), kWave - 1); }

// 64-bit sum via two 32-bit halves' inclusive scans is unnecessary here: the
// callers reduce counts that fit in int32 per row.

// Value of `x` held by lane `src` (0..63), per lane (ds_bpermute).
__device__ __forceinline__ uint32_t shfl(uint32_t x, int src)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)x);
}
__device__ __forceinline__ uint8_t shfl(uint8_t x, int src) { return (uint8_t)shfl((uint32_t)x, src); }
__device__ __forceinline__ uint16_t shfl(uint16_t x, int src) { return (uint16_t)shfl((uint32_t)x, src); }
__device__ __forceinline__ uint64_t shfl(uint64_t x, int src)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}


// Sums 64 values per lane over the wave: lane l returns the sum over all
// lanes of v[l].  Six exchange-and-add halvings -- lane bit 5 keeps the upper
// or lower 32 values and trades the other half with lane ^ 32, then bit 4
// with lane ^ 16, ... -- so 63 exchanges and adds in all instead of 64 wave
// reductions.  v is clobbered.
__device__ __forceinline__ int wave_sum64(int (&v)[64])
{
    const int lane = lane_id();
#pragma unroll
    for (int st = 0; st < 6; ++st) {
        const int m = 32 >> st, half = 32 >> st;
        const bool hi = (lane & m) != 0;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const int a = v[i], b = v[i + half];
            const int send = hi ? a : b, keep = hi ? b : a;
            v[i] = keep + (int)shfl((uint32_t)send, lane ^ m);
        }
    }
    return v[0];
}
} // namespace ibp
