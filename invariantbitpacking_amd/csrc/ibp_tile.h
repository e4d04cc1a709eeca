// ibp_tile.h -- 64-row tiles of a row-major table copied to LDS, shared by
// the row-per-lane kernels (preprocess scoring, pack).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ibp_wave.h"

namespace ibp {

// Rows of up to 512 bytes of whole 16-byte chunks.  Lane r reads row r of
// the tile 16 bytes at a time; eight lanes' reads share the LDS banks once
// each when the row stride spreads them.  A stride that is a multiple of 128
// bytes would put all of them in one bank group, so there chunk c of tile row
// r is stored at position c ^ (r & 7) of its row (a permutation inside each
// 128-byte group of the row: every copy instruction still reads 1 KB of
// consecutive bytes).  Strides of 64 bytes modulo 128 stay out (four lanes per
// bank group).
__host__ __device__ constexpr bool tile_swizzled(int64_t row_bytes) { return row_bytes % 128 == 0; }

constexpr int kTileMaxRowBytes = 512;
__host__ __device__ constexpr bool tile_rows_ok(int64_t row_bytes)
{
    return row_bytes % 16 == 0 && row_bytes <= kTileMaxRowBytes && row_bytes % 128 != 64;
}

// The XOR applied to the chunk index of tile row r (0 without the swizzle).
__device__ __forceinline__ int tile_swz(int r, bool swz) { return swz ? (r & 7) : 0; }

// The copy of tile tl (rows tl*64 ...) into `buf` by wave `wib` of WPT: 16-byte
// chunks wib, wib + WPT, ... of the tile, each instruction 1 KB of consecutive
// bytes straight to LDS (global_load_lds_dwordx4: no registers, tracked by
// vmcnt).  C = row bytes / 16; swz: the placement above (then C % 8 == 0, and
// the lane of LDS chunk g loads chunk g ^ (row & 7) of the same row).
template <typename T, int WPT>
__device__ __forceinline__ void copy_tile(const T *data, int64_t num_vecs, int V, int C, int64_t tl, int wib, int lane,
                                          unsigned char *buf, bool swz)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(data + tl * kWave * (int64_t)V);
    const int64_t rows_left = num_vecs - tl * kWave;
    // the last tile may be short: its missing rows re-read the last chunk
    const int64_t last = (rows_left < kWave ? rows_left : kWave) * C - 1;
    const float rc = 1.0f / (float)C;
    for (int k = wib; k < C; k += WPT) {
        int g = k * kWave + lane;
        if (swz) g ^= (int)(((float)g + 0.5f) * rc) & 7; // row g / C (exact: g < 2048, C <= 32)
        __builtin_amdgcn_global_load_lds(static_cast<const void *>(src + (g < last ? g : last)),
                                         (__attribute__((address_space(3))) void *)(buf + (size_t)k * (kWave * 16)),
                                         16, 0, 0);
    }
}

// The same copy staged through registers: wave `wib` of WPT loads chunks wib,
// wib + WPT, ... of tile tl (the placement of copy_tile) into r[] with plain
// vector loads, and store_tile_regs writes them to LDS later.  Loads into
// registers can be in flight while the waves read the tile before it from
// LDS; the compiler orders a read of LDS after a copy_tile into it by waiting
// for all vector memory, which would serialise the two.
typedef uint32_t tile_chunk_t __attribute__((ext_vector_type(4))); // 16 bytes, kept in VGPRs

template <typename T, int WPT, int NR>
__device__ __forceinline__ void load_tile_regs(const T *data, int64_t num_vecs, int V, int C, int64_t tl, int wib,
                                               int lane, bool swz, tile_chunk_t (&r)[NR])
{
    const tile_chunk_t *src = reinterpret_cast<const tile_chunk_t *>(data + tl * kWave * (int64_t)V);
    const int64_t rows_left = num_vecs - tl * kWave;
    const int64_t last = (rows_left < kWave ? rows_left : kWave) * C - 1;
    const float rc = 1.0f / (float)C;
    // every r[i] is loaded (slots past the row's last chunk load that chunk
    // again and are never stored), so r[] stays in registers
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int k = wib + i * WPT;
        int g = (k < C ? k : C - 1) * kWave + lane;
        if (swz) g ^= (int)(((float)g + 0.5f) * rc) & 7;
        r[i] = src[g < last ? g : last];
    }
}

template <int WPT, int NR>
__device__ __forceinline__ void store_tile_regs(unsigned char *buf, int C, int wib, int lane,
                                                const tile_chunk_t (&r)[NR])
{
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int k = wib + i * WPT;
        if (k < C) reinterpret_cast<tile_chunk_t *>(buf)[k * kWave + lane] = r[i];
    }
}

// Element e of a 16-byte chunk of 4- or 8-byte words.
template <typename T>
__device__ __forceinline__ T chunk_elem(const uint4 &v, int e)
{
    if constexpr (sizeof(T) == 8) {
        return e == 0 ? (T)(((uint64_t)v.y << 32) | v.x) : (T)(((uint64_t)v.w << 32) | v.z);
    } else {
        return (T)(e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w);
    }
}

// lost + (d & m != 0 ? popc(m) : 0) as AND, (AND-OR), MIN, MAD_U24 with the
// mask and its popcount in SGPRs (one scalar operand per instruction).
// ORDERED keeps the steps in program order with the other volatile asm of the
// loop, so the scalar loads of the next column stay ahead of them and their
// wait behind them.
#define IBP_LOST_STEP64                                                                                                \
    "v_and_b32 %0, %2, %3\n\t"                                                                                          \
    "v_and_or_b32 %0, %4, %5, %0\n\t"                                                                                   \
    "v_min_u32 %0, 1, %0\n\t"                                                                                           \
    "v_mad_u32_u24 %1, %0, %6, %1"
#define IBP_LOST_STEP32                                                                                                \
    "v_and_b32 %0, %2, %3\n\t"                                                                                          \
    "v_min_u32 %0, 1, %0\n\t"                                                                                           \
    "v_mad_u32_u24 %1, %0, %4, %1"
template <bool ORDERED = false>
__device__ __forceinline__ uint32_t lost_step(uint64_t d, uint64_t m, uint32_t pc, uint32_t lost)
{
    uint32_t z;
    if constexpr (ORDERED)
        asm volatile(IBP_LOST_STEP64
                     : "=&v"(z), "+v"(lost)
                     : "s"((uint32_t)m), "v"((uint32_t)d), "v"((uint32_t)(d >> 32)), "s"((uint32_t)(m >> 32)), "s"(pc));
    else
        asm(IBP_LOST_STEP64
            : "=&v"(z), "+v"(lost)
            : "s"((uint32_t)m), "v"((uint32_t)d), "v"((uint32_t)(d >> 32)), "s"((uint32_t)(m >> 32)), "s"(pc));
    return lost;
}
template <bool ORDERED = false>
__device__ __forceinline__ uint32_t lost_step(uint32_t d, uint32_t m, uint32_t pc, uint32_t lost)
{
    uint32_t z;
    if constexpr (ORDERED)
        asm volatile(IBP_LOST_STEP32 : "=&v"(z), "+v"(lost) : "s"(m), "v"(d), "s"(pc));
    else
        asm(IBP_LOST_STEP32 : "=&v"(z), "+v"(lost) : "s"(m), "v"(d), "s"(pc));
    return lost;
}

// x ^ b (b uniform) as volatile asm, so it stays in order with the steps.
__device__ __forceinline__ uint64_t xor_ordered(uint64_t x, uint64_t b)
{
    uint32_t lo, hi;
    asm volatile("v_xor_b32 %0, %2, %3\n\tv_xor_b32 %1, %4, %5"
                 : "=&v"(lo), "=&v"(hi)
                 : "s"((uint32_t)b), "v"((uint32_t)x), "s"((uint32_t)(b >> 32)), "v"((uint32_t)(x >> 32)));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t xor_ordered(uint32_t x, uint32_t b)
{
    uint32_t d;
    asm volatile("v_xor_b32 %0, %1, %2" : "=v"(d) : "s"(b), "v"(x));
    return d;
}

} // namespace ibp
