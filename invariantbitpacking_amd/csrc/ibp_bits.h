// ibp_bits.h -- element-level bit arithmetic of the IBP slot codec.
//
// Compiled for the device (gfx950) and for the host (self-test).  T is
// uint8_t, uint16_t, uint32_t or uint64_t (B = 8, 16, 32 or 64 bits).  Every
// expression on a narrow T is cast back to T, since C++ promotes it to int.
//
// Shift semantics: the reference relies on PTX `shl`/`shr` returning 0 for
// counts >= B (ibp_compress_dev.cuh:47,53-55,80; ibp_decompress_dev.cuh:315,
// 327-328).  gfx950 v_lshlrev_b64 uses the count modulo 64, so every shift
// whose count can reach B goes through shl()/shr() below.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ibp {

template <typename T>
struct Bits {
    static constexpr int B = 8 * (int)sizeof(T);
};

__host__ __device__ __forceinline__ int popc(uint8_t v) { return __builtin_popcount((uint32_t)v); }
__host__ __device__ __forceinline__ int popc(uint16_t v) { return __builtin_popcount((uint32_t)v); }
__host__ __device__ __forceinline__ int popc(uint32_t v) { return __builtin_popcount(v); }
__host__ __device__ __forceinline__ int popc(uint64_t v) { return __builtin_popcountll(v); }
__host__ __device__ __forceinline__ int clz(uint32_t v) { return v ? __builtin_clz(v) : 32; }
__host__ __device__ __forceinline__ int clz(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
__host__ __device__ __forceinline__ int clz(uint8_t v) { return clz((uint32_t)v) - 24; }
__host__ __device__ __forceinline__ int clz(uint16_t v) { return clz((uint32_t)v) - 16; }

// Unsigned type of a w-byte element.
template <int W> struct UWord;
template <> struct UWord<1> { using type = uint8_t; };
template <> struct UWord<2> { using type = uint16_t; };
template <> struct UWord<4> { using type = uint32_t; };
template <> struct UWord<8> { using type = uint64_t; };

// v << s and v >> s with the PTX clamp: counts >= B give 0.
template <typename T>
__host__ __device__ __forceinline__ T shl(T v, int s)
{
    return s >= Bits<T>::B ? T(0) : T(v << s);
}
template <typename T>
__host__ __device__ __forceinline__ T shr(T v, int s)
{
    return s >= Bits<T>::B ? T(0) : T(v >> s);
}
// Top n bits set (n in [0, B]).
template <typename T>
__host__ __device__ __forceinline__ T himask(int n)
{
    return n <= 0 ? T(0) : T(T(~T(0)) << (Bits<T>::B - n));
}

// Gathers the bits of x at the set positions of `free` (the zero positions of
// the column mask), highest first, into the top of the result; bits below the
// gathered ones are zero.  This is the value compress_and_write builds in
// `compressed_val` (ibp_compress_dev.cuh:39-58), one run of free bits per step.
template <typename T>
__host__ __device__ __forceinline__ T extract_msb(T x, T free)
{
    T out = 0;
    int filled = 0;
    while (free) {
        const int s = clz(free);             // masked bits above this run
        const int r = clz(T(~T(free << s))); // run length (free << s != 0, s < B)
        const T run = shr(himask<T>(r), s);
        out |= shr(shl(T(x & run), s), filled);
        filled += r;
        free &= ~run;
    }
    return out;
}

// Inverse: drops the top bits of v, highest first, into the set positions of
// `free`, highest position first (the decode loop of decompress_fetch_cpu,
// ibp_decompress_dev.cuh:309-334).  Bits of v beyond popc(free) are ignored.
template <typename T>
__host__ __device__ __forceinline__ T deposit_msb(T v, T free)
{
    T out = 0;
    while (free) {
        const int s = clz(free);
        const int r = clz(T(~T(free << s)));
        const T run = shr(himask<T>(r), s);
        out |= shr(v, s) & run;
        v = shl(v, r);
        free &= ~run;
    }
    return out;
}

// The top B bits of the bit stream starting at bit o (0 <= o < B) of word w0,
// continuing into w1: (w0 << o) | (w1 >> (B - o)).
// (w1 >> 1) >> (B - 1 - o) equals w1 >> (B - o) for o in [1, B) and gives 0
// for o = 0 without a range check.
template <typename T>
__host__ __device__ __forceinline__ T stream_bits(T w0, T w1, int o)
{
    return T(w0 << o) | T(T(w1 >> 1) >> (Bits<T>::B - 1 - o));
}

// Header and slot geometry (ibp_compress_dev.cuh:13-15; ibp_misc_dev.cuh:33-42).
__host__ __device__ __forceinline__ int64_t header_bytes(int64_t vec_size, int w)
{
    return ((vec_size + 7) / 8 + w - 1) / w * w;
}

// s(r) = min(V*w, align_w(ceil(V/8)) + align_w(V*w - floor(S/8))).
__host__ __device__ __forceinline__ int64_t slot_size(int64_t vec_size, int w, int64_t saved_bits)
{
    const int64_t vec_bytes = vec_size * w;
    const int64_t data = (vec_bytes - saved_bits / 8 + w - 1) / w * w;
    const int64_t comp = header_bytes(vec_size, w) + data;
    return comp < vec_bytes ? comp : vec_bytes;
}


// Column-bit counting over up to 15 words per call: the words are added into
// bit-sliced 4-bit counters (ones, twos, fours, eights) with a carry-save
// adder -- 7 logic operations per word instead of B extract-and-adds -- and
// the slices are folded into the B per-bit counters once.  W is uint32_t or
// uint64_t; bits above B are ignored.
template <typename W, int B, int CHUNK>
__device__ __forceinline__ void csa_count(const W (&x)[CHUNK], int32_t (&c)[B])
{
    static_assert(CHUNK <= 15, "4-bit counters hold at most 15 words");
    W ones = 0, twos = 0, fours = 0, eights = 0;
#pragma unroll
    for (int q = 0; q < CHUNK; ++q) {
        const W c1 = ones & x[q];
        ones ^= x[q];
        const W c2 = twos & c1;
        twos ^= c1;
        const W c4 = fours & c2;
        fours ^= c2;
        eights |= c4; // at most 15 words: nothing carries out of eights
    }
#pragma unroll
    for (int k = 0; k < B; ++k)
        c[k] += (int32_t)((ones >> k) & 1) + ((int32_t)((twos >> k) & 1) << 1) + ((int32_t)((fours >> k) & 1) << 2) +
                ((int32_t)((eights >> k) & 1) << 3);
}


// Vertical bit counters in carry-save form for long runs of words (the
// Harley-Seal scheme): 16 words at a time go through 15 full adders --
// sum = a ^ b ^ c and carry = majority(a, b, c), one three-input logic
// operation each on gfx950 -- into the bit slices ones..eights, and the
// carry out (sixteens: one bit per position per 16 words) into a second set
// of slices that count groups of 16.  Those are folded into the B int32
// counters every 15 groups (240 words), the rest once at the end: about 6
// vector operations per 64-bit word instead of the 40 of csa_count's fold
// every 15 words.
template <typename W, int B>
struct VCounter {
    W ones = 0, twos = 0, fours = 0, eights = 0; // words below 16, bit-sliced
    W g1 = 0, g2 = 0, g4 = 0, g8 = 0;            // groups of 16, bit-sliced
    int groups = 0;

    static __device__ __forceinline__ void csa(W &carry, W &acc, W b, W c)
    {
        const W a = acc;
        acc = a ^ b ^ c;
        carry = (a & b) | (a & c) | (b & c);
    }

    __device__ __forceinline__ void add16(const W (&x)[16], int32_t (&cnt)[B])
    {
        W twosA, twosB, foursA, foursB, eightsA, eightsB, sixteens;
        csa(twosA, ones, x[0], x[1]);
        csa(twosB, ones, x[2], x[3]);
        csa(foursA, twos, twosA, twosB);
        csa(twosA, ones, x[4], x[5]);
        csa(twosB, ones, x[6], x[7]);
        csa(foursB, twos, twosA, twosB);
        csa(eightsA, fours, foursA, foursB);
        csa(twosA, ones, x[8], x[9]);
        csa(twosB, ones, x[10], x[11]);
        csa(foursA, twos, twosA, twosB);
        csa(twosA, ones, x[12], x[13]);
        csa(twosB, ones, x[14], x[15]);
        csa(foursB, twos, twosA, twosB);
        csa(eightsB, fours, foursA, foursB);
        csa(sixteens, eights, eightsA, eightsB);
        // one more group of 16 in the second slices (at most 15 held)
        const W c1 = g1 & sixteens;
        g1 ^= sixteens;
        const W c2 = g2 & c1;
        g2 ^= c1;
        const W c4 = g4 & c2;
        g4 ^= c2;
        g8 |= c4;
        if (++groups == 15) {
            fold_groups(cnt);
            groups = 0;
        }
    }

    __device__ __forceinline__ void fold_groups(int32_t (&cnt)[B])
    {
#pragma unroll
        for (int k = 0; k < B; ++k)
            cnt[k] += (((int32_t)((g1 >> k) & 1)) + ((int32_t)((g2 >> k) & 1) << 1) +
                       ((int32_t)((g4 >> k) & 1) << 2) + ((int32_t)((g8 >> k) & 1) << 3))
                      << 4;
        g1 = g2 = g4 = g8 = 0;
    }

    // Adds everything still held into cnt (once, at the end).
    __device__ __forceinline__ void flush(int32_t (&cnt)[B])
    {
        fold_groups(cnt);
#pragma unroll
        for (int k = 0; k < B; ++k)
            cnt[k] += (int32_t)((ones >> k) & 1) + ((int32_t)((twos >> k) & 1) << 1) +
                      ((int32_t)((fours >> k) & 1) << 2) + ((int32_t)((eights >> k) & 1) << 3);
        ones = twos = fours = eights = 0;
        groups = 0;
    }
};

} // namespace ibp
