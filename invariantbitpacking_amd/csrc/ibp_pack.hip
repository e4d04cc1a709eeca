// ibp_pack.hip -- compressed sizes, the in-place / slot pack and the condensed
// stream.
//
// get_compress_size:  src/compress.cu:25-174 -> check_compress_size_kernel
//   (include/misc/ibp_misc_kernels.cuh:9-32, ibp_misc_dev.cuh:20-46).
// compress_inplace:   src/compress.cu:176-236 -> compress_inplace_kernel
//   (include/compress/ibp_compress_kernel.cuh:10-56, ibp_compress_dev.cuh:8-97).
// compress_condensed: src/compress.cu:238-313 -> sizes, thrust::inclusive_scan,
//   compress_condensed_kernel (ibp_compress_kernel.cuh:59-94).
//
// MI355X mapping: one wave64 per row, lane j owns element j of each 64-element
// chunk; a wave owns 32 consecutive output rows so it writes their bitmask
// word itself (no atomics, no pre-zeroing).  The packed slot is assembled in a
// per-wave LDS buffer (ds_or for the payload, ballot for the header words)
// and then written out, so in-place packing never overwrites input the wave
// still has to read.  Words of 1 or 2 bytes are OR-ed through the 32-bit word
// that contains them (LDS has no sub-dword atomics).
#include <mutex>
#include <type_traits>

#include "ibp_bits.h"
#include "ibp_fetch_dev.h"
#include "ibp_internal.h"
#include "ibp_tile.h"
#include "ibp_wave.h"

namespace ibp {

template <typename IdxT>
__device__ __forceinline__ int64_t row_of(const void *index, int64_t i)
{
    return index ? (int64_t)((const IdxT *)index)[i] : i;
}

// Saved bits S of one row: sum of popc(mask[j]) over flagged elements.
template <typename T>
__device__ __forceinline__ int row_saved_bits(const T *row, const T *mask, const T *bitval, int64_t V)
{
    const int lane = lane_id();
    int saved = 0;
    for (int64_t j = lane; j < V; j += kWave) {
        const T x = row[j], m = mask[j], b = bitval[j];
        if ((T)(x & m) == b) saved += popc((typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type)m);
    }
    return wave_sum(saved);
}

// Mask row of source row r (clusters NULL: row 0); -1 if the id is invalid.
__device__ __forceinline__ int64_t mask_row_of(const int32_t *clusters, int64_t num_clusters, int64_t r)
{
    if (!clusters) return 0;
    const int64_t c = readfirstlane((int)clusters[r]);
    return (c >= 0 && c < num_clusters) ? c : -1;
}

template <typename T, typename IdxT>
__global__ __launch_bounds__(256) void size_kernel(const T *data, int64_t num_vecs, int64_t V, const T *mask,
                                                   const T *bitval, const int32_t *clusters,
                                                   int64_t num_clusters, const void *index, int64_t n,
                                                   int64_t *sizes, unsigned long long *total,
                                                   unsigned long long *err)
{
    const int lane = lane_id();
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    unsigned long long acc = 0, bad = 0;
    if (V < kWave / 2 && clusters == nullptr) {
        // Rows of fewer than 32 elements with one mask: G = 64 / V rows per
        // step, lane l on column l % V of row l / V (consecutive rows: one
        // load reads G * V consecutive elements).  A row's saved bits are the
        // wave's inclusive scan at its last lane minus the scan before its
        // first; the row's first lane stores its size.
        using W = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;
        const int Vi = (int)V;
        const int G = kWave / Vi;
        const int lrow = lane / Vi, lcol = lane % Vi;
        const bool on = lrow < G;
        const T m = on ? mask[lcol] : T(0), b = on ? bitval[lcol] : T(1);
        const int pc = on ? popc((W)m) : 0; // lanes past the rows never match (b has a bit outside m)
        const int first = lrow * Vi, last = first + Vi - 1;
        for (int64_t i0 = wave * G; i0 < n; i0 += nwaves * G) {
            const int64_t i = i0 + lrow;
            const bool live = on && i < n;
            const int64_t r = live ? row_of<IdxT>(index, i) : -1;
            const bool ok = r >= 0 && r < num_vecs;
            const T x = ok ? data[r * V + lcol] : T(0);
            const int incl = wave_incl_scan(ok && (T)(x & m) == b ? pc : 0);
            // both exchanges run in every lane: ds_bpermute returns 0 from a
            // lane that is not active
            const int hi = (int)shfl((uint32_t)incl, last & (kWave - 1));
            const int before = (int)shfl((uint32_t)incl, (first + kWave - 1) & (kWave - 1));
            const int lo = first > 0 ? before : 0;
            if (live && lcol == 0) {
                const int64_t size = ok ? slot_size(V, (int)sizeof(T), hi - lo) : 0;
                sizes[i] = size;
                acc += (unsigned long long)size;
                bad += ok ? 0 : 1;
            }
        }
        // the wave's totals: 24-bit low parts and the rest, each summed in 32 bits
        const unsigned long long a =
            (unsigned long long)(uint32_t)wave_sum((int)(uint32_t)(acc & 0xffffffull)) +
            ((unsigned long long)(uint32_t)wave_sum((int)(uint32_t)(acc >> 24)) << 24);
        const unsigned long long e = (unsigned long long)(uint32_t)wave_sum((int)bad);
        if (lane == 0) {
            if (total && a) atomicAdd(total, a);
            if (err && e) atomicAdd(err, e);
        }
        return;
    }
    if (V <= kWave && clusters == nullptr) {
        // Rows of <= 64 elements with one mask: the lane's column constants
        // stay in registers and R rows are in flight per wave.
        constexpr int R = 4;
        using W = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;
        const bool col = lane < V;
        const T m = col ? mask[lane] : T(0), b = col ? bitval[lane] : T(1);
        const int pc = col ? popc((W)m) : 0; // lanes past the row never match (b has a bit outside m)
        for (int64_t i0 = wave * R; i0 < n; i0 += nwaves * R) {
            int64_t r[R];
            T x[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                r[q] = i0 + q < n ? readfirstlane(row_of<IdxT>(index, i0 + q)) : -1;
                const bool ok = r[q] >= 0 && r[q] < num_vecs;
                x[q] = (ok && col) ? data[r[q] * V + lane] : T(0);
            }
            // Two rows per 32-bit reduction: a row saves at most 64 * 64 < 2^16 bits.
            int saved[R];
#pragma unroll
            for (int q = 0; q < R; q += 2) {
                const int s2 = wave_sum(((T)(x[q] & m) == b ? pc : 0) | (((T)(x[q + 1] & m) == b ? pc : 0) << 16));
                saved[q] = s2 & 0xffff;
                saved[q + 1] = s2 >> 16;
            }
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (i0 + q >= n) break; // uniform
                int64_t size = 0;
                if (r[q] >= 0 && r[q] < num_vecs)
                    size = slot_size(V, (int)sizeof(T), saved[q]);
                else
                    ++bad;
                if (lane == 0) sizes[i0 + q] = size;
                acc += (unsigned long long)size;
            }
        }
    } else {
        for (int64_t i = wave; i < n; i += nwaves) {
            const int64_t r = readfirstlane(row_of<IdxT>(index, i));
            int64_t size = 0;
            const int64_t c = (r >= 0 && r < num_vecs) ? mask_row_of(clusters, num_clusters, r) : -1;
            if (c >= 0) {
                const int saved = row_saved_bits<T>(data + r * V, mask + c * V, bitval + c * V, V);
                size = slot_size(V, (int)sizeof(T), saved);
            } else {
                ++bad;
            }
            if (lane == 0) sizes[i] = size;
            acc += (unsigned long long)size;
        }
    }
    if (lane == 0) {
        if (total && acc) atomicAdd(total, acc);
        if (err && bad) atomicAdd(err, bad);
    }
}

// get_compress_size of a whole table (no index, one mask) in 64-row tiles
// copied to LDS: lane r scores row r of the tile.  A row saves the mask
// popcounts of its flagged elements, i.e. the sum of all popcounts minus
// those of the elements that are not flagged (lost_step).  Mask and bitval
// are uniform scalar loads straight from the arguments, a chunk ahead (the
// last chunk's look-ahead re-reads it), popcounts on the scalar unit.
template <typename T>
__device__ __forceinline__ void size_score_chunk(const uint4 &v, const T (&m)[16 / sizeof(T)],
                                                 const T (&b)[16 / sizeof(T)], uint32_t &lost)
{
#pragma unroll
    for (int e = 0; e < (int)(16 / sizeof(T)); ++e)
        lost = lost_step<true>(xor_ordered(chunk_elem<T>(v, e), b[e]), m[e], (uint32_t)popc(m[e]), lost);
}

template <typename T>
__device__ __forceinline__ void size_load_chunk(const T *__restrict__ mask, const T *__restrict__ bitval, int c,
                                                T (&m)[16 / sizeof(T)], T (&b)[16 / sizeof(T)])
{
    constexpr int EPC = 16 / (int)sizeof(T);
#pragma unroll
    for (int e = 0; e < EPC; ++e) {
        m[e] = mask[c * EPC + e];
        b[e] = bitval[c * EPC + e];
    }
}

template <typename T>
__device__ __forceinline__ void size_need_chunk(const T (&m)[16 / sizeof(T)], const T (&b)[16 / sizeof(T)],
                                                const uint4 &v)
{
#pragma unroll
    for (int e = 0; e < (int)(16 / sizeof(T)); ++e) asm volatile("" ::"s"(m[e]), "s"(b[e]));
    asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
}

template <typename T>
__global__ __launch_bounds__(kWave) void size_tile_kernel(const T *__restrict__ data, int64_t num_vecs, int V,
                                                          const T *__restrict__ mask, const T *__restrict__ bitval,
                                                          int64_t *__restrict__ sizes, unsigned long long *total)
{
    constexpr int W = (int)sizeof(T);
    constexpr int EPC = 16 / W;
    extern __shared__ __align__(16) unsigned char tile[];
    const int lane = lane_id();
    const int C = V * W / 16;
    const int64_t ntiles = (num_vecs + kWave - 1) / kWave;
    int pcs = 0;
    for (int j = lane; j < V; j += kWave) pcs += popc(mask[j]);
    const uint32_t pcsum = (uint32_t)wave_sum(pcs);
    const uint4 *rowq = reinterpret_cast<const uint4 *>(tile + (size_t)lane * V * W);
    const bool swz = tile_swizzled((int64_t)V * W);
    const int sw = tile_swz(lane, swz); // chunk c of this lane's row sits at c ^ sw
    unsigned long long acc = 0;
    for (int64_t tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
        copy_tile<T, 1>(data, num_vecs, V, C, tl, 0, lane, tile, swz);
        __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0): the tile has landed
        __builtin_amdgcn_wave_barrier();
        uint32_t lost = 0;
        T ma[EPC], ba[EPC], mb[EPC], bb[EPC];
        size_load_chunk<T>(mask, bitval, 0, ma, ba);
        uint4 va = rowq[sw], vb;
        int c = 0;
        for (; c + 1 < C; c += 2) {
            size_need_chunk<T>(ma, ba, va);
            size_load_chunk<T>(mask, bitval, c + 1, mb, bb);
            vb = rowq[(c + 1) ^ sw];
            size_score_chunk<T>(va, ma, ba, lost);
            size_need_chunk<T>(mb, bb, vb);
            const int cn = c + 2 < C ? c + 2 : c + 1;
            size_load_chunk<T>(mask, bitval, cn, ma, ba);
            va = rowq[cn ^ sw];
            size_score_chunk<T>(vb, mb, bb, lost);
        }
        if (c < C) { // odd C: the last chunk
            size_need_chunk<T>(ma, ba, va);
            size_score_chunk<T>(va, ma, ba, lost);
        }
        const int64_t r = tl * kWave + lane;
        if (r < num_vecs) {
            const int64_t sz = slot_size(V, W, (int64_t)(pcsum - lost));
            sizes[r] = sz;
            acc += (unsigned long long)sz;
        }
        // the tile's LDS reads are done before the next copy lands
        __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0)
        __builtin_amdgcn_wave_barrier();
    }
    if (total) {
        // one atomic per wave: the lanes' sums as 24-bit low parts and the
        // rest (each of 64 fits 32 bits while a lane's sum is below 2^49)
        const uint32_t lo = (uint32_t)wave_sum((int)(uint32_t)(acc & 0xffffffull));
        const uint32_t hi = (uint32_t)wave_sum((int)(uint32_t)(acc >> 24));
        const unsigned long long tot = (unsigned long long)lo + ((unsigned long long)hi << 24);
        if (lane == 0 && tot) atomicAdd(total, tot);
    }
}

// Orders one wave's accesses to its slot buffer (LDS: wave scope suffices;
// device memory: agent scope, so stores and L2 atomics are seen in order).
template <bool BUF_LDS>
__device__ __forceinline__ void slot_fence()
{
    if constexpr (BUF_LDS)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    else
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
}

// buf[q] |= v.  Words narrower than 32 bits are OR-ed through their 32-bit
// container (little-endian: byte b of the container is bits 8b..8b+7); the
// buffer is 4-byte aligned and padded to a multiple of 16 bytes.
template <typename T>
__device__ __forceinline__ void or_word(T *buf, int64_t q, T v)
{
    if constexpr (sizeof(T) >= 4) {
        __hip_atomic_fetch_or(&buf[q], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        const int64_t byte = q * (int64_t)sizeof(T);
        uint32_t *d = (uint32_t *)((unsigned char *)buf + (byte & ~int64_t(3)));
        __hip_atomic_fetch_or(d, (uint32_t)v << (8 * (int)(byte & 3)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Free bits of a flagged element gathered MSB first (the inverse of
// small_deposit): run k of the column's free bits moves up by shift[k].
template <typename T>
__device__ __forceinline__ T small_extract(const SmallCols<T> &c, T x)
{
    if (c.wave_runs > kRunCap) return extract_msb<T>(x, c.free);
    T v = T(0);
#pragma unroll
    for (int k = 0; k < kRunCap; ++k)
        if (k < c.wave_runs) v |= T(T(x & c.run[k]) << c.shift[k]);
    return v;
}

// Bytes of one wave's slot buffer: V words padded to 16 bytes.
__host__ __device__ __forceinline__ int64_t slot_buf_bytes(int64_t V, int w) { return (V * w + 15) / 16 * 16; }

struct PackParams {
    const void *in;
    void *out;
    int64_t num_vecs;
    int64_t V;
    const void *mask;
    const void *bitval;
    const int32_t *clusters; // mask row per source row, or nullptr
    int64_t num_clusters;
    const int64_t *offsets;  // condensed output: inclusive end byte offset per output row
    int64_t stream_bytes;
    const void *index;
    int64_t n;
    uint32_t *bitmask; // may be nullptr with a condensed output
    void *gbuf;        // per-wave slot buffers when they do not fit LDS
    unsigned long long *err;
    int split = 1;     // waves of a block sharing one 32-row group (1: a wave per group)
};

// Pack kernel.  BUF_LDS: the per-wave slot buffer is in LDS (dynamic shared
// memory); otherwise in p.gbuf.  CONDENSED: output row i goes to bytes
// [offsets[i-1], offsets[i]) of the stream, cut at its size; otherwise to the
// V-element slot i.
//
// Groups of 32 rows (one bitmask word each) go to one wave, or, with
// p.split > 1 (few groups, wide rows), to the p.split waves of a block, which
// take every p.split-th row and combine the group's bitmask word in LDS.
template <typename T, typename IdxT, bool BUF_LDS, bool CONDENSED>
__global__ __launch_bounds__(1024) void pack_kernel(PackParams p)
{
    constexpr int B = Bits<T>::B;
    constexpr int W = (int)sizeof(T);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = lane_id();
    const int wave_in_block = wave_in_block_id();
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int64_t V = p.V;
    const int64_t buf_bytes = slot_buf_bytes(V, W);
    unsigned char *bufb = BUF_LDS ? lds_raw + (int64_t)wave_in_block * buf_bytes
                                  : (unsigned char *)p.gbuf + wave * buf_bytes;
    T *buf = (T *)bufb;
    const T *in = (const T *)p.in;
    const int64_t Hw = header_bytes(V, W) / W;
    const int64_t vec_bytes = V * W;
    const bool in_place = !CONDENSED && p.in == (const void *)p.out && p.index == nullptr;
    const int64_t ngroups = (p.n + 31) / 32;
    unsigned long long bad = 0;
    // Rows of <= 64 elements with one mask: the lane's column constants and
    // run table stay in registers, and each row is read once.
    const bool small = V <= kWave && p.clusters == nullptr;
    SmallCols<T> col{};
    if (small) col = load_small_cols<T>(p.mask, p.bitval, (int)V);
    const int len_raw = col.valid ? B : 0;

    const int split = p.split;
    __shared__ uint32_t group_bits;
    // LDS slot buffers of 4- and 8-byte words are zeroed once: the read-back
    // of each packed row exchanges every word with zero, leaving the buffer
    // clean for the next row.  Otherwise every packed row zeroes it first.
    constexpr bool XCHG = BUF_LDS && sizeof(T) >= 4;
    if constexpr (XCHG) {
        for (int64_t k = lane; k < buf_bytes / 4; k += kWave) ((uint32_t *)bufb)[k] = 0u;
        slot_fence<BUF_LDS>();
    }
    for (int64_t g = split > 1 ? (int64_t)blockIdx.x : wave; g < ngroups; g += split > 1 ? (int64_t)gridDim.x : nwaves) {
        if (split > 1) {
            if (threadIdx.x == 0) group_bits = 0;
            __syncthreads();
        }
        uint32_t bits = 0;
        const int64_t iend = (g + 1) * 32 < p.n ? (g + 1) * 32 : p.n;
        for (int64_t i = g * 32 + (split > 1 ? wave_in_block : 0); i < iend; i += split) {
            const int64_t r = readfirstlane(row_of<IdxT>(p.index, i));
            const int64_t c = (r >= 0 && r < p.num_vecs) ? mask_row_of(p.clusters, p.num_clusters, r) : -1;
            if (c < 0) {
                ++bad;
                continue;
            }
            const T *src = in + r * V;
            const T *mask = (const T *)p.mask + c * V;
            const T *bitval = (const T *)p.bitval + c * V;
            T xs = T(0);
            bool fs = false;
            int incl = 0; // small rows: inclusive scan of the payload lengths
            int64_t size;
            if (small) {
                xs = col.valid ? src[lane] : T(0);
                fs = col.valid && (T)(xs & T(~col.free)) == col.bitval;
                // The payload scan gives the size as well: the bits saved are
                // V*B minus the payload length.
                incl = wave_incl_scan(fs ? col.len_flagged : len_raw);
                size = slot_size(V, W, (int)V * B - readlane(incl, kWave - 1));
            } else {
                // Wide rows are packed into the slot buffer in one pass over
                // the row; the size follows from the payload length, and a row
                // that does not come out smaller is dropped afterwards.  So
                // every row is read once (a raw row copied out of place twice),
                // which halves the reads of host-resident tables over PCIe.
                if constexpr (!XCHG) {
                    for (int64_t k = lane; k < buf_bytes / 4; k += kWave) ((uint32_t *)bufb)[k] = 0u;
                    slot_fence<BUF_LDS>();
                }
                int64_t cursor = 0;
                // chunk c0 + 64's loads are issued before chunk c0 is packed
                T xn = lane < V ? src[lane] : T(0), mn = lane < V ? mask[lane] : T(0),
                  bn = lane < V ? bitval[lane] : T(0);
                for (int64_t c0 = 0; c0 < V; c0 += kWave) {
                    const int64_t j = c0 + lane;
                    const bool valid = j < V;
                    const T x = xn, m = mn, bvj = bn;
                    if (c0 + kWave < V) {
                        const int64_t jn = j + kWave;
                        xn = jn < V ? src[jn] : T(0);
                        mn = jn < V ? mask[jn] : T(0);
                        bn = jn < V ? bitval[jn] : T(0);
                    }
                    const bool f = valid && (T)(x & m) == bvj;
                    // Header flags of this chunk, LSB first (ibp_compress_dev.cuh:60-71):
                    // 64 / B header words, word k holding flags k*B .. k*B+B-1.
                    const uint64_t hb = ballot(f);
                    if (lane < kWave / B && c0 + lane * B < V) {
                        const T hw = (T)(hb >> (lane * B));
                        if (hw) or_word<T>(buf, c0 / B + lane, hw);
                    }
                    const int L = valid ? (f ? B - popc(m) : B) : 0;
                    const int incl = wave_incl_scan(L);
                    const int tot = readlane(incl, kWave - 1);
                    const int64_t pbit = cursor + incl - L;
                    // Payload: free bits MSB first (flagged) or the whole element.
                    // A packed row ends before word V; words from V on belong
                    // to a row that will be dropped and stay unwritten (the
                    // buffer holds V words).
                    const T v = f ? extract_msb<T>(x, T(~m)) : x;
                    const int64_t q = Hw + pbit / B;
                    const int o = (int)(pbit % B);
                    if (L > 0 && q < V) {
                        or_word<T>(buf, q, shr<T>(v, o));
                        if (o + L > B && q + 1 < V) or_word<T>(buf, q + 1, shl<T>(v, B - o));
                    }
                    cursor += tot;
                }
                size = slot_size(V, W, (int)(V * B - cursor));
            }
            // Size test (ibp_compress_kernel.cuh:28-31): pack only if smaller.
            T *dst = nullptr;
            bool row_ok = true;
            if constexpr (CONDENSED) {
                const int64_t lo = i > 0 ? readfirstlane(p.offsets[i - 1]) : 0;
                const int64_t hi = readfirstlane(p.offsets[i]);
                // The stored length must be this row's size: offsets from
                // another dataset or mask never make a write leave the row.
                if (lo < 0 || hi - lo != size || hi > p.stream_bytes || (lo % W) != 0) {
                    ++bad;
                    row_ok = false;
                } else {
                    dst = (T *)((unsigned char *)p.out + lo);
                }
            } else {
                dst = (T *)p.out + i * V;
            }
            if (!row_ok || size >= vec_bytes) {
                if (XCHG && !small) { // undo the dropped row's words: the buffer stays zero
                    slot_fence<BUF_LDS>();
                    for (int64_t k = lane; k < buf_bytes / 16; k += kWave) ((uint4 *)bufb)[k] = make_uint4(0u, 0u, 0u, 0u);
                    slot_fence<BUF_LDS>();
                }
                if (row_ok && !in_place) {
                    if (small) {
                        if (col.valid) dst[lane] = xs;
                    } else {
                        for (int64_t k = lane; k < V; k += kWave) dst[k] = src[k];
                    }
                }
                continue;
            }
            if (small) {
                if constexpr (!XCHG) {
                    for (int64_t k = lane; k < buf_bytes / 4; k += kWave) ((uint32_t *)bufb)[k] = 0u;
                    slot_fence<BUF_LDS>();
                }
                const uint64_t hb = ballot(fs);
                if (lane < kWave / B && lane * B < V) {
                    const T hw = (T)(hb >> (lane * B));
                    if (hw) or_word<T>(buf, lane, hw);
                }
                const int L = fs ? col.len_flagged : len_raw;
                const unsigned pbit = (unsigned)(incl - L);
                const T v = fs ? small_extract<T>(col, xs) : xs;
                const int64_t q = Hw + pbit / B;
                const int o = (int)(pbit % B);
                if (L > 0) {
                    or_word<T>(buf, q, shr<T>(v, o));
                    if (o + L > B) or_word<T>(buf, q + 1, shl<T>(v, B - o));
                }
            }
            slot_fence<BUF_LDS>();
            const int64_t words = size / W; // the rest of the slot is zero
            for (int64_t k = lane; k < (CONDENSED ? words : V); k += kWave) {
                if constexpr (XCHG)
                    dst[k] = __hip_atomic_exchange(&buf[k], T(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    dst[k] = __hip_atomic_load(&buf[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if constexpr (XCHG) slot_fence<BUF_LDS>(); // the exchanges land before the next row's ORs
            bits |= 1u << (i - g * 32);
        }
        if (split > 1) {
            if (lane == 0 && bits) atomicOr(&group_bits, bits);
            __syncthreads();
            if (threadIdx.x == 0 && p.bitmask) p.bitmask[g] = group_bits;
            __syncthreads();
        } else if (lane == 0 && p.bitmask) {
            p.bitmask[g] = bits;
        }
    }
    if (p.err && bad && lane == 0) atomicAdd(p.err, bad);
}

// Pack of rows of <= 64 elements of 4 or 8 bytes with one mask (the GNN and
// DLRM shapes).  A wave takes whole 32-row groups (one bitmask word each) in
// batches of PB rows: while batch b is packed, the rows of batch b + 1 --
// and their indices and stream offsets -- are already being loaded, so a
// wave keeps up to 2 * PB rows in flight instead of one.  Each row is packed
// as in pack_kernel: lane j holds element j; one DPP scan of the payload
// lengths gives every lane its bit position and the row's size; the slot is
// OR-ed together in the wave's LDS buffer and read back by exchanging every
// word with zero, which leaves the buffer clean for the next row.
constexpr int kPackBatch = 8;

template <typename T, typename IdxT, bool CONDENSED>
__global__ __launch_bounds__(256) void pack_small_kernel(PackParams p)
{
    static_assert(sizeof(T) >= 4, "4- and 8-byte words");
    constexpr int B = Bits<T>::B;
    constexpr int W = (int)sizeof(T);
    constexpr int PB = kPackBatch;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = lane_id();
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int64_t V = p.V;
    const int64_t buf_bytes = slot_buf_bytes(V, W);
    T *buf = (T *)(lds_raw + (int64_t)wave_in_block_id() * buf_bytes);
    const T *in = (const T *)p.in;
    const int64_t Hw = header_bytes(V, W) / W;
    const int64_t vec_bytes = V * W;
    const bool in_place = !CONDENSED && p.in == (const void *)p.out && p.index == nullptr;
    const int64_t ngroups = (p.n + 31) / 32;
    unsigned long long bad = 0;
    const SmallCols<T> col = load_small_cols<T>(p.mask, p.bitval, (int)V);
    const int len_raw = col.valid ? B : 0;
    const int64_t colj = col.valid ? lane : 0;
    for (int64_t k = lane; k < buf_bytes / 4; k += kWave) ((uint32_t *)buf)[k] = 0u;
    slot_fence<true>();
    // This wave's batches: group g = wave + (b / 4) * nwaves, rows
    // g * 32 + (b % 4) * PB + q.
    const int64_t my_groups = wave < ngroups ? (ngroups - 1 - wave) / nwaves + 1 : 0;
    const int64_t nbatch = my_groups * (32 / PB);
    // Per batch, lane q < PB holds row q's source row (-1: invalid or past n)
    // and, for a condensed output, its byte range [lo, hi).
    auto batch_meta = [&](int64_t b, int64_t &src_row, int64_t &lo, int64_t &hi) {
        const int64_t i = (wave + (b / (32 / PB)) * nwaves) * 32 + (b % (32 / PB)) * PB + (lane % PB);
        src_row = -1;
        lo = hi = 0;
        if (i < p.n) {
            const int64_t r = row_of<IdxT>(p.index, i);
            src_row = (r >= 0 && r < p.num_vecs) ? r : -1;
            if constexpr (CONDENSED) {
                lo = i > 0 ? p.offsets[i - 1] : 0;
                hi = p.offsets[i];
            }
        }
    };
    int64_t rc = -1, loc = 0, hic = 0, rn = -1, lon = 0, hin = 0;
    T xc[PB], xn[PB];
    if (nbatch > 0) {
        batch_meta(0, rc, loc, hic);
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int64_t r = readlane(rc, q);
            xc[q] = in[(r >= 0 ? r : 0) * V + colj];
        }
    }
    uint32_t bits = 0;
    for (int64_t b = 0; b < nbatch; ++b) {
        // prefetch the next batch (the last batch re-reads itself)
        batch_meta(b + 1 < nbatch ? b + 1 : b, rn, lon, hin);
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int64_t r = readlane(rn, q);
            xn[q] = in[(r >= 0 ? r : 0) * V + colj];
        }
        const int64_t g = wave + (b / (32 / PB)) * nwaves;
        const int64_t i0 = g * 32 + (b % (32 / PB)) * PB;
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int64_t i = i0 + q;
            if (i >= p.n) break; // uniform
            const int64_t r = readlane(rc, q);
            if (r < 0) {
                ++bad;
                continue;
            }
            const T xs = col.valid ? xc[q] : T(0);
            const bool fs = col.valid && (T)(xs & T(~col.free)) == col.bitval;
            const int incl = wave_incl_scan(fs ? col.len_flagged : len_raw);
            const int saved = (int)V * B - readlane(incl, kWave - 1);
            const int64_t size = slot_size(V, W, saved);
            T *dst;
            if constexpr (CONDENSED) {
                const int64_t lo = readlane(loc, q), hi = readlane(hic, q);
                // The stored length must be this row's size: offsets from
                // another dataset or mask never make a write leave the row.
                if (lo < 0 || hi - lo != size || hi > p.stream_bytes || (lo % W) != 0) {
                    ++bad;
                    continue;
                }
                dst = (T *)((unsigned char *)p.out + lo);
            } else {
                dst = (T *)p.out + i * V;
            }
            if (size >= vec_bytes) {
                if (!in_place && col.valid) dst[lane] = xs;
                continue;
            }
            const uint64_t hb = ballot(fs);
            if (lane < kWave / B && lane * B < V) {
                const T hw = (T)(hb >> (lane * B));
                if (hw) or_word<T>(buf, lane, hw);
            }
            const int L = fs ? col.len_flagged : len_raw;
            const unsigned pbit = (unsigned)(incl - L);
            const T v = fs ? small_extract<T>(col, xs) : xs;
            const int64_t qw = Hw + pbit / B;
            const int o = (int)(pbit % B);
            if (L > 0) {
                or_word<T>(buf, qw, shr<T>(v, o));
                if (o + L > B) or_word<T>(buf, qw + 1, shl<T>(v, B - o));
            }
            slot_fence<true>();
            const int64_t words = CONDENSED ? size / W : V; // the rest of the slot is zero
            if (lane < words) dst[lane] = __hip_atomic_exchange(&buf[lane], T(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else if (lane < V) (void)__hip_atomic_exchange(&buf[lane], T(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            slot_fence<true>(); // the exchanges land before the next row's ORs
            bits |= 1u << (i - g * 32);
        }
        if ((b % (32 / PB)) == 32 / PB - 1) { // the group is complete
            if (lane == 0 && p.bitmask) p.bitmask[g] = bits;
            bits = 0;
        }
        rc = rn;
        loc = lon;
        hic = hin;
#pragma unroll
        for (int q = 0; q < PB; ++q) xc[q] = xn[q];
    }
    if (p.err && bad && lane == 0) atomicAdd(p.err, bad);
}

// ---- Row-per-lane pack of 64-row tiles, in place, one mask.
//
// The tile is copied to LDS (global_load_lds), lane r packs row r there in
// place, and the packed slots go back to the table with coalesced 16-byte
// stores.  Per element a lane tests the flag, gathers the free bits of a
// flagged element with the column's run table, and appends them to a word
// accumulator; a completed word is written to the row in LDS.  Column
// constants are uniform scalar loads, a chunk ahead.  In place is safe: after
// element j at most j + 1 payload words are complete, so the words written
// while chunk c is packed end inside chunk c + 1 (the header is at most one
// chunk), which was read before chunk c's first write; a raw row's words past
// its end stay in its own last word.
//
// Compared with a wave per row (pack_small_kernel: a wave scan, a ballot and
// LDS atomics per row, about 78 vector and 76 scalar instructions per row on
// the GNN table) a lane per row needs about 30 vector instructions per
// element-lane, i.e. per row of 64 rows a wave packs together.

// Column constants of the pack: mask and bitval, the runs of free bits of a
// flagged element (run k moves up by sh[k] to its place in the MSB-first
// payload), the payload bits of a flagged element, and the run count (more
// than kRunCap: the generic gather).  Laid out [chunk][element], with a zero
// chunk after the last.
template <typename T>
struct PackCol {
    T m;
    T b;
    T run[kRunCap];
    uint32_t sh[kRunCap];
    uint32_t len;
    uint32_t nruns;
};

// One block: thread j < (C + 1) * EPC builds column j's record; *max_runs is
// the largest run count.
template <typename T>
__global__ __launch_bounds__(256) void pack_cols_kernel(const T *__restrict__ mask, const T *__restrict__ bitval,
                                                        int64_t V, int ncols, PackCol<T> *cols, int *max_runs)
{
    constexpr int B = Bits<T>::B;
    __shared__ int mx;
    if (threadIdx.x == 0) mx = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < ncols; j += blockDim.x) {
        PackCol<T> c;
        const T m = j < V ? mask[j] : T(0);
        c.m = m;
        c.b = j < V ? bitval[j] : T(0);
        c.len = j < V ? (uint32_t)(B - popc(m)) : 0u;
        for (int k = 0; k < kRunCap; ++k) {
            c.run[k] = T(0);
            c.sh[k] = 0;
        }
        T z = j < V ? T(~m) : T(0);
        int filled = 0, runs = 0;
        while (z) {
            const int sft = clz(z);
            const int r = clz(T(~T(z << sft)));
            const T run = shr(himask<T>(r), sft);
            if (runs < kRunCap) {
                c.run[runs] = run;
                c.sh[runs] = (uint32_t)(sft - filled);
            }
            filled += r;
            z &= ~run;
            ++runs;
        }
        c.nruns = (uint32_t)runs;
        cols[j] = c;
        atomicMax(&mx, runs);
    }
    __syncthreads();
    if (threadIdx.x == 0) *max_runs = mx;
}

// The constants of one chunk in SGPRs (R runs per element; R < 0: the generic
// gather, which needs only the mask).
template <typename T, int R>
struct PackChunkCols {
    static constexpr int EPC = 16 / (int)sizeof(T);
    static constexpr int NR = R > 0 ? R : 1;
    T m[EPC], b[EPC], run[EPC][NR];
    uint32_t sh[EPC][NR], len[EPC];
};

template <typename T, int R>
__device__ __forceinline__ void load_pack_cols(const PackCol<T> *__restrict__ c, PackChunkCols<T, R> &k)
{
    constexpr int EPC = PackChunkCols<T, R>::EPC;
#pragma unroll
    for (int e = 0; e < EPC; ++e) {
        const T *f = reinterpret_cast<const T *>(c + e); // m, b, run[kRunCap]
        const uint32_t *g = reinterpret_cast<const uint32_t *>(f + 2 + kRunCap); // sh[kRunCap], len
        k.m[e] = f[0];
        k.b[e] = f[1];
#pragma unroll
        for (int q = 0; q < (R > 0 ? R : 0); ++q) {
            k.run[e][q] = f[2 + q];
            k.sh[e][q] = g[q];
        }
        k.len[e] = g[kRunCap];
    }
}

// Empty asm that needs the chunk's constants in SGPRs (places the wait for
// their scalar loads).
template <typename T, int R>
__device__ __forceinline__ void need_pack_cols(const PackChunkCols<T, R> &k)
{
    constexpr int EPC = PackChunkCols<T, R>::EPC;
#pragma unroll
    for (int e = 0; e < EPC; ++e) {
        asm volatile("" ::"s"(k.m[e]), "s"(k.b[e]), "s"(k.len[e]));
#pragma unroll
        for (int q = 0; q < (R > 0 ? R : 0); ++q) asm volatile("" ::"s"(k.run[e][q]), "s"(k.sh[e][q]));
    }
}

// Per-lane state of one row being packed.
template <typename T>
struct RowPacker {
    T acc;        // the payload word being filled, MSB first
    int nb;       // bits already in acc (< B)
    int wpos;     // slot word the next completed word goes to
    uint64_t fl;  // flags, bit j for element j
};

// Packs the EPC elements of chunk c (16 bytes v) of a row: completed words go
// to row[wpos ^ swe] (swe: the row's chunk XOR in words), or to the lane's
// dummy word when none completes (no branch).  A row whose words reach V is
// raw and never written back: its words past V - 1 go to its own last word,
// so no lane writes into another lane's row.  wet = false sends every word
// to the dummy: the row is measured, not changed.
template <typename T, int R>
__device__ __forceinline__ void pack_chunk(const uint4 &v, const PackChunkCols<T, R> &k, int c, T *row, int swe,
                                           int vlast, T *dummy, RowPacker<T> &st, bool wet = true)
{
    constexpr int B = Bits<T>::B;
    constexpr int EPC = PackChunkCols<T, R>::EPC;
    uint32_t cf = 0; // the chunk's flags, bit e for element e, shifted into fl once
#pragma unroll
    for (int e = 0; e < EPC; ++e) {
        const T x = chunk_elem<T>(v, e);
        const bool fs = (T)(x & k.m[e]) == k.b[e];
        cf |= fs ? (1u << e) : 0u;
        const int L = fs ? (int)k.len[e] : B;
        T ext;
        if constexpr (R < 0) {
            ext = extract_msb<T>(x, T(~k.m[e]));
        } else {
            ext = T(0);
#pragma unroll
            for (int q = 0; q < R; ++q) ext |= T(T(x & k.run[e][q]) << k.sh[e][q]);
        }
        const T val = fs ? ext : x;
        st.acc |= T(val >> st.nb);
        const int t = st.nb + L;
        const bool full = t >= B;
        *(full && wet ? row + ((st.wpos < vlast ? st.wpos : vlast) ^ swe) : dummy) = st.acc;
        // val << (B - nb), which is 0 for nb = 0
        const T rem = T(T(val << 1) << (B - 1 - st.nb));
        st.acc = full ? rem : st.acc;
        st.wpos += full ? 1 : 0;
        st.nb = t & (B - 1);
    }
    st.fl |= (uint64_t)cf << (c * EPC);
}

struct PackTileParams {
    int64_t num_vecs;
    int V;
    const int *max_runs;   // the largest run count of the column records
    uint32_t *bitmask;     // may be nullptr with a condensed output
    // condensed output only: row i goes to bytes [offsets[i-1], offsets[i])
    const int64_t *offsets;
    int64_t stream_bytes;
    unsigned long long *err; // rows whose stored length is not their size
};

// Packs lane's row of the tile, in LDS at `row` (chunk c at rowq[c ^ sw],
// word w at w ^ swe), in place: header flags, then the payload words.
// Returns the slot's length in words, V when the row does not get smaller.
// wet = false measures the row without changing it.
template <typename T, int R>
__device__ __forceinline__ int pack_lane_row(const PackCol<T> *__restrict__ cols, T *row, const uint4 *rowq, int sw,
                                             int swe, int V, int C, int H, T *dummy, bool wet)
{
    RowPacker<T> st;
    st.acc = T(0);
    st.nb = 0;
    st.wpos = H;
    st.fl = 0;
    PackChunkCols<T, R> ka, kb;
    load_pack_cols<T, R>(cols, ka);
    uint4 va = rowq[sw], vb;
    int c = 0;
    for (; c + 1 < C; c += 2) {
        need_pack_cols<T, R>(ka);
        load_pack_cols<T, R>(cols + (size_t)(c + 1) * PackChunkCols<T, R>::EPC, kb);
        vb = rowq[(c + 1) ^ sw];
        pack_chunk<T, R>(va, ka, c, row, swe, V - 1, dummy, st, wet);
        need_pack_cols<T, R>(kb);
        load_pack_cols<T, R>(cols + (size_t)(c + 2) * PackChunkCols<T, R>::EPC, ka); // chunk C is the zero one
        va = rowq[(c + 2 < C ? c + 2 : c + 1) ^ sw];
        pack_chunk<T, R>(vb, kb, c + 1, row, swe, V - 1, dummy, st, wet);
    }
    if (c < C) { // odd C: the last chunk
        need_pack_cols<T, R>(ka);
        pack_chunk<T, R>(va, ka, c, row, swe, V - 1, dummy, st, wet);
    }
    *(wet && st.nb > 0 ? row + ((st.wpos < V - 1 ? st.wpos : V - 1) ^ swe) : dummy) = st.acc; // the last, partial word
    st.wpos += st.nb > 0 ? 1 : 0;
    // header: flag j in word j / B, LSB first (at most two words)
    *(wet ? row + (0 ^ swe) : dummy) = (T)st.fl;
    if (H > 1) *(wet ? row + (1 ^ swe) : dummy) = (T)(st.fl >> 32);
    return st.wpos < V ? st.wpos : V;
}

template <typename T, int R>
__device__ __forceinline__ void pack_tile_body(const PackTileParams &p, T *__restrict__ data,
                                               const PackCol<T> *__restrict__ cols, unsigned char *tile, int *rowlen,
                                               T *dummies)
{
    constexpr int W = (int)sizeof(T);
    constexpr int EPC = 16 / W;
    const int lane = lane_id();
    const int V = p.V;
    const int C = V * W / 16;
    const int H = (int)(header_bytes(V, W) / W);
    const int64_t ntiles = (p.num_vecs + kWave - 1) / kWave;
    const int64_t ngroups = (p.num_vecs + 31) / 32;
    T *const row = reinterpret_cast<T *>(tile) + (size_t)lane * V;
    const uint4 *rowq = reinterpret_cast<const uint4 *>(row);
    T *const dummy = dummies + lane;
    const bool swz = tile_swizzled((int64_t)V * W);
    const int sw = tile_swz(lane, swz); // chunk c of this lane's row sits at c ^ sw
    const int swe = sw * EPC;           // ... and word w at w ^ swe
    // write-back walk: chunk g = lane + 64 i of the tile is (g / C, g % C)
    const int q64 = kWave / C, r64 = kWave % C;
    const int wr0 = lane / C, wc0 = lane - (lane / C) * C;
    for (int64_t tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
        copy_tile<T, 1>(data, p.num_vecs, V, C, tl, 0, lane, tile, swz);
        __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0): the tile has landed
        __builtin_amdgcn_wave_barrier();
        const int64_t r = tl * kWave + lane;
        const int words = pack_lane_row<T, R>(cols, row, rowq, sw, swe, V, C, H, dummy, true);
        const bool packed = r < p.num_vecs && words < V;
        rowlen[lane] = packed ? words : 0;
        const uint64_t pk = ballot(packed);
        if (lane == 0) {
            p.bitmask[2 * tl] = (uint32_t)pk;
            if (2 * tl + 1 < ngroups) p.bitmask[2 * tl + 1] = (uint32_t)(pk >> 32);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // packed slots back to the table: whole slots, zero past the row's
        // words; raw rows stay as they are
        uint4 *out = reinterpret_cast<uint4 *>(data + tl * kWave * (int64_t)V);
        const int64_t rows_left = p.num_vecs - tl * kWave;
        const int nch = (int)((rows_left < kWave ? rows_left : kWave) * C);
        int wr = wr0, wc = wc0;
        for (int g = lane; g < nch; g += kWave) {
            const int len = rowlen[wr];
            if (len > 0) {
                uint4 v = reinterpret_cast<const uint4 *>(tile)[g];
                const int ws = tile_swz(wr, swz); // LDS chunk g is chunk wc ^ ws of row wr
                const int w0 = (wc ^ ws) * EPC;   // first word of the chunk
                if constexpr (W == 8) {
                    if (w0 >= len) v.x = v.y = 0u;
                    if (w0 + 1 >= len) v.z = v.w = 0u;
                } else {
                    if (w0 >= len) v.x = 0u;
                    if (w0 + 1 >= len) v.y = 0u;
                    if (w0 + 2 >= len) v.z = 0u;
                    if (w0 + 3 >= len) v.w = 0u;
                }
                out[g ^ ws] = v;
            }
            wr += q64;
            wc += r64;
            if (wc >= C) {
                wc -= C;
                ++wr;
            }
        }
        // every LDS read of this tile is done before the next copy lands
        __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0)
        __builtin_amdgcn_wave_barrier();
    }
}

// data and cols are separate restrict arguments so that the stores to the
// table cannot alias the column records, which then stay scalar loads.
template <typename T>
__global__ __launch_bounds__(kWave) void pack_tile_kernel(PackTileParams p, T *__restrict__ data,
                                                          const PackCol<T> *__restrict__ cols)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const size_t tile_bytes = (size_t)kWave * p.V * sizeof(T);
    int *rowlen = reinterpret_cast<int *>(lds_raw + tile_bytes);
    T *dummies = reinterpret_cast<T *>(lds_raw + tile_bytes + kWave * sizeof(int));
    const int R = __builtin_amdgcn_readfirstlane(*p.max_runs);
    switch (R) {
    case 0: pack_tile_body<T, 0>(p, data, cols, lds_raw, rowlen, dummies); break;
    case 1: pack_tile_body<T, 1>(p, data, cols, lds_raw, rowlen, dummies); break;
    case 2: pack_tile_body<T, 2>(p, data, cols, lds_raw, rowlen, dummies); break;
    case 3: pack_tile_body<T, 3>(p, data, cols, lds_raw, rowlen, dummies); break;
    case 4: pack_tile_body<T, 4>(p, data, cols, lds_raw, rowlen, dummies); break;
    default: pack_tile_body<T, -1>(p, data, cols, lds_raw, rowlen, dummies); break;
    }
}

// The condensed form of pack_tile_kernel (compress_condensed of a whole table
// with one mask, rows the LDS tiles take): the tile is packed in LDS as
// there, except that a row whose offsets say it is stored raw is only
// measured, so its LDS copy stays the original row.  A row is written only
// when its stored length (offsets[i] - offsets[i-1]) is its size; otherwise
// it is counted in p.err.  Row t's words then go to the stream at
// offsets[t-1], four rows per step: their LDS words are read together, then
// stored, lane k taking word k of each row (rows of <= 64 elements).
template <typename T, int R>
__device__ __forceinline__ void pack_tile_condensed_body(const PackTileParams &p, const T *__restrict__ data,
                                                         T *__restrict__ stream, const PackCol<T> *__restrict__ cols,
                                                         unsigned char *tile, T *dummies)
{
    constexpr int W = (int)sizeof(T);
    constexpr int EPC = 16 / W;
    constexpr int kRowsPerStep = 4;
    const int lane = lane_id();
    const int V = p.V;
    const int C = V * W / 16;
    const int H = (int)(header_bytes(V, W) / W);
    const int64_t ntiles = (p.num_vecs + kWave - 1) / kWave;
    const int64_t ngroups = (p.num_vecs + 31) / 32;
    T *const row = reinterpret_cast<T *>(tile) + (size_t)lane * V;
    const uint4 *rowq = reinterpret_cast<const uint4 *>(row);
    T *const dummy = dummies + lane;
    const bool swz = tile_swizzled((int64_t)V * W);
    const int sw = tile_swz(lane, swz);
    const int swe = sw * EPC;
    int bad = 0;
    for (int64_t tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
        copy_tile<T, 1>(data, p.num_vecs, V, C, tl, 0, lane, tile, swz);
        const int64_t r = tl * kWave + lane;
        const bool in = r < p.num_vecs;
        const int64_t lo = in && r > 0 ? p.offsets[r - 1] : 0;
        const int64_t hi = in ? p.offsets[r] : 0;
        __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0): the tile and the offsets have landed
        __builtin_amdgcn_wave_barrier();
        const int64_t stored = hi - lo;
        const bool range_ok = in && lo >= 0 && hi <= p.stream_bytes && (lo % W) == 0;
        const bool wet = range_ok && stored != (int64_t)V * W; // stored as a packed slot
        const int words = pack_lane_row<T, R>(cols, row, rowq, sw, swe, V, C, H, dummy, wet);
        const bool ok = range_ok && stored == (int64_t)words * W;
        bad += in && !ok ? 1 : 0;
        if (p.bitmask) {
            const uint64_t pk = ballot(ok && words < V);
            if (lane == 0) {
                p.bitmask[2 * tl] = (uint32_t)pk;
                if (2 * tl + 1 < ngroups) p.bitmask[2 * tl + 1] = (uint32_t)(pk >> 32);
            }
        }
        const int len = ok ? words : 0;
        const int64_t dw = lo / W; // the row's first word in the stream
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int t0 = 0; t0 < kWave; t0 += kRowsPerStep) {
            T v[kRowsPerStep];
            int lt[kRowsPerStep];
            int64_t dt[kRowsPerStep];
#pragma unroll
            for (int u = 0; u < kRowsPerStep; ++u) {
                const int t = t0 + u;
                lt[u] = readlane(len, t);
                dt[u] = readlane(dw, t);
                const T *rt = reinterpret_cast<const T *>(tile) + (size_t)t * V;
                v[u] = rt[(lane < lt[u] ? lane : 0) ^ (tile_swz(t, swz) * EPC)];
            }
#pragma unroll
            for (int u = 0; u < kRowsPerStep; ++u)
                if (lane < lt[u]) stream[dt[u] + lane] = v[u];
        }
        // every LDS read of this tile is done before the next copy lands
        __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0)
        __builtin_amdgcn_wave_barrier();
    }
    const int nbad = wave_sum(bad);
    if (p.err && nbad && lane == 0) atomicAdd(p.err, (unsigned long long)nbad);
}

template <typename T>
__global__ __launch_bounds__(kWave) void pack_tile_condensed_kernel(PackTileParams p, const T *__restrict__ data,
                                                                    T *__restrict__ stream,
                                                                    const PackCol<T> *__restrict__ cols)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const size_t tile_bytes = (size_t)kWave * p.V * sizeof(T);
    T *dummies = reinterpret_cast<T *>(lds_raw + tile_bytes + kWave * sizeof(int));
    const int R = __builtin_amdgcn_readfirstlane(*p.max_runs);
    switch (R) {
    case 0: pack_tile_condensed_body<T, 0>(p, data, stream, cols, lds_raw, dummies); break;
    case 1: pack_tile_condensed_body<T, 1>(p, data, stream, cols, lds_raw, dummies); break;
    case 2: pack_tile_condensed_body<T, 2>(p, data, stream, cols, lds_raw, dummies); break;
    case 3: pack_tile_condensed_body<T, 3>(p, data, stream, cols, lds_raw, dummies); break;
    case 4: pack_tile_condensed_body<T, 4>(p, data, stream, cols, lds_raw, dummies); break;
    default: pack_tile_condensed_body<T, -1>(p, data, stream, cols, lds_raw, dummies); break;
    }
}

// ---- inclusive scan of int64 (the thrust::inclusive_scan of compress_condensed,
// src/compress.cu:284-285): tile sums, an exclusive scan of the tile sums in
// one block, then each tile scanned with its carry-in.

constexpr int kScanPer = 16;                  // consecutive values per thread
constexpr int kScanTile = 256 * kScanPer;     // values per 256-thread block

// Inclusive scan over the block's threads (blockDim.x <= 1024) through LDS.
__device__ __forceinline__ int64_t block_incl_scan(int64_t x, int64_t *lds)
{
    const int t = threadIdx.x;
    lds[t] = x;
    __syncthreads();
    for (int d = 1; d < (int)blockDim.x; d <<= 1) {
        const int64_t y = t >= d ? lds[t - d] : 0;
        __syncthreads();
        x += y;
        lds[t] = x;
        __syncthreads();
    }
    return x;
}

__global__ __launch_bounds__(256) void tile_sum_kernel(const int64_t *v, int64_t n, int64_t *sums)
{
    __shared__ int64_t lds[256];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int64_t acc = 0;
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < n) acc += v[base + k];
    const int64_t tot = block_incl_scan(acc, lds);
    if (threadIdx.x == blockDim.x - 1) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(int64_t *sums, int64_t ntiles)
{
    __shared__ int64_t lds[1024];
    int64_t carry = 0;
    for (int64_t b = 0; b < ntiles; b += blockDim.x) {
        const int64_t k = b + threadIdx.x;
        const int64_t x = k < ntiles ? sums[k] : 0;
        const int64_t incl = block_incl_scan(x, lds);
        if (k < ntiles) sums[k] = carry + incl - x; // exclusive
        carry += lds[blockDim.x - 1];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void tile_scan_kernel(int64_t *v, int64_t n, const int64_t *carry_in)
{
    __shared__ int64_t lds[256];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPer;
    int64_t x[kScanPer];
    int64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        x[k] = base + k < n ? v[base + k] : 0;
        acc += x[k];
    }
    const int64_t incl = block_incl_scan(acc, lds);
    int64_t run = carry_in[blockIdx.x] + incl - acc;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        run += x[k];
        if (base + k < n) v[base + k] = run;
    }
}

namespace {

template <typename T, typename IdxT>
int size_typed(const void *data, int64_t num_vecs, int64_t V, const void *mask, const void *bitval,
               const int32_t *clusters, int64_t num_clusters, const void *index, int64_t n, int64_t *sizes,
               uint64_t *total, hipStream_t stream)
{
    const DeviceInfo &dev = device_info();
    if constexpr (sizeof(T) >= 4) {
        // a whole table with one mask, rows the LDS tiles take: a lane per row
        if (!index && !clusters && n == num_vecs && n > 0 && tile_rows_ok(V * (int64_t)sizeof(T)) &&
            ((uintptr_t)data % 16) == 0) {
            const size_t lds = (size_t)kWave * V * sizeof(T);
            int64_t per_cu = (int64_t)(160 * 1024) / (int64_t)lds;
            if (per_cu > 8) per_cu = 8;
            const int64_t ntiles = (n + kWave - 1) / kWave;
            int64_t grid = (int64_t)dev.num_cus * per_cu;
            if (grid > ntiles) grid = ntiles;
            hipLaunchKernelGGL((size_tile_kernel<T>), dim3((unsigned)grid), dim3(kWave), lds, stream, (const T *)data,
                               num_vecs, (int)V, (const T *)mask, (const T *)bitval, sizes,
                               (unsigned long long *)total);
            IBP_LAUNCH_CHECK();
            return IBP_OK;
        }
    }
    // 16 blocks of four waves per CU (8: 13% slower on 32768 x 1024 words)
    int64_t grid = (n * kWave + 255) / 256;
    if (grid > (int64_t)dev.num_cus * 16) grid = (int64_t)dev.num_cus * 16;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL((size_kernel<T, IdxT>), dim3((unsigned)grid), dim3(256), 0, stream, (const T *)data,
                       num_vecs, V, (const T *)mask, (const T *)bitval, clusters, num_clusters, index, n, sizes,
                       (unsigned long long *)total,
                       (index || clusters) ? index_error_counter() : nullptr);
    IBP_LAUNCH_CHECK();
    return IBP_OK;
}

template <typename T>
int size_dispatch(int index_bytes, const void *data, int64_t num_vecs, int64_t V, const void *mask,
                  const void *bitval, const int32_t *clusters, int64_t num_clusters, const void *index, int64_t n,
                  int64_t *sizes, uint64_t *total, hipStream_t stream)
{
    if (index && index_bytes == 4)
        return size_typed<T, int32_t>(data, num_vecs, V, mask, bitval, clusters, num_clusters, index, n, sizes,
                                      total, stream);
    return size_typed<T, int64_t>(data, num_vecs, V, mask, bitval, clusters, num_clusters, index, n, sizes, total,
                                  stream);
}

int size_any(int elem_bytes, int index_bytes, const void *data, int64_t num_vecs, int64_t V, const void *mask,
             const void *bitval, const int32_t *clusters, int64_t num_clusters, const void *index, int64_t n,
             int64_t *sizes, uint64_t *total, hipStream_t s)
{
    switch (elem_bytes) {
    case 1: return size_dispatch<uint8_t>(index_bytes, data, num_vecs, V, mask, bitval, clusters, num_clusters, index, n, sizes, total, s);
    case 2: return size_dispatch<uint16_t>(index_bytes, data, num_vecs, V, mask, bitval, clusters, num_clusters, index, n, sizes, total, s);
    case 4: return size_dispatch<uint32_t>(index_bytes, data, num_vecs, V, mask, bitval, clusters, num_clusters, index, n, sizes, total, s);
    default: return size_dispatch<uint64_t>(index_bytes, data, num_vecs, V, mask, bitval, clusters, num_clusters, index, n, sizes, total, s);
    }
}

// In-place inclusive scan of v[n] (device memory) on `stream`.  The tile
// sums (one int64 per kScanTile values) go to `tile_sums` when given (device
// memory of at least that many values), else to a stream-ordered allocation.
// Every tile sum is written before it is read: nothing needs clearing.
int scan_inclusive(int64_t *v, int64_t n, hipStream_t stream, int64_t *tile_sums = nullptr)
{
    if (n <= 0) return IBP_OK;
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    int64_t *sums = tile_sums;
    if (!sums) IBP_HIP(hipMallocAsync((void **)&sums, (size_t)ntiles * sizeof(int64_t), stream));
    hipLaunchKernelGGL(tile_sum_kernel, dim3((unsigned)ntiles), dim3(256), 0, stream, (const int64_t *)v, n, sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, stream, sums, ntiles);
    hipLaunchKernelGGL(tile_scan_kernel, dim3((unsigned)ntiles), dim3(256), 0, stream, v, n, (const int64_t *)sums);
    hipError_t e = hipGetLastError();
    if (!tile_sums) IBP_HIP(hipFreeAsync(sums, stream));
    if (e != hipSuccess) return set_error(IBP_ERR_HIP, "scan kernel launch failed: %s", hipGetErrorString(e));
    return IBP_OK;
}

// Scratch of ibp_condensed_offsets_noalloc: the tile sums of its offsets
// scan.  kNoallocScanTiles values of 8 bytes (512 KiB), so up to
// kNoallocScanTiles * kScanTile = 268,435,456 rows per call.  Each device holds
// its own copy for the life of the process; calls on one device share it.
constexpr int64_t kNoallocScanTiles = 65536;
static __device__ __align__(256) int64_t scratch[kNoallocScanTiles];

std::mutex g_scratch_mu;
int64_t *g_scratch_ptr[64] = {nullptr};

// The device address of `scratch` on the current device, looked up once per
// device with that device current.  (In host code `scratch` names a host-side
// stand-in, never the device memory.)  Sets the error and returns nullptr on
// failure.
int64_t *noalloc_scratch()
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64) {
        set_error(IBP_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(IBP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = (int64_t *)ptr;
    }
    return g_scratch_ptr[dev];
}

// LDS bytes of one pack_tile_kernel block: the tile, the rows' word counts
// and the lanes' dummy words.
inline size_t pack_tile_lds(int64_t V, int w) { return (size_t)kWave * V * w + kWave * sizeof(int) + kWave * w; }

// Pack of a whole table of rows of <= 64 elements in 64-row tiles, in place
// (pack_tile_kernel) or into a condensed stream (pack_tile_condensed_kernel);
// the column records live in a stream-ordered allocation.
template <typename T>
int pack_tile_launch(const PackParams &p, hipStream_t stream, bool condensed = false)
{
    const DeviceInfo &dev = device_info();
    constexpr int W = (int)sizeof(T);
    const int V = (int)p.V;
    const int C = V * W / 16;
    const int ncols = (C + 1) * (16 / W);
    const size_t cols_bytes = ((size_t)ncols * sizeof(PackCol<T>) + 255) / 256 * 256;
    void *buf = nullptr;
    IBP_HIP(hipMallocAsync(&buf, cols_bytes + 256, stream));
    PackCol<T> *cols = (PackCol<T> *)buf;
    int *max_runs = (int *)((unsigned char *)buf + cols_bytes);
    hipLaunchKernelGGL((pack_cols_kernel<T>), dim3(1), dim3(256), 0, stream, (const T *)p.mask, (const T *)p.bitval,
                       p.V, ncols, cols, max_runs);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        PackTileParams tp;
        tp.num_vecs = p.num_vecs;
        tp.V = V;
        tp.max_runs = max_runs;
        tp.bitmask = p.bitmask;
        tp.offsets = p.offsets;
        tp.stream_bytes = p.stream_bytes;
        tp.err = p.err;
        const size_t lds = pack_tile_lds(V, W);
        int64_t per_cu = (int64_t)(160 * 1024) / (int64_t)lds;
        if (per_cu > 8) per_cu = 8;
        if (per_cu < 1) per_cu = 1;
        const int64_t ntiles = (p.num_vecs + kWave - 1) / kWave;
        int64_t grid = (int64_t)dev.num_cus * per_cu;
        if (grid > ntiles) grid = ntiles;
        if (condensed)
            hipLaunchKernelGGL((pack_tile_condensed_kernel<T>), dim3((unsigned)grid), dim3(kWave), lds, stream, tp,
                               (const T *)p.in, (T *)p.out, (const PackCol<T> *)cols);
        else
            hipLaunchKernelGGL((pack_tile_kernel<T>), dim3((unsigned)grid), dim3(kWave), lds, stream, tp, (T *)p.out,
                               (const PackCol<T> *)cols);
        e = hipGetLastError();
    }
    IBP_HIP(hipFreeAsync(buf, stream));
    if (e != hipSuccess) return set_error(IBP_ERR_HIP, "pack kernel launch failed: %s", hipGetErrorString(e));
    return IBP_OK;
}

template <typename T, typename IdxT, bool CONDENSED>
int pack_typed(PackParams p, hipStream_t stream)
{
    const DeviceInfo &dev = device_info();
    const int64_t buf_bytes = slot_buf_bytes(p.V, (int)sizeof(T));
    const int64_t lds_budget = 64 * 1024;
    const int64_t ngroups = (p.n + 31) / 32;
    if constexpr (sizeof(T) >= 4 && !CONDENSED) {
        // in place, no index, one mask, rows the LDS tiles take, six tiles
        // per CU at least (the pack hides its latency with resident waves:
        // 512-byte rows, four tiles per CU, pack slower than with a wave per
        // row): a lane per row
        if (p.in == (const void *)p.out && p.index == nullptr && p.clusters == nullptr && p.bitmask != nullptr &&
            p.n == p.num_vecs && p.n > 0 && p.V <= kWave && tile_rows_ok(p.V * (int64_t)sizeof(T)) &&
            pack_tile_lds(p.V, (int)sizeof(T)) * 6 <= 160 * 1024 && ((uintptr_t)p.in % 16) == 0)
            return pack_tile_launch<T>(p, stream);
    }
    if constexpr (sizeof(T) >= 4 && CONDENSED) {
        // the condensed stream of a whole table with one mask, same tiles
        if (p.index == nullptr && p.clusters == nullptr && p.n == p.num_vecs && p.n > 0 && p.V <= kWave &&
            tile_rows_ok(p.V * (int64_t)sizeof(T)) && pack_tile_lds(p.V, (int)sizeof(T)) * 6 <= 160 * 1024 &&
            ((uintptr_t)p.in % 16) == 0)
            return pack_tile_launch<T>(p, stream, true);
    }
    if (ngroups < (int64_t)dev.num_cus * 8) {
        // Too few 32-row groups for a wave each: the waves of a block share a
        // group, with slot buffers in LDS when four or more fit one CU's LDS
        // (as many waves as fit, up to 16), else 16 waves with buffers in
        // device memory.
        const int64_t lds_cu = 160 * 1024;
        const bool lds = buf_bytes * 4 <= lds_cu;
        int wpb = lds ? (int)((buf_bytes * 8 <= lds_budget ? lds_budget : lds_cu) / buf_bytes) : 16;
        if (wpb > 16) wpb = 16;
        p.split = wpb;
        const int64_t grid = ngroups > 0 ? ngroups : 1;
        if (lds) {
            p.gbuf = nullptr;
            hipLaunchKernelGGL((pack_kernel<T, IdxT, true, CONDENSED>), dim3((unsigned)grid), dim3(wpb * kWave),
                               (size_t)(wpb * buf_bytes), stream, p);
            IBP_LAUNCH_CHECK();
            return IBP_OK;
        }
        IBP_HIP(hipMallocAsync(&p.gbuf, (size_t)(grid * wpb * buf_bytes), stream));
        hipLaunchKernelGGL((pack_kernel<T, IdxT, false, CONDENSED>), dim3((unsigned)grid), dim3(wpb * kWave), 0,
                           stream, p);
        hipError_t e = hipGetLastError();
        IBP_HIP(hipFreeAsync(p.gbuf, stream));
        if (e != hipSuccess) return set_error(IBP_ERR_HIP, "pack kernel launch failed: %s", hipGetErrorString(e));
        return IBP_OK;
    }
    if constexpr (sizeof(T) >= 4) {
        if (p.V <= kWave && p.clusters == nullptr) {
            // rows of <= 64 elements with one mask: batched loads, 8 blocks of
            // four waves per CU
            int64_t grid = (ngroups + 3) / 4;
            if (grid > (int64_t)dev.num_cus * 8) grid = (int64_t)dev.num_cus * 8;
            if (grid < 1) grid = 1;
            p.gbuf = nullptr;
            hipLaunchKernelGGL((pack_small_kernel<T, IdxT, CONDENSED>), dim3((unsigned)grid), dim3(4 * kWave),
                               (size_t)(4 * buf_bytes), stream, p);
            IBP_LAUNCH_CHECK();
            return IBP_OK;
        }
    }
    if (buf_bytes <= lds_budget) {
        int wpb = (int)(lds_budget / buf_bytes);
        if (wpb > 4) wpb = 4;
        // Up to 64 blocks per CU: a packed row is a long chain of scan, LDS
        // and fence steps, so more resident waves hide more of it (GNN table
        // 0.845 -> 0.816 ms, largest DLRM table 3.51 -> 3.23 ms against 16).
        int64_t grid = (ngroups + wpb - 1) / wpb;
        if (grid > (int64_t)dev.num_cus * 64) grid = (int64_t)dev.num_cus * 64;
        if (grid < 1) grid = 1;
        p.gbuf = nullptr;
        hipLaunchKernelGGL((pack_kernel<T, IdxT, true, CONDENSED>), dim3((unsigned)grid), dim3(wpb * kWave),
                           (size_t)(wpb * buf_bytes), stream, p);
        IBP_LAUNCH_CHECK();
        return IBP_OK;
    }
    // Rows beyond the LDS budget: per-wave slot buffers in device memory.
    int64_t nwaves = ngroups < (int64_t)dev.num_cus * 4 ? ngroups : (int64_t)dev.num_cus * 4;
    if (nwaves < 1) nwaves = 1;
    const int64_t grid = (nwaves + 3) / 4;
    IBP_HIP(hipMallocAsync(&p.gbuf, (size_t)(grid * 4 * buf_bytes), stream));
    hipLaunchKernelGGL((pack_kernel<T, IdxT, false, CONDENSED>), dim3((unsigned)grid), dim3(256), 0, stream, p);
    hipError_t e = hipGetLastError();
    IBP_HIP(hipFreeAsync(p.gbuf, stream));
    if (e != hipSuccess) return set_error(IBP_ERR_HIP, "pack kernel launch failed: %s", hipGetErrorString(e));
    return IBP_OK;
}

template <typename T>
int pack_width(const PackParams &p, int index_bytes, hipStream_t s)
{
    const bool i32 = p.index && index_bytes == 4;
    if (p.offsets)
        return i32 ? pack_typed<T, int32_t, true>(p, s) : pack_typed<T, int64_t, true>(p, s);
    return i32 ? pack_typed<T, int32_t, false>(p, s) : pack_typed<T, int64_t, false>(p, s);
}

bool ranges_overlap(const void *a, int64_t an, const void *b, int64_t bn)
{
    const uintptr_t a0 = (uintptr_t)a, b0 = (uintptr_t)b;
    return a0 < b0 + (uintptr_t)bn && b0 < a0 + (uintptr_t)an;
}

} // namespace

} // namespace ibp

using namespace ibp;

#define IBP_CHECK_WIDTH(w)                                                                                  \
    IBP_CHECK((w) == 1 || (w) == 2 || (w) == 4 || (w) == 8, "Input tensor must be 1, 2, 4 or 8-byte datatype (got %d)", (w))

extern "C" int ibp_get_compress_size_ex(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                                        const void *mask, const void *bitval, const void *index, int index_bytes,
                                        int64_t n, int64_t *sizes, uint64_t *total, const ibp_layout *layout,
                                        void *stream)
{
    IBP_TRACE("ibp_get_compress_size_ex");
    IBP_CHECK_WIDTH(elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0 && n >= 0, "invalid shape");
    IBP_CHECK(index == nullptr || index_bytes == 4 || index_bytes == 8, "index must be int32 or int64");
    IBP_CHECK(index != nullptr || n <= num_vecs, "without an index n must not exceed num_vecs");
    IBP_CHECK(layout == nullptr || layout->offsets == nullptr, "sizes take no condensed offsets");
    const int32_t *clusters = layout ? layout->clusters : nullptr;
    const int64_t k = layout && clusters ? layout->num_clusters : 1;
    IBP_CHECK(k >= 1, "num_clusters must be >= 1");
    if (n == 0) return IBP_OK;
    IBP_CHECK(data && mask && bitval && sizes, "null pointer argument");
    return size_any(elem_bytes, index_bytes, data, num_vecs, vec_size, mask, bitval, clusters, k, index, n, sizes,
                    total, as_stream(stream));
}

extern "C" int ibp_get_compress_size(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                                     const void *mask, const void *bitval, const void *index, int index_bytes,
                                     int64_t n, int64_t *sizes, uint64_t *total, void *stream)
{
    return ibp_get_compress_size_ex(elem_bytes, data, num_vecs, vec_size, mask, bitval, index, index_bytes, n,
                                    sizes, total, nullptr, stream);
}

extern "C" int ibp_condensed_offsets(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                                     const void *mask, const void *bitval, const void *index, int index_bytes,
                                     int64_t n, int64_t *offsets, const ibp_layout *layout, void *stream)
{
    IBP_TRACE("ibp_condensed_offsets");
    IBP_CHECK_WIDTH(elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0 && n >= 0, "invalid shape");
    IBP_CHECK(index == nullptr || index_bytes == 4 || index_bytes == 8, "index must be int32 or int64");
    IBP_CHECK(index != nullptr || n <= num_vecs, "without an index n must not exceed num_vecs");
    IBP_CHECK(layout == nullptr || layout->offsets == nullptr, "layout->offsets is the output here: pass it as offsets");
    const int32_t *clusters = layout ? layout->clusters : nullptr;
    const int64_t k = layout && clusters ? layout->num_clusters : 1;
    IBP_CHECK(k >= 1, "num_clusters must be >= 1");
    if (n == 0) return IBP_OK;
    IBP_CHECK(data && mask && bitval && offsets, "null pointer argument");
    hipStream_t s = as_stream(stream);
    // Sizes are written and scanned in device memory; host (pinned) offsets
    // receive one copy at the end.
    const bool host = pointer_is_host(offsets);
    int64_t *dev = offsets;
    if (host) IBP_HIP(hipMallocAsync((void **)&dev, (size_t)n * sizeof(int64_t), s));
    int st = size_any(elem_bytes, index_bytes, data, num_vecs, vec_size, mask, bitval, clusters, k, index, n, dev,
                      nullptr, s);
    if (st == IBP_OK) st = scan_inclusive(dev, n, s);
    if (host) {
        if (st == IBP_OK) IBP_HIP(hipMemcpyAsync(offsets, dev, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        IBP_HIP(hipFreeAsync(dev, s));
    }
    return st;
}

extern "C" int ibp_condensed_offsets_noalloc(int elem_bytes, const void *data, int64_t num_vecs,
                                             int64_t vec_size, const void *mask, const void *bitval,
                                             const void *index, int index_bytes, int64_t n, int64_t *offsets,
                                             const ibp_layout *layout, void *stream)
{
    IBP_TRACE("ibp_condensed_offsets_noalloc");
    IBP_CHECK_WIDTH(elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0 && n >= 0, "invalid shape");
    IBP_CHECK(index == nullptr || index_bytes == 4 || index_bytes == 8, "index must be int32 or int64");
    IBP_CHECK(index != nullptr || n <= num_vecs, "without an index n must not exceed num_vecs");
    IBP_CHECK(layout == nullptr || layout->offsets == nullptr, "layout->offsets is the output here: pass it as offsets");
    IBP_CHECK(n <= kNoallocScanTiles * kScanTile,
              "n = %lld rows exceeds the allocation-free limit of %lld rows (use ibp_condensed_offsets)",
              (long long)n, (long long)(kNoallocScanTiles * kScanTile));
    const int32_t *clusters = layout ? layout->clusters : nullptr;
    const int64_t k = layout && clusters ? layout->num_clusters : 1;
    IBP_CHECK(k >= 1, "num_clusters must be >= 1");
    // Resolved before any early return, so the first call on a device sets up
    // its scratch address whatever its size.
    int64_t *sums = noalloc_scratch();
    if (!sums) return IBP_ERR_HIP;
    if (n == 0) return IBP_OK;
    IBP_CHECK(data && mask && bitval && offsets, "null pointer argument");
    IBP_CHECK(!pointer_is_host(offsets),
              "offsets must be device memory here (host offsets need a staging buffer: use ibp_condensed_offsets)");
    hipStream_t s = as_stream(stream);
    int st = size_any(elem_bytes, index_bytes, data, num_vecs, vec_size, mask, bitval, clusters, k, index, n,
                      offsets, nullptr, s);
    if (st == IBP_OK) st = scan_inclusive(offsets, n, s, sums);
    return st;
}

extern "C" int ibp_compress_ex(int elem_bytes, const void *input, void *output, int64_t num_vecs, int64_t vec_size,
                               const void *mask, const void *bitval, const void *index, int index_bytes, int64_t n,
                               int32_t *bitmask, const ibp_layout *layout, void *stream)
{
    IBP_TRACE("ibp_compress_ex");
    IBP_CHECK_WIDTH(elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0 && n >= 0, "invalid shape");
    IBP_CHECK(index == nullptr || index_bytes == 4 || index_bytes == 8, "index must be int32 or int64");
    IBP_CHECK(index != nullptr || n <= num_vecs, "without an index n must not exceed num_vecs");
    const bool condensed = layout && layout->offsets;
    const int32_t *clusters = layout ? layout->clusters : nullptr;
    const int64_t k = clusters ? layout->num_clusters : 1;
    IBP_CHECK(k >= 1, "num_clusters must be >= 1");
    IBP_CHECK(condensed || index == nullptr || input != output,
              "an index with in-place output aliases rows (unsupported)");
    if (condensed) {
        IBP_CHECK(layout->stream_bytes >= 0, "stream_bytes must be >= 0");
        IBP_CHECK(!ranges_overlap(output, layout->stream_bytes, input, num_vecs * vec_size * elem_bytes),
                  "the condensed stream must not overlap the input");
    }
    if (n == 0) return IBP_OK;
    IBP_CHECK(input && output && mask && bitval && (bitmask || condensed), "null pointer argument");
    PackParams p;
    p.in = input;
    p.out = output;
    p.num_vecs = num_vecs;
    p.V = vec_size;
    p.mask = mask;
    p.bitval = bitval;
    p.clusters = clusters;
    p.num_clusters = k;
    p.offsets = condensed ? layout->offsets : nullptr;
    p.stream_bytes = condensed ? layout->stream_bytes : 0;
    p.index = index;
    p.n = n;
    p.bitmask = (uint32_t *)bitmask;
    p.gbuf = nullptr;
    p.err = (index || clusters || condensed) ? index_error_counter() : nullptr;
    hipStream_t s = as_stream(stream);
    switch (elem_bytes) {
    case 1: return pack_width<uint8_t>(p, index_bytes, s);
    case 2: return pack_width<uint16_t>(p, index_bytes, s);
    case 4: return pack_width<uint32_t>(p, index_bytes, s);
    default: return pack_width<uint64_t>(p, index_bytes, s);
    }
}

extern "C" int ibp_compress(int elem_bytes, const void *input, void *output, int64_t num_vecs, int64_t vec_size,
                            const void *mask, const void *bitval, const void *index, int index_bytes, int64_t n,
                            int32_t *bitmask, void *stream)
{
    return ibp_compress_ex(elem_bytes, input, output, num_vecs, vec_size, mask, bitval, index, index_bytes, n,
                           bitmask, nullptr, stream);
}
