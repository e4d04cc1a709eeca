// ibp_fetch_dev.h -- device routines that fetch and unpack rows of the IBP
// slot format.  Shared by the plain fetch kernels (ibp_fetch_launch.h) and the
// batched kernels (ibp_batch.hip).
//
// Slot format (reference: ibp_compress_dev.cuh:13-15,61-92; decode
// ibp_decompress_dev.cuh:279-336): V words of T.  A packed slot starts with
// Hw = align_w(ceil(V/8))/w header words whose bit j (LSB first, word j/B)
// flags element j; the payload that follows is an MSB-first bit stream of
// B - popc(mask[j]) free bits per flagged element and B raw bits otherwise;
// the rest of the slot is zero.  A row whose bitmask bit is clear is stored raw.
//
// Row layouts (template parameter LAYOUT, bit flags):
//  * slots (0): stored row s is the V-element slot at element s*V; packed
//    when bitmask bit s is set (compress_inplace, src/compress.cu:176-236);
//  * condensed (kLayoutCondensed): stored row s is bytes [offsets[s-1],
//    offsets[s]) of a stream (offsets[-1] = 0), the slot cut at its size;
//    packed when that length is not V*w (compress_condensed,
//    src/compress.cu:238-313, ibp_compress_kernel.cuh:59-94);
//  * clustered (kLayoutClustered): row s is coded with mask/bitval row
//    clusters[s] of a [K x V] table (the reference applies row 0 to every
//    row, src/compress.cu:221-232).
//
// Mapping: one wave64 per row.  Lane j owns element j of each 64-element chunk.
// For w = 8 the 64-bit header word c holds exactly the flags of chunk c.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "ibp_bits.h"
#include "ibp_wave.h"

namespace ibp {

enum : int { kLayoutSlots = 0, kLayoutCondensed = 1, kLayoutClustered = 2 };

struct FetchParams {
    void *out = nullptr;
    const void *comp = nullptr;
    const void *mask = nullptr;
    const void *bitval = nullptr;
    const uint32_t *bitmask = nullptr; // slots layout
    const void *index = nullptr;       // nullptr: identity
    int64_t num_vecs = 0;              // stored rows
    int64_t vec_size = 0;
    int64_t n = 0;
    int32_t first_words = 0; // words read before the header is known (small rows)
    int32_t pad = 0;
    unsigned long long *err = nullptr;
    const int64_t *offsets = nullptr;  // condensed layout
    const int32_t *clusters = nullptr; // clustered layout
    int64_t stream_bytes = 0;          // condensed: bytes of the stream (reads stay inside)
    int64_t num_clusters = 1;          // clustered: rows of mask/bitval
};

// Output stores.  Rows that are a whole number of 128-byte lines are written
// with non-temporal stores: nothing is written twice and the outputs do not
// evict hot compressed rows from L2/MALL (DLRM, 512-byte rows: 0.385 vs
// 0.466 ms per batch).  Other rows share a line with their neighbour; there
// ordinary stores let L2 merge the halves, where non-temporal ones would write
// the line twice (GNN minibatch, 400-byte rows: 6% faster).
template <typename T, bool NT>
__device__ __forceinline__ void store_row_word(T *p, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Whether rows of this table take non-temporal stores (see store_row_word).
__device__ __forceinline__ bool rows_fill_lines(int64_t vec_size, int elem_bytes)
{
    return ((vec_size * elem_bytes) & 127) == 0;
}

template <typename IdxT>
__device__ __forceinline__ int64_t load_index(const void *index, int64_t i)
{
    return index ? (int64_t)((const IdxT *)index)[i] : i;
}

// Column constants of lane j for rows of at most 64 elements.  The zero runs
// of the column mask are tabulated once per kernel: a flagged element's free
// bits (top-aligned in v) land in run k as (v >> shift[k]) & run[k], where
// shift[k] counts the mask bits above run k.  Masks with more than kRunCap
// runs in some lane of the wave fall back to the generic deposit loop.
constexpr int kRunCap = 4;

template <typename T>
struct SmallCols {
    T bitval;
    T free; // ~mask: positions carried in the payload of a flagged element
    int len_flagged; // B - popc(mask)
    bool valid;      // j < V
    int wave_runs;   // max zero runs over the wave's lanes (uniform)
    int shift[kRunCap];
    T run[kRunCap];
};

template <typename T>
__device__ __forceinline__ SmallCols<T> load_small_cols(const void *mask, const void *bitval, int V)
{
    const int lane = lane_id();
    SmallCols<T> c;
    c.valid = lane < V;
    const T m = c.valid ? ((const T *)mask)[lane] : T(0);
    c.bitval = c.valid ? ((const T *)bitval)[lane] : T(0);
    c.free = c.valid ? T(~m) : T(0);
    c.len_flagged = Bits<T>::B - popc(m);
    T z = c.free;
    int filled = 0, runs = 0;
#pragma unroll
    for (int k = 0; k < kRunCap; ++k) {
        c.shift[k] = 0;
        c.run[k] = T(0);
    }
    while (z) {
        const int sft = clz(z);
        const int r = clz(T(~T(z << sft)));
        const T run = shr(himask<T>(r), sft);
#pragma unroll
        for (int k = 0; k < kRunCap; ++k)
            if (k == runs) {
                c.shift[k] = sft - filled;
                c.run[k] = run;
            }
        filled += r;
        z &= ~run;
        ++runs;
    }
    int wr = 0;
    for (int k = 1; k <= kRunCap + 1; ++k)
        if (ballot(runs >= k)) wr = k;
    c.wave_runs = wr;
    return c;
}

// Free bits of a flagged element (top-aligned in v) dropped into the zero
// positions of the column mask.  RUNS is the wave's run count (lanes with
// fewer runs have zero entries), or -1 for the generic loop.
template <typename T, int RUNS>
__device__ __forceinline__ T small_deposit(const SmallCols<T> &c, T v)
{
    if constexpr (RUNS < 0) {
        return deposit_msb<T>(v, c.free);
    } else {
        T d = T(0);
#pragma unroll
        for (int k = 0; k < RUNS; ++k) d |= T(v >> c.shift[k]) & c.run[k];
        return d;
    }
}

// Calls f(std::integral_constant<int, RUNS>, std::bool_constant<NT>) for the
// wave's run count and the table's store policy, so the row loops are
// compiled once per combination with no per-row selection.
template <typename T, typename F>
__device__ __forceinline__ void with_runs(const SmallCols<T> &c, bool nt, F &&f)
{
    auto go = [&](auto runs) {
        if (nt)
            f(runs, std::true_type());
        else
            f(runs, std::false_type());
    };
    switch (c.wave_runs) {
    case 0: go(std::integral_constant<int, 0>()); break;
    case 1: go(std::integral_constant<int, 1>()); break;
    case 2: go(std::integral_constant<int, 2>()); break;
    case 3: go(std::integral_constant<int, 3>()); break;
    case 4: go(std::integral_constant<int, 4>()); break;
    default: go(std::integral_constant<int, -1>()); break;
    }
}

// Header flag of lane `lane` for a row of <= 64 elements, given the first
// slot words held one per lane in `w`.
template <typename T>
__device__ __forceinline__ bool small_flag(T w, int lane)
{
    if constexpr (sizeof(T) == 8) {
        const uint64_t h = readlane((uint64_t)w, 0);
        return (h >> lane) & 1ull;
    } else if constexpr (sizeof(T) == 4) {
        const uint32_t h0 = readlane((uint32_t)w, 0);
        const uint32_t h1 = readlane((uint32_t)w, 1);
        return ((lane < 32 ? h0 : h1) >> (lane & 31)) & 1u;
    } else { // 8- or 16-bit words: flag j is bit j % B of header word j / B
        const T h = shfl(w, lane / Bits<T>::B);
        return (h >> (lane % Bits<T>::B)) & 1;
    }
}

// Per-row metadata of the condensed and clustered layouts, as loaded (raw):
// decoded only where used, so the loads of the next group stay in flight.
struct RowMetaRaw {
    int64_t lo, hi; // condensed: byte range of the stored row
    int32_t cl;     // clustered: mask row
    uint32_t bm;    // slots: bitmask word of the row
};

template <int LAYOUT>
__device__ __forceinline__ RowMetaRaw load_meta(const FetchParams &p, int64_t s, bool live)
{
    RowMetaRaw m{0, 0, 0, 0u};
    if (live && s >= 0 && s < p.num_vecs) {
        if constexpr ((LAYOUT & kLayoutCondensed) != 0) {
            m.lo = s > 0 ? p.offsets[s - 1] : 0;
            m.hi = p.offsets[s];
        } else {
            m.bm = p.bitmask[s >> 5];
        }
        if constexpr ((LAYOUT & kLayoutClustered) != 0) m.cl = p.clusters[s];
    }
    return m;
}

struct RowMeta {
    int64_t word; // first element of the stored row in comp
    int64_t len;  // elements stored (V for a raw row)
    int32_t cl;   // mask row
    bool ok;      // index in range, byte range inside the stream, cluster in range
    bool packed;
};

template <typename T, int LAYOUT>
__device__ __forceinline__ RowMeta decode_meta(const FetchParams &p, int64_t s, bool live, const RowMetaRaw &r)
{
    constexpr int64_t W = (int64_t)sizeof(T);
    const int64_t V = p.vec_size;
    RowMeta m{0, 0, 0, false, false};
    bool ok = live && s >= 0 && s < p.num_vecs;
    if constexpr ((LAYOUT & kLayoutCondensed) != 0) {
        const int64_t len = r.hi - r.lo;
        ok = ok && r.lo >= 0 && len >= 0 && r.hi <= p.stream_bytes && len <= V * W && (r.lo % W) == 0 &&
             (len % W) == 0;
        m.word = r.lo / W;
        m.len = len / W;
        m.packed = len != V * W;
    } else {
        m.word = s * V;
        m.len = V;
        m.packed = (r.bm >> (s & 31)) & 1u;
    }
    if constexpr ((LAYOUT & kLayoutClustered) != 0) {
        ok = ok && r.cl >= 0 && r.cl < p.num_clusters;
        m.cl = r.cl;
    }
    m.ok = ok;
    if (!ok) {
        m.word = 0;
        m.len = 0;
        m.packed = false;
    }
    return m;
}

// Stages 2 and 3 of fetch_small_range_r for one group of rows whose slot
// prefixes are in w: header -> flags -> bit offsets -> words needed, the rest
// of the rows read, then unpack and store.  ALL_PACKED: every row of the
// group is valid and packed, so nothing branches per row.
template <typename T, int ROWS, int RUNS, bool NT, bool ALL_PACKED>
__device__ __forceinline__ void small_unpack_rows(const FetchParams &p, const SmallCols<T> &col, T (&w)[ROWS],
                                                  int64_t cur, int64_t base, int nrows, uint64_t okmask,
                                                  uint64_t pkmask, int len_flag, int len_raw, int Hw, int E,
                                                  const T *comp, T *out)
{
    constexpr unsigned B = (unsigned)Bits<T>::B;
    constexpr unsigned LOGB = sizeof(T) == 8 ? 6 : sizeof(T) == 4 ? 5 : sizeof(T) == 2 ? 4 : 3;
    const int lane = lane_id();
    const int V = (int)p.vec_size;
    const unsigned hw_addr = (unsigned)Hw << 2;

    unsigned pos[ROWS];
    bool flag[ROWS];
    int need[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        flag[r] = false;
        pos[r] = 0;
        need[r] = (ALL_PACKED || ((okmask >> r) & 1ull)) ? V : 0;
        if (ALL_PACKED || ((pkmask >> r) & 1ull)) { // wave-uniform
            flag[r] = small_flag<T>(w[r], lane) && col.valid;
            const int L = flag[r] ? len_flag : len_raw;
            const int incl = wave_incl_scan(L);
            const unsigned total = (unsigned)readlane(incl, kWave - 1);
            pos[r] = (unsigned)(incl - L);
            // Clamped to the slot: a corrupt header cannot make the read leave the row.
            const int nw = Hw + (int)((total + B - 1) >> LOGB);
            need[r] = nw < V ? nw : V;
        }
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
        if (need[r] > E) { // uniform; rare for packed rows
            const T *row = comp + readlane(cur, r) * V;
            if (lane >= E && lane < need[r]) w[r] = row[lane];
        }
    // Every load of the group (and the index prefetch) lands here, once:
    // the stores below then wait on nothing (vmcnt counts stores too).
    __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0) expcnt(7) lgkmcnt(15)

    // Unpack and store.  Payload bits never pass the slot's last word, so
    // q + 1 needs no clamp (ds_bpermute stays inside the wave).
    T *orow = out + base * V;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (ALL_PACKED || r < nrows) {
            T val = T(0);
            if (ALL_PACKED || ((okmask >> r) & 1ull)) {
                val = w[r];
                if (ALL_PACKED || ((pkmask >> r) & 1ull)) {
                    const unsigned a0 = hw_addr + ((pos[r] >> LOGB) << 2);
                    const int o = (int)(pos[r] & (B - 1));
                    const T w0 = shfl(w[r], (int)(a0 >> 2));
                    const T w1 = shfl(w[r], (int)(a0 >> 2) + 1);
                    const T v = stream_bits<T>(w0, w1, o);
                    val = flag[r] ? T(col.bitval | small_deposit<T, RUNS>(col, v)) : v;
                }
            }
            if (lane < V) store_row_word<T, NT>(&orow[lane], val);
        }
        orow += V;
    }
}

// Fetches and unpacks rows of a table whose rows have at most 64 elements:
// groups of ROWS consecutive output rows starting at begin, begin + step, ...
// below end.  Per group, every row's slot read is issued at once (a prefix of
// first_words words, or the whole raw row) in parallel with the bitmask words,
// and the next group's indices are prefetched, so a wave keeps ROWS rows and
// one index load in flight.  Row pointers are wave-uniform (scalar base +
// lane offset addressing) and bit positions unsigned.  Wave-uniform.
template <typename T, typename IdxT, int ROWS, int RUNS, bool NT>
__device__ __forceinline__ void fetch_small_range_r(const FetchParams &p, const SmallCols<T> &col, int64_t begin,
                                                    int64_t end, int64_t step)
{
    const int lane = lane_id();
    const int V = (int)p.vec_size;
    const int Hw = (int)(header_bytes(V, (int)sizeof(T)) / (int)sizeof(T));
    const int E = p.first_words;
    const T *comp = (const T *)p.comp;
    T *out = (T *)p.out;
    const bool has_index = p.index != nullptr;
    // Per-lane payload length of a flagged and of a raw element (zero past the row).
    const int len_flag = col.valid ? col.len_flagged : 0;
    const int len_raw = col.valid ? Bits<T>::B : 0;

    int64_t nxt = 0;
    if (has_index && lane < ROWS && begin + lane < end) nxt = load_index<IdxT>(p.index, begin + lane);
    for (int64_t base = begin; base < end; base += step) {
        // Rows of this group (uniform): the last group may be partial.
        const int nrows = end - base < ROWS ? (int)(end - base) : ROWS;
        const bool live = lane < nrows;
        const int64_t cur = has_index ? nxt : base + lane;
        const bool my_ok = live && cur >= 0 && cur < p.num_vecs;
        const uint64_t okmask = ballot(my_ok);

        // Stage 1: slot prefixes of every row, bitmask words, next indices.
        T w[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            w[r] = T(0);
            if ((okmask >> r) & 1ull) {
                const T *row = comp + readlane(cur, r) * V;
                if (lane < E) w[r] = row[lane];
            }
        }
        uint32_t my_bm = 0;
        if (my_ok) my_bm = p.bitmask[cur >> 5];
        if (has_index) {
            const int64_t nb = base + step;
            nxt = 0;
            if (lane < ROWS && nb + lane < end) nxt = load_index<IdxT>(p.index, nb + lane);
        }
        const uint64_t pkmask = ballot(my_ok && ((my_bm >> (cur & 31)) & 1u));

        // Stages 2 and 3.  A full group whose rows are all valid and packed
        // (the common case) runs without per-row branches, so the rows' scans
        // and exchanges interleave; any other group takes the general form.
        const uint64_t all = nrows == ROWS ? ((1ull << ROWS) - 1) : 0ull;
        if (all && (okmask & pkmask & all) == all)
            small_unpack_rows<T, ROWS, RUNS, NT, true>(p, col, w, cur, base, ROWS, okmask, pkmask, len_flag, len_raw, Hw, E,
                                                   comp, out);
        else
            small_unpack_rows<T, ROWS, RUNS, NT, false>(p, col, w, cur, base, nrows, okmask, pkmask, len_flag, len_raw, Hw,
                                                    E, comp, out);
        // The ballot needs the whole wave active, so it is taken before the lane-0 branch.
        const uint64_t livemask = ballot(live);
        if (p.err && lane == 0) {
            const int bad = __builtin_popcountll(livemask & ~okmask);
            if (bad) atomicAdd(p.err, (unsigned long long)bad);
        }
    }
}

template <typename T, typename IdxT, int ROWS>
__device__ __forceinline__ void fetch_small_range(const FetchParams &p, const SmallCols<T> &col, int64_t begin,
                                                  int64_t end, int64_t step)
{
    with_runs<T>(col, rows_fill_lines(p.vec_size, (int)sizeof(T)), [&](auto runs, auto nt) {
        fetch_small_range_r<T, IdxT, ROWS, decltype(runs)::value, decltype(nt)::value>(p, col, begin, end, step);
    });
}

// fetch_small_range for the condensed and clustered layouts.  A row's
// metadata (its byte range, its mask row) depends on its index, so indices
// are loaded two groups ahead and metadata one group ahead; with a mask row
// per row the column constants are loaded per row and the generic deposit is
// used.  Rows whose index, byte range or mask row is invalid give zero rows
// and are counted.  Wave-uniform.
template <typename T, typename IdxT, int ROWS, int LAYOUT, bool NT, int RUNS>
__device__ __forceinline__ void fetch_small_range_meta(const FetchParams &p, const SmallCols<T> &col,
                                                       int64_t begin, int64_t end, int64_t step)
{
    constexpr unsigned B = (unsigned)Bits<T>::B;
    constexpr unsigned LOGB = sizeof(T) == 8 ? 6 : sizeof(T) == 4 ? 5 : sizeof(T) == 2 ? 4 : 3;
    constexpr bool CLUS = (LAYOUT & kLayoutClustered) != 0;
    const int lane = lane_id();
    const int V = (int)p.vec_size;
    const int Hw = (int)(header_bytes(V, (int)sizeof(T)) / (int)sizeof(T));
    const int E = p.first_words;
    const T *comp = (const T *)p.comp;
    const T *mask = (const T *)p.mask;
    const T *bitval = (const T *)p.bitval;
    T *out = (T *)p.out;
    const bool colv = lane < V;
    const int len_raw = colv ? (int)B : 0;
    const unsigned hw_addr = (unsigned)Hw << 2;

    auto live_at = [&](int64_t b) { return lane < ROWS && b + lane < end; };
    auto id_at = [&](int64_t b) -> int64_t { return live_at(b) ? load_index<IdxT>(p.index, b + lane) : 0; };

    int64_t cur = id_at(begin);
    RowMetaRaw craw = load_meta<LAYOUT>(p, cur, live_at(begin));
    int64_t nxt = id_at(begin + step);
    for (int64_t base = begin; base < end; base += step) {
        const int nrows = end - base < ROWS ? (int)(end - base) : ROWS; // uniform
        const bool live = lane < nrows;
        const RowMeta me = decode_meta<T, LAYOUT>(p, cur, live, craw);
        const uint64_t okmask = ballot(me.ok);
        const uint64_t pkmask = ballot(me.ok && me.packed);

        // Stage 1: stored-row prefixes (and mask rows), the next group's
        // metadata, the index of the group after it.
        T w[ROWS], cm[ROWS], cb[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            w[r] = cm[r] = cb[r] = T(0);
            if ((okmask >> r) & 1ull) { // uniform
                const T *row = comp + readlane(me.word, r);
                const int len = (int)readlane(me.len, r);
                if (lane < (E < len ? E : len)) w[r] = row[lane];
                if constexpr (CLUS) {
                    const int64_t mrow = (int64_t)readlane(me.cl, r) * V;
                    if (colv) {
                        cm[r] = mask[mrow + lane];
                        cb[r] = bitval[mrow + lane];
                    }
                }
            }
        }
        const RowMetaRaw nraw = load_meta<LAYOUT>(p, nxt, live_at(base + step));
        const int64_t nxt2 = id_at(base + 2 * step);

        // Stage 2: header -> flags -> bit offsets -> words needed; read the rest.
        unsigned pos[ROWS];
        bool flag[ROWS];
        int need[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            flag[r] = false;
            pos[r] = 0;
            need[r] = ((okmask >> r) & 1ull) ? (int)readlane(me.len, r) : 0;
            if ((pkmask >> r) & 1ull) { // uniform
                flag[r] = small_flag<T>(w[r], lane) && colv;
                int lf;
                if constexpr (CLUS)
                    lf = (int)B - popc(cm[r]);
                else
                    lf = col.len_flagged;
                const int L = flag[r] ? lf : len_raw;
                const int incl = wave_incl_scan(L);
                const unsigned total = (unsigned)readlane(incl, kWave - 1);
                pos[r] = (unsigned)(incl - L);
                const int nw = Hw + (int)((total + B - 1) >> LOGB);
                need[r] = nw < need[r] ? nw : need[r]; // never past the stored row
            }
        }
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
            if (need[r] > E) { // uniform
                const T *row = comp + readlane(me.word, r);
                if (lane >= E && lane < need[r]) w[r] = row[lane];
            }
        // Every load of the group lands here, once: the stores below wait on
        // nothing (vmcnt counts stores too).  The next group's metadata and
        // index loads are issued before this point and land here as well.
        __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0) expcnt(7) lgkmcnt(15)

        // Stage 3: unpack and store.
        T *orow = out + base * V;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            if (r < nrows) {
                T val = T(0);
                if ((okmask >> r) & 1ull) {
                    val = w[r];
                    if ((pkmask >> r) & 1ull) {
                        const unsigned a0 = hw_addr + ((pos[r] >> LOGB) << 2);
                        const int o = (int)(pos[r] & (B - 1));
                        const T w0 = shfl(w[r], (int)(a0 >> 2));
                        const T w1 = shfl(w[r], (int)(a0 >> 2) + 1);
                        const T v = stream_bits<T>(w0, w1, o);
                        if constexpr (CLUS)
                            val = flag[r] ? T(cb[r] | deposit_msb<T>(v, T(~cm[r]))) : v;
                        else
                            val = flag[r] ? T(col.bitval | small_deposit<T, RUNS>(col, v)) : v;
                    }
                }
                if (lane < V) store_row_word<T, NT>(&orow[lane], val);
            }
            orow += V;
        }
        const uint64_t livemask = ballot(live);
        if (p.err && lane == 0) {
            const int bad = __builtin_popcountll(livemask & ~okmask);
            if (bad) atomicAdd(p.err, (unsigned long long)bad);
        }
        cur = nxt;
        craw = nraw;
        nxt = nxt2;
    }
}

// Fetches and unpacks output row i (stored row s, already loaded) of a table
// with rows of any length, one 64-element chunk at a time.  Wave-uniform.
template <typename T, int LAYOUT>
__device__ __forceinline__ void fetch_large_row_from(const FetchParams &p, int64_t i, int64_t s)
{
    constexpr int B = Bits<T>::B;
    constexpr int FLAGS_PER_WORD = B;
    const int lane = lane_id();
    const int64_t V = p.vec_size;
    const int64_t Hw = header_bytes(V, (int)sizeof(T)) / (int64_t)sizeof(T);
    T *dst = (T *)p.out + i * V;

    const RowMeta me = decode_meta<T, LAYOUT>(p, s, true, load_meta<LAYOUT>(p, s, true));
    if (!me.ok) {
        for (int64_t k = lane; k < V; k += kWave) dst[k] = T(0);
        if (p.err && lane == 0) atomicAdd(p.err, 1ull);
        return;
    }
    const T *slot = (const T *)p.comp + me.word;
    const T *mask = (const T *)p.mask + (int64_t)me.cl * V;
    const T *bitval = (const T *)p.bitval + (int64_t)me.cl * V;
    if (!me.packed) {
        // Raw row: four independent loads in flight per lane.
        for (int64_t k0 = lane; k0 < V; k0 += 4 * kWave) {
            T a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t k = k0 + u * kWave;
                a[u] = k < V ? slot[k] : T(0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t k = k0 + u * kWave;
                if (k < V) store_row_word<T, true>(&dst[k], a[u]);
            }
        }
        return;
    }
    const int64_t hlim = Hw < me.len ? Hw : me.len;
    const T *payload = slot + Hw;
    int64_t hdr_group = -1;
    T hdr = T(0);
    int64_t cursor = 0; // bit position in the payload
    const int64_t nchunks = (V + kWave - 1) / kWave;
    for (int64_t c = 0; c < nchunks; ++c) {
        const int64_t j = c * kWave + lane;
        const bool valid = j < V;
        const T m = valid ? mask[j] : T(0);
        const T b = valid ? bitval[j] : T(0);
        // Header words covering this chunk: word (c*64 + lane) / B.
        const int64_t wlo = (c * kWave) / FLAGS_PER_WORD;
        const int64_t whi = (c * kWave + kWave - 1) / FLAGS_PER_WORD;
        if (whi / kWave != hdr_group) { // header words are staged 64 at a time
            hdr_group = whi / kWave;
            const int64_t hk = hdr_group * kWave + lane;
            hdr = hk < hlim ? slot[hk] : T(0);
        }
        T hw;
        if constexpr (sizeof(T) == 8) {
            hw = readlane(hdr, (int)(wlo % kWave));
        } else if constexpr (sizeof(T) == 4) { // two header words per chunk
            const T h0 = readlane(hdr, (int)(wlo % kWave));
            const T h1 = readlane(hdr, (int)(whi % kWave));
            hw = lane < FLAGS_PER_WORD ? h0 : h1;
        } else { // 4 or 8 header words per chunk, all in the staged group
            hw = shfl(hdr, (int)(((c * kWave + lane) / FLAGS_PER_WORD) % kWave));
        }
        const bool f = valid && ((hw >> (j % FLAGS_PER_WORD)) & 1);
        const int L = valid ? (f ? B - popc(m) : B) : 0;
        const int incl = wave_incl_scan(L);
        const int tot = readlane(incl, kWave - 1);
        const int64_t p_bit = cursor + incl - L;
        const int64_t w0 = cursor / B;
        const int64_t cnt = (cursor + tot + B - 1) / B - w0; // <= 65
        // Reads stay inside the stored row even for a corrupt header.
        const int64_t room = me.len - Hw - w0;
        const T A = (lane < cnt && lane < room) ? payload[w0 + lane] : T(0);
        const T X = (cnt > kWave && kWave < room) ? payload[w0 + kWave] : T(0);
        const int q = (int)(p_bit / B - w0);
        const int o = (int)(p_bit % B);
        const T s0 = shfl(A, q & (kWave - 1));
        const T s1 = shfl(A, (q + 1) & (kWave - 1));
        const T a0 = q < kWave ? s0 : X;
        const T a1 = q + 1 < kWave ? s1 : X;
        const T v = stream_bits<T>(a0, a1, o);
        const T val = f ? T(b | deposit_msb<T>(v, T(~m))) : v;
        if (valid) store_row_word<T, true>(&dst[j], val);
        cursor += tot;
    }
}

// Output rows [b, e), step `step`, of a wide-row table.  The next row's
// index is loaded while the current row is unpacked.
template <typename T, typename IdxT, int LAYOUT = kLayoutSlots>
__device__ __forceinline__ void fetch_large_range(const FetchParams &p, int64_t b, int64_t e, int64_t step)
{
    if (b >= e) return;
    int64_t s = readfirstlane(load_index<IdxT>(p.index, b));
    for (int64_t i = b; i < e; i += step) {
        const int64_t nxt = i + step < e ? readfirstlane(load_index<IdxT>(p.index, i + step)) : 0;
        fetch_large_row_from<T, LAYOUT>(p, i, s);
        s = nxt;
    }
}

} // namespace ibp
