// ibp_kmeans.hip -- bitwise k-means mask selection (preprocess_kmeans).
//
// Replaces src/preprocess.cu:88-139 and ibp::preproc_kmeans
// (include/preproc/ibp_preproc_host.cuh:119-284) with its kernels
// cluster_vecs, compute_new_centroids, construct_bitmasks and create_mask_many
// (include/preproc/ibp_preproc_kmeans.cuh:106-297).
//
// Semantics kept from the reference:
//  * centroid 0 = row 0; centroids 1..min(K,N)-1 = rows at positions 1.. of a
//    libstdc++-style random_shuffle of 0..N-1 driven by rand();
//  * 10 Lloyd iterations of Hamming-distance assignment and majority-bit
//    centroids, where a 32-lane accumulator is never cleared between the
//    columns l, l+32, l+64, ... it visits (ibp_preproc_kmeans.cuh:226-233);
//  * the returned masks are those of the last sweep threshold (1.0f): a bit is
//    in mask c iff it is constant over cluster c; an empty cluster gets all ones;
//  * the threshold argument is ignored.
// Defined here where the reference is undefined: ties in the assignment go to
// the highest centroid index (the reference races on atomicExch); for 8-byte
// elements every thread owns 64 counters (the reference overruns its
// 8192-float shared array); centroids beyond N rows are zero.
#include <stdlib.h>

#include <vector>

#include "ibp_bits.h"
#include "ibp_internal.h"
#include "ibp_wave.h"

namespace ibp {

template <typename T>
using KW = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;

// Centroids of the lane-per-row assignment (K <= kRowsK) are also kept by
// column: cent_cols[j * kRowsK + c] = centroid c's word j (zero for c >= K).
constexpr int kRowsK = 16;

template <typename T>
__global__ void km_gather_kernel(const T *data, int64_t V, const int64_t *rows, int K, T *centroids, T *cent_cols)
{
    const int64_t total = (int64_t)K * V;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = g / V, j = g % V;
        const int64_t r = rows[c];
        const T v = r >= 0 ? data[r * V + j] : T(0);
        centroids[g] = v;
        if (cent_cols) cent_cols[j * kRowsK + c] = v;
    }
    // the by-column copy's centroids past K are zero
    if (cent_cols && K < kRowsK) {
        const int pad = kRowsK - K;
        for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < pad * V;
             g += (int64_t)gridDim.x * blockDim.x)
            cent_cols[(g / pad) * kRowsK + K + g % pad] = T(0);
    }
}

// Nearest centroid by popc(x ^ c); equal distances go to the highest index.
// Blocks of four waves.  SPLIT = 1: a wave per group of kAssignRows rows.
// SPLIT = 4 (rows of at least 256 elements): the four waves share one group
// and wave w takes the columns j = lane + 64 (w + 4k), which gives a few
// thousand long rows four times the waves (their assignment is
// latency-bound).  Each centroid word is loaded once per group and compared
// with the group's elements, 16 centroids at a time: 16 x kAssignRows = 64
// distance accumulators per lane, summed over the wave in one exchange-and-add
// pass (wave_sum64: lane l gets (centroid l / 4, row l % 4)).  With SPLIT = 4
// the four waves' partial sums meet in LDS.  The nearest of the 16 centroids
// for each row is then a lane minimum of (distance, 15 - centroid) over the
// 16 lanes of that row, so lane l tracks row l % 4's best so far.
constexpr int kAssignGroup = 16;
constexpr int kAssignRows = 4;
constexpr int kAssignWaves = 4;
static_assert(kAssignGroup * kAssignRows == 64, "one distance per lane after the wave sum");

// d + popc(v), one v_bcnt_u32_b32 per 32-bit half with d as its addend.
__device__ __forceinline__ int popc_add(uint64_t v, int d)
{
    d += __builtin_popcount((uint32_t)v);
    return d + __builtin_popcount((uint32_t)(v >> 32));
}
__device__ __forceinline__ int popc_add(uint32_t v, int d) { return d + __builtin_popcount(v); }

// The W at byte `off` of `base`.
template <typename W>
__device__ __forceinline__ W load_at(const void *base, uint32_t off)
{
    return *reinterpret_cast<const W *>(static_cast<const char *>(base) + off);
}

template <typename T, int SPLIT>
__global__ __launch_bounds__(kWave * kAssignWaves) void km_assign_kernel(const T *data, int64_t N, int64_t V,
                                                                         const T *centroids, int K, int32_t *cluster)
{
    constexpr int GPB = kAssignWaves / SPLIT; // row groups per block
    __shared__ int part[SPLIT][kWave];
    const int lane = lane_id();
    const int wv = wave_in_block_id();
    const int sl = wv % SPLIT; // this wave's column slice
    // lane l ends with the distance of row rr = l % kAssignRows to centroid gl of the group
    const int gl = lane / kAssignRows;
    for (int64_t i0 = ((int64_t)blockIdx.x * GPB + wv / SPLIT) * kAssignRows; i0 < N;
         i0 += (int64_t)gridDim.x * GPB * kAssignRows) {
        int best = 0, best_d = 0x7fffffff; // of row rr
        // rows past N repeat row i0 (valid memory); their result is dropped
        int64_t row[kAssignRows];
#pragma unroll
        for (int r = 0; r < kAssignRows; ++r) row[r] = i0 + r < N ? i0 + r : i0;
        for (int c0 = 0; c0 < K; c0 += kAssignGroup) {
            int d[kAssignGroup * kAssignRows];
#pragma unroll
            for (int q = 0; q < kAssignGroup * kAssignRows; ++q) d[q] = 0;
            // Addresses are a row's or a centroid's base (the same for the
            // whole wave, so scalar) plus the lane's 32-bit byte offset: the
            // host keeps V * sizeof(T) below 2^24.
            const T *cb = centroids + (int64_t)c0 * V;
            const uint32_t row_bytes = (uint32_t)(V * (int64_t)sizeof(T));
            const uint32_t off0 = (uint32_t)((sl * kWave + lane) * (int)sizeof(T));
            constexpr uint32_t kStep = kWave * SPLIT * sizeof(T);
            if (c0 + kAssignGroup <= K) {
                // a whole group: the loads of a column step are issued before
                // its arithmetic
                for (uint32_t off = off0; off < row_bytes; off += kStep) {
                    KW<T> x[kAssignRows], cw[kAssignGroup];
#pragma unroll
                    for (int r = 0; r < kAssignRows; ++r) x[r] = load_at<KW<T>>(data + row[r] * V, off);
#pragma unroll
                    for (int g = 0; g < kAssignGroup; ++g) cw[g] = load_at<KW<T>>(cb + (int64_t)g * V, off);
#pragma unroll
                    for (int g = 0; g < kAssignGroup; ++g)
#pragma unroll
                        for (int r = 0; r < kAssignRows; ++r)
                            d[g * kAssignRows + r] = popc_add((KW<T>)(x[r] ^ cw[g]), d[g * kAssignRows + r]);
                }
            } else {
                for (uint32_t off = off0; off < row_bytes; off += kStep) {
                    KW<T> x[kAssignRows];
#pragma unroll
                    for (int r = 0; r < kAssignRows; ++r) x[r] = load_at<KW<T>>(data + row[r] * V, off);
#pragma unroll
                    for (int g = 0; g < kAssignGroup; ++g) {
                        if (c0 + g < K) { // uniform
                            const KW<T> cw = load_at<KW<T>>(cb + (int64_t)g * V, off);
#pragma unroll
                            for (int r = 0; r < kAssignRows; ++r)
                                d[g * kAssignRows + r] = popc_add((KW<T>)(x[r] ^ cw), d[g * kAssignRows + r]);
                        }
                    }
                }
            }
            int dist = wave_sum64(d);
            if constexpr (SPLIT > 1) {
                part[sl][lane] = dist;
                __syncthreads();
                dist = 0;
#pragma unroll
                for (int w = 0; w < SPLIT; ++w) dist += part[w][lane];
                __syncthreads(); // the partials are read before the next group's
            }
            // nearest of the group for row rr: minimum of (dist, 15 - g)
            // over the 16 lanes of the row (lane bits 2..5)
            int key = c0 + gl < K ? (dist << 4) | (kAssignGroup - 1 - gl) : 0x7fffffff;
#pragma unroll
            for (int m = kAssignRows; m < kWave; m <<= 1) {
                const int o = (int)shfl((uint32_t)key, lane ^ m);
                key = o < key ? o : key;
            }
            const int gd = key >> 4;
            if (gd <= best_d) {
                best_d = gd;
                best = c0 + kAssignGroup - 1 - (key & 15);
            }
        }
        if (sl == 0 && lane < kAssignRows && i0 + lane < N) cluster[i0 + lane] = best;
    }
}

// Assignment with a lane per row, for K <= kRowsK centroids (kRowsK
// distances are computed; those past K are never chosen).
// Block (t, g) of a (row tiles) x (column groups) grid takes rows
// 64 t .. 64 t + 63 and the column group's words.  The block copies
// 64 x kRowsTileBytes of its rows at a time to LDS (each load instruction
// reads four 128-byte row pieces; the next tile is loaded into registers
// while the current one is used), then each of its four waves takes a quarter
// of the tile's words: lane r reads row r's word from LDS and the sixteen
// centroid words of that column, consecutive in cent_cols, are the same for
// the whole wave, so they are scalar loads used as instruction operands.  A
// word costs an XOR and a popcount-accumulate per 32 bits and centroid; no
// lane exchange is needed.  The four waves' distances are summed in LDS.
// With one column group the block picks each row's nearest centroid;
// otherwise it stores its partial distances to part[g][row][c] for
// km_argmin_kernel.
constexpr int kRowsWaves = 4;
constexpr int kRowsTileBytes = 128; // of each row per tile

// d + popc(v) as one v_bcnt_u32_b32 with d as its addend.
__device__ __forceinline__ int bcnt_acc(uint32_t v, int d)
{
    asm("v_bcnt_u32_b32 %0, %1, %0" : "+v"(d) : "v"(v));
    return d;
}
__device__ __forceinline__ int hamming_acc(uint64_t x, uint64_t c, int d)
{
    return bcnt_acc((uint32_t)(x >> 32) ^ (uint32_t)(c >> 32), bcnt_acc((uint32_t)x ^ (uint32_t)c, d));
}
__device__ __forceinline__ int hamming_acc(uint32_t x, uint32_t c, int d) { return bcnt_acc(x ^ c, d); }

// The kRowsK centroid words of one column (consecutive in cent_cols), read
// as 64-byte vectors so that a uniform address becomes s_load_dwordx16.
typedef uint32_t dword16_t __attribute__((ext_vector_type(16)));
template <typename W>
struct CentColumn {
    static constexpr int kVecs = kRowsK * (int)sizeof(W) / 64;
    dword16_t v[kVecs];
    __device__ __forceinline__ explicit CentColumn(const W *col)
    {
#pragma unroll
        for (int i = 0; i < kVecs; ++i) v[i] = reinterpret_cast<const dword16_t *>(col)[i];
    }
    __device__ __forceinline__ W operator[](int c) const
    {
        if constexpr (sizeof(W) == 8) {
            const dword16_t &h = v[c / 8];
            return ((uint64_t)h[2 * (c % 8) + 1] << 32) | h[2 * (c % 8)];
        } else {
            return v[0][c];
        }
    }
};

// Nearest of the first K distances; equal distances go to the highest index.
__device__ __forceinline__ int32_t nearest(const int (&d)[kRowsK], int K)
{
    int best = 0, best_d = 0x7fffffff;
#pragma unroll
    for (int c = 0; c < kRowsK; ++c)
        if (c < K && d[c] <= best_d) {
            best_d = d[c];
            best = c;
        }
    return best;
}

template <typename T>
__global__ __launch_bounds__(kWave * kRowsWaves) void km_assign_rows_kernel(const T *__restrict__ data, int64_t N,
                                                                            int64_t V, const T *__restrict__ cent_cols,
                                                                            int K, int64_t group_words,
                                                                            int32_t *__restrict__ cluster,
                                                                            int32_t *__restrict__ part)
{
    using W = KW<T>;
    constexpr int TW = kRowsTileBytes / (int)sizeof(W); // words of a row per tile
    constexpr int WPW = TW / kRowsWaves;                // of them per wave
    constexpr int LOADS = kWave * TW / (kWave * kRowsWaves);
    __shared__ W tile[kWave][TW + 1];
    __shared__ int sum[kRowsWaves - 1][kRowsK][kWave];
    const W *__restrict__ cent = reinterpret_cast<const W *>(cent_cols);
    const int lane = lane_id();
    const int wv = wave_in_block_id();
    const int64_t row0 = (int64_t)blockIdx.x * kWave;
    const int64_t col0 = (int64_t)blockIdx.y * group_words;
    const int64_t col1 = col0 + group_words < V ? col0 + group_words : V;
    // the words this thread copies: tile row q / TW, word q % TW, q = tid + i * block
    const W *src[LOADS];
    int dst_row[LOADS], dst_word[LOADS];
#pragma unroll
    for (int i = 0; i < LOADS; ++i) {
        const int q = (int)threadIdx.x + i * kWave * kRowsWaves;
        dst_row[i] = q / TW;
        dst_word[i] = q % TW;
        int64_t r = row0 + dst_row[i];
        if (r >= N) r = N - 1; // rows past N copy the last row; their result is dropped
        src[i] = reinterpret_cast<const W *>(data) + r * V + dst_word[i];
    }
    int d[kRowsK];
#pragma unroll
    for (int c = 0; c < kRowsK; ++c) d[c] = 0;
    W next[LOADS];
    {
        const int valid = (int)(col1 - col0 < TW ? col1 - col0 : TW);
#pragma unroll
        for (int i = 0; i < LOADS; ++i) next[i] = dst_word[i] < valid ? src[i][col0] : W(0);
    }
    for (int64_t j0 = col0; j0 < col1; j0 += TW) {
        const int valid = (int)(col1 - j0 < TW ? col1 - j0 : TW);
        __syncthreads(); // the previous tile has been read
#pragma unroll
        for (int i = 0; i < LOADS; ++i) tile[dst_row[i]][dst_word[i]] = next[i];
        __syncthreads();
        if (j0 + TW < col1) {
            const int64_t j1 = j0 + TW;
            const int valid1 = (int)(col1 - j1 < TW ? col1 - j1 : TW);
#pragma unroll
            for (int i = 0; i < LOADS; ++i) next[i] = dst_word[i] < valid1 ? src[i][j1] : W(0);
        }
        const int w0 = wv * WPW;
        const int w1 = w0 + WPW < valid ? w0 + WPW : valid;
        int w = w0;
        for (; w + 2 <= w1; w += 2) { // two columns' loads in flight together
            const W x0 = tile[lane][w], x1 = tile[lane][w + 1];
            const CentColumn<W> c0(cent + (j0 + w) * kRowsK), c1(cent + (j0 + w + 1) * kRowsK);
#pragma unroll
            for (int c = 0; c < kRowsK; ++c) d[c] = hamming_acc(x0, c0[c], d[c]);
#pragma unroll
            for (int c = 0; c < kRowsK; ++c) d[c] = hamming_acc(x1, c1[c], d[c]);
        }
        if (w < w1) {
            const W x = tile[lane][w];
            const CentColumn<W> cw(cent + (j0 + w) * kRowsK);
#pragma unroll
            for (int c = 0; c < kRowsK; ++c) d[c] = hamming_acc(x, cw[c], d[c]);
        }
    }
    // waves 1..3 hand their distances to wave 0
    if (wv > 0) {
#pragma unroll
        for (int c = 0; c < kRowsK; ++c) sum[wv - 1][c][lane] = d[c];
    }
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int w = 0; w < kRowsWaves - 1; ++w)
#pragma unroll
        for (int c = 0; c < kRowsK; ++c) d[c] += sum[w][c][lane];
    const int64_t r = row0 + lane;
    if (r >= N) return;
    if (gridDim.y == 1) {
        cluster[r] = nearest(d, K);
    } else {
        int4 *p = reinterpret_cast<int4 *>(part + ((int64_t)blockIdx.y * gridDim.x * kWave + r) * kRowsK);
#pragma unroll
        for (int c = 0; c < kRowsK; c += 4) p[c / 4] = make_int4(d[c], d[c + 1], d[c + 2], d[c + 3]);
    }
}

// Sums the column groups' partial distances of each row and picks its
// nearest centroid (km_assign_rows_kernel with more than one group): thread
// (row, c) sums distance c over the groups (the 16 threads of a row read 64
// consecutive bytes per group), then the row's 16 lanes take the minimum of
// (distance, 15 - c), so equal distances go to the highest index.
__global__ __launch_bounds__(256) void km_argmin_kernel(int64_t N, int K, int groups, int64_t group_stride,
                                                        const int32_t *__restrict__ part, int32_t *__restrict__ cluster)
{
    const int c = (int)(threadIdx.x % kRowsK);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N * kRowsK;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / kRowsK;
        const int32_t *p = part + r * kRowsK + c;
        const int64_t gs = group_stride * kRowsK;
        int a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        int g = 0;
        for (; g + 4 <= groups; g += 4) {
            a0 += p[(int64_t)g * gs];
            a1 += p[(int64_t)(g + 1) * gs];
            a2 += p[(int64_t)(g + 2) * gs];
            a3 += p[(int64_t)(g + 3) * gs];
        }
        for (; g < groups; ++g) a0 += p[(int64_t)g * gs];
        const int dist = (a0 + a1) + (a2 + a3);
        int key = c < K ? (dist << 4) | (kRowsK - 1 - c) : 0x7fffffff;
#pragma unroll
        for (int m = 1; m < kRowsK; m <<= 1) {
            const int o = (int)shfl((uint32_t)key, lane_id() ^ m);
            key = o < key ? o : key;
        }
        if (c == 0) cluster[r] = kRowsK - 1 - (key & (kRowsK - 1));
    }
}

// Rows of each cluster, grouped (a counting sort of the assignment) in one
// block of 1024 threads with the histogram in LDS: count[c], offset[c] and
// perm[offset[c] ..) = the rows of cluster c (in no particular order).  For
// K <= kSortMaxK; larger K use the three-kernel form below.
constexpr int kSortMaxK = 4096;
constexpr int kSortUnroll = 4;
//
// It also cuts every cluster into chunks of at most `chunk_rows` rows for the
// bit count: chunk e = (cluster, first row, end row) in chunks[3e..3e+2],
// cluster c's chunks are chunk_base[c] .. chunk_base[c] + chunk_count[c] - 1,
// and nchunks[0] is the total.
// Counting-sort steps for the lanes of one wave holding cluster ids c (-1:
// none).  With few clusters (K <= 64) many lanes may share an id, and
// atomics on one LDS word are serialised: up to kSortPeel times the wave
// takes its lowest pending lane's id, ballots the lanes with it, and one lane
// adds their number; lanes still pending then add their own.  (Peeling every
// id costs more than it saves when a wave holds many different ids, and the
// sort runs on one CU.)
constexpr int kSortPeel = 4;
__device__ __forceinline__ void wave_count(int32_t *hist, int c, bool aggregate)
{
    bool mine = c >= 0;
    if (aggregate) {
        uint64_t todo = ballot(mine);
        for (int it = 0; it < kSortPeel && todo; ++it) {
            const int leader = __builtin_ctzll(todo);
            const int v = readlane(c, leader);
            const uint64_t same = ballot(c == v) & todo;
            if (lane_id() == leader) atomicAdd(&hist[v], (int)__builtin_popcountll(same));
            if (c == v) mine = false;
            todo &= ~same;
        }
    }
    if (mine) atomicAdd(&hist[c], 1);
}

// base[c] + a slot among cluster c's rows (fill[c] counts the slots taken).
__device__ __forceinline__ int wave_claim(const int32_t *base, int32_t *fill, int c, bool aggregate)
{
    int slot = -1;
    bool mine = c >= 0;
    if (aggregate) {
        const int lane = lane_id();
        const uint64_t below = (1ull << lane) - 1;
        uint64_t todo = ballot(mine);
        for (int it = 0; it < kSortPeel && todo; ++it) {
            const int leader = __builtin_ctzll(todo);
            const int v = readlane(c, leader);
            const uint64_t same = ballot(c == v) & todo;
            int first = 0;
            if (lane == leader) first = base[v] + atomicAdd(&fill[v], (int)__builtin_popcountll(same));
            first = readlane(first, leader);
            if (c == v) {
                slot = first + (int)__builtin_popcountll(same & below);
                mine = false;
            }
            todo &= ~same;
        }
    }
    if (mine) slot = base[c] + atomicAdd(&fill[c], 1);
    return slot;
}

__global__ __launch_bounds__(1024) void km_sort_kernel(const int32_t *cluster, int64_t N, int K, int32_t *count,
                                                       int32_t *offset, int32_t *perm, int chunk_rows,
                                                       int32_t *chunks, int32_t *chunk_base, int32_t *chunk_count,
                                                       int32_t *nchunks)
{
    __shared__ int32_t hist[kSortMaxK];
    __shared__ int32_t base[kSortMaxK];
    // Rows are read kSortUnroll at a time (their loads in flight together);
    // the first kSortUnroll * blockDim.x rows' clusters stay in registers for
    // the scatter.
    constexpr int U = kSortUnroll;
    const int64_t step = (int64_t)U * blockDim.x;
    int first[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = threadIdx.x + (int64_t)u * blockDim.x;
        first[u] = i < N ? cluster[i] : -1;
    }
    const bool aggregate = K <= kWave;
    for (int c = threadIdx.x; c < K; c += blockDim.x) hist[c] = 0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) wave_count(hist, first[u], aggregate);
    // the block's threads run the same number of iterations (whole waves
    // take part in the ballots)
    for (int64_t b0 = step; b0 < N; b0 += step) {
        const int64_t i0 = b0 + threadIdx.x;
        int cu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            cu[u] = i < N ? cluster[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) wave_count(hist, cu[u], aggregate);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t acc = 0, e = 0;
        for (int c = 0; c < K; ++c) {
            base[c] = acc;
            const int32_t n = hist[c];
            chunk_base[c] = e;
            for (int32_t b = 0; b < n; b += chunk_rows, ++e) {
                chunks[3 * e] = c;
                chunks[3 * e + 1] = acc + b;
                chunks[3 * e + 2] = acc + (b + chunk_rows < n ? b + chunk_rows : n);
            }
            chunk_count[c] = e - chunk_base[c];
            acc += n;
        }
        nchunks[0] = e;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < K; c += blockDim.x) {
        count[c] = hist[c];
        offset[c] = base[c];
        hist[c] = 0; // becomes the fill counter
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int slot = wave_claim(base, hist, first[u], aggregate);
        if (slot >= 0) perm[slot] = (int32_t)(threadIdx.x + u * blockDim.x);
    }
    for (int64_t b0 = step; b0 < N; b0 += step) {
        const int64_t i0 = b0 + threadIdx.x;
        int cu[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            cu[u] = i < N ? cluster[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int slot = wave_claim(base, hist, cu[u], aggregate);
            if (slot >= 0) perm[slot] = (int32_t)(i0 + (int64_t)u * blockDim.x);
        }
    }
}

// Beyond the one-block sort (more than 2^20 rows): histogram and scatter
// over many blocks.  Each block counts its rows per cluster in LDS (wave-
// aggregated atomics) and adds one global atomic per cluster it holds, so a
// few clusters do not serialise millions of global atomics on one word.
// Blocks of 256 threads; K <= kSortMaxK (larger K take the per-row form).
__global__ __launch_bounds__(256) void km_hist_kernel(const int32_t *cluster, int64_t N, int K, int32_t *count)
{
    __shared__ int32_t local[kSortMaxK];
    const bool in_lds = K <= kSortMaxK;
    const bool aggregate = K <= kWave;
    if (in_lds) {
        for (int c = threadIdx.x; c < K; c += blockDim.x) local[c] = 0;
        __syncthreads();
    }
    // whole waves take part in every step (the ballots of wave_count)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < N; b0 += stride) {
        const int64_t i = b0 + threadIdx.x;
        const int c = i < N ? cluster[i] : -1;
        if (in_lds)
            wave_count(local, c, aggregate);
        else if (c >= 0)
            atomicAdd(&count[c], 1);
    }
    if (!in_lds) return;
    __syncthreads();
    for (int c = threadIdx.x; c < K; c += blockDim.x)
        if (local[c]) atomicAdd(&count[c], local[c]);
}

// offset[c] (exclusive scan of the counts), fill[c] = 0, and with
// chunk_base: cluster c's chunks of at most chunk_rows rows are
// chunk_base[c] .. chunk_base[c] + chunk_count[c] - 1, nchunks[0] in all.
__global__ void km_offsets_kernel(const int32_t *count, int K, int32_t *offset, int32_t *fill, int chunk_rows,
                                  int32_t *chunk_base, int32_t *chunk_count, int32_t *nchunks)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int32_t acc = 0, e = 0;
    for (int c = 0; c < K; ++c) {
        offset[c] = acc;
        fill[c] = 0;
        acc += count[c];
        if (chunk_base) {
            const int32_t nc = (int32_t)(((int64_t)count[c] + chunk_rows - 1) / chunk_rows);
            chunk_base[c] = e;
            chunk_count[c] = nc;
            e += nc;
        }
    }
    if (chunk_base) nchunks[0] = e;
}

// The chunk list of km_sort_kernel's form (chunks[3e..3e+2] = cluster, first
// and end position in perm) from km_offsets_kernel's chunk counts: thread e
// finds its cluster, the last one whose chunks start at or before e.
__global__ void km_chunks_kernel(const int32_t *count, const int32_t *offset, const int32_t *chunk_base,
                                 const int32_t *nchunks, int K, int chunk_rows, int64_t max_chunks, int32_t *chunks)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= max_chunks || e >= nchunks[0]) return;
    int lo = 0, hi = K - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (chunk_base[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int c = lo;
    const int64_t b = (e - chunk_base[c]) * (int64_t)chunk_rows;
    const int64_t n = count[c];
    chunks[3 * e] = c;
    chunks[3 * e + 1] = (int32_t)(offset[c] + b);
    chunks[3 * e + 2] = (int32_t)(offset[c] + (b + chunk_rows < n ? b + chunk_rows : n));
}

// perm[offset[c] ..) = the rows of cluster c.  A block counts its rows per
// cluster in LDS, reserves one range per cluster with a global atomic, then
// places its rows in that range (a second pass over its cluster ids).
__global__ __launch_bounds__(256) void km_scatter_kernel(const int32_t *cluster, int64_t N, int K,
                                                         const int32_t *offset, int32_t *fill, int32_t *perm)
{
    __shared__ int32_t local[kSortMaxK];
    __shared__ int32_t base[kSortMaxK];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (K > kSortMaxK) {
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
            const int c = cluster[i];
            perm[offset[c] + atomicAdd(&fill[c], 1)] = (int32_t)i;
        }
        return;
    }
    const bool aggregate = K <= kWave;
    for (int c = threadIdx.x; c < K; c += blockDim.x) local[c] = 0;
    __syncthreads();
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < N; b0 += stride) {
        const int64_t i = b0 + threadIdx.x;
        wave_count(local, i < N ? cluster[i] : -1, aggregate);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < K; c += blockDim.x) {
        base[c] = local[c] ? offset[c] + atomicAdd(&fill[c], local[c]) : 0;
        local[c] = 0; // becomes the block's fill counter
    }
    __syncthreads();
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < N; b0 += stride) {
        const int64_t i = b0 + threadIdx.x;
        const int slot = wave_claim(base, local, i < N ? cluster[i] : -1, aggregate);
        if (slot >= 0) perm[slot] = (int32_t)i;
    }
}

// bits_set[(c*V + j)*B + k] = members of cluster c whose column j has bit k set.
// grid = (column tiles, K, row splits); lane <-> column, the block's waves
// split the cluster's rows, 16 rows at a time (16 gathered loads in flight)
// added with carry-save adders (VCounter).  The waves' counters are summed
// in LDS; with one row split the block owns its counters and stores them,
// otherwise it adds them to counters zeroed beforehand.
template <typename T>
__global__ __launch_bounds__(256) void km_bitcount_kernel(const T *data, int64_t V, const int32_t *perm,
                                                          const int32_t *offset, const int32_t *count,
                                                          int32_t *bits_set)
{
    constexpr int B = 8 * (int)sizeof(T);
    constexpr int CHUNK = 16;
    using W = KW<T>;
    __shared__ int32_t sum[B * kWave];
    const int lane = lane_id();
    const int wv = wave_in_block_id();
    const int c = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * kWave + lane;
    const bool col = j < V;
    const int waves = blockDim.x / kWave;
    const int64_t stride = (int64_t)waves * gridDim.z;
    const int32_t *rows = perm + offset[c];
    const int64_t nrows = count[c];
    for (int i = threadIdx.x; i < B * kWave; i += blockDim.x) sum[i] = 0;
    int32_t acc[B];
#pragma unroll
    for (int k = 0; k < B; ++k) acc[k] = 0;
    VCounter<W, B> vc;
    for (int64_t q = (int64_t)blockIdx.z * waves + wv; q < nrows; q += (int64_t)CHUNK * stride) {
        W x[CHUNK];
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) {
            const int64_t qi = q + (int64_t)u * stride;
            x[u] = (col && qi < nrows) ? (W)data[(int64_t)rows[qi] * V + j] : W(0);
        }
        vc.add16(x, acc);
    }
    vc.flush(acc);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < B; ++k)
        if (acc[k]) atomicAdd(&sum[k * kWave + lane], acc[k]);
    __syncthreads();
    // thread i -> column tile*64 + i / B, bit i % B: consecutive threads write
    // consecutive counters
    for (int i = threadIdx.x; i < B * kWave; i += blockDim.x) {
        const int jl = i / B, k = i % B;
        const int64_t jj = (int64_t)blockIdx.x * kWave + jl;
        if (jj >= V) continue;
        int32_t *dst = bits_set + ((int64_t)c * V + jj) * B + k;
        const int32_t v = sum[k * kWave + jl];
        if (gridDim.z == 1) *dst = v;
        else if (v) atomicAdd(dst, v);
    }
}

// Bit counts of one chunk of a cluster's rows (km_sort_kernel's list):
// grid = (column tiles, maximum chunk count); blocks past the list's length
// return at once.  Same counting as km_bitcount_kernel; the block's counters
// go to part[e][j*B + k], summed per cluster by km_reduce_bits_kernel.
// Chunks of equal size keep every block's work alike however uneven the
// clusters are.
template <typename T>
__global__ __launch_bounds__(256) void km_bitcount_chunks_kernel(const T *data, int64_t V, const int32_t *perm,
                                                                 const int32_t *chunks, const int32_t *nchunks,
                                                                 int32_t *part)
{
    constexpr int B = 8 * (int)sizeof(T);
    constexpr int CHUNK = 16;
    using W = KW<T>;
    __shared__ int32_t sum[B * kWave];
    const int e = blockIdx.y;
    if (e >= nchunks[0]) return; // uniform for the block
    const int lane = lane_id();
    const int wv = wave_in_block_id();
    const int64_t j = (int64_t)blockIdx.x * kWave + lane;
    const bool col = j < V;
    const int waves = blockDim.x / kWave;
    const int64_t q0 = chunks[3 * e + 1], q1 = chunks[3 * e + 2];
    for (int i = threadIdx.x; i < B * kWave; i += blockDim.x) sum[i] = 0;
    int32_t acc[B];
#pragma unroll
    for (int k = 0; k < B; ++k) acc[k] = 0;
    VCounter<W, B> vc;
    // The wave's rows are q0 + wv, q0 + wv + waves, ...: their ids come 64 at a
    // time in one vector load (lane l holds row l), and 32 rows' words are
    // loaded together before they are counted.
    for (int64_t qw = q0 + wv; qw < q1; qw += (int64_t)kWave * waves) {
        const int64_t ql = qw + (int64_t)lane * waves;
        const int32_t rid = ql < q1 ? perm[ql] : 0;
        const int64_t left = (q1 - qw + waves - 1) / waves;
        const int nr = left < kWave ? (int)left : kWave;
        for (int b = 0; b < nr; b += 2 * CHUNK) {
            W xa[CHUNK], xb[CHUNK];
#pragma unroll
            for (int u = 0; u < CHUNK; ++u) {
                const int64_t ra = readlane(rid, b + u), rb = readlane(rid, b + CHUNK + u);
                xa[u] = (col && b + u < nr) ? (W)data[ra * V + j] : W(0);
                xb[u] = (col && b + CHUNK + u < nr) ? (W)data[rb * V + j] : W(0);
            }
            vc.add16(xa, acc);
            vc.add16(xb, acc);
        }
    }
    vc.flush(acc);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < B; ++k)
        if (acc[k]) atomicAdd(&sum[k * kWave + lane], acc[k]);
    __syncthreads();
    int32_t *dst = part + (int64_t)e * V * B;
    for (int i = threadIdx.x; i < B * kWave; i += blockDim.x) {
        const int jl = i / B, k = i % B;
        const int64_t jj = (int64_t)blockIdx.x * kWave + jl;
        if (jj < V) dst[jj * B + k] = sum[k * kWave + jl];
    }
}

// bits_set[(c*V + j)*B + k] = sum of part[e][j*B + k] over cluster c's chunks.
__global__ void km_reduce_bits_kernel(int64_t VB, int K, const int32_t *chunk_base, const int32_t *chunk_count,
                                      const int32_t *part, int32_t *bits_set)
{
    const int64_t total = (int64_t)K * VB;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(g / VB);
        const int64_t i = g % VB;
        const int32_t e0 = chunk_base[c], ne = chunk_count[c];
        int32_t acc = 0;
        for (int32_t e = 0; e < ne; ++e) acc += part[(int64_t)(e0 + e) * VB + i];
        bits_set[g] = acc;
    }
}

// km_reduce_bits_kernel and km_majority_kernel in one pass: a wave sums the
// chunk counts of 64 consecutive (column, bit) counters of one cluster (one
// 8-byte column or two 4-byte ones), lane l the counter l, and the majority
// words are one ballot of (count > members / 2).  The summed counts are
// stored only when `store_bits` (the last iteration, whose counts give the
// masks); the other iterations need only the majority words.
template <typename T>
__global__ __launch_bounds__(256) void km_reduce_majority_kernel(int64_t V, int K, const int32_t *count,
                                                                 const int32_t *chunk_base,
                                                                 const int32_t *chunk_count, const int32_t *part,
                                                                 int store_bits, int32_t *bits_set, T *maj)
{
    constexpr int B = 8 * (int)sizeof(T);
    const int64_t VB = V * B;
    const int64_t waves_per_cluster = VB / kWave; // V * B is a multiple of 64 (B = 32 needs even V: see launch)
    const int64_t total = (int64_t)K * waves_per_cluster;
    const int lane = lane_id();
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t gw = global_wave_id(); gw < total; gw += nwaves) {
        const int c = (int)(gw / waves_per_cluster);
        const int64_t i = (gw % waves_per_cluster) * kWave + lane; // counter (column j = i / B, bit i % B)
        const int32_t e0 = chunk_base[c], ne = chunk_count[c];
        const int32_t *p = part + (int64_t)e0 * VB + i;
        int32_t acc = 0;
        int32_t e = 0;
        for (; e + 4 <= ne; e += 4) {
            const int32_t a0 = p[0], a1 = p[VB], a2 = p[2 * VB], a3 = p[3 * VB];
            acc += (a0 + a1) + (a2 + a3);
            p += 4 * VB;
        }
        for (; e < ne; ++e, p += VB) acc += p[0];
        if (store_bits) bits_set[(int64_t)c * VB + i] = acc;
        const uint64_t m = ballot(acc > count[c] / 2);
        if (lane % B == 0) maj[((int64_t)c * VB + i) / B] = (T)(m >> lane);
    }
}

// Majority word of cluster c, column j: bit k set iff more than half of the
// cluster's members have it (the per-column part of construct_bitmasks).
template <typename T>
__global__ void km_majority_kernel(int64_t V, int K, const int32_t *count, const int32_t *bits_set, T *maj)
{
    constexpr int B = 8 * (int)sizeof(T);
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)K * V) return;
    const int32_t half = count[g / V] / 2;
    const int32_t *bs = bits_set + g * B;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < B; ++k)
        if (bs[k] > half) acc |= 1ull << k;
    maj[g] = (T)acc;
}

// construct_bitmasks (ibp_preproc_kmeans.cuh:219-236): chain l < 32 of
// centroid c walks columns l, l+32, ... and ORs majority bits into one
// accumulator that is never cleared, so
// centroid[c][l + 32i] = OR of maj[c][l + 32i'], i' <= i (a prefix OR).
// A block per centroid: thread (p, l) takes part p of 32 consecutive parts of
// chain l, ORs its words, the parts' ORs meet in LDS, and the thread writes
// its part's prefix ORs starting from the OR of the parts before it.  The
// majority words come from km_majority_kernel or km_reduce_majority_kernel;
// with cent_cols the centroid is also stored by column.
template <typename T>
__global__ __launch_bounds__(1024) void km_construct_kernel(int64_t V, int K, const T *maj, T *centroids,
                                                            T *cent_cols)
{
    __shared__ T part_or[32][32];
    const int c = blockIdx.x;
    const int l = (int)threadIdx.x % 32, p = (int)threadIdx.x / 32;
    const int64_t chain = (V + 31) / 32;
    const int64_t per = (chain + 31) / 32;
    const int64_t i0 = p * per, i1 = i0 + per < chain ? i0 + per : chain;
    const T *src = maj + (int64_t)c * V;
    T *dst = centroids + (int64_t)c * V;
    T acc = 0;
    for (int64_t i = i0; i < i1; ++i) {
        const int64_t j = l + 32 * i;
        if (j < V) acc = (T)(acc | src[j]);
    }
    part_or[p][l] = acc;
    __syncthreads();
    T run = 0;
    for (int q = 0; q < p; ++q) run = (T)(run | part_or[q][l]);
    for (int64_t i = i0; i < i1; ++i) {
        const int64_t j = l + 32 * i;
        if (j < V) {
            run = (T)(run | src[j]);
            dst[j] = run;
            if (cent_cols) cent_cols[j * kRowsK + c] = run;
        }
    }
}

// create_mask_many at threshold t (ibp_preproc_kmeans.cuh:273-293):
// float(bits) >= t * float(members) -> mask and bitval;
// double(bits) <= (1.0 - t) * members -> mask only.
template <typename T>
__global__ void km_masks_kernel(int64_t V, int K, float t, const int32_t *count, const int32_t *bits_set, T *masks,
                                T *bitvals)
{
    constexpr int B = 8 * (int)sizeof(T);
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)K * V) return;
    const int c = (int)(g / V);
    const float hi = __fmul_rn(t, (float)count[c]);
    const double lo = __dmul_rn(1.0 - (double)t, (double)count[c]);
    const int32_t *bs = bits_set + g * B;
    uint64_t val = 0, m = 0;
    for (int k = 0; k < B; ++k) {
        const float f = (float)bs[k];
        if (f >= hi) {
            val |= 1ull << k;
            m |= 1ull << k;
        } else if ((double)f <= lo) {
            m |= 1ull << k;
        }
    }
    masks[g] = (T)m;
    bitvals[g] = (T)val;
}

__global__ void km_zero_kernel(int32_t *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

float sweep_last_threshold()
{
    float last = 0.f;
    for (volatile float t = 0.7f; t <= 1.0f; t = (float)((double)t + 0.05)) last = t;
    return last;
}

// The sequence rand() returns after srand(seed) in glibc, without the
// process-wide generator and its lock: glibc's additive feedback generator
// (31 words, taps 31 and 3, seeded by the Park-Miller recurrence
// s[i] = 16807 s[i-1] mod (2^31 - 1), with the first 310 outputs dropped),
// each output the new word shifted right by one.  ibp_selftest_rand checks it
// against rand() itself.
class LibcRandSequence {
public:
    explicit LibcRandSequence(unsigned seed)
    {
        int32_t s = seed ? (int32_t)seed : 1;
        ring_[0] = (uint32_t)s;
        for (int i = 1; i < 31; ++i) {
            const int64_t hi = s / 127773, lo = s % 127773;
            int64_t w = 16807 * lo - 2836 * hi; // Schrage's form of 16807 s mod (2^31 - 1)
            if (w < 0) w += 2147483647;
            s = (int32_t)w;
            ring_[i] = (uint32_t)s;
        }
        for (int i = 0; i < 310; ++i) (void)next();
    }
    int next()
    {
        ring_[front_] += ring_[rear_];
        const uint32_t out = ring_[front_] >> 1;
        front_ = front_ == 30 ? 0 : front_ + 1;
        rear_ = rear_ == 30 ? 0 : rear_ + 1;
        return (int)out;
    }

private:
    uint32_t ring_[31];
    int front_ = 3, rear_ = 0;
};

// Positions 0..m-1 of libstdc++'s random_shuffle(first, last) over 0..n-1
// with the generator `rnd` (rand() or its sequence): every draw of the whole
// shuffle is made, but only the first m positions are kept.  Step i swaps
// positions i and j <= i, and position i still holds i then (no earlier step
// reaches past its own index), so for i >= m the step only sets position j
// to i when j < m.
template <typename Rnd>
std::vector<int32_t> random_shuffle_head(int64_t n, int64_t m, Rnd &&rnd)
{
    std::vector<int32_t> a((size_t)m);
    for (int64_t i = 0; i < m; ++i) a[(size_t)i] = (int32_t)i;
    for (int64_t i = 1; i < n; ++i) {
        const int64_t j = (int64_t)(rnd() % (int)(i + 1));
        if (i < m) {
            if (i != j) std::swap(a[(size_t)i], a[(size_t)j]);
        } else if (j < m) {
            a[(size_t)j] = (int32_t)i;
        }
    }
    return a;
}

namespace {

unsigned blocks_for(int64_t n, int64_t cap = 4096)
{
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

template <typename T>
int kmeans_typed(const void *data_v, int64_t N, int64_t V, int K, int64_t seed, void *masks, void *bitvals,
                 int32_t *clusters_out, hipStream_t s)
{
    constexpr int B = 8 * (int)sizeof(T);
    const DeviceInfo &dev = device_info();
    const T *data = (const T *)data_v;
    // Host side initialisation (ibp_preproc_host.cuh:150-151,176-188).
    std::vector<int64_t> init((size_t)K, -1);
    if (N > 0) init[0] = 0;
    {
        // seed >= 0: the draws of srand(seed), from a private copy of the
        // sequence (the process-wide generator is left as it is); otherwise
        // rand() itself, continuing the process-wide state like the reference
        const int64_t m = K < N ? K : N;
        std::vector<int32_t> head;
        if (seed >= 0) {
            LibcRandSequence seq((unsigned)seed);
            head = random_shuffle_head(N, m, [&]() { return seq.next(); });
        } else {
            head = random_shuffle_head(N, m, []() { return rand(); });
        }
        for (int64_t c = 1; c < m; ++c) init[(size_t)c] = head[(size_t)c];
    }
    T *cent = nullptr, *maj = nullptr, *cent_cols = nullptr;
    int32_t *dist = nullptr;
    int32_t *cluster = nullptr, *count = nullptr, *offset = nullptr, *fill = nullptr, *perm = nullptr, *bits = nullptr;
    int32_t *chunks = nullptr, *part = nullptr;
    int64_t *rows = nullptr;
    const int64_t nbits = (int64_t)K * V * B;
    // The LDS counting sort (one block) serves K <= kSortMaxK and up to 2^20
    // rows, a histogram and scatter over many blocks more rows; both cut the
    // clusters into chunks of chunk_rows rows whose partial bit counts (about
    // 256 MiB of them at most) are summed per cluster.
    const bool sort_in_lds = K <= kSortMaxK && N <= ((int64_t)1 << 20);
    const int64_t per_chunk = V * B * (int64_t)sizeof(int32_t);
    // (the chunk index is blockIdx.y of the bit count: below 65535)
    int64_t max_chunks_mem = ((int64_t)256 << 20) / per_chunk;
    if (max_chunks_mem > 65534) max_chunks_mem = 65534;
    const bool chunked = K <= kSortMaxK && max_chunks_mem > K;
    // 256-row chunks for the one-block sort; beyond it, about eight bit-count
    // blocks per CU (more, smaller chunks cost more in the per-cluster sums
    // than they gain)
    int64_t chunk_rows = 256;
    if (chunked && !sort_in_lds) {
        const int64_t tiles64 = (V + kWave - 1) / kWave;
        int64_t want = (int64_t)dev.num_cus * 8 / tiles64;
        if (want < K + 1) want = K + 1;
        const int64_t rows = (N + want - 1) / want;
        if (rows > chunk_rows) chunk_rows = rows;
    }
    if (chunked && N / chunk_rows + K + 1 > max_chunks_mem) chunk_rows = (N + max_chunks_mem - K - 1) / (max_chunks_mem - K);
    const int64_t max_chunks = chunked ? N / chunk_rows + K + 1 : 0;
    hipError_t e = hipSuccess;
    // Lane-per-row assignment (K <= kRowsK): row tiles x column groups, about
    // eight blocks per CU (2 to 8 measured alike on 4096-row blocks), and the
    // centroids also by column.
    const bool rows_assign = K <= kRowsK;
    constexpr int64_t kTileWords = kRowsTileBytes / (int64_t)sizeof(KW<T>);
    const int64_t row_tiles = (N + kWave - 1) / kWave;
    int64_t col_groups = 1, group_words = V;
    if (rows_assign && row_tiles > 0) {
        const int64_t max_groups = (V + kTileWords - 1) / kTileWords;
        col_groups = ((int64_t)dev.num_cus * 8 + row_tiles - 1) / row_tiles;
        if (col_groups > max_groups) col_groups = max_groups;
        if (col_groups < 1) col_groups = 1;
        group_words = ((V + col_groups - 1) / col_groups + kTileWords - 1) / kTileWords * kTileWords;
        col_groups = (V + group_words - 1) / group_words;
    }
    // One workspace allocation, cut into 256-byte aligned buffers.
    const int64_t n1 = N > 0 ? N : 1;
    const size_t sz_cent = sizeof(T) * (size_t)(K * V);
    const size_t sz_cols = rows_assign ? sizeof(T) * (size_t)(kRowsK * V) : 0;
    const size_t sz_dist = col_groups > 1 ? sizeof(int32_t) * (size_t)(col_groups * row_tiles * kWave * kRowsK) : 0;
    const size_t sz_cluster = sizeof(int32_t) * (size_t)n1;
    const size_t sz_count = sizeof(int32_t) * 3 * (size_t)K;
    const size_t sz_bits = sizeof(int32_t) * (size_t)nbits;
    const size_t sz_rows = sizeof(int64_t) * (size_t)K;
    const size_t sz_chunks = chunked ? sizeof(int32_t) * (size_t)(3 * max_chunks + 2 * K + 1) : 0;
    const size_t sz_part = chunked ? (size_t)(per_chunk * max_chunks) : 0;
    const size_t sizes[] = {sz_cent, sz_cols, sz_dist, sz_cent, sz_cluster, sz_count, sz_cluster, sz_bits, sz_rows,
                            sz_chunks, sz_part};
    size_t total = 0;
    for (size_t b : sizes) total += (b + 255) / 256 * 256;
    unsigned char *ws = nullptr;
    if ((e = hipMallocAsync((void **)&ws, total, s)) != hipSuccess)
        return set_error(IBP_ERR_NOMEM, "kmeans workspace allocation failed: %s", hipGetErrorString(e));
    {
        size_t at = 0;
        auto take = [&](size_t b) -> void * {
            void *p = b ? ws + at : nullptr;
            at += (b + 255) / 256 * 256;
            return p;
        };
        cent = (T *)take(sz_cent);
        cent_cols = (T *)take(sz_cols);
        dist = (int32_t *)take(sz_dist);
        maj = (T *)take(sz_cent);
        cluster = (int32_t *)take(sz_cluster);
        count = (int32_t *)take(sz_count);
        perm = (int32_t *)take(sz_cluster);
        bits = (int32_t *)take(sz_bits);
        rows = (int64_t *)take(sz_rows);
        chunks = (int32_t *)take(sz_chunks);
        part = (int32_t *)take(sz_part);
    }
    auto cleanup = [&]() { (void)hipFreeAsync(ws, s); };
    offset = count + K;
    fill = count + 2 * K;
    if ((e = hipMemcpyAsync(rows, init.data(), sizeof(int64_t) * (size_t)K, hipMemcpyHostToDevice, s)) != hipSuccess) {
        cleanup();
        return set_error(IBP_ERR_HIP, "kmeans initialisation failed: %s", hipGetErrorString(e));
    }
    hipLaunchKernelGGL((km_gather_kernel<T>), dim3(blocks_for((int64_t)K * V)), dim3(256), 0, s, data, V, rows, K, cent,
                       cent_cols);
    hipLaunchKernelGGL(km_zero_kernel, dim3(blocks_for(N)), dim3(256), 0, s, cluster, N);
    if ((e = hipGetLastError()) != hipSuccess) {
        cleanup();
        return set_error(IBP_ERR_HIP, "kmeans initialisation failed: %s", hipGetErrorString(e));
    }

    const bool assign_split = V >= kWave * kAssignWaves;
    const int64_t groups = (N + kAssignRows - 1) / kAssignRows;
    int64_t assign_grid = assign_split ? groups : (groups + kAssignWaves - 1) / kAssignWaves;
    if (assign_grid > (int64_t)dev.num_cus * 8) assign_grid = (int64_t)dev.num_cus * 8;
    if (assign_grid < 1) assign_grid = 1;
    const unsigned tiles = (unsigned)((V + kWave - 1) / kWave);
    // Without the chunk list (K > kSortMaxK, or rows so wide that K chunks of
    // partial counts exceed the 256 MiB bound): row splits of the bit count,
    // about 64 rows of a cluster per wave (four waves per block); one split
    // lets each block store its counters.
    const int64_t per_cluster = N / (K > 0 ? K : 1);
    int64_t splits64 = (per_cluster + 255) / 256;
    if (splits64 < 1) splits64 = 1;
    if (splits64 > 64) splits64 = 64;
    const unsigned splits = (unsigned)splits64;
    int32_t *chunk_base = chunked ? chunks + 3 * max_chunks : nullptr;
    int32_t *chunk_count = chunked ? chunk_base + K : nullptr;
    int32_t *nchunks = chunked ? chunk_count + K : nullptr;
    // Blocks of the many-block histogram and scatter: each adds one global
    // atomic per cluster it holds, and those on one cluster's counter are
    // serialised, so two blocks per CU (4096 blocks: 338 us per scatter at
    // K = 4 on 2.45M rows)
    const unsigned group_blocks = blocks_for(N, (int64_t)dev.num_cus * 2);
    // chunk sums and majority words in one kernel when every wave's 64
    // counters lie in one cluster
    const bool fused_majority = chunked && (V * B) % kWave == 0;
    for (int it = 0; it < 10; ++it) {
        if (rows_assign) {
            if (N > 0)
                hipLaunchKernelGGL((km_assign_rows_kernel<T>), dim3((unsigned)row_tiles, (unsigned)col_groups),
                                   dim3(kWave * kRowsWaves), 0, s, data, N, V, (const T *)cent_cols, K, group_words,
                                   cluster, dist);
            if (col_groups > 1)
                hipLaunchKernelGGL(km_argmin_kernel, dim3(blocks_for(N * kRowsK)), dim3(256), 0, s, N, K, (int)col_groups,
                                   row_tiles * kWave, (const int32_t *)dist, cluster);
        } else if (assign_split)
            hipLaunchKernelGGL((km_assign_kernel<T, kAssignWaves>), dim3((unsigned)assign_grid),
                               dim3(kWave * kAssignWaves), 0, s, data, N, V, cent, K, cluster);
        else
            hipLaunchKernelGGL((km_assign_kernel<T, 1>), dim3((unsigned)assign_grid), dim3(kWave * kAssignWaves), 0,
                               s, data, N, V, cent, K, cluster);
        if (chunked) {
            if (sort_in_lds) {
                hipLaunchKernelGGL(km_sort_kernel, dim3(1), dim3(1024), 0, s, cluster, N, K, count, offset, perm,
                                   (int)chunk_rows, chunks, chunk_base, chunk_count, nchunks);
            } else {
                // more than 2^20 rows: the same grouping and chunk list over
                // many blocks
                hipLaunchKernelGGL(km_zero_kernel, dim3(1), dim3(256), 0, s, count, (int64_t)K);
                hipLaunchKernelGGL(km_hist_kernel, dim3(group_blocks), dim3(256), 0, s, cluster, N, K, count);
                hipLaunchKernelGGL(km_offsets_kernel, dim3(1), dim3(64), 0, s, count, K, offset, fill,
                                   (int)chunk_rows, chunk_base, chunk_count, nchunks);
                hipLaunchKernelGGL(km_chunks_kernel, dim3(blocks_for(max_chunks, 1 << 20)), dim3(256), 0, s, count,
                                   offset, chunk_base, nchunks, K, (int)chunk_rows, max_chunks, chunks);
                hipLaunchKernelGGL(km_scatter_kernel, dim3(group_blocks), dim3(256), 0, s, cluster, N, K, offset,
                                   fill, perm);
            }
            hipLaunchKernelGGL((km_bitcount_chunks_kernel<T>), dim3(tiles, (unsigned)max_chunks), dim3(256), 0, s,
                               data, V, perm, (const int32_t *)chunks, (const int32_t *)nchunks, part);
            if (fused_majority) {
                hipLaunchKernelGGL((km_reduce_majority_kernel<T>), dim3(blocks_for(nbits, 1 << 16)), dim3(256), 0,
                                   s, V, K, (const int32_t *)count, (const int32_t *)chunk_base,
                                   (const int32_t *)chunk_count, (const int32_t *)part, it == 9 ? 1 : 0, bits, maj);
            } else {
                hipLaunchKernelGGL(km_reduce_bits_kernel, dim3(blocks_for(nbits, 1 << 16)), dim3(256), 0, s, V * B,
                                   K, (const int32_t *)chunk_base, (const int32_t *)chunk_count,
                                   (const int32_t *)part, bits);
            }
        } else {
            hipLaunchKernelGGL(km_zero_kernel, dim3(1), dim3(256), 0, s, count, (int64_t)K);
            hipLaunchKernelGGL(km_hist_kernel, dim3(blocks_for(N)), dim3(256), 0, s, cluster, N, K, count);
            hipLaunchKernelGGL(km_offsets_kernel, dim3(1), dim3(64), 0, s, count, K, offset, fill, 1,
                               (int32_t *)nullptr, (int32_t *)nullptr, (int32_t *)nullptr);
            hipLaunchKernelGGL(km_scatter_kernel, dim3(blocks_for(N)), dim3(256), 0, s, cluster, N, K, offset, fill,
                               perm);
            if (splits > 1) hipLaunchKernelGGL(km_zero_kernel, dim3(blocks_for(nbits)), dim3(256), 0, s, bits, nbits);
            hipLaunchKernelGGL((km_bitcount_kernel<T>), dim3(tiles, (unsigned)K, splits), dim3(256), 0, s, data, V,
                               perm, offset, count, bits);
        }
        if (!fused_majority)
            hipLaunchKernelGGL((km_majority_kernel<T>), dim3(blocks_for((int64_t)K * V, 1 << 20)), dim3(256), 0, s, V,
                               K, count, bits, maj);
        hipLaunchKernelGGL((km_construct_kernel<T>), dim3((unsigned)K), dim3(1024), 0, s, V, K, (const T *)maj, cent,
                           cent_cols);
        e = hipGetLastError();
        if (e != hipSuccess) {
            cleanup();
            return set_error(IBP_ERR_HIP, "kmeans iteration launch failed: %s", hipGetErrorString(e));
        }
    }
    // Final masks from the last assignment's counts (still in count/bits).
    hipLaunchKernelGGL((km_masks_kernel<T>), dim3(blocks_for((int64_t)K * V, 1 << 20)), dim3(256), 0, s, V, K,
                       sweep_last_threshold(), count, bits, (T *)masks, (T *)bitvals);
    e = hipGetLastError();
    if (e == hipSuccess && N > 0)
        e = hipMemcpyAsync(clusters_out, cluster, sizeof(int32_t) * (size_t)N, hipMemcpyDefault, s);
    cleanup();
    if (e != hipSuccess) return set_error(IBP_ERR_HIP, "kmeans finish failed: %s", hipGetErrorString(e));
    IBP_HIP(hipStreamSynchronize(s));
    return IBP_OK;
}

} // namespace

} // namespace ibp

using namespace ibp;

extern "C" int ibp_preprocess_kmeans(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                                     int num_clusters, float threshold, int64_t seed, void *masks, void *bitvals,
                                     int32_t *clusters, void *stream)
{
    IBP_TRACE("ibp_preprocess_kmeans");
    (void)threshold; // unused by the reference as well (ibp_preproc_host.cuh:119-284)
    IBP_CHECK(elem_bytes == 4 || elem_bytes == 8, "Input tensor must be 4-byte or 8-byte datatype (got %d)", elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0, "invalid shape");
    IBP_CHECK(num_clusters >= 1 && num_clusters <= 65535, "num_clusters must be in [1, 65535]");
    IBP_CHECK(num_vecs < (int64_t)1 << 31, "num_vecs must fit int32 cluster ids and counters");
    // distances (at most 8 * vec_size * elem_bytes bits) carry a 4-bit
    // centroid tag in an int32
    IBP_CHECK(vec_size * elem_bytes < (int64_t)1 << 24, "rows of 16 MiB or more are not supported");
    IBP_CHECK(masks && bitvals && ((data && clusters) || num_vecs == 0), "null pointer argument");
    const int st = elem_bytes == 8
                       ? kmeans_typed<uint64_t>(data, num_vecs, vec_size, num_clusters, seed, masks, bitvals, clusters,
                                                as_stream(stream))
                       : kmeans_typed<uint32_t>(data, num_vecs, vec_size, num_clusters, seed, masks, bitvals, clusters,
                                                as_stream(stream));
    if (st == IBP_OK) dprintf("[ibp] kmeans: K=%d V=%lld N=%lld\n", num_clusters, (long long)vec_size, (long long)num_vecs);
    if (st == IBP_OK && debug_enabled()) {
        // The per-centroid report of the reference's debug build
        // (ibp_preproc_host.cuh:221-232): invariant bits of each used
        // centroid's mask and its member count.  The call has synchronised.
        std::vector<int32_t> cl((size_t)num_vecs);
        std::vector<unsigned char> mk((size_t)num_clusters * vec_size * elem_bytes);
        if (hipMemcpy(cl.data(), clusters, cl.size() * sizeof(int32_t), hipMemcpyDefault) == hipSuccess &&
            hipMemcpy(mk.data(), masks, mk.size(), hipMemcpyDefault) == hipSuccess) {
            std::vector<int64_t> members((size_t)num_clusters, 0);
            for (int32_t c : cl)
                if (c >= 0 && c < num_clusters) ++members[(size_t)c];
            for (int c = 0; c < num_clusters; ++c) {
                if (!members[(size_t)c]) continue;
                int64_t pc = 0;
                for (int64_t j = 0; j < vec_size * elem_bytes; ++j)
                    pc += __builtin_popcount(mk[(size_t)c * vec_size * elem_bytes + j]);
                const double bits = (double)vec_size * 8.0 * elem_bytes;
                dprintf("Centroid %d: Set bits %lld of %.0f (%f%%) for %lld nodes\n", c, (long long)pc, bits,
                        (double)pc * 100.0 / bits, (long long)members[(size_t)c]);
            }
        } else {
            (void)hipGetLastError();
        }
        dprintf("Finished clustering\n");
    }
    return st;
}

extern "C" int ibp_selftest_rand(uint32_t seed, int64_t n)
{
    IBP_CHECK(n >= 0, "n must be >= 0");
    LibcRandSequence seq(seed);
    srand(seed);
    for (int64_t i = 0; i < n; ++i) {
        const int want = rand(), got = seq.next();
        IBP_CHECK(want == got, "draw %lld after srand(%u): rand() gave %d, the private sequence %d", (long long)i,
                  seed, want, got);
    }
    return IBP_OK;
}
