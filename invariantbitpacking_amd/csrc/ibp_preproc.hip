// ibp_preproc.hip -- mask/bitval selection (preprocess).
//
// Replaces src/preprocess.cu:17-76 and ibp::preproc_data
// (include/preproc/ibp_preproc_host.cuh:24-115) with its kernels count_bit_kernel,
// create_mask and check_feats (include/preproc/ibp_preproc_kernels.cuh:8-77).
//
// Differences by design (values are identical):
//  * all thresholds of the sweep are scored in ONE pass over the data instead
//    of one pass plus a host synchronisation per threshold;
//  * the choice (first strict maximum, compared as double) is made on the
//    device, so the call is asynchronous and never copies to the host.
#include <type_traits>
#include <vector>

#include "ibp_bits.h"
#include "ibp_internal.h"
#include "ibp_tile.h"
#include "ibp_wave.h"

namespace ibp {

constexpr int kMaxThresholds = 8;

struct Thresholds {
    float t[kMaxThresholds];
    int count;
};

template <typename T>
using WideT = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;

// Column-bit counts of the row block blockIdx.x and the 64-column tile
// blockIdx.y: part[blockIdx.x][k*V + j] = rows of the block whose column j has
// bit k set.  Lane <-> column; the block's waves split the rows; every thread
// keeps its B counters in registers.  Rows of fewer than 32 elements take
// G = 64 / V rows per load (lane l: column l % V of row l / V, so a load reads
// G * V consecutive elements), and the G lanes of a column meet in the LDS sum.  Rows are taken 16 at a time (16 loads in
// flight) and added into bit-sliced counters (VCounter), folded into the B
// counters every 240 rows.  The waves' counters are summed in LDS and stored
// once per block (no global atomics: every block writes its own partial
// counts, reduce_counts_kernel adds them up).
template <typename T>
__global__ __launch_bounds__(256) void count_bits_kernel(const T *data, int64_t num_vecs, int64_t V,
                                                         int64_t rows_per_block, int32_t *part)
{
    constexpr int B = 8 * (int)sizeof(T);
    constexpr int CHUNK = 16;
    using W = WideT<T>;
    __shared__ int32_t sum[B * kWave];
    const int lane = lane_id();
    const int wave_in_block = wave_in_block_id();
    const int waves = blockDim.x / kWave;
    const int G = V < kWave / 2 ? (int)(kWave / V) : 1; // rows per load
    const int lcol = G > 1 ? lane % (int)V : lane;
    const int lrow = G > 1 ? lane / (int)V : 0;
    const int64_t j = (int64_t)blockIdx.y * kWave + lcol;
    const bool col = j < V && lrow < G;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    int64_t r1 = r0 + rows_per_block;
    if (r1 > num_vecs) r1 = num_vecs;
    for (int i = threadIdx.x; i < B * kWave; i += blockDim.x) sum[i] = 0;
    int32_t c[B];
#pragma unroll
    for (int k = 0; k < B; ++k) c[k] = 0;
    VCounter<W, B> vc;
    const int64_t step = (int64_t)waves * G; // rows between a wave's loads
    for (int64_t r = r0 + (int64_t)wave_in_block * G; r < r1; r += CHUNK * step) {
        W x[CHUNK];
#pragma unroll
        for (int q = 0; q < CHUNK; ++q) {
            const int64_t rr = r + (int64_t)q * step + lrow;
            x[q] = (col && rr < r1) ? (W)data[rr * V + j] : W(0);
        }
        vc.add16(x, c);
    }
    vc.flush(c);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < B; ++k)
        if (c[k]) atomicAdd(&sum[k * kWave + lcol], c[k]);
    __syncthreads();
    int32_t *out = part + (int64_t)blockIdx.x * V * B;
    for (int i = threadIdx.x; i < B * kWave; i += blockDim.x) {
        const int64_t jj = (int64_t)blockIdx.y * kWave + (i & (kWave - 1));
        if (jj < V) out[(int64_t)(i / kWave) * V + jj] = sum[i];
    }
}

// cnt[j*B + k] = sum over the row blocks of part[b][k*V + j].  A block of
// 1024 threads owns 64 consecutive counts (lane = count): its 16 waves each
// sum every 16th row block, 4 loads in flight, and the 16 partial sums are
// added in LDS.  (One thread per count walking all row blocks serially was
// latency-bound: 294 us for 3200 counts x 1024 blocks.)
constexpr int kReduceWaves = 16;
__global__ __launch_bounds__(1024) void reduce_counts_kernel(const int32_t *part, int64_t blocks, int64_t V, int B,
                                                             int32_t *cnt)
{
    __shared__ int32_t sums[kReduceWaves][kWave];
    const int64_t n = V * B;
    const int lane = lane_id();
    const int wv = wave_in_block_id();
    const int64_t i = (int64_t)blockIdx.x * kWave + lane;
    const bool ok = i < n;
    int32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int64_t b = wv;
    for (; b + 3 * kReduceWaves < blocks; b += 4 * kReduceWaves) {
        if (ok) {
            a0 += part[b * n + i];
            a1 += part[(b + kReduceWaves) * n + i];
            a2 += part[(b + 2 * kReduceWaves) * n + i];
            a3 += part[(b + 3 * kReduceWaves) * n + i];
        }
    }
    for (; b < blocks; b += kReduceWaves)
        if (ok) a0 += part[b * n + i];
    sums[wv][lane] = a0 + a1 + a2 + a3;
    __syncthreads();
    if (wv == 0 && ok) {
        int32_t acc = 0;
#pragma unroll
        for (int q = 0; q < kReduceWaves; ++q) acc += sums[q][lane];
        const int64_t k = i / V, j = i % V;
        cnt[j * B + k] = acc;
    }
}

// masks[t*V + j], bitvals[t*V + j] for every threshold t
// (ibp_preproc_kernels.cuh:34-51): bit in mask and bitval if
// float(cnt) > float(t) * float(N); in the mask only if
// double(cnt) < (1.0 - double(t)) * double(N).
//
// Also writes the column-major copy colm/colb[j*kMaxThresholds + t] that the
// row-per-lane score kernel reads with scalar loads.
template <typename T>
__global__ void create_masks_kernel(const int32_t *cnt, int64_t num_vecs, int64_t V, Thresholds th, T *masks,
                                    T *bitvals, T *colm, T *colb, uint32_t *colp)
{
    constexpr int B = 8 * (int)sizeof(T);
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < kMaxThresholds) { // the zero column V after the last
        colm[V * kMaxThresholds + gid] = T(0);
        colb[V * kMaxThresholds + gid] = T(0);
        colp[V * kMaxThresholds + gid] = 0;
    }
    if (gid >= V * kMaxThresholds) return;
    const int t = (int)(gid / V);
    const int64_t j = gid % V;
    if (t >= th.count) { // unused thresholds: zero masks in the column-major copy
        colm[j * kMaxThresholds + t] = T(0);
        colb[j * kMaxThresholds + t] = T(0);
        colp[j * kMaxThresholds + t] = 0;
        return;
    }
    const float thr = th.t[t];
    const float hi = __fmul_rn(thr, (float)(uint64_t)num_vecs);
    const double lo = __dmul_rn(1.0 - (double)thr, (double)(uint64_t)num_vecs);
    uint64_t val = 0, masker = 0;
    for (int k = 0; k < B; ++k) {
        const int32_t c = cnt[j * B + k];
        if ((float)c > hi) {
            val |= 1ull << k;
            masker |= 1ull << k;
        } else if ((double)c < lo) {
            masker |= 1ull << k;
        }
    }
    masks[gid] = (T)masker;
    bitvals[gid] = (T)val;
    colm[j * kMaxThresholds + t] = (T)masker;
    colb[j * kMaxThresholds + t] = (T)val;
    colp[j * kMaxThresholds + t] = (uint32_t)__builtin_popcountll((uint64_t)(T)masker);
}

// scores[t] += sum over rows of max(0, S_t(row) - V), S_t = bits of the
// elements flagged under threshold t's mask (ibp_preproc_kernels.cuh:54-77).
template <typename T, int NT>
__global__ __launch_bounds__(256) void score_kernel(const T *data, int64_t num_vecs, int64_t V, const T *masks,
                                                    const T *bitvals, int nt, unsigned long long *scores)
{
    const int lane = lane_id();
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    unsigned long long saved[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) saved[t] = 0;
    if (V <= kWave) {
        // Rows of <= 64 elements: every lane keeps its column's masks in
        // registers; 4 rows in flight per wave; two thresholds share one
        // 32-bit reduction (a row's sum is at most 64 * 64 < 2^16).
        constexpr int R = 4;
        constexpr int NP = (NT + 1) / 2;
        T m_reg[NT], b_reg[NT];
        uint32_t pc_reg[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bool ok = t < nt && lane < V;
            m_reg[t] = ok ? masks[t * V + lane] : T(0);
            b_reg[t] = ok ? bitvals[t * V + lane] : T(0);
            // a lane past the row (or a threshold past nt) never adds anything
            pc_reg[t] = ok ? (uint32_t)popc((WideT<T>)m_reg[t]) << (16 * (t & 1)) : 0u;
        }
        for (int64_t r = wave * R; r < num_vecs; r += nwaves * R) {
            T x[R];
#pragma unroll
            for (int i = 0; i < R; ++i) x[i] = (r + i < num_vecs && lane < V) ? data[(r + i) * V + lane] : T(0);
#pragma unroll
            for (int i = 0; i < R; ++i) {
                if (r + i >= num_vecs) break; // uniform
                uint32_t acc[NP];
#pragma unroll
                for (int q = 0; q < NP; ++q) acc[q] = 0;
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    if ((T)(x[i] & m_reg[t]) == b_reg[t]) acc[t >> 1] += pc_reg[t];
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const uint32_t sum = (uint32_t)wave_sum((int)acc[q]);
                    const int64_t s0 = sum & 0xffffu, s1 = sum >> 16;
                    if (s0 > V) saved[2 * q] += (unsigned long long)(s0 - V);
                    if (2 * q + 1 < NT && s1 > V) saved[2 * q + 1] += (unsigned long long)(s1 - V);
                }
            }
        }
    } else {
        for (int64_t r = wave; r < num_vecs; r += nwaves) {
            int acc[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = 0;
            for (int64_t j = lane; j < V; j += kWave) {
                const T x = data[r * V + j];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    if (t >= nt) break;
                    const T m = masks[t * V + j];
                    if ((T)(x & m) == bitvals[t * V + j]) acc[t] += popc((WideT<T>)m);
                }
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (t >= nt) break;
                const int64_t s = wave_sum(acc[t]);
                if (s > V) saved[t] += (unsigned long long)(s - V);
            }
        }
    }
    if (lane == 0)
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (t < nt && saved[t]) atomicAdd(&scores[t], saved[t]);
}

// The same scores with lane <-> row (4- and 8-byte elements, rows of a whole
// number of 16-byte chunks): each lane streams its own row through registers,
// four chunks at a time with the next four in flight, and a column's masks
// are uniform scalar loads (colm/colb, column-major).  Per row and threshold
// it accumulates the popcounts of the columns that are NOT flagged, so
// S_t = (sum of popcounts) - lost_t needs no cross-lane reduction.  A
// threshold whose mask equals the previous one reuses its test.
// NESTED: every threshold's bitval is bitval[0] & mask[t] (true for one
// threshold, and for an ascending list of thresholds >= 0.5, where the masks
// shrink as t grows), so one XOR per element serves all thresholds.
template <typename T>
__device__ __forceinline__ uint32_t any_bit(T u)
{
    if constexpr (sizeof(T) == 8) return (uint32_t)u | (uint32_t)((uint64_t)u >> 32);
    else return (uint32_t)u;
}

template <typename T, int NT, bool NESTED>
__global__ __launch_bounds__(256) void score_rows_kernel(const T *__restrict__ data, int64_t num_vecs, int V,
                                                         const T *__restrict__ colm, const T *__restrict__ colb,
                                                         int nt, unsigned long long *scores)
{
    constexpr int EPC = 16 / (int)sizeof(T); // elements per 16-byte chunk
    constexpr int G = 4;                     // chunks per group
    __shared__ unsigned long long bsum[NT];
    if (threadIdx.x < NT) bsum[threadIdx.x] = 0;
    __syncthreads();
    const int lane = lane_id();
    const int64_t wave = global_wave_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    const int chunks = V / EPC;
    // pcsum[t] = flagged bits of a row whose every column is flagged
    uint32_t pcsum[NT];
    unsigned long long saved[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        pcsum[t] = 0;
        saved[t] = 0;
    }
    for (int j = lane; j < V; j += kWave)
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (t < nt) pcsum[t] += (uint32_t)popc((WideT<T>)colm[(int64_t)j * kMaxThresholds + t]);
#pragma unroll
    for (int t = 0; t < NT; ++t) pcsum[t] = (uint32_t)wave_sum((int)pcsum[t]);
    for (int64_t r0 = wave * kWave; r0 < num_vecs; r0 += nwaves * kWave) {
        const int64_t r = r0 + lane;
        const bool ok = r < num_vecs;
        const uint4 *row = reinterpret_cast<const uint4 *>(data + (ok ? r : r0) * (int64_t)V);
        uint32_t lost[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) lost[t] = 0;
        // Chunk loads are unconditional (indices clamped to the row) so the
        // next group's loads stay in flight while this group is scored.
        uint4 cur[G], nxt[G];
#pragma unroll
        for (int c = 0; c < G; ++c) cur[c] = row[min(c, chunks - 1)];
        for (int c0 = 0; c0 < chunks; c0 += G) {
#pragma unroll
            for (int c = 0; c < G; ++c) nxt[c] = row[min(c0 + G + c, chunks - 1)];
#pragma unroll
            for (int c = 0; c < G; ++c) {
                if (c0 + c >= chunks) break; // uniform
#pragma unroll
                for (int e = 0; e < EPC; ++e) {
                    const int64_t j = (int64_t)(c0 + c) * EPC + e;
                    const T x = chunk_elem<T>(cur[c], e);
                    // the column's masks (zero past nt) and bitvals: scalar
                    // loads issued together
                    const T *cm = colm + j * kMaxThresholds;
                    const T *cb = colb + j * kMaxThresholds;
                    T m[NT];
#pragma unroll
                    for (int t = 0; t < NT; ++t) m[t] = cm[t];
                    const T d0 = (T)(x ^ cb[0]);
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        if (m[t] == 0) continue; // uniform: no bits, nothing to lose
                        const T d = NESTED ? d0 : (T)(x ^ cb[t]);
                        lost[t] = lost_step(d, m[t], (uint32_t)popc((WideT<T>)m[t]), lost[t]);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < G; ++c) cur[c] = nxt[c];
        }
        if (ok) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const uint32_t st = pcsum[t] - lost[t];
                if (t < nt && st > (uint32_t)V) saved[t] += (unsigned long long)(st - (uint32_t)V);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
        if (t < nt && saved[t]) atomicAdd(&bsum[t], saved[t]);
    __syncthreads();
    if ((int)threadIdx.x < nt && threadIdx.x < NT && bsum[threadIdx.x]) atomicAdd(&scores[threadIdx.x], bsum[threadIdx.x]);
}

// Row-per-lane scoring from LDS copies of 64-row tiles.  A tile's rows are
// contiguous in memory, so it is copied to LDS with C = row bytes / 16
// instructions of global_load_lds_dwordx4, each moving 1 KB of consecutive
// bytes straight to LDS (no registers; tracked by vmcnt).  Lane r then reads
// row r from LDS chunk by chunk, with the column masks uniform scalar loads.
// (Loading row r straight into lane r makes every load instruction touch 64
// rows 400 bytes apart, so each 128-byte line is fetched again by later
// loads.)
//
// The WPT waves of a block share the tile and split the thresholds, TPW
// each: a tile holds 64 rows whatever the number of waves scoring it, so
// four waves per tile give four times the waves per CU that one wave per
// tile would, and each does a quarter of the vector work (GNN table, 7
// thresholds: 237 us at two per wave, 280 at four, 309 at seven).  The next
// tile is loaded into registers (about 7 chunks per lane) while the current
// one is scored, and stored to LDS after it: one LDS tile per block, and the
// loads overlap the scoring (a second LDS tile filled with copy_tile did not
// overlap: the compiler waits for every pending LDS copy before any LDS read).
// Used where the row stride in dwords is not a multiple of 16 modulo 32, so
// eight lanes reading 16 bytes each hit at most two bank groups twice.

// The column constants one wave of score_tile_kernel needs for one 16-byte
// chunk of a row (EPC elements): masks and their popcounts for its TPW
// thresholds, and the bitvals (one per element when the thresholds are
// nested, else one per threshold).  Records are laid out [wave][chunk], with a
// zero record after the last chunk, so a chunk's constants are three uniform
// scalar loads from one pointer.
template <typename T, int TPW, int NB>
struct ScoreRec {
    static constexpr int EPC = 16 / (int)sizeof(T);
    T m[EPC][TPW];
    uint32_t p[EPC][TPW];
    T b[EPC][NB];
};

// Fills the records of WPT waves for C chunks (plus the zero record) from the
// column-major tables; thread = (wave, chunk, element, threshold slot).
template <typename T, int TPW, int NB>
__global__ void score_recs_kernel(const T *__restrict__ colm, const T *__restrict__ colb,
                                  const uint32_t *__restrict__ colp, int64_t V, int C, int WPT,
                                  ScoreRec<T, TPW, NB> *recs)
{
    constexpr int EPC = ScoreRec<T, TPW, NB>::EPC;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)WPT * (C + 1) * EPC * TPW) return;
    const int t = (int)(gid % TPW);
    const int e = (int)((gid / TPW) % EPC);
    const int64_t c = (gid / ((int64_t)TPW * EPC)) % (C + 1);
    const int w = (int)(gid / ((int64_t)TPW * EPC * (C + 1)));
    const int64_t j = c * EPC + e;
    const int tt = w * TPW + t;
    const bool ok = j < V && tt < kMaxThresholds;
    ScoreRec<T, TPW, NB> &r = recs[(int64_t)w * (C + 1) + c];
    r.m[e][t] = ok ? colm[j * kMaxThresholds + tt] : T(0);
    r.p[e][t] = ok ? colp[j * kMaxThresholds + tt] : 0u;
    if (NB == 1) {
        if (t == 0) r.b[e][0] = j < V ? colb[j * kMaxThresholds] : T(0);
    } else {
        r.b[e][NB == 1 ? 0 : t] = ok ? colb[j * kMaxThresholds + tt] : T(0);
    }
}

// Loads one record as plain element loads from the kernel's read-only
// argument (a struct copy would be lowered to vector loads).
template <typename T, int TPW, int NB>
__device__ __forceinline__ void load_rec(const ScoreRec<T, TPW, NB> *__restrict__ r, T (&m)[ScoreRec<T, TPW, NB>::EPC][TPW],
                                         uint32_t (&p)[ScoreRec<T, TPW, NB>::EPC][TPW],
                                         T (&b)[ScoreRec<T, TPW, NB>::EPC][NB])
{
    constexpr int EPC = ScoreRec<T, TPW, NB>::EPC;
    const T *rm = reinterpret_cast<const T *>(r);
    const uint32_t *rp = reinterpret_cast<const uint32_t *>(rm + EPC * TPW);
    const T *rb = reinterpret_cast<const T *>(rp + EPC * TPW);
#pragma unroll
    for (int e = 0; e < EPC; ++e)
#pragma unroll
        for (int t = 0; t < TPW; ++t) m[e][t] = rm[e * TPW + t];
#pragma unroll
    for (int e = 0; e < EPC; ++e)
#pragma unroll
        for (int t = 0; t < TPW; ++t) p[e][t] = rp[e * TPW + t];
#pragma unroll
    for (int e = 0; e < EPC; ++e)
#pragma unroll
        for (int t = 0; t < NB; ++t) b[e][t] = rb[e * NB + t];
}

// Empty asm that needs the record's values in SGPRs: the wait for their
// scalar loads is placed here.
template <typename T, int TPW, int NB>
__device__ __forceinline__ void need_rec(const T (&m)[ScoreRec<T, TPW, NB>::EPC][TPW],
                                         const uint32_t (&p)[ScoreRec<T, TPW, NB>::EPC][TPW],
                                         const T (&b)[ScoreRec<T, TPW, NB>::EPC][NB])
{
    constexpr int EPC = ScoreRec<T, TPW, NB>::EPC;
#pragma unroll
    for (int e = 0; e < EPC; ++e) {
#pragma unroll
        for (int t = 0; t < TPW; ++t) asm volatile("" ::"s"(m[e][t]), "s"(p[e][t]));
#pragma unroll
        for (int t = 0; t < NB; ++t) asm volatile("" ::"s"(b[e][t]));
    }
}

// TPW thresholds per wave (t0 = wave * TPW), WPT waves per tile; RP: the
// next tile is loaded into registers while the current one is scored from
// LDS, then stored to LDS (otherwise each tile is copied to LDS directly and
// waited for before it is scored).  Every threshold of a wave is scored
// without a branch (a zero mask loses nothing), with the mask popcounts from
// the records.
template <typename T, int TPW, int WPT, bool RP, bool NESTED>
__global__ __launch_bounds__(64 * WPT) void score_tile_kernel(const T *__restrict__ data, int64_t num_vecs, int V,
                                                              const ScoreRec<T, TPW, NESTED ? 1 : TPW> *__restrict__ recs,
                                                              const uint32_t *__restrict__ colp, int nt,
                                                              unsigned long long *scores)
{
    constexpr int EPC = 16 / (int)sizeof(T);
    constexpr int NB = NESTED ? 1 : TPW;
    extern __shared__ __align__(16) unsigned char tiles[];
    const int lane = lane_id();
    const int wib = wave_in_block_id();
    const int t0 = wib * TPW; // this wave's first threshold
    const int C = V * (int)sizeof(T) / 16;
    const int64_t ntiles = (num_vecs + kWave - 1) / kWave;
    const int64_t nblocks = gridDim.x;
    const bool swz = tile_swizzled((int64_t)C * 16);
    const int sw = tile_swz(lane, swz); // chunk c of this lane's row sits at c ^ sw
    const ScoreRec<T, TPW, NB> *rec = recs + (size_t)wib * (C + 1);
    uint32_t pcsum[TPW];
    unsigned long long saved[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        pcsum[t] = 0;
        saved[t] = 0;
    }
    for (int j = lane; j < V; j += kWave)
#pragma unroll
        for (int t = 0; t < TPW; ++t)
            if (t0 + t < nt) pcsum[t] += colp[(int64_t)j * kMaxThresholds + t0 + t];
#pragma unroll
    for (int t = 0; t < TPW; ++t) pcsum[t] = (uint32_t)wave_sum((int)pcsum[t]);
    const bool active = t0 < nt; // waves past the last threshold only copy
    constexpr int NPRE = RP ? (kTileMaxRowBytes / 16 + WPT - 1) / WPT : 1;
    tile_chunk_t pre[NPRE];
    int64_t tl = blockIdx.x;
    if (RP && tl < ntiles) load_tile_regs<T, WPT, NPRE>(data, num_vecs, V, C, tl, wib, lane, swz, pre);
    for (; tl < ntiles; tl += nblocks) {
        if (RP) {
            store_tile_regs<WPT, NPRE>(tiles, C, wib, lane, pre);
            __syncthreads(); // every wave's part of the tile is in LDS
            if (tl + nblocks < ntiles)
                load_tile_regs<T, WPT, NPRE>(data, num_vecs, V, C, tl + nblocks, wib, lane, swz, pre);
        } else {
            copy_tile<T, WPT>(data, num_vecs, V, C, tl, wib, lane, tiles, swz);
            __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0): this wave's part of the tile has landed
            __syncthreads();                    // ... and every other wave's part
        }
        const unsigned char *buf = tiles;
        if (active) {
            const uint4 *mine = reinterpret_cast<const uint4 *>(buf) + lane * C;
            const int64_t row = tl * kWave + lane;
            uint32_t lost[TPW];
#pragma unroll
            for (int t = 0; t < TPW; ++t) lost[t] = 0;
            // Chunk c + 1's constants (three scalar loads) and its 16 bytes
            // of the row (one LDS read) are issued before chunk c is scored.
            // Scalar loads complete in any order, so their only wait is
            // lgkmcnt(0): the empty asm statements of need_rec place it at
            // the start of chunk c + 1, and the ordered steps keep chunk c's
            // work ahead of it.  Two register sets alternate, so nothing is
            // moved between chunks.
            T ma[EPC][TPW], mb[EPC][TPW], ba[EPC][NB], bb[EPC][NB];
            uint32_t pa[EPC][TPW], pb[EPC][TPW];
            auto score = [&](const uint4 &v, const T(&m)[EPC][TPW], const uint32_t(&pc)[EPC][TPW],
                             const T(&bv)[EPC][NB]) {
#pragma unroll
                for (int e = 0; e < EPC; ++e) {
                    const T x = chunk_elem<T>(v, e);
                    const T d0 = xor_ordered(x, bv[e][0]);
#pragma unroll
                    for (int t = 0; t < TPW; ++t) {
                        const T d = NESTED ? d0 : xor_ordered(x, bv[e][NB == 1 ? 0 : t]);
                        lost[t] = lost_step<true>(d, m[e][t], pc[e][t], lost[t]);
                    }
                }
            };
            load_rec<T, TPW, NB>(rec, ma, pa, ba);
            uint4 va = mine[sw], vb;
            int c = 0;
            for (; c + 1 < C; c += 2) {
                need_rec<T, TPW, NB>(ma, pa, ba);
                asm volatile("" ::"v"(va.x), "v"(va.y), "v"(va.z), "v"(va.w));
                load_rec<T, TPW, NB>(rec + c + 1, mb, pb, bb);
                vb = mine[(c + 1) ^ sw];
                score(va, ma, pa, ba);
                need_rec<T, TPW, NB>(mb, pb, bb);
                asm volatile("" ::"v"(vb.x), "v"(vb.y), "v"(vb.z), "v"(vb.w));
                load_rec<T, TPW, NB>(rec + c + 2, ma, pa, ba); // record C is the zero one
                va = mine[(c + 2 < C ? c + 2 : c + 1) ^ sw];
                score(vb, mb, pb, bb);
            }
            if (c < C) { // odd C: the last chunk
                need_rec<T, TPW, NB>(ma, pa, ba);
                score(va, ma, pa, ba);
            }
            if (row < num_vecs) {
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    const uint32_t sv = pcsum[t] - lost[t];
                    if (t0 + t < nt && sv > (uint32_t)V) saved[t] += (unsigned long long)(sv - (uint32_t)V);
                }
            }
        }
        __syncthreads(); // this tile's reads are done before the next one is stored
    }
    // One atomic per threshold and wave: the lanes' savings are summed as
    // 24-bit low parts and the rest (each sum of 64 fits 32 bits as long as a
    // lane's saving is below 2^49).
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const uint32_t lo = (uint32_t)wave_sum((int)(uint32_t)(saved[t] & 0xffffffull));
        const uint32_t hi = (uint32_t)wave_sum((int)(uint32_t)(saved[t] >> 24));
        const unsigned long long tot = (unsigned long long)lo + ((unsigned long long)hi << 24);
        if (lane == 0 && t0 + t < nt && tot) atomicAdd(&scores[t0 + t], tot);
    }
}

// Launch of score_tile_kernel for `nt` thresholds (1 or a sweep of up to
// kMaxThresholds): IBP_SCORE_SWEEP_TPW thresholds per wave and as many waves
// per tile as the sweep needs, one wave for one threshold.  The records go to
// `recbuf` (score_recs_bytes), written by score_recs_kernel first.
#ifndef IBP_SCORE_SWEEP_TPW
#define IBP_SCORE_SWEEP_TPW 2
#endif
#ifndef IBP_SCORE_RP
#define IBP_SCORE_RP 1
#endif
constexpr int kScoreTPW = IBP_SCORE_SWEEP_TPW;
constexpr int kScoreWPT = (kMaxThresholds - 1 + kScoreTPW - 1) / kScoreTPW;

template <typename T, int TPW, int WPT, bool NESTED>
void launch_score_tile_cfg(const T *data, int64_t N, int64_t V, const T *colm, const T *colb, const uint32_t *colp,
                           int nt, unsigned long long *scores, void *recbuf, int64_t grid, size_t lds,
                           hipStream_t stream)
{
    constexpr int NB = NESTED ? 1 : TPW;
    constexpr bool RP = IBP_SCORE_RP != 0;
    using Rec = ScoreRec<T, TPW, NB>;
    const int C = (int)(V * (int64_t)sizeof(T) / 16);
    const int64_t threads = (int64_t)WPT * (C + 1) * Rec::EPC * TPW;
    hipLaunchKernelGGL((score_recs_kernel<T, TPW, NB>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream,
                       colm, colb, colp, V, C, WPT, (Rec *)recbuf);
    hipLaunchKernelGGL((score_tile_kernel<T, TPW, WPT, RP, NESTED>), dim3((unsigned)grid), dim3(kWave * WPT), lds,
                       stream, data, N, (int)V, (const Rec *)recbuf, colp, nt, scores);
}

// Bytes of records the launch below needs for rows of V elements of w bytes.
size_t score_recs_bytes(int w, int64_t V)
{
    auto need = [&](auto tag) -> size_t {
        using T = decltype(tag);
        const int64_t C = (V * (int64_t)sizeof(T) + 15) / 16;
        const size_t one = (size_t)(C + 1) * sizeof(ScoreRec<T, 1, 1>);
        const size_t sweep = (size_t)kScoreWPT * (C + 1) * sizeof(ScoreRec<T, kScoreTPW, kScoreTPW>);
        return one > sweep ? one : sweep;
    };
    if (w == 8) return need(uint64_t());
    if (w == 4) return need(uint32_t());
    return 0;
}

template <typename T>
void launch_score_tile(const T *data, int64_t N, int64_t V, const T *colm, const T *colb, const uint32_t *colp,
                       int nt, bool nested, unsigned long long *scores, void *recbuf, int num_cus, hipStream_t stream)
{
    const size_t lds = (size_t)kWave * (size_t)V * sizeof(T);
    const int64_t ntiles = (N + kWave - 1) / kWave;
    int64_t per_cu = (int64_t)(160 * 1024) / (int64_t)lds;
    if (per_cu > 8) per_cu = 8;
    if (per_cu < 1) per_cu = 1;
    int64_t grid = (int64_t)num_cus * per_cu;
    if (grid > ntiles) grid = ntiles;
    if (nt <= 1)
        launch_score_tile_cfg<T, 1, 1, true>(data, N, V, colm, colb, colp, nt, scores, recbuf, grid, lds, stream);
    else if (nested)
        launch_score_tile_cfg<T, kScoreTPW, kScoreWPT, true>(data, N, V, colm, colb, colp, nt, scores, recbuf, grid,
                                                             lds, stream);
    else
        launch_score_tile_cfg<T, kScoreTPW, kScoreWPT, false>(data, N, V, colm, colb, colp, nt, scores, recbuf, grid,
                                                              lds, stream);
}

// First strict maximum of the scores compared as double
// (ibp_preproc_host.cuh:64,92-99); copies that threshold's mask/bitval out.
template <typename T>
__global__ void select_kernel(const unsigned long long *scores, int nt, const T *masks, const T *bitvals,
                              int64_t V, T *mask_out, T *bitval_out, int32_t *chosen)
{
    int best = -1;
    double max_saved = 0;
    for (int t = 0; t < nt; ++t)
        if ((double)scores[t] > max_saved) {
            max_saved = (double)scores[t];
            best = t;
        }
    if (chosen && blockIdx.x == 0 && threadIdx.x == 0) *chosen = best;
    if (best < 0) return;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < V; j += (int64_t)gridDim.x * blockDim.x) {
        mask_out[j] = masks[best * V + j];
        bitval_out[j] = bitvals[best * V + j];
    }
}

__global__ void zero_u64_kernel(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The thresholds the reference tries (ibp_preproc_host.cuh:61-67): only
// `threshold` if it is >= 0, else 0.7f, then t = float(double(t) + 0.05) while
// t <= 1.0f.
Thresholds threshold_list(float threshold)
{
    Thresholds th;
    th.count = 0;
    const float lo = threshold < 0 ? 0.7f : threshold;
    const float hi = threshold < 0 ? 1.0f : threshold;
    for (volatile float t = lo; t <= hi && th.count < kMaxThresholds; t = (float)((double)t + 0.05))
        th.t[th.count++] = t;
    return th;
}

// One threshold, or an ascending list of thresholds >= 0.5: then a bit in
// threshold t's mask is in every earlier mask with the same value (a count
// above t*N is above every smaller t*N; one below (1-t)*N <= N/2 is below
// every larger (1-t)*N and cannot exceed t'*N >= N/2), so
// bitval[t] = bitval[0] & mask[t].
bool thresholds_nested(const Thresholds &th)
{
    if (th.count <= 1) return true;
    for (int i = 0; i < th.count; ++i)
        if (th.t[i] < 0.5f || (i > 0 && th.t[i] <= th.t[i - 1])) return false;
    return true;
}

namespace {

struct PreWorkspace {
    size_t cnt_off, masks_off, bitvals_off, colm_off, colb_off, colp_off, recs_off, scores_off, part_off, total;
};

// Row blocks of count_bits_kernel: about 1024 blocks in all (four per CU of a
// 256-CU device), at least one per 64-column tile, at least 64 rows each.
constexpr int64_t kCountBlocks = 1024;
int64_t count_row_blocks(int64_t N, int64_t V, int64_t *rows_per_block)
{
    const int64_t tiles = (V + kWave - 1) / kWave;
    int64_t want = kCountBlocks / tiles;
    if (want < 1) want = 1;
    int64_t rpb = (N + want - 1) / want;
    if (rpb < 64) rpb = 64;
    if (rows_per_block) *rows_per_block = rpb;
    return N > 0 ? (N + rpb - 1) / rpb : 0;
}
int64_t max_count_row_blocks(int64_t V)
{
    const int64_t tiles = (V + kWave - 1) / kWave;
    return kCountBlocks / tiles > 1 ? kCountBlocks / tiles : 1;
}

PreWorkspace layout(int w, int64_t V, int nt)
{
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    PreWorkspace L;
    L.cnt_off = 0;
    L.masks_off = al(L.cnt_off + sizeof(int32_t) * (size_t)V * 8 * w);
    L.bitvals_off = al(L.masks_off + (size_t)w * V * nt);
    L.colm_off = al(L.bitvals_off + (size_t)w * V * nt);
    L.colb_off = al(L.colm_off + (size_t)w * (V + 1) * kMaxThresholds);
    L.colp_off = al(L.colb_off + (size_t)w * (V + 1) * kMaxThresholds);
    L.recs_off = al(L.colp_off + sizeof(uint32_t) * (size_t)(V + 1) * kMaxThresholds);
    L.scores_off = al(L.recs_off + score_recs_bytes(w, V));
    L.part_off = al(L.scores_off + sizeof(unsigned long long) * kMaxThresholds);
    L.total = al(L.part_off + sizeof(int32_t) * (size_t)max_count_row_blocks(V) * V * 8 * w);
    return L;
}

template <typename T>
int preprocess_typed(const void *data, int64_t N, int64_t V, const Thresholds &th, void *mask, void *bitval,
                     int32_t *chosen, unsigned char *ws, hipStream_t stream)
{
    const DeviceInfo &dev = device_info();
    const PreWorkspace L = layout((int)sizeof(T), V, th.count);
    int32_t *cnt = (int32_t *)(ws + L.cnt_off);
    T *masks = (T *)(ws + L.masks_off);
    T *bitvals = (T *)(ws + L.bitvals_off);
    T *colm = (T *)(ws + L.colm_off);
    T *colb = (T *)(ws + L.colb_off);
    uint32_t *colp = (uint32_t *)(ws + L.colp_off);
    unsigned long long *scores = (unsigned long long *)(ws + L.scores_off);
    int32_t *part = (int32_t *)(ws + L.part_off);
    const int64_t ncnt = V * 8 * (int64_t)sizeof(T);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, stream, scores, kMaxThresholds);
    IBP_LAUNCH_CHECK();
    int64_t rows_per_block = 0;
    const int64_t row_blocks = count_row_blocks(N, V, &rows_per_block);
    if (row_blocks > 0) {
        const int64_t tiles = (V + kWave - 1) / kWave;
        hipLaunchKernelGGL((count_bits_kernel<T>), dim3((unsigned)row_blocks, (unsigned)tiles), dim3(256), 0, stream,
                           (const T *)data, N, V, rows_per_block, part);
        IBP_LAUNCH_CHECK();
    }
    const int64_t rgrid = (ncnt + kWave - 1) / kWave;
    hipLaunchKernelGGL(reduce_counts_kernel, dim3((unsigned)rgrid), dim3(kWave * kReduceWaves), 0, stream,
                       (const int32_t *)part, row_blocks, V, 8 * (int)sizeof(T), cnt);
    IBP_LAUNCH_CHECK();
    const int64_t mthreads = V * kMaxThresholds;
    hipLaunchKernelGGL((create_masks_kernel<T>), dim3((unsigned)((mthreads + 255) / 256)), dim3(256), 0, stream, cnt, N,
                       V, th, masks, bitvals, colm, colb, colp);
    IBP_LAUNCH_CHECK();
    // Rows of whole 16-byte chunks of 4- or 8-byte elements: the row-per-lane
    // kernel (about 5x fewer vector instructions per row than lane-per-column
    // with a wave reduction per row).
    bool scored = N == 0;
    if constexpr (sizeof(T) >= 4) {
        // ... for a threshold sweep, or for one threshold when there are rows
        // enough for four 64-row tiles per CU.  With one threshold and fewer,
        // wider rows a wave per row is faster (32768 x 1024 words: 0.25 vs
        // 0.54 ms; the 7-threshold sweep on it: 0.96 vs 0.49 ms).
        const bool rows_path = (V * (int64_t)sizeof(T)) % 16 == 0 && V <= (int64_t)1 << 20 &&
                               ((uintptr_t)data % 16) == 0 &&
                               (th.count > 1 || N >= (int64_t)dev.num_cus * 4 * kWave);
        // Rows of up to 512 bytes whose stride spreads over the LDS banks:
        // 64-row tiles copied to LDS (25.6 KB at 400-byte rows).
        const int64_t row_bytes = V * (int64_t)sizeof(T);
        if (!scored && rows_path && tile_rows_ok(row_bytes) && th.count <= 7) {
            launch_score_tile<T>((const T *)data, N, V, (const T *)colm, (const T *)colb, colp, th.count,
                                 thresholds_nested(th), scores, ws + L.recs_off, dev.num_cus, stream);
            IBP_LAUNCH_CHECK();
            scored = true;
        }
        if (!scored && rows_path) {
            int64_t grid = ((N + kWave - 1) / kWave + 3) / 4;
            if (grid > (int64_t)dev.num_cus * 8) grid = (int64_t)dev.num_cus * 8;
            const bool nested = thresholds_nested(th);
            if (th.count <= 1)
                hipLaunchKernelGGL((score_rows_kernel<T, 1, true>), dim3((unsigned)grid), dim3(256), 0, stream,
                                   (const T *)data, N, (int)V, (const T *)colm, (const T *)colb, th.count, scores);
            else if (nested)
                hipLaunchKernelGGL((score_rows_kernel<T, kMaxThresholds, true>), dim3((unsigned)grid), dim3(256), 0,
                                   stream, (const T *)data, N, (int)V, (const T *)colm, (const T *)colb, th.count,
                                   scores);
            else
                hipLaunchKernelGGL((score_rows_kernel<T, kMaxThresholds, false>), dim3((unsigned)grid), dim3(256), 0,
                                   stream, (const T *)data, N, (int)V, (const T *)colm, (const T *)colb, th.count,
                                   scores);
            IBP_LAUNCH_CHECK();
            scored = true;
        }
    }
    if (!scored) {
        int64_t grid = (N * kWave + 255) / 256;
        if (grid > (int64_t)dev.num_cus * 8) grid = (int64_t)dev.num_cus * 8;
        if (th.count <= 1)
            hipLaunchKernelGGL((score_kernel<T, 1>), dim3((unsigned)grid), dim3(256), 0, stream, (const T *)data, N, V,
                               masks, bitvals, th.count, scores);
        else
            hipLaunchKernelGGL((score_kernel<T, kMaxThresholds>), dim3((unsigned)grid), dim3(256), 0, stream,
                               (const T *)data, N, V, masks, bitvals, th.count, scores);
        IBP_LAUNCH_CHECK();
    }
    unsigned sel_grid = (unsigned)((V + 255) / 256);
    if (sel_grid > 256) sel_grid = 256;
    hipLaunchKernelGGL((select_kernel<T>), dim3(sel_grid), dim3(256), 0, stream, scores, th.count, masks, bitvals, V,
                       (T *)mask, (T *)bitval, chosen);
    IBP_LAUNCH_CHECK();
    return IBP_OK;
}

} // namespace

} // namespace ibp

using namespace ibp;

extern "C" size_t ibp_preprocess_workspace_size(int elem_bytes, int64_t vec_size, float threshold)
{
    return layout(elem_bytes, vec_size, threshold_list(threshold).count).total;
}

extern "C" int ibp_preprocess(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size, float threshold,
                              void *mask, void *bitval, int32_t *chosen_step, void *workspace,
                              size_t workspace_bytes, void *stream)
{
    IBP_TRACE("ibp_preprocess");
    IBP_CHECK(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8,
              "Input tensor must be 1, 2, 4, or 8-byte datatype (got %d)", elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0, "invalid shape");
    IBP_CHECK(mask && bitval && (data || num_vecs == 0), "null pointer argument");
    IBP_CHECK(num_vecs < (int64_t)1 << 31, "num_vecs must fit the int32 bit counters");
    const Thresholds th = threshold_list(threshold);
    const size_t need = layout(elem_bytes, vec_size, th.count).total;
    hipStream_t s = as_stream(stream);
    unsigned char *ws = (unsigned char *)workspace;
    bool own = false;
    if (ws == nullptr) {
        IBP_HIP(hipMallocAsync((void **)&ws, need, s));
        own = true;
    } else {
        IBP_CHECK(workspace_bytes >= need, "workspace too small (%zu < %zu)", workspace_bytes, need);
    }
    int st;
    switch (elem_bytes) {
    case 1: st = preprocess_typed<uint8_t>(data, num_vecs, vec_size, th, mask, bitval, chosen_step, ws, s); break;
    case 2: st = preprocess_typed<uint16_t>(data, num_vecs, vec_size, th, mask, bitval, chosen_step, ws, s); break;
    case 4: st = preprocess_typed<uint32_t>(data, num_vecs, vec_size, th, mask, bitval, chosen_step, ws, s); break;
    default: st = preprocess_typed<uint64_t>(data, num_vecs, vec_size, th, mask, bitval, chosen_step, ws, s); break;
    }
    if (st == IBP_OK && debug_enabled()) {
        // Debug mode reports what the reference's sweep prints
        // (ibp_preproc_host.cuh:66,89-98): the bits each threshold saves, and
        // each new best with its expected compressed length in elements.
        // This synchronises the stream; without debug the call stays
        // asynchronous.
        unsigned long long sc[kMaxThresholds] = {0};
        const PreWorkspace L = layout(elem_bytes, vec_size, th.count);
        if (hipMemcpyAsync(sc, ws + L.scores_off, sizeof(sc), hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess) {
            const double total = (double)num_vecs * (double)vec_size * 8.0 * elem_bytes;
            dprintf("Num elems: %lld\n", (long long)num_vecs);
            double best = 0;
            for (int t = 0; t < th.count; ++t) {
                dprintf("Threshold %.2f: Saved bits per element: %llu (Total %.0f, %.3f%%)\n", th.t[t], sc[t], total,
                        total > 0 ? (double)sc[t] * 100.0 / total : 0.0);
                if ((double)sc[t] > best) {
                    best = (double)sc[t];
                    const int len = total > 0 ? (int)((double)vec_size * (total - (double)sc[t]) / total) + 1 : 0;
                    dprintf("Selected threshold %.2f; compress_len %d\n", th.t[t], len);
                }
            }
        } else {
            (void)hipGetLastError();
        }
    }
    if (own) {
        hipError_t e = hipFreeAsync(ws, s);
        if (st == IBP_OK && e != hipSuccess) return set_error(IBP_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(e));
    }
    if (st == IBP_OK) dprintf("[ibp] preprocess: %d threshold(s), V=%lld, N=%lld\n", th.count, (long long)vec_size,
                              (long long)num_vecs);
    return st;
}
