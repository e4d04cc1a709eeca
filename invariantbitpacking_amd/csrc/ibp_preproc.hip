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
#include "ibp_wave.h"

namespace ibp {

constexpr int kMaxThresholds = 8;

struct Thresholds {
    float t[kMaxThresholds];
    int count;
};

template <typename T>
using WideT = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;

// cnt[j*B + k] += rows of [r0, r1) whose column j has bit k set.  Lane <-> column
// of a 64-column tile (blockIdx.y); the block's waves split the rows; every
// thread keeps its B counters in registers.
template <typename T>
__global__ __launch_bounds__(256) void count_bits_kernel(const T *data, int64_t num_vecs, int64_t V,
                                                         int64_t rows_per_block, int32_t *cnt)
{
    constexpr int B = 8 * (int)sizeof(T);
    const int lane = lane_id();
    const int wave_in_block = wave_in_block_id();
    const int waves = blockDim.x / kWave;
    const int64_t j = (int64_t)blockIdx.y * kWave + lane;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    int64_t r1 = r0 + rows_per_block;
    if (r1 > num_vecs) r1 = num_vecs;
    if (j >= V) return;
    int32_t c[B];
#pragma unroll
    for (int k = 0; k < B; ++k) c[k] = 0;
    int64_t r = r0 + wave_in_block;
    // Two rows per iteration keep two loads in flight per lane.
    for (; r + waves < r1; r += 2 * waves) {
        const WideT<T> x0 = (WideT<T>)data[r * V + j];
        const WideT<T> x1 = (WideT<T>)data[(r + waves) * V + j];
#pragma unroll
        for (int k = 0; k < B; ++k) c[k] += (int32_t)((x0 >> k) & 1) + (int32_t)((x1 >> k) & 1);
    }
    for (; r < r1; r += waves) {
        const WideT<T> x = (WideT<T>)data[r * V + j];
#pragma unroll
        for (int k = 0; k < B; ++k) c[k] += (int32_t)((x >> k) & 1);
    }
#pragma unroll
    for (int k = 0; k < B; ++k)
        if (c[k]) atomicAdd(&cnt[j * B + k], c[k]);
}

// masks[t*V + j], bitvals[t*V + j] for every threshold t
// (ibp_preproc_kernels.cuh:34-51): bit in mask and bitval if
// float(cnt) > float(t) * float(N); in the mask only if
// double(cnt) < (1.0 - double(t)) * double(N).
template <typename T>
__global__ void create_masks_kernel(const int32_t *cnt, int64_t num_vecs, int64_t V, Thresholds th, T *masks,
                                    T *bitvals)
{
    constexpr int B = 8 * (int)sizeof(T);
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= V * th.count) return;
    const int t = (int)(gid / V);
    const int64_t j = gid % V;
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
    // Rows of <= 64 elements: every lane keeps its column's masks in registers.
    T m_reg[NT], b_reg[NT];
    int pc_reg[NT];
    const bool small = V <= kWave;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const bool ok = small && t < nt && lane < V;
        m_reg[t] = ok ? masks[t * V + lane] : T(0);
        b_reg[t] = ok ? bitvals[t * V + lane] : T(0);
        pc_reg[t] = popc((WideT<T>)m_reg[t]);
    }
    for (int64_t r = wave; r < num_vecs; r += nwaves) {
        int acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = 0;
        for (int64_t j = lane; j < V; j += kWave) {
            const T x = data[r * V + j];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (t >= nt) break;
                const T m = small ? m_reg[t] : masks[t * V + j];
                const T b = small ? b_reg[t] : bitvals[t * V + j];
                const int pc = small ? pc_reg[t] : popc((WideT<T>)m);
                if ((T)(x & m) == b) acc[t] += pc;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (t >= nt) break;
            const int64_t s = wave_sum(acc[t]);
            if (s > V) saved[t] += (unsigned long long)(s - V);
        }
    }
    if (lane == 0)
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if (t < nt && saved[t]) atomicAdd(&scores[t], saved[t]);
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

__global__ void zero_i32_kernel(int32_t *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
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

namespace {

struct PreWorkspace {
    size_t cnt_off, masks_off, bitvals_off, scores_off, total;
};

PreWorkspace layout(int w, int64_t V, int nt)
{
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    PreWorkspace L;
    L.cnt_off = 0;
    L.masks_off = al(L.cnt_off + sizeof(int32_t) * (size_t)V * 8 * w);
    L.bitvals_off = al(L.masks_off + (size_t)w * V * nt);
    L.scores_off = al(L.bitvals_off + (size_t)w * V * nt);
    L.total = al(L.scores_off + sizeof(unsigned long long) * kMaxThresholds);
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
    unsigned long long *scores = (unsigned long long *)(ws + L.scores_off);
    const int64_t ncnt = V * 8 * (int64_t)sizeof(T);
    hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)((ncnt + 255) / 256 < 1024 ? (ncnt + 255) / 256 : 1024)),
                       dim3(256), 0, stream, cnt, ncnt);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, stream, scores, kMaxThresholds);
    IBP_LAUNCH_CHECK();
    if (N > 0) {
        // Enough row blocks to cover every CU a few times per column tile.
        const int64_t tiles = (V + kWave - 1) / kWave;
        int64_t want_blocks = (int64_t)dev.num_cus * 4 / tiles;
        if (want_blocks < 1) want_blocks = 1;
        int64_t rows_per_block = (N + want_blocks - 1) / want_blocks;
        if (rows_per_block < 64) rows_per_block = 64;
        const int64_t row_blocks = (N + rows_per_block - 1) / rows_per_block;
        hipLaunchKernelGGL((count_bits_kernel<T>), dim3((unsigned)row_blocks, (unsigned)tiles), dim3(256), 0, stream,
                           (const T *)data, N, V, rows_per_block, cnt);
        IBP_LAUNCH_CHECK();
    }
    const int64_t mthreads = V * th.count;
    hipLaunchKernelGGL((create_masks_kernel<T>), dim3((unsigned)((mthreads + 255) / 256)), dim3(256), 0, stream, cnt, N,
                       V, th, masks, bitvals);
    IBP_LAUNCH_CHECK();
    if (N > 0) {
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
    if (own) {
        hipError_t e = hipFreeAsync(ws, s);
        if (st == IBP_OK && e != hipSuccess) return set_error(IBP_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(e));
    }
    if (st == IBP_OK) dprintf("[ibp] preprocess: %d threshold(s), V=%lld, N=%lld\n", th.count, (long long)vec_size,
                              (long long)num_vecs);
    return st;
}
