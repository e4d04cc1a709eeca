// ibp_pack.hip -- compressed-size scan and the (in-place) pack.
//
// get_compress_size: src/compress.cu:25-174 -> check_compress_size_kernel
//   (include/misc/ibp_misc_kernels.cuh:9-32, ibp_misc_dev.cuh:20-46).
// compress_inplace:  src/compress.cu:176-236 -> compress_inplace_kernel
//   (include/compress/ibp_compress_kernel.cuh:10-56, ibp_compress_dev.cuh:8-97).
//
// MI355X mapping: one wave64 per row, lane j owns element j of each 64-element
// chunk; a wave owns 32 consecutive output rows so it writes their bitmask
// word itself (no atomics, no pre-zeroing).  The packed slot is assembled in a
// per-wave LDS buffer (64-bit ds_or for the payload, ballot for the header
// words) and then written out whole, so in-place packing never overwrites
// input the wave still has to read.
#include <type_traits>

#include "ibp_bits.h"
#include "ibp_internal.h"
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

template <typename T, typename IdxT>
__global__ __launch_bounds__(256) void size_kernel(const T *data, int64_t num_vecs, int64_t V, const T *mask,
                                                   const T *bitval, const void *index, int64_t n,
                                                   int64_t *sizes, unsigned long long *total,
                                                   unsigned long long *err)
{
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    unsigned long long acc = 0, bad = 0;
    for (int64_t i = wave; i < n; i += nwaves) {
        const int64_t r = readfirstlane(row_of<IdxT>(index, i));
        int64_t size = 0;
        if (r >= 0 && r < num_vecs) {
            const int saved = row_saved_bits<T>(data + r * V, mask, bitval, V);
            size = slot_size(V, (int)sizeof(T), saved);
        } else {
            ++bad;
        }
        if (lane == 0) sizes[i] = size;
        acc += (unsigned long long)size;
    }
    if (lane == 0) {
        if (total && acc) atomicAdd(total, acc);
        if (err && bad) atomicAdd(err, bad);
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

// Pack kernel.  BUF_LDS: the per-wave slot buffer is in LDS (dynamic shared
// memory, V words per wave); otherwise it is `gbuf` in device memory.
template <typename T, typename IdxT, bool BUF_LDS>
__global__ __launch_bounds__(256) void pack_kernel(const T *in, T *out, int64_t num_vecs, int64_t V,
                                                   const T *mask, const T *bitval, const void *index,
                                                   int64_t n, uint32_t *bitmask, T *gbuf,
                                                   unsigned long long *err)
{
    constexpr int B = Bits<T>::B;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = lane_id();
    const int wave_in_block = threadIdx.x / kWave;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
    T *buf = BUF_LDS ? (T *)lds_raw + (int64_t)wave_in_block * V : gbuf + wave * V;
    const int64_t Hw = header_bytes(V, (int)sizeof(T)) / (int64_t)sizeof(T);
    const int64_t vec_bytes = V * (int64_t)sizeof(T);
    const bool in_place = (const void *)in == (const void *)out && index == nullptr;
    const int64_t ngroups = (n + 31) / 32;
    unsigned long long bad = 0;

    for (int64_t g = wave; g < ngroups; g += nwaves) {
        uint32_t bits = 0;
        const int64_t iend = (g + 1) * 32 < n ? (g + 1) * 32 : n;
        for (int64_t i = g * 32; i < iend; ++i) {
            const int64_t r = readfirstlane(row_of<IdxT>(index, i));
            if (r < 0 || r >= num_vecs) {
                ++bad;
                continue;
            }
            const T *src = in + r * V;
            T *dst = out + i * V;
            // Size test (ibp_compress_kernel.cuh:28-31): pack only if smaller.
            const int saved = row_saved_bits<T>(src, mask, bitval, V);
            if (slot_size(V, (int)sizeof(T), saved) >= vec_bytes) {
                if (!in_place)
                    for (int64_t k = lane; k < V; k += kWave) dst[k] = src[k];
                continue;
            }
            for (int64_t k = lane; k < V; k += kWave) buf[k] = T(0);
            slot_fence<BUF_LDS>();
            int64_t cursor = 0;
            for (int64_t c0 = 0; c0 < V; c0 += kWave) {
                const int64_t j = c0 + lane;
                const bool valid = j < V;
                const T x = valid ? src[j] : T(0);
                const T m = valid ? mask[j] : T(0);
                const bool f = valid && (T)(x & m) == (valid ? bitval[j] : T(0));
                // Header flags of this chunk, LSB first (ibp_compress_dev.cuh:60-71).
                const uint64_t hb = ballot(f);
                if constexpr (sizeof(T) == 8) {
                    if (lane == 0) buf[c0 / 64] = (T)hb;
                } else {
                    if (lane == 0) buf[c0 / 32] = (T)(uint32_t)hb;
                    if (lane == 1 && c0 + 32 < V) buf[c0 / 32 + 1] = (T)(uint32_t)(hb >> 32);
                }
                const int L = valid ? (f ? B - popc(m) : B) : 0;
                const int incl = wave_incl_scan(L);
                const int tot = readlane(incl, kWave - 1);
                const int64_t pbit = cursor + incl - L;
                // Payload: free bits MSB first (flagged) or the whole element.
                const T v = f ? extract_msb<T>(x, T(~m)) : x;
                const int64_t q = Hw + pbit / B;
                const int o = (int)(pbit % B);
                if (L > 0) {
                    __hip_atomic_fetch_or(&buf[q], shr<T>(v, o), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (o + L > B)
                        __hip_atomic_fetch_or(&buf[q + 1], shl<T>(v, B - o), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
                }
                cursor += tot;
            }
            slot_fence<BUF_LDS>();
            for (int64_t k = lane; k < V; k += kWave)
                dst[k] = __hip_atomic_load(&buf[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bits |= 1u << (i - g * 32);
        }
        if (lane == 0) bitmask[g] = bits;
    }
    if (err && bad && lane == 0) atomicAdd(err, bad);
}

namespace {

template <typename T, typename IdxT>
int size_typed(const void *data, int64_t num_vecs, int64_t V, const void *mask, const void *bitval,
               const void *index, int64_t n, int64_t *sizes, uint64_t *total, hipStream_t stream)
{
    const DeviceInfo &dev = device_info();
    int64_t grid = (n * kWave + 255) / 256;
    if (grid > (int64_t)dev.num_cus * 8) grid = (int64_t)dev.num_cus * 8;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL((size_kernel<T, IdxT>), dim3((unsigned)grid), dim3(256), 0, stream, (const T *)data,
                       num_vecs, V, (const T *)mask, (const T *)bitval, index, n, sizes,
                       (unsigned long long *)total, index ? index_error_counter() : nullptr);
    IBP_LAUNCH_CHECK();
    return IBP_OK;
}

template <typename T>
int size_dispatch(int index_bytes, const void *data, int64_t num_vecs, int64_t V, const void *mask,
                  const void *bitval, const void *index, int64_t n, int64_t *sizes, uint64_t *total,
                  hipStream_t stream)
{
    if (index && index_bytes == 4)
        return size_typed<T, int32_t>(data, num_vecs, V, mask, bitval, index, n, sizes, total, stream);
    return size_typed<T, int64_t>(data, num_vecs, V, mask, bitval, index, n, sizes, total, stream);
}

template <typename T, typename IdxT>
int pack_typed(const void *input, void *output, int64_t num_vecs, int64_t V, const void *mask,
               const void *bitval, const void *index, int64_t n, int32_t *bitmask, hipStream_t stream)
{
    const DeviceInfo &dev = device_info();
    const int64_t slot_bytes = V * (int64_t)sizeof(T);
    const int64_t lds_budget = 64 * 1024;
    const int64_t ngroups = (n + 31) / 32;
    unsigned long long *err = index ? index_error_counter() : nullptr;
    if (slot_bytes <= lds_budget) {
        int wpb = (int)(lds_budget / slot_bytes);
        if (wpb > 4) wpb = 4;
        int64_t grid = (ngroups + wpb - 1) / wpb;
        if (grid > (int64_t)dev.num_cus * 16) grid = (int64_t)dev.num_cus * 16;
        if (grid < 1) grid = 1;
        hipLaunchKernelGGL((pack_kernel<T, IdxT, true>), dim3((unsigned)grid), dim3(wpb * kWave),
                           (size_t)(wpb * slot_bytes), stream, (const T *)input, (T *)output, num_vecs, V,
                           (const T *)mask, (const T *)bitval, index, n, (uint32_t *)bitmask, (T *)nullptr, err);
        IBP_LAUNCH_CHECK();
        return IBP_OK;
    }
    // Rows beyond the LDS budget: per-wave slot buffers in device memory.
    int64_t nwaves = ngroups < (int64_t)dev.num_cus * 4 ? ngroups : (int64_t)dev.num_cus * 4;
    if (nwaves < 1) nwaves = 1;
    const int64_t grid = (nwaves + 3) / 4;
    void *gbuf = nullptr;
    IBP_HIP(hipMallocAsync(&gbuf, (size_t)(grid * 4 * slot_bytes), stream));
    hipLaunchKernelGGL((pack_kernel<T, IdxT, false>), dim3((unsigned)grid), dim3(256), 0, stream,
                       (const T *)input, (T *)output, num_vecs, V, (const T *)mask, (const T *)bitval, index, n,
                       (uint32_t *)bitmask, (T *)gbuf, err);
    hipError_t e = hipGetLastError();
    IBP_HIP(hipFreeAsync(gbuf, stream));
    if (e != hipSuccess) return set_error(IBP_ERR_HIP, "pack kernel launch failed: %s", hipGetErrorString(e));
    return IBP_OK;
}

} // namespace

} // namespace ibp

using namespace ibp;

extern "C" int ibp_get_compress_size(int elem_bytes, const void *data, int64_t num_vecs, int64_t vec_size,
                                     const void *mask, const void *bitval, const void *index, int index_bytes,
                                     int64_t n, int64_t *sizes, uint64_t *total, void *stream)
{
    IBP_CHECK(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8,
              "Input tensor must be 1, 2, 4, or 8-byte datatype (got %d)", elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0 && n >= 0, "invalid shape");
    IBP_CHECK(index == nullptr || index_bytes == 4 || index_bytes == 8, "index must be int32 or int64");
    IBP_CHECK(index != nullptr || n <= num_vecs, "without an index n must not exceed num_vecs");
    if (n == 0) return IBP_OK;
    IBP_CHECK(data && mask && bitval && sizes, "null pointer argument");
    hipStream_t s = as_stream(stream);
    switch (elem_bytes) {
    case 1: return size_dispatch<uint8_t>(index_bytes, data, num_vecs, vec_size, mask, bitval, index, n, sizes, total, s);
    case 2: return size_dispatch<uint16_t>(index_bytes, data, num_vecs, vec_size, mask, bitval, index, n, sizes, total, s);
    case 4: return size_dispatch<uint32_t>(index_bytes, data, num_vecs, vec_size, mask, bitval, index, n, sizes, total, s);
    default: return size_dispatch<uint64_t>(index_bytes, data, num_vecs, vec_size, mask, bitval, index, n, sizes, total, s);
    }
}

extern "C" int ibp_compress(int elem_bytes, const void *input, void *output, int64_t num_vecs, int64_t vec_size,
                            const void *mask, const void *bitval, const void *index, int index_bytes, int64_t n,
                            int32_t *bitmask, void *stream)
{
    IBP_CHECK(elem_bytes == 4 || elem_bytes == 8, "Input tensor must be 4-byte or 8-byte datatype (got %d)", elem_bytes);
    IBP_CHECK(vec_size > 0 && num_vecs >= 0 && n >= 0, "invalid shape");
    IBP_CHECK(index == nullptr || index_bytes == 4 || index_bytes == 8, "index must be int32 or int64");
    IBP_CHECK(index != nullptr || n <= num_vecs, "without an index n must not exceed num_vecs");
    IBP_CHECK(index == nullptr || input != output, "an index with in-place output aliases rows (unsupported)");
    if (n == 0) return IBP_OK;
    IBP_CHECK(input && output && mask && bitval && bitmask, "null pointer argument");
    hipStream_t s = as_stream(stream);
    const bool i32 = index && index_bytes == 4;
    if (elem_bytes == 8)
        return i32 ? pack_typed<uint64_t, int32_t>(input, output, num_vecs, vec_size, mask, bitval, index, n, bitmask, s)
                   : pack_typed<uint64_t, int64_t>(input, output, num_vecs, vec_size, mask, bitval, index, n, bitmask, s);
    return i32 ? pack_typed<uint32_t, int32_t>(input, output, num_vecs, vec_size, mask, bitval, index, n, bitmask, s)
               : pack_typed<uint32_t, int64_t>(input, output, num_vecs, vec_size, mask, bitval, index, n, bitmask, s);
}
