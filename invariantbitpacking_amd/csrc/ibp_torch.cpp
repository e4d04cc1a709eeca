// ibp_cuda host side: the reference's torch extension module
// (src/ibp_api.cu:124-135; its functions src/preprocess.cu, src/compress.cu,
// src/decompress.cu) as C++ over the C ABI of libibp.so (include/ibp.h).
//
// Each binding checks its tensors as the reference's TORCH_CHECKs do (same
// conditions, same messages, RuntimeError in Python), allocates its outputs
// with torch (device memory, pinned host memory, the caching allocator's
// stream order), and issues one C ABI call on the current HIP stream.  No
// kernel is launched from here and nothing runs on the CPU in place of a
// kernel: every value comes from libibp's HIP code.
//
// Positional arguments are the reference's, in its order and with its
// optional (None) arguments.  Keyword-only arguments are extensions:
// clusters= (per-row mask rows from preprocess_kmeans), seed= (k-means
// initialisation), threads= (read_shared).  Library failures raise
// invariantbitpacking_amd._capi.IBPError (a RuntimeError) with the status.
#include <torch/extension.h>
#include <torch/python.h>

#include <ATen/MapAllocator.h>
#include <c10/hip/HIPStream.h>

#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "ibp.h"

namespace py = pybind11;
using at::Tensor;
using OptTensor = std::optional<Tensor>;

namespace {

// ---- errors -------------------------------------------------------------------

struct StatusError : std::runtime_error {
    int status;
    StatusError(int st, const std::string &msg) : std::runtime_error(msg), status(st) {}
};

const char *status_name(int st)
{
    switch (st) {
    case IBP_OK: return "IBP_OK";
    case IBP_ERR_INVALID: return "IBP_ERR_INVALID";
    case IBP_ERR_HIP: return "IBP_ERR_HIP";
    case IBP_ERR_NOMEM: return "IBP_ERR_NOMEM";
    case IBP_ERR_UNSUPPORTED: return "IBP_ERR_UNSUPPORTED";
    case IBP_ERR_BUSY: return "IBP_ERR_BUSY";
    default: return "IBP_ERR";
    }
}

void check_status(int st)
{
    if (st != IBP_OK) throw StatusError(st, std::string(status_name(st)) + ": " + ibp_last_error());
}

// The reference's TORCH_CHECK: a RuntimeError carrying exactly `msg`.
inline void need(bool cond, const char *msg)
{
    if (!cond) throw std::runtime_error(msg);
}

// ---- argument helpers -----------------------------------------------------------

void *ptr(const Tensor &t) { return t.defined() ? t.data_ptr() : nullptr; }

// On a HIP device or in pinned host memory.
bool accessible(const Tensor &t) { return t.is_cuda() || (t.device().is_cpu() && t.is_pinned()); }

bool same(const Tensor &a, const Tensor &b) { return a.unsafeGetTensorImpl() == b.unsafeGetTensorImpl(); }

// `t`, converted from `orig` (dtype or layout): a host copy of a pinned tensor
// is pinned again, since the kernels read host arguments in place and a
// pageable copy is not visible to the device.
Tensor kernel_readable(const Tensor &orig, Tensor t)
{
    if (!same(t, orig) && !t.is_cuda() && !t.is_pinned()) t = t.pin_memory();
    return t;
}

// A converted argument in host memory must outlive the asynchronous kernels
// that read it: the call then synchronises its stream before returning.
bool host_temp(const OptTensor &orig, const Tensor &used)
{
    return orig.has_value() && used.defined() && !same(*orig, used) && !used.is_cuda();
}

void *cur_stream() { return c10::hip::getCurrentHIPStream().stream(); }

// Makes the GPU that holds `t` (or `d`) current for the rest of the call, so
// outputs, workspaces, the stream and the kernels are that GPU's; host-resident
// (pinned) data leaves the current GPU in place.
struct CallDevice {
    c10::OptionalDeviceGuard g;
    explicit CallDevice(const Tensor &t)
    {
        if (t.is_cuda()) g.reset_device(t.device());
    }
    CallDevice(const c10::Device &d, const Tensor &t)
    {
        if (d.is_cuda() && d.has_index())
            g.reset_device(d);
        else if (t.is_cuda())
            g.reset_device(t.device());
    }
};

// Waits for the current stream with the GIL released (other Python threads
// keep running, as they do while a ctypes call waits).
void sync_stream()
{
    py::gil_scoped_release nogil;
    c10::hip::getCurrentHIPStream().synchronize();
}

at::ScalarType signed_type(int64_t w)
{
    switch (w) {
    case 1: return at::kChar;
    case 2: return at::kShort;
    case 4: return at::kInt;
    default: return at::kLong;
    }
}

bool width_ok(int64_t w) { return w == 1 || w == 2 || w == 4 || w == 8; }

Tensor pinned_empty(at::IntArrayRef shape, at::ScalarType t)
{
    return at::empty(shape, at::TensorOptions().dtype(t).pinned_memory(true));
}

at::TensorOptions on_gpu(at::ScalarType t) { return at::TensorOptions().dtype(t).device(at::kCUDA); }

struct Index {
    Tensor t;
    int bytes = 0;
};

Index index_arg(const OptTensor &index, bool require_cuda)
{
    Index r;
    if (!index.has_value()) return r;
    const Tensor &ix = *index;
    if (require_cuda)
        need(ix.is_cuda(), "Index array must accessible by CUDA device (GPU or pinned memory)");
    else
        need(accessible(ix), "Index array must accessible by CUDA device (GPU or pinned memory)");
    need(ix.dim() == 1, "Index array should be 1D");
    Tensor t = ix;
    if (t.scalar_type() != at::kInt && t.scalar_type() != at::kLong) t = t.to(at::kLong);
    r.t = kernel_readable(ix, t.contiguous());
    r.bytes = (int)r.t.element_size();
    return r;
}

struct Clusters {
    Tensor t;
    int64_t k = 1;
};

Clusters clusters_arg(const OptTensor &clusters, const Tensor &mask, int64_t rows)
{
    Clusters r;
    if (!clusters.has_value()) return r;
    const Tensor &c = *clusters;
    need(accessible(c), "Clusters tensor must accessible by CUDA device (GPU or pinned memory)");
    need(c.dim() == 1 && c.numel() >= rows, "Clusters must be 1D with one entry per row");
    need(mask.dim() == 2, "With clusters, mask and bitval must be 2D [num_clusters x vec_size]");
    Tensor t = c.scalar_type() == at::kInt ? c : c.to(at::kInt);
    r.t = kernel_readable(c, t.contiguous());
    r.k = mask.size(0);
    return r;
}

// The layout of the *_ex entry points; nullptr for the reference's slots with
// mask row 0.
struct Layout {
    ibp_layout l{};
    bool used = false;
    Layout(const Tensor *offsets, int64_t stream_bytes, const Clusters &cl)
    {
        if (offsets) {
            l.offsets = static_cast<const int64_t *>(offsets->data_ptr());
            l.stream_bytes = stream_bytes;
            used = true;
        }
        if (cl.t.defined()) {
            l.clusters = static_cast<const int32_t *>(cl.t.data_ptr());
            l.num_clusters = cl.k;
            used = true;
        } else {
            l.num_clusters = 1;
        }
    }
    const ibp_layout *get() const { return used ? &l : nullptr; }
};

void dataset_checks(const Tensor &d, const char *width_msg)
{
    need(accessible(d), "Input tensor must accessible by CUDA device (GPU or pinned memory)");
    need(d.dim() == 1 || d.dim() == 2, "Input tensor must be 1D [vec_size] or 2D [num_vecs x vec_size]");
    need(d.is_contiguous(), "Input tensor must be contiguous");
    need(width_ok(d.element_size()), width_msg);
}

void mask_checks(const Tensor &d, const Tensor &mask, const Tensor &bitval, int64_t v)
{
    need(accessible(mask), "Mask tensor must accessible by CUDA device (GPU or pinned memory)");
    need(accessible(bitval), "Bitval tensor must accessible by CUDA device (GPU or pinned memory)");
    const int64_t w = d.element_size();
    need(mask.element_size() == w && bitval.element_size() == w,
         "Mask and bitval tensors must have the same datatype size as input dataset");
    need(mask.is_contiguous() && bitval.is_contiguous(), "Mask and bitval tensors must be contiguous");
    need(mask.numel() >= v && bitval.numel() >= v, "Mask and bitval must hold at least vec_size elements");
    need(mask.dim() == 1 || mask.size(-1) == v, "Mask rows must have vec_size elements");
    need(mask.sizes() == bitval.sizes(), "Mask and bitval must have the same shape");
}

std::pair<int64_t, int64_t> rows_cols(const Tensor &d)
{
    return d.dim() == 1 ? std::make_pair<int64_t, int64_t>(1, d.size(0)) : std::make_pair(d.size(0), d.size(1));
}

int hint(const std::optional<int64_t> &comp_len)
{
    if (!comp_len.has_value()) return 0;
    return *comp_len > 0 ? (int)*comp_len : 0;
}

int opt_int(const std::optional<int64_t> &v, int none) { return v.has_value() ? (int)*v : none; }

int loc(const Tensor &t) { return t.is_cuda() ? IBP_LOC_DEVICE : IBP_LOC_HOST; }

Tensor fetch_output(at::ScalarType dtype, int64_t n, int64_t v, const c10::Device &device, const OptTensor &output)
{
    if (output.has_value()) {
        const Tensor &o = *output;
        need(o.device() == device || (device.is_cuda() && o.is_cuda() && !device.has_index()),
             "Output tensor device does not match specified output device");
        need(o.scalar_type() == dtype, "Output tensor dtype does not match input dataset dtype");
        need(o.dim() == 2 && o.size(0) == n && o.size(1) == v,
             "Output tensor shape does not match expected decompressed shape");
        need(o.is_contiguous(), "Output tensor must be contiguous");
        if (!o.is_cuda()) need(o.is_pinned(), "Output tensor on the host must be pinned");
        return o;
    }
    if (device.is_cuda()) return at::empty({n, v}, at::TensorOptions().dtype(dtype).device(device));
    return pinned_empty({n, v}, dtype);
}

// Passes that read a table several times (preprocess: the count and the
// score; k-means: two per iteration, twenty in all) read pinned host data
// over PCIe each time.  When the table fits in half the free device memory
// it is copied to the GPU once, on the current stream, and the passes read
// that copy from HBM; the values do not change.  The copy is released on
// return, after the stream-ordered kernels that read it were queued.
Tensor staged_for_passes(const Tensor &dataset)
{
    if (dataset.is_cuda() || dataset.numel() == 0) return dataset;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return dataset;
    }
    const size_t bytes = (size_t)dataset.numel() * (size_t)dataset.element_size();
    if (bytes > free_b / 2) return dataset;
    return dataset.to(at::TensorOptions().device(at::kCUDA), /*non_blocking=*/true);
}

// ---- bindings -----------------------------------------------------------------

// print_debug_msg (src/ibp_api.cu:9-12).
void print_debug(bool flag) { ibp_set_debug(flag ? 1 : 0); }

// preprocess (src/preprocess.cu:17-76): mask, bitval [V] on the GPU, zero when
// no threshold saves bits.  threshold None sweeps 0.7..1.0.  Asynchronous.
std::tuple<Tensor, Tensor> preprocess(const Tensor &dataset, std::optional<double> threshold)
{
    need(accessible(dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)");
    need(dataset.dim() == 2, "Input tensor must be 2D [num_vecs x vec_size]");
    const int64_t w = dataset.element_size();
    need(width_ok(w), "Input tensor must be 1, 2, 4, or 8-byte datatype");
    need(dataset.is_contiguous(), "Input tensor must be contiguous");
    CallDevice dg(dataset);
    const float t = threshold.has_value() ? (float)*threshold : -1.0f;
    const int64_t n = dataset.size(0), v = dataset.size(1);
    Tensor mask = at::zeros({v}, on_gpu(signed_type(w)));
    Tensor bitval = at::zeros({v}, on_gpu(signed_type(w)));
    const size_t ws_bytes = ibp_preprocess_workspace_size((int)w, v, t);
    Tensor ws = at::empty({(int64_t)ws_bytes}, on_gpu(at::kByte));
    const Tensor data = staged_for_passes(dataset);
    check_status(ibp_preprocess((int)w, ptr(data), n, v, t, ptr(mask), ptr(bitval), nullptr, ptr(ws), ws_bytes,
                                cur_stream()));
    return {mask, bitval};
}

// preprocess_kmeans (src/preprocess.cu:88-139): masks, bitvals [K x V] on the
// GPU and clusters [N] int32 in pinned host memory.  threshold is ignored, as
// in the reference.  seed (extension) draws the initialisation from the
// sequence of srand(seed) without touching libc's generator; None continues
// the process-wide rand() state.  Synchronises.
std::tuple<Tensor, Tensor, Tensor> preprocess_kmeans(const Tensor &dataset, int64_t num_clusters,
                                                     std::optional<double> threshold, std::optional<int64_t> seed)
{
    need(accessible(dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)");
    need(dataset.dim() == 2, "Input tensor must be 2D [num_vecs x vec_size]");
    const int64_t w = dataset.element_size();
    need(w == 4 || w == 8, "Input tensor must be 4-byte or 8-byte datatype");
    need(dataset.is_contiguous(), "Input tensor must be contiguous");
    CallDevice dg(dataset);
    const int64_t n = dataset.size(0), v = dataset.size(1);
    Tensor masks = at::zeros({num_clusters, v}, on_gpu(signed_type(w)));
    Tensor bitvals = at::zeros({num_clusters, v}, on_gpu(signed_type(w)));
    Tensor clusters = pinned_empty({n}, at::kInt).zero_();
    const float t = threshold.has_value() ? (float)*threshold : -1.0f;
    const Tensor data = staged_for_passes(dataset);
    void *st = cur_stream();
    int rc;
    {
        py::gil_scoped_release nogil;  // the call synchronises its stream
        rc = ibp_preprocess_kmeans((int)w, ptr(data), n, v, (int)num_clusters, t, seed.has_value() ? *seed : -1,
                                   ptr(masks), ptr(bitvals), static_cast<int32_t *>(clusters.data_ptr()), st);
    }
    check_status(rc);
    return {masks, bitvals, clusters};
}

// get_compress_size (src/compress.cu:25-174): compressed bytes of each row,
// int64 [n] next to the dataset.  compress_ctr (1 element) is incremented by
// the total, never reset.  Asynchronous unless the counter or a converted
// argument lives in host memory.
Tensor get_compress_size(const Tensor &dataset, const Tensor &mask, const Tensor &bitval, const OptTensor &index,
                         const OptTensor &compress_ctr, const OptTensor &clusters)
{
    dataset_checks(dataset, "Input tensor must be 1, 2, 4, or 8-byte datatype");
    const int64_t w = dataset.element_size();
    auto [num_vecs, v] = rows_cols(dataset);
    mask_checks(dataset, mask, bitval, v);
    CallDevice dg(dataset);
    Index ix = index_arg(index, false);
    const int64_t n = ix.t.defined() ? ix.t.size(0) : num_vecs;
    Tensor total;
    if (compress_ctr.has_value()) {
        const Tensor &c = *compress_ctr;
        need(accessible(c), "Compress ctr must accessible by CUDA device (GPU or pinned memory)");
        need(c.dim() == 1 && c.size(0) == 1, "Compress ctr should be a single element");
        total = (c.is_cuda() && c.scalar_type() == at::kLong) ? c : at::zeros({1}, on_gpu(at::kLong));
    }
    Clusters cl = clusters_arg(clusters, mask, num_vecs);
    Tensor sizes = dataset.is_cuda() ? at::empty({n}, at::TensorOptions().dtype(at::kLong).device(dataset.device()))
                                     : pinned_empty({n}, at::kLong);
    Layout lay(nullptr, 0, cl);
    check_status(ibp_get_compress_size_ex((int)w, ptr(dataset), num_vecs, v, ptr(mask), ptr(bitval), ptr(ix.t),
                                          ix.bytes, n, static_cast<int64_t *>(sizes.data_ptr()),
                                          static_cast<uint64_t *>(ptr(total)), lay.get(), cur_stream()));
    if (compress_ctr.has_value() && !same(total, *compress_ctr)) {
        Tensor c = *compress_ctr;
        c.add_(total.to(c.device()).to(c.scalar_type()));
    }
    if (host_temp(index, ix.t) || host_temp(clusters, cl.t)) sync_stream();
    return sizes;
}

// compress_inplace (src/compress.cu:176-236): packs rows in place, returns the
// int32 bitmask [ceil(n/32)] of packed rows.  With an index, slot i receives
// row index[i]; the rows are gathered first, so the result does not depend on
// the order rows are packed in.  Synchronises.
Tensor compress_inplace(const Tensor &dataset, const Tensor &mask, const Tensor &bitval, const OptTensor &index,
                        const OptTensor &clusters)
{
    dataset_checks(dataset, "Input tensor must be 1, 2, 4 or 8-byte datatype");
    const int64_t w = dataset.element_size();
    auto [num_vecs, v] = rows_cols(dataset);
    mask_checks(dataset, mask, bitval, v);
    CallDevice dg(dataset);
    Index ix = index_arg(index, false);
    Clusters cl = clusters_arg(clusters, mask, num_vecs);
    const int64_t n = ix.t.defined() ? ix.t.size(0) : num_vecs;
    Tensor bitmask = at::empty({(n + 31) / 32}, on_gpu(at::kInt));
    Tensor src = dataset;
    if (ix.t.defined()) {
        need(n <= num_vecs, "index_arr longer than the dataset");
        Tensor rows = dataset.view({num_vecs, v});
        src = rows.index_select(0, ix.t.to(rows.device()).to(at::kLong)).contiguous();
        if (!src.is_cuda()) src = src.pin_memory();
        if (cl.t.defined()) {  // the gathered rows keep their own mask rows
            cl.t = cl.t.index_select(0, ix.t.to(cl.t.device()).to(at::kLong)).contiguous();
            if (!cl.t.is_cuda()) cl.t = cl.t.pin_memory();
        }
    }
    Layout lay(nullptr, 0, cl);
    check_status(ibp_compress_ex((int)w, ptr(src), ptr(dataset), n, v, ptr(mask), ptr(bitval), nullptr, 0, n,
                                 static_cast<int32_t *>(bitmask.data_ptr()), lay.get(), cur_stream()));
    sync_stream();  // in place, like the reference; keeps the temporaries alive
    return bitmask;
}

// compress_condensed (src/compress.cu:238-313): (stream, byte_offsets).  Row
// index[i] (or i) occupies bytes [offsets[i-1], offsets[i]) of the stream; a
// row of V*w bytes is stored raw, any other as its slot cut at its size.  The
// stream sits with the dataset (pinned when the dataset is on the host).
// Synchronises (the stream's length is read from the offsets).
std::tuple<Tensor, Tensor> compress_condensed(const Tensor &dataset, const Tensor &mask, const Tensor &bitval,
                                              const OptTensor &index, const OptTensor &clusters)
{
    dataset_checks(dataset, "Input tensor must be 1, 2, 4 or 8-byte datatype");
    const int64_t w = dataset.element_size();
    auto [num_vecs, v] = rows_cols(dataset);
    mask_checks(dataset, mask, bitval, v);
    CallDevice dg(dataset);
    Index ix = index_arg(index, false);
    Clusters cl = clusters_arg(clusters, mask, num_vecs);
    const int64_t n = ix.t.defined() ? ix.t.size(0) : num_vecs;
    const bool on_dev = dataset.is_cuda();
    Tensor offsets = on_dev ? at::empty({n}, at::TensorOptions().dtype(at::kLong).device(dataset.device()))
                            : pinned_empty({n}, at::kLong);
    Layout lay(nullptr, 0, cl);
    check_status(ibp_condensed_offsets((int)w, ptr(dataset), num_vecs, v, ptr(mask), ptr(bitval), ptr(ix.t),
                                       ix.bytes, n, static_cast<int64_t *>(offsets.data_ptr()), lay.get(),
                                       cur_stream()));
    sync_stream();
    const int64_t total = n > 0 ? offsets[n - 1].item<int64_t>() : 0;
    need(total % w == 0, "internal error: stream length is not a multiple of the element size");
    Tensor stream = on_dev ? at::empty({total / w}, at::TensorOptions().dtype(dataset.scalar_type())
                                                         .device(dataset.device()))
                           : pinned_empty({total / w}, dataset.scalar_type());
    if (n > 0) {
        Layout full(&offsets, total, cl);
        check_status(ibp_compress_ex((int)w, ptr(dataset), ptr(stream), num_vecs, v, ptr(mask), ptr(bitval),
                                     ptr(ix.t), ix.bytes, n, nullptr, full.get(), cur_stream()));
    }
    if (host_temp(index, ix.t) || host_temp(clusters, cl.t)) sync_stream();
    return {stream, offsets};
}

// decompress_fetch (src/decompress.cu:10-156): out[k] = row index[k] (every
// row in order without an index), unpacked.  The compressed rows may be in
// pinned host memory (read over PCIe inside the kernel) or on the GPU.
// comp_len: expected compressed row length in elements (a hint); impl: the
// reference's variants (-1/0/1/5 unpack, 2 raw gather, 3/4 transfer-only
// baselines).  Asynchronous on the current stream.
Tensor decompress_fetch(const Tensor &comp, const Tensor &mask, const Tensor &bitval, const Tensor &bitmask,
                        const c10::Device &out_device, std::optional<int64_t> comp_len, const OptTensor &index,
                        std::optional<int64_t> nblks, std::optional<int64_t> nthreads, std::optional<int64_t> impl,
                        const OptTensor &output, const OptTensor &clusters)
{
    need(accessible(comp), "Input tensor must accessible by CUDA device (GPU or pinned memory)");
    need(comp.dim() == 1 || comp.dim() == 2, "Input tensor must be 1D [vec_size] or 2D [num_vecs x vec_size]");
    need(comp.is_contiguous(), "Input tensor must be contiguous");
    const int64_t w = comp.element_size();
    need(width_ok(w), "Input tensor must be 1, 2, 4 or 8-byte datatype");
    auto [num_vecs, v] = rows_cols(comp);
    mask_checks(comp, mask, bitval, v);
    CallDevice dg(out_device, comp);
    need(bitmask.is_cuda() || accessible(bitmask), "Bitmask tensor must accessible by CUDA device");
    need(bitmask.element_size() == 4 && bitmask.is_contiguous(), "Bitmask must be a contiguous int32 tensor");
    need(bitmask.numel() >= (num_vecs + 31) / 32, "Bitmask too short for the dataset");
    Index ix = index_arg(index, true);
    const int64_t n = ix.t.defined() ? ix.t.size(0) : num_vecs;
    Clusters cl = clusters_arg(clusters, mask, num_vecs);
    Tensor out = fetch_output(comp.scalar_type(), n, v, out_device, output);
    Layout lay(nullptr, 0, cl);
    check_status(ibp_decompress_fetch_ex((int)w, ptr(out), ptr(comp), num_vecs, v, ptr(mask), ptr(bitval),
                                         static_cast<const int32_t *>(bitmask.data_ptr()), ptr(ix.t), ix.bytes, n,
                                         hint(comp_len), opt_int(nblks, 0), opt_int(nthreads, 0), opt_int(impl, -1),
                                         loc(comp), lay.get(), cur_stream()));
    if (host_temp(clusters, cl.t)) sync_stream();
    return out;
}

// The condensed stream's decompress, the inverse of compress_condensed: the
// reference declares it and never built it (src/ibp_api.cu:134).  stream: 1D
// (GPU or pinned); byte_offsets: its int64 offsets; vec_size: elements per
// row.  Asynchronous on the current stream.
Tensor decompress(const Tensor &stream, const Tensor &offsets, const Tensor &mask, const Tensor &bitval,
                  const c10::Device &device, int64_t vec_size, const OptTensor &index, const OptTensor &output,
                  std::optional<int64_t> comp_len, const OptTensor &clusters)
{
    need(accessible(stream), "Input tensor must accessible by CUDA device (GPU or pinned memory)");
    need(stream.dim() == 1 && stream.is_contiguous(), "The condensed stream must be a contiguous 1D tensor");
    need(accessible(offsets), "Offsets must accessible by CUDA device (GPU or pinned memory)");
    need(offsets.scalar_type() == at::kLong && offsets.dim() == 1 && offsets.is_contiguous(),
         "Offsets must be a contiguous 1D int64 tensor");
    const int64_t w = stream.element_size();
    need(width_ok(w), "Input tensor must be 1, 2, 4 or 8-byte datatype");
    need(vec_size > 0, "vec_size must be positive");
    mask_checks(stream, mask, bitval, vec_size);
    CallDevice dg(device, stream);
    const int64_t num_rows = offsets.numel();
    Index ix = index_arg(index, true);
    const int64_t n = ix.t.defined() ? ix.t.size(0) : num_rows;
    Clusters cl = clusters_arg(clusters, mask, num_rows);
    Tensor out = fetch_output(stream.scalar_type(), n, vec_size, device, output);
    Layout lay(&offsets, stream.numel() * w, cl);
    check_status(ibp_decompress_fetch_ex((int)w, ptr(out), ptr(stream), num_rows, vec_size, ptr(mask), ptr(bitval),
                                         nullptr, ptr(ix.t), ix.bytes, n, hint(comp_len), 0, 0, -1, loc(stream),
                                         lay.get(), cur_stream()));
    if (host_temp(clusters, cl.t)) sync_stream();
    return out;
}

// read_shared (src/ibp_api.cu:52-122): a tensor of `shape` and `dtype` in
// process-shared memory holding the file's bytes, read by `threads` threads
// (0: 16).  Bytes past the file's end are zero; a file longer than the tensor
// is an error.
Tensor read_shared(const std::string &filename, const std::vector<int64_t> &shape, const py::object &dtype,
                   int64_t threads)
{
    const at::ScalarType type = torch::python::detail::py_object_to_dtype(dtype);
    int64_t numel = 1;
    for (int64_t d : shape) {
        need(d >= 0, "read_shared: negative dimension");
        numel *= d;
    }
    const size_t nbytes = (size_t)numel * c10::elementSize(type);
    Tensor t;
    if (nbytes == 0) {
        t = at::empty(shape, at::TensorOptions().dtype(type));
    } else {
        // A storage of its own in a process-wide shared memory segment (the
        // segment is unlinked at once and kept alive by its descriptor).
        const int flags = at::ALLOCATOR_MAPPED_SHAREDMEM | at::ALLOCATOR_MAPPED_EXCLUSIVE |
                          at::ALLOCATOR_MAPPED_KEEPFD | at::ALLOCATOR_MAPPED_UNLINK;
        const std::string handle = at::NewProcessWideShmHandle();
        at::Storage storage(c10::make_intrusive<c10::StorageImpl>(
            c10::StorageImpl::use_byte_size_t(), nbytes,
            at::MapAllocator::makeDataPtr(handle.c_str(), flags, nbytes, nullptr), nullptr, false));
        t = at::empty({0}, at::TensorOptions().dtype(type)).set_(storage, 0, shape);
    }
    size_t got = 0;
    void *dst = nbytes ? t.data_ptr() : nullptr;
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = ibp_read_file(filename.c_str(), dst, nbytes, (int)threads, &got);
    }
    check_status(rc);
    return t;
}

// Rows kernels skipped because their index, offsets or cluster id was out of
// range (synchronises the device).
uint64_t index_errors(bool reset)
{
    uint64_t v = 0;
    int rc;
    {
        py::gil_scoped_release nogil;  // synchronises the device
        rc = ibp_index_errors(&v, reset ? 1 : 0);
    }
    check_status(rc);
    return v;
}

std::string version() { return ibp_version(); }

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "InvariantBitPacking for MI355X: the ibp_cuda bindings over libibp.so";
    // Library failures: invariantbitpacking_amd._capi.IBPError(status, message).
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const StatusError &e) {
            py::object cls = py::module_::import("invariantbitpacking_amd._capi").attr("IBPError");
            py::object err = cls(e.status, e.what());
            PyErr_SetObject(cls.ptr(), err.ptr());
        }
    });
    using py::arg;
    const py::none None;
    m.def("print_debug", &print_debug, "Print debug messages", arg("print"));
    m.def("preprocess", &preprocess, "Mask and bitval of a dataset", arg("dataset"), arg("threshold_") = None);
    m.def("preprocess_kmeans", &preprocess_kmeans, "Per-cluster masks and bitvals of a dataset", arg("dataset"),
          arg("num_clusters"), arg("threshold_") = None, py::kw_only(), arg("seed") = None);
    m.def("get_compress_size", &get_compress_size, "Compressed size of each row", arg("dataset"), arg("mask"),
          arg("bitval"), arg("index_array_") = None, arg("compress_ctr_") = None, py::kw_only(),
          arg("clusters") = None);
    m.def("compress_inplace", &compress_inplace, "Pack rows in place (same space)", arg("dataset"), arg("mask"),
          arg("bitval"), arg("index_array_") = None, py::kw_only(), arg("clusters") = None);
    m.def("compress_condensed", &compress_condensed, "Condensed stream of the rows and its byte offsets",
          arg("dataset"), arg("mask"), arg("bitval"), arg("index_array_") = None, py::kw_only(),
          arg("clusters") = None);
    m.def("decompress_fetch", &decompress_fetch, "Fetch and unpack rows into a new or given tensor",
          arg("comp_dataset"), arg("mask"), arg("bitval"), arg("bitmask"), arg("out_device"),
          arg("comp_len_") = None, arg("index_array_") = None, arg("nblks_") = None, arg("nthreads_") = None,
          arg("impl_") = None, arg("output_tensor") = None, py::kw_only(), arg("clusters") = None);
    m.def("read_shared", &read_shared, "Read a file into a tensor in shared memory", arg("filename"), arg("shape"),
          arg("dtype"), py::kw_only(), arg("threads") = 0);
    m.def("decompress", &decompress, "Unpack rows of a condensed stream", arg("comp_dataset"),
          arg("byte_offsets"), arg("mask"), arg("bitval"), arg("device"), arg("vec_size"),
          arg("index_arr") = None, arg("output_tensor") = None, arg("comp_len") = None, py::kw_only(),
          arg("clusters") = None);
    m.def("index_errors", &index_errors, "Rows skipped for out-of-range indices", arg("reset") = false);
    m.def("version", &version, "libibp version");
}
