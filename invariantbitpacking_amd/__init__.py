"""Invariant Bit Packing on MI355X: the `ibp` Python API.

Drop-in for the reference module `ibp` (ibp/__init__.py:1-41) and its
extension `ibp_cuda` (src/ibp_api.cu:124-135): same function names, argument
order, defaults, return values and error type (RuntimeError).  Every call
goes through the C ABI of libibp.so (include/ibp.h); torch supplies device
memory, pinned host memory and the current HIP stream only.

    import invariantbitpacking_amd as ibp
    mask, bitval = ibp.preprocess(data, 0.9)
    bitmask = ibp.compress_inplace(pinned, mask, bitval)
    rows = ibp.decompress_fetch(pinned, mask, bitval, bitmask, torch.device("cuda"), index_arr=idx)

Beyond the reference (keyword-only, so positional calls are unchanged):
clusters= on the size, compress and fetch calls codes each row with its own
preprocess_kmeans mask row; compress_condensed / compress and decompress
give and read the condensed stream; 1- and 2-byte elements are accepted by
every call.
"""
import ctypes

import torch

from . import _capi
from ._capi import IBPError  # noqa: F401  (re-exported)

__all__ = ["print_debug", "preprocess", "preprocess_kmeans", "get_compress_size", "get_single_compress_len",
           "compress_inplace", "compress_condensed", "compress", "decompress_fetch", "decompress", "read_shared",
           "IBPError", "version", "index_errors"]

_SIGNED = {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device_index=None):
    """Raw handle of the current HIP stream of the current (or given) device."""
    if device_index is None:
        device_index = torch.cuda.current_device()
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(device_index))


def _accessible(t):
    """On a HIP device or in pinned host memory (the reference's TORCH_CHECK)."""
    return t.is_cuda or (t.device.type == "cpu" and t.is_pinned())


def _check(cond, msg):
    if not cond:
        raise RuntimeError(msg)


def _loc(t):
    return 1 if t.is_cuda else 2


def _kernel_readable(orig, t):
    """`t` converted from `orig` (dtype or layout): a host copy of a pinned
    tensor is pinned again, since the kernels read host arguments in place
    and a pageable copy is not visible to the device."""
    if t is not orig and not t.is_cuda and not t.is_pinned():
        t = t.pin_memory()
    return t


def _index_arg(index_arr, require_cuda):
    """Returns (tensor kept alive, element bytes) for an index array."""
    if index_arr is None:
        return None, 0
    if require_cuda:
        _check(index_arr.is_cuda, "Index array must accessible by CUDA device (GPU or pinned memory)")
    else:
        _check(_accessible(index_arr), "Index array must accessible by CUDA device (GPU or pinned memory)")
    _check(index_arr.dim() == 1, "Index array should be 1D")
    idx = index_arr
    if idx.dtype not in (torch.int32, torch.int64):
        idx = idx.to(torch.int64)
    return _kernel_readable(index_arr, idx.contiguous()), idx.element_size()


def _clusters_arg(clusters, mask, rows):
    """(int32 tensor kept alive, number of mask rows) for a clusters argument."""
    if clusters is None:
        return None, 1
    _check(_accessible(clusters), "Clusters tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(clusters.dim() == 1 and clusters.numel() >= rows, "Clusters must be 1D with one entry per row")
    _check(mask.dim() == 2, "With clusters, mask and bitval must be 2D [num_clusters x vec_size]")
    cl = clusters if clusters.dtype == torch.int32 else clusters.to(torch.int32)
    return _kernel_readable(clusters, cl.contiguous()), mask.size(0)


def _layout(offsets=None, stream_bytes=0, clusters=None, num_clusters=1):
    if offsets is None and clusters is None:
        return None
    return ctypes.byref(_capi.Layout(None if offsets is None else offsets.data_ptr(), int(stream_bytes),
                                     None if clusters is None else clusters.data_ptr(), int(num_clusters)))


def _host_temps(*pairs):
    """True if some argument was converted into a temporary in host memory:
    such a temporary must outlive the asynchronous kernels that read it."""
    return any(a is not None and b is not None and a is not b and not b.is_cuda for a, b in pairs)


def version():
    return _capi.lib().ibp_version().decode()


def print_debug(flag):
    """Turn the library's debug messages on or off (ibp/__init__.py:8-9)."""
    _capi.lib().ibp_set_debug(1 if flag else 0)


def index_errors(reset=False):
    """Rows skipped by kernels because their index was out of range (synchronises)."""
    v = ctypes.c_uint64(0)
    _capi.call("ibp_index_errors", ctypes.byref(v), 1 if reset else 0)
    return v.value


def preprocess(dataset, threshold=None):
    """mask, bitval [V] (signed int of the element width, on the GPU).

    src/preprocess.cu:17-76.  threshold None sweeps 0.7..1.0 and keeps the
    threshold that saves the most bits; both outputs stay zero if none saves any.
    Asynchronous, on the current stream.
    """
    _check(_accessible(dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(dataset.dim() == 2, "Input tensor must be 2D [num_vecs x vec_size]")
    w = dataset.element_size()
    _check(w in (1, 2, 4, 8), "Input tensor must be 1, 2, 4, or 8-byte datatype")
    _check(dataset.is_contiguous(), "Input tensor must be contiguous")
    t = -1.0 if threshold is None else float(threshold)
    n, v = dataset.shape
    mask = torch.zeros(v, dtype=_SIGNED[w], device="cuda")
    bitval = torch.zeros(v, dtype=_SIGNED[w], device="cuda")
    ws_bytes = _capi.lib().ibp_preprocess_workspace_size(w, v, t)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    _capi.call("ibp_preprocess", w, _ptr(dataset), n, v, t, _ptr(mask), _ptr(bitval), None, _ptr(ws), ws_bytes,
               _stream())
    return mask, bitval


def preprocess_kmeans(dataset, centroids, threshold=None, *, seed=None):
    """masks [K,V], bitvals [K,V] (GPU) and clusters [N] int32 (pinned host).

    src/preprocess.cu:88-139.  `threshold` is ignored, as in the reference.
    `seed` (extension): re-seed libc rand() before the random initialisation;
    None continues the process-wide rand() state like the reference.
    """
    _check(_accessible(dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(dataset.dim() == 2, "Input tensor must be 2D [num_vecs x vec_size]")
    w = dataset.element_size()
    _check(w in (4, 8), "Input tensor must be 4-byte or 8-byte datatype")
    _check(dataset.is_contiguous(), "Input tensor must be contiguous")
    k = int(centroids)
    n, v = dataset.shape
    masks = torch.zeros((k, v), dtype=_SIGNED[w], device="cuda")
    bitvals = torch.zeros((k, v), dtype=_SIGNED[w], device="cuda")
    clusters = torch.zeros(n, dtype=torch.int32).pin_memory()
    t = -1.0 if threshold is None else float(threshold)
    _capi.call("ibp_preprocess_kmeans", w, _ptr(dataset), n, v, k, t, -1 if seed is None else int(seed),
               _ptr(masks), _ptr(bitvals), _ptr(clusters), _stream())
    return masks, bitvals, clusters


def _mask_check(dataset, mask, bitval, v=None):
    _check(_accessible(mask), "Mask tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(_accessible(bitval), "Bitval tensor must accessible by CUDA device (GPU or pinned memory)")
    w = dataset.element_size()
    _check(mask.element_size() == w and bitval.element_size() == w,
           "Mask and bitval tensors must have the same datatype size as input dataset")
    _check(mask.is_contiguous() and bitval.is_contiguous(), "Mask and bitval tensors must be contiguous")
    if v is None:
        v = dataset.size(0) if dataset.dim() == 1 else dataset.size(1)
    _check(mask.numel() >= v and bitval.numel() >= v, "Mask and bitval must hold at least vec_size elements")
    _check(mask.dim() == 1 or mask.size(-1) == v, "Mask rows must have vec_size elements")
    _check(mask.shape == bitval.shape, "Mask and bitval must have the same shape")


def _shape(dataset):
    return (1, dataset.size(0)) if dataset.dim() == 1 else (dataset.size(0), dataset.size(1))


def get_compress_size(dataset, mask, bitval, index_arr=None, compress_total=None, *, clusters=None):
    """Compressed bytes of each row (int64 [n]; src/compress.cu:25-174).

    compress_total (1-element int64 tensor) is incremented by the total and
    never reset.  clusters (extension): mask row of each dataset row, mask and
    bitval [K, V].  Asynchronous unless compress_total lives in host memory or
    an argument had to be converted in host memory.
    """
    _check(_accessible(dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(dataset.dim() in (1, 2), "Input tensor must be 1D [vec_size] or 2D [num_vecs x vec_size]")
    _check(dataset.is_contiguous(), "Input tensor must be contiguous")
    w = dataset.element_size()
    _check(w in (1, 2, 4, 8), "Input tensor must be 1, 2, 4, or 8-byte datatype")
    _mask_check(dataset, mask, bitval)
    num_vecs, v = _shape(dataset)
    idx, ib = _index_arg(index_arr, require_cuda=False)
    n = num_vecs if idx is None else idx.size(0)
    total_dev = None
    if compress_total is not None:
        _check(_accessible(compress_total), "Compress ctr must accessible by CUDA device (GPU or pinned memory)")
        _check(compress_total.dim() == 1 and compress_total.size(0) == 1, "Compress ctr should be a single element")
        if compress_total.is_cuda and compress_total.dtype == torch.int64:
            total_dev = compress_total
        else:
            total_dev = torch.zeros(1, dtype=torch.int64, device="cuda")
    cl, k = _clusters_arg(clusters, mask, num_vecs)
    if dataset.is_cuda:
        sizes = torch.empty(n, dtype=torch.int64, device=dataset.device)
    else:
        sizes = torch.empty(n, dtype=torch.int64, pin_memory=True)
    _capi.call("ibp_get_compress_size_ex", w, _ptr(dataset), num_vecs, v, _ptr(mask), _ptr(bitval), _ptr(idx), ib,
               n, _ptr(sizes), _ptr(total_dev), _layout(clusters=cl, num_clusters=k), _stream())
    if compress_total is not None and total_dev is not compress_total:
        compress_total += total_dev.to(compress_total.device).to(compress_total.dtype)
    if _host_temps((index_arr, idx), (clusters, cl)):
        torch.cuda.current_stream().synchronize()
    return sizes


def get_single_compress_len(dataset, mask, bitval, index_arr=None):
    """Average compressed row length in elements (ibp/__init__.py:22-29).
    index_arr is accepted and ignored, as in the reference."""
    comp_sizes = get_compress_size(dataset, mask, bitval)
    torch.cuda.synchronize()
    og_size = dataset.numel() * dataset.element_size()
    comp_size = int(torch.sum(comp_sizes))
    return int(comp_size / og_size * dataset.size(1))


def compress_inplace(dataset, mask, bitval, index_arr=None, *, clusters=None):
    """Packs rows in place; returns the int32 bitmask [ceil(n/32)] of packed rows
    (src/compress.cu:176-236).  With index_arr, slot i receives row index_arr[i]
    (rows are gathered first, so the result does not depend on execution order).
    clusters (extension): mask row of each dataset row, mask and bitval [K, V].
    Synchronises the current stream.
    """
    _check(_accessible(dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(dataset.dim() in (1, 2), "Input tensor must be 1D [vec_size] or 2D [num_vecs x vec_size]")
    _check(dataset.is_contiguous(), "Input tensor must be contiguous")
    w = dataset.element_size()
    _check(w in (1, 2, 4, 8), "Input tensor must be 1, 2, 4 or 8-byte datatype")
    _mask_check(dataset, mask, bitval)
    num_vecs, v = _shape(dataset)
    idx, _ = _index_arg(index_arr, require_cuda=False)
    cl, k = _clusters_arg(clusters, mask, num_vecs)
    n = num_vecs if idx is None else idx.size(0)
    bitmask = torch.empty((n + 31) // 32, dtype=torch.int32, device="cuda")
    src = dataset
    if idx is not None:
        _check(n <= num_vecs, "index_arr longer than the dataset")
        rows = dataset.view(num_vecs, v)
        src = rows.index_select(0, idx.to(rows.device).long()).contiguous()
        if not src.is_cuda:
            src = src.pin_memory()
        if cl is not None:  # the gathered rows keep their own mask rows
            cl = cl.index_select(0, idx.to(cl.device).long()).contiguous()
            if not cl.is_cuda:
                cl = cl.pin_memory()
    _capi.call("ibp_compress_ex", w, _ptr(src), _ptr(dataset), n, v, _ptr(mask), _ptr(bitval), None, 0, n,
               _ptr(bitmask), _layout(clusters=cl, num_clusters=k), _stream())
    torch.cuda.current_stream().synchronize()  # in place, like the reference; keeps temporaries alive
    return bitmask


def compress_condensed(dataset, mask, bitval, index_arr=None, *, clusters=None):
    """Condensed stream of the rows and their byte offsets
    (src/compress.cu:238-313; binding src/ibp_api.cu:131).

    Returns (stream, byte_offsets): stream is a 1D tensor of the dataset's
    dtype on the dataset's device (pinned if the dataset is on the host) holding
    row index_arr[i] (or i) in bytes [byte_offsets[i-1], byte_offsets[i]);
    byte_offsets (int64 [n], same placement) is the inclusive scan of the row
    sizes.  A row of size V*w is stored raw, any other as its slot cut at its
    size.  clusters (extension): mask row of each dataset row.  Synchronises
    (the stream's length is read from the offsets, as in the reference).
    """
    _check(_accessible(dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(dataset.dim() in (1, 2), "Input tensor must be 1D [vec_size] or 2D [num_vecs x vec_size]")
    _check(dataset.is_contiguous(), "Input tensor must be contiguous")
    w = dataset.element_size()
    _check(w in (1, 2, 4, 8), "Input tensor must be 1, 2, 4 or 8-byte datatype")
    _mask_check(dataset, mask, bitval)
    num_vecs, v = _shape(dataset)
    idx, ib = _index_arg(index_arr, require_cuda=False)
    cl, k = _clusters_arg(clusters, mask, num_vecs)
    n = num_vecs if idx is None else idx.size(0)
    on_dev = dataset.is_cuda
    offsets = (torch.empty(n, dtype=torch.int64, device=dataset.device) if on_dev
               else torch.empty(n, dtype=torch.int64, pin_memory=True))
    lay = _layout(clusters=cl, num_clusters=k)
    _capi.call("ibp_condensed_offsets", w, _ptr(dataset), num_vecs, v, _ptr(mask), _ptr(bitval), _ptr(idx), ib, n,
               _ptr(offsets), lay, _stream())
    torch.cuda.current_stream().synchronize()
    total = int(offsets[-1]) if n > 0 else 0
    _check(total % w == 0, "internal error: stream length %d is not a multiple of %d" % (total, w))
    stream = (torch.empty(total // w, dtype=dataset.dtype, device=dataset.device) if on_dev
              else torch.empty(total // w, dtype=dataset.dtype, pin_memory=True))
    if n > 0:
        _capi.call("ibp_compress_ex", w, _ptr(dataset), _ptr(stream), num_vecs, v, _ptr(mask), _ptr(bitval),
                   _ptr(idx), ib, n, None, _layout(offsets, total, cl, k), _stream())
    if _host_temps((index_arr, idx), (clusters, cl)):
        torch.cuda.current_stream().synchronize()
    return stream, offsets


def compress(dataset, mask, bitval, index_arr=None, *, clusters=None):
    """`ibp.compress` (ibp/__init__.py:35-36) calls `ibp_cuda.compress`, which the
    reference never binds; the binding it means is compress_condensed
    (src/ibp_api.cu:131), which this calls."""
    return compress_condensed(dataset, mask, bitval, index_arr, clusters=clusters)


def decompress_fetch(comp_dataset, mask, bitval, bitmask, device, comp_len=None, index_arr=None, nblks=None,
                     nthds=None, imp=None, output_tensor=None, *, clusters=None):
    """Rows of the compressed dataset, unpacked (src/decompress.cu:10-156).

    out[k] = row index_arr[k] (all rows in order without an index).  The
    compressed data may be pinned host memory (read over PCIe inside the
    kernel) or GPU memory.  comp_len is a hint (expected compressed row length
    in elements); imp selects the reference's variants (-1/0/1/5 unpack, 2 raw
    gather, 3/4 transfer-only baselines).  clusters (extension): mask row of
    each stored row, mask and bitval [K, V].  Asynchronous on the current stream.
    """
    if clusters is None and nblks is None and nthds is None and output_tensor is not None and \
            _fetch_fast(comp_dataset, mask, bitval, bitmask, device, comp_len, index_arr, imp, output_tensor):
        return output_tensor
    _check(_accessible(comp_dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(comp_dataset.dim() in (1, 2),
           "Input tensor must be 1D [vec_size] or 2D [num_vecs x vec_size] got %d" % comp_dataset.dim())
    _check(comp_dataset.is_contiguous(), "Input tensor must be contiguous")
    w = comp_dataset.element_size()
    _check(w in (1, 2, 4, 8), "Input tensor must be 1, 2, 4 or 8-byte datatype")
    _mask_check(comp_dataset, mask, bitval)
    _check(bitmask.is_cuda or _accessible(bitmask), "Bitmask tensor must accessible by CUDA device")
    _check(bitmask.element_size() == 4 and bitmask.is_contiguous(), "Bitmask must be a contiguous int32 tensor")
    num_vecs, v = _shape(comp_dataset)
    _check(bitmask.numel() >= (num_vecs + 31) // 32, "Bitmask too short for the dataset")
    idx, ib = _index_arg(index_arr, require_cuda=True)
    n = num_vecs if idx is None else idx.size(0)
    clu, k = _clusters_arg(clusters, mask, num_vecs)
    out = _fetch_output(comp_dataset.dtype, n, v, device, output_tensor)
    _capi.call("ibp_decompress_fetch_ex", w, _ptr(out), _ptr(comp_dataset), num_vecs, v, _ptr(mask), _ptr(bitval),
               _ptr(bitmask), _ptr(idx), ib, n, _hint(comp_len), 0 if nblks is None else int(nblks),
               0 if nthds is None else int(nthds), -1 if imp is None else int(imp), _loc(comp_dataset),
               _layout(clusters=clu, num_clusters=k), _stream())
    if _host_temps((clusters, clu)):
        torch.cuda.current_stream().synchronize()
    return out


_FETCH_FN = []


def _fetch_fast(comp, mask, bitval, bitmask, device, comp_len, index_arr, imp, out):
    """The common call -- every tensor on a HIP device and contiguous, 1D
    mask/bitval, an int32/int64 index, an output tensor of the right shape --
    with the checks of the full path folded into one expression, the launch
    issued straight through the cached C entry point.  Returns False without
    doing anything when one of them does not hold; the full path then runs
    and raises its own message (about 4 us less host time per call)."""
    if not (comp.is_cuda and mask.is_cuda and bitval.is_cuda and bitmask.is_cuda and out.is_cuda and
            index_arr is not None and index_arr.is_cuda and comp.dim() == 2 and mask.dim() == 1 and
            bitval.dim() == 1 and index_arr.dim() == 1 and out.dim() == 2 and comp.is_contiguous() and
            mask.is_contiguous() and bitval.is_contiguous() and bitmask.is_contiguous() and
            index_arr.is_contiguous() and out.is_contiguous() and out.dtype == comp.dtype and
            index_arr.dtype in (torch.int32, torch.int64) and type(device) is torch.device and
            device.type == "cuda" and (device.index is None or device.index == out.get_device())):
        return False
    num_vecs, v = comp.shape
    w = comp.element_size()
    n = index_arr.size(0)
    if not (w in (1, 2, 4, 8) and mask.element_size() == w and bitval.element_size() == w and
            mask.numel() >= v and mask.numel() == bitval.numel() and bitmask.element_size() == 4 and
            bitmask.numel() >= (num_vecs + 31) // 32 and out.size(0) == n and out.size(1) == v):
        return False
    if not _FETCH_FN:
        _FETCH_FN.append(_capi.lib().ibp_decompress_fetch_ex)
    st = _FETCH_FN[0](w, out.data_ptr(), comp.data_ptr(), num_vecs, v, mask.data_ptr(), bitval.data_ptr(),
                      bitmask.data_ptr(), index_arr.data_ptr(), index_arr.element_size(), n, _hint(comp_len), 0, 0,
                      -1 if imp is None else int(imp), 1, None,
                      torch._C._cuda_getCurrentRawStream(torch.cuda.current_device()))
    if st != 0:
        _capi.check(st)
    return True


def _hint(comp_len):
    if comp_len is None:
        return 0
    cl = int(comp_len)
    return cl if cl > 0 else 0


def _fetch_output(dtype, n, v, device, output_tensor):
    if not isinstance(device, torch.device):
        device = torch.device(device)
    if output_tensor is not None:
        _check(output_tensor.device == device or (device.type == "cuda" and output_tensor.is_cuda and
                                                  device.index is None),
               "Output tensor device does not match specified output device")
        _check(output_tensor.dtype == dtype, "Output tensor dtype does not match input dataset dtype")
        _check(output_tensor.dim() == 2 and output_tensor.size(0) == n and output_tensor.size(1) == v,
               "Output tensor shape does not match expected decompressed shape")
        _check(output_tensor.is_contiguous(), "Output tensor must be contiguous")
        out = output_tensor
        if not out.is_cuda:
            _check(out.is_pinned(), "Output tensor on the host must be pinned")
    elif device.type == "cuda":
        out = torch.empty((n, v), dtype=dtype, device=device)
    else:
        out = torch.empty((n, v), dtype=dtype, pin_memory=True)
    return out


def decompress(comp_dataset, byte_offsets, mask, bitval, device, vec_size, index_arr=None, output_tensor=None,
               comp_len=None, *, clusters=None):
    """Rows of a condensed stream, unpacked: the inverse of compress_condensed.

    The reference declares this binding but never built it (src/ibp_api.cu:134).
    comp_dataset: the 1D stream (GPU or pinned); byte_offsets: its int64
    offsets (GPU or pinned); vec_size: elements per row.  out[k] = stored row
    index_arr[k] (all rows in order without an index).  clusters: mask row of
    each stored row.  Asynchronous on the current stream.
    """
    _check(_accessible(comp_dataset), "Input tensor must accessible by CUDA device (GPU or pinned memory)")
    _check(comp_dataset.dim() == 1 and comp_dataset.is_contiguous(), "The condensed stream must be a contiguous 1D tensor")
    _check(_accessible(byte_offsets), "Offsets must accessible by CUDA device (GPU or pinned memory)")
    _check(byte_offsets.dtype == torch.int64 and byte_offsets.dim() == 1 and byte_offsets.is_contiguous(),
           "Offsets must be a contiguous 1D int64 tensor")
    w = comp_dataset.element_size()
    _check(w in (1, 2, 4, 8), "Input tensor must be 1, 2, 4 or 8-byte datatype")
    v = int(vec_size)
    _check(v > 0, "vec_size must be positive")
    _mask_check(comp_dataset, mask, bitval, v)
    num_rows = byte_offsets.numel()
    idx, ib = _index_arg(index_arr, require_cuda=True)
    n = num_rows if idx is None else idx.size(0)
    clu, k = _clusters_arg(clusters, mask, num_rows)
    out = _fetch_output(comp_dataset.dtype, n, v, device, output_tensor)
    stream_bytes = comp_dataset.numel() * w
    _capi.call("ibp_decompress_fetch_ex", w, _ptr(out), _ptr(comp_dataset), num_rows, v, _ptr(mask), _ptr(bitval),
               None, _ptr(idx), ib, n, _hint(comp_len), 0, 0, -1, _loc(comp_dataset),
               _layout(byte_offsets, stream_bytes, clu, k), _stream())
    if _host_temps((clusters, clu)):
        torch.cuda.current_stream().synchronize()
    return out


def read_shared(filename, shape, dtype, *, threads=0):
    """A tensor of `shape` and `dtype` in process-shared memory holding the
    contents of `filename` (src/ibp_api.cu:52-122).  The file is read by
    `threads` threads (0: 16); bytes past its end are zero; a file larger than
    the tensor raises (the reference copies past the tensor's end)."""
    shape = tuple(int(d) for d in shape)
    nbytes = torch.Size(shape).numel() * torch.empty(0, dtype=dtype).element_size()
    storage = torch.UntypedStorage._new_shared(nbytes)
    t = torch.empty(0, dtype=dtype).set_(storage, 0, shape)
    got = ctypes.c_size_t(0)
    _capi.call("ibp_read_file", str(filename).encode(), ctypes.c_void_p(t.data_ptr()) if nbytes else None, nbytes,
               int(threads), ctypes.byref(got))
    return t


from .batched import BatchedFetch  # noqa: E402
from .serve import Server  # noqa: E402

__all__ += ["BatchedFetch", "Server"]
