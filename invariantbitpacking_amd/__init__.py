"""Invariant Bit Packing on MI355X: the `ibp` Python API.

Drop-in for the reference module `ibp` (ibp/__init__.py:1-41) and its
extension `ibp_cuda` (src/ibp_api.cu:124-135): same function names, argument
order, defaults, return values and error type (RuntimeError).  Each function
forwards to the C++ torch extension `_C` (csrc/ibp_torch.cpp), which checks
the tensors, allocates the outputs with torch and calls the C ABI of
libibp.so (include/ibp.h) on the current HIP stream.

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
import os

import torch

from ._capi import IBPError  # noqa: F401  (re-exported)

_HERE = os.path.dirname(os.path.abspath(__file__))


class _Unbuilt:
    """Stands in for the extension when it is not built, so that the package
    (and its build script, invariantbitpacking_amd.csrc.build) can still be
    imported; every call raises.  There is no fallback implementation."""

    def __init__(self, err):
        self._err = err

    def __getattr__(self, name):
        raise ImportError("the ibp_cuda extension (invariantbitpacking_amd/_C) is not built or cannot load "
                          "libibp.so from %s; run `python -m invariantbitpacking_amd.csrc.build`: %s"
                          % (_HERE, self._err))


try:
    from . import _C
except ImportError as _e:
    _C = _Unbuilt(_e)

__all__ = ["print_debug", "preprocess", "preprocess_kmeans", "get_compress_size", "get_single_compress_len",
           "compress_inplace", "compress_condensed", "compress", "decompress_fetch", "decompress", "read_shared",
           "IBPError", "version", "index_errors"]


def _device(device):
    return device if type(device) is torch.device else torch.device(device)


def version():
    return _C.version()


def print_debug(flag):
    """Turn the library's debug messages on or off (ibp/__init__.py:8-9)."""
    _C.print_debug(bool(flag))


def index_errors(reset=False):
    """Rows skipped by kernels because their index was out of range (synchronises)."""
    return _C.index_errors(bool(reset))


def preprocess(dataset, threshold=None):
    """mask, bitval [V] (signed int of the element width, on the GPU).

    src/preprocess.cu:17-76.  threshold None sweeps 0.7..1.0 and keeps the
    threshold that saves the most bits; both outputs stay zero if none saves any.
    Asynchronous, on the current stream.
    """
    return _C.preprocess(dataset, threshold)


def preprocess_kmeans(dataset, centroids, threshold=None, *, seed=None):
    """masks [K,V], bitvals [K,V] (GPU) and clusters [N] int32 (pinned host).

    src/preprocess.cu:88-139.  `threshold` is ignored, as in the reference.
    `seed` (extension): the random initialisation draws the numbers rand()
    gives after srand(seed), from a private copy of that sequence (libc's
    process-wide generator is left untouched); None continues the
    process-wide rand() state like the reference.
    """
    return _C.preprocess_kmeans(dataset, int(centroids), threshold, seed=seed)


def get_compress_size(dataset, mask, bitval, index_arr=None, compress_total=None, *, clusters=None):
    """Compressed bytes of each row (int64 [n]; src/compress.cu:25-174).

    compress_total (1-element int64 tensor) is incremented by the total and
    never reset.  clusters (extension): mask row of each dataset row, mask and
    bitval [K, V].  Asynchronous unless compress_total lives in host memory or
    an argument had to be converted in host memory.
    """
    return _C.get_compress_size(dataset, mask, bitval, index_arr, compress_total, clusters=clusters)


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
    return _C.compress_inplace(dataset, mask, bitval, index_arr, clusters=clusters)


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
    return _C.compress_condensed(dataset, mask, bitval, index_arr, clusters=clusters)


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
    if type(device) is not torch.device:
        device = torch.device(device)
    if clusters is None:
        return _C.decompress_fetch(comp_dataset, mask, bitval, bitmask, device, comp_len, index_arr, nblks, nthds,
                                   imp, output_tensor)
    return _C.decompress_fetch(comp_dataset, mask, bitval, bitmask, device, comp_len, index_arr, nblks, nthds, imp,
                               output_tensor, clusters=clusters)


def decompress(comp_dataset, byte_offsets, mask, bitval, device, vec_size, index_arr=None, output_tensor=None,
               comp_len=None, *, clusters=None):
    """Rows of a condensed stream, unpacked: the inverse of compress_condensed.

    The reference declares this binding but never built it (src/ibp_api.cu:134).
    comp_dataset: the 1D stream (GPU or pinned); byte_offsets: its int64
    offsets (GPU or pinned); vec_size: elements per row.  out[k] = stored row
    index_arr[k] (all rows in order without an index).  clusters: mask row of
    each stored row.  Asynchronous on the current stream.
    """
    return _C.decompress(comp_dataset, byte_offsets, mask, bitval, _device(device), int(vec_size), index_arr,
                         output_tensor, comp_len, clusters=clusters)


def read_shared(filename, shape, dtype, *, threads=0):
    """A tensor of `shape` and `dtype` in process-shared memory holding the
    contents of `filename` (src/ibp_api.cu:52-122).  The file is read by
    `threads` threads (0: 16); bytes past its end are zero; a file larger than
    the tensor raises (the reference copies past the tensor's end)."""
    return _C.read_shared(str(filename), [int(d) for d in shape], dtype, threads=int(threads))


from .batched import BatchedFetch  # noqa: E402
from .serve import Server  # noqa: E402

__all__ += ["BatchedFetch", "Server"]
