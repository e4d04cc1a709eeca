"""`import ibp_cuda` for callers of the reference's extension module
(src/ibp_api.cu:124-135; C++ signatures src/ibp_api.cu:20-40).

The reference binds these with pybind11 and no argument names, so callers pass
every argument positionally; optional (c10::optional) arguments take None.
Each function here takes the same arguments in the same order and forwards to
invariantbitpacking_amd, which calls libibp.so.  The reference also declares a
`decompress` binding that it never built (src/ibp_api.cu:134); the condensed
stream's decompress is invariantbitpacking_amd.decompress.
"""
import torch  # noqa: F401

import invariantbitpacking_amd as _ibp

__all__ = ["print_debug", "preprocess", "preprocess_kmeans", "get_compress_size", "compress_inplace",
           "compress_condensed", "decompress_fetch", "read_shared"]


def print_debug(print_):
    """print_debug_msg(bool) (src/ibp_api.cu:9-12)."""
    _ibp.print_debug(print_)


def preprocess(dataset, threshold_=None):
    """preprocess(Tensor, float?) -> (mask, bitval) (src/preprocess.cu:17-76)."""
    return _ibp.preprocess(dataset, threshold_)


def preprocess_kmeans(dataset, num_clusters, threshold_=None):
    """preprocess_kmeans(Tensor, int, float?) -> (masks, bitvals, clusters) (src/preprocess.cu:88-139)."""
    return _ibp.preprocess_kmeans(dataset, num_clusters, threshold_)


def get_compress_size(dataset, mask, bitval, index_array_=None, compress_ctr_=None):
    """get_compress_size(Tensor, Tensor, Tensor, Tensor?, Tensor?) -> sizes (src/compress.cu:25-174)."""
    return _ibp.get_compress_size(dataset, mask, bitval, index_array_, compress_ctr_)


def compress_inplace(dataset, mask, bitval, index_array_=None):
    """compress_inplace(Tensor, Tensor, Tensor, Tensor?) -> bitmask (src/compress.cu:176-236)."""
    return _ibp.compress_inplace(dataset, mask, bitval, index_array_)


def compress_condensed(dataset, mask, bitval, index_array_=None):
    """compress_condensed(Tensor, Tensor, Tensor, Tensor?) -> (stream, byte_offsets) (src/compress.cu:238-313)."""
    return _ibp.compress_condensed(dataset, mask, bitval, index_array_)


def decompress_fetch(comp_dataset, mask, bitval, bitmask, out_device, comp_len_=None, index_array_=None,
                     nblks_=None, nthreads_=None, impl_=None, output_tensor=None):
    """decompress_fetch(Tensor x4, Device, int?, Tensor?, int?, int?, int?, Tensor?) -> rows
    (src/decompress.cu:10-156)."""
    return _ibp.decompress_fetch(comp_dataset, mask, bitval, bitmask, out_device, comp_len_, index_array_, nblks_,
                                 nthreads_, impl_, output_tensor)


def read_shared(filename, shape, dtype):
    """read_shared(str, list[int], dtype) -> Tensor in process-shared memory (src/ibp_api.cu:52-122)."""
    return _ibp.read_shared(filename, shape, dtype)
