"""`import ibp_cuda` for callers of the reference's extension module
(src/ibp_api.cu:124-135; C++ signatures src/ibp_api.cu:20-40).

These are the compiled bindings themselves: invariantbitpacking_amd._C, a
C++ torch extension (invariantbitpacking_amd/csrc/ibp_torch.cpp) over the C
ABI of libibp.so.  Positional arguments, their order and their optional
(None) trailing arguments are the reference's; keyword-only extensions
(clusters=, seed=, threads=) come after them.  The reference also declares a
`decompress` binding that it never built (src/ibp_api.cu:134); the condensed
stream's decompress is exported here as well.
"""
import torch  # noqa: F401  (loads the PyTorch libraries the extension links to)

from invariantbitpacking_amd import _C

print_debug = _C.print_debug
preprocess = _C.preprocess
preprocess_kmeans = _C.preprocess_kmeans
get_compress_size = _C.get_compress_size
compress_inplace = _C.compress_inplace
compress_condensed = _C.compress_condensed
decompress_fetch = _C.decompress_fetch
read_shared = _C.read_shared
decompress = _C.decompress

__all__ = ["print_debug", "preprocess", "preprocess_kmeans", "get_compress_size", "compress_inplace",
           "compress_condensed", "decompress_fetch", "read_shared", "decompress"]
