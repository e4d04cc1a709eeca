"""`import ibp` for callers of the reference module (ibp/__init__.py:1-41).

The reference's `ibp` package wraps its CUDA extension `ibp_cuda`; here the
same names come from invariantbitpacking_amd, whose every call goes through
libibp.so (hand-written HIP for gfx950).  Names, positional order, defaults,
return values and the RuntimeError on bad arguments are the reference's:

    print_debug(flag)
    preprocess(dataset, threshold=None) -> (mask, bitval)
    preprocess_kmeans(dataset, centroids, threshold=None) -> (masks, bitvals, clusters)
    get_compress_size(dataset, mask, bitval, index_arr=None, compress_total=None) -> sizes
    get_single_compress_len(dataset, mask, bitval, index_arr=None) -> int
    compress_inplace(dataset, mask, bitval, index_arr=None) -> bitmask
    compress(dataset, mask, bitval, index_arr=None) -> (stream, byte_offsets)
    decompress_fetch(comp_dataset, mask, bitval, bitmask, device, comp_len=None,
                     index_arr=None, nblks=None, nthds=None, imp=None, output_tensor=None)

`compress` returns the condensed stream: the reference's `ibp.compress` calls
`ibp_cuda.compress`, which it never binds (ibp/__init__.py:35-36 against
src/ibp_api.cu:131); the binding it names is compress_condensed.  The
keyword-only extensions of invariantbitpacking_amd (clusters=, decompress,
BatchedFetch, Server, ...) are re-exported as well.
"""
import torch  # noqa: F401  (the reference imports torch before its extension)

from invariantbitpacking_amd import *  # noqa: F401,F403
from invariantbitpacking_amd import (IBPError, compress, compress_inplace, decompress_fetch,  # noqa: F401
                                     get_compress_size, get_single_compress_len, preprocess, preprocess_kmeans,
                                     print_debug)
from invariantbitpacking_amd import __all__  # noqa: F401

# The reference turns decode messages off at import (ibp/__init__.py:5-6).
print_debug(False)
