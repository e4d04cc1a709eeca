"""`import ibp` and `import ibp_cuda` work for the reference's callers: the
module names, the function names and their positional parameter order are the
reference's (ibp/__init__.py:1-41; src/ibp_api.cu:20-40,124-135), and both
modules forward to the HIP library.  No GPU needed."""
import inspect

import pytest

# Positional parameters of the reference's Python module, in order (ibp/__init__.py).
IBP_MODULE = {
    "print_debug": ["flag"],
    "preprocess": ["dataset", "threshold"],
    "preprocess_kmeans": ["dataset", "centroids", "threshold"],
    "get_compress_size": ["dataset", "mask", "bitval", "index_arr", "compress_total"],
    "get_single_compress_len": ["dataset", "mask", "bitval", "index_arr"],
    "compress_inplace": ["dataset", "mask", "bitval", "index_arr"],
    "compress": ["dataset", "mask", "bitval", "index_arr"],
    "decompress_fetch": ["comp_dataset", "mask", "bitval", "bitmask", "device", "comp_len", "index_arr", "nblks",
                         "nthds", "imp", "output_tensor"],
}

# Arity of the extension's bindings and which trailing arguments are optional
# (c10::optional in src/ibp_api.cu:20-40; read_shared src/ibp_api.cu:52).
IBP_CUDA_ARITY = {
    "print_debug": (1, 0),
    "preprocess": (2, 1),
    "preprocess_kmeans": (3, 1),
    "get_compress_size": (5, 2),
    "compress_inplace": (4, 1),
    "compress_condensed": (4, 1),
    "decompress_fetch": (11, 6),
    "read_shared": (3, 0),
}


def positional(fn):
    return [p for p in inspect.signature(fn).parameters.values()
            if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]


def test_ibp_module_names_and_order():
    import ibp
    for name, params in IBP_MODULE.items():
        fn = getattr(ibp, name)
        got = positional(fn)
        assert [p.name for p in got] == params, name
        # defaults: None for every optional argument, as in the reference
        for p in got:
            if p.name in ("threshold", "index_arr", "compress_total", "comp_len", "nblks", "nthds", "imp",
                          "output_tensor"):
                assert p.default is None, (name, p.name)
            else:
                assert p.default is inspect.Parameter.empty, (name, p.name)


def test_ibp_cuda_module_bindings():
    import ibp_cuda
    assert sorted(ibp_cuda.__all__) == sorted(IBP_CUDA_ARITY)
    for name, (arity, optional) in IBP_CUDA_ARITY.items():
        got = positional(getattr(ibp_cuda, name))
        assert len(got) == arity, name
        assert sum(p.default is None for p in got) == optional, name
        assert all(p.default is None for p in got[arity - optional:]), name


def test_modules_forward_to_the_library():
    """Both modules reach libibp.so: argument errors are the library's
    RuntimeError, raised before any GPU work."""
    import torch

    import ibp
    import ibp_cuda
    x = torch.zeros((4, 3), dtype=torch.int64)  # not pinned, not on a GPU
    for mod in (ibp, ibp_cuda):
        with pytest.raises(RuntimeError, match="accessible by CUDA device"):
            mod.preprocess(x, None)
        with pytest.raises(RuntimeError, match="accessible by CUDA device"):
            mod.get_compress_size(x, x[0], x[0], None, None)
    assert ibp.IBPError is not None
    assert "gfx950" in ibp.version()
