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


def binding_params(fn):
    """(name, has_default, default text) of each positional parameter of a
    compiled (pybind11) binding, read from the signature line pybind11 writes
    at the top of its docstring; keyword-only parameters (after `*`) are left
    out."""
    line = fn.__doc__.strip().splitlines()[0]
    inner = line[line.index("(") + 1:line.rindex(")")]
    parts, depth, cur = [], 0, ""
    for ch in inner:
        if ch in "[(":
            depth += 1
        elif ch in "])":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    out = []
    for part in parts:
        if part == "*":
            break
        name = part.split(":")[0].strip()
        default = part.split("=", 1)[1].strip() if "=" in part else None
        out.append((name, default is not None, default))
    return out


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
    """ibp_cuda's functions are compiled bindings (a C++ torch extension) with
    the reference's positional arity and optional trailing arguments; one
    positional argument more is a TypeError, as for the reference's bindings."""
    import torch

    import ibp_cuda
    assert set(IBP_CUDA_ARITY) <= set(ibp_cuda.__all__)
    assert set(ibp_cuda.__all__) - set(IBP_CUDA_ARITY) == {"decompress"}  # declared, never built by the reference
    for name, (arity, optional) in IBP_CUDA_ARITY.items():
        fn = getattr(ibp_cuda, name)
        assert type(fn).__name__ == "builtin_function_or_method", name
        got = binding_params(fn)
        assert len(got) == arity, (name, got)
        assert [d for _, d, _ in got] == [False] * (arity - optional) + [True] * optional, (name, got)
        assert all(text == "None" for _, d, text in got if d), (name, got)
        with pytest.raises(TypeError):
            fn(*([None] * (arity + 1)))
    # the reference's positional call of decompress_fetch reaches the argument checks
    x = torch.zeros((4, 3), dtype=torch.int64)
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp_cuda.decompress_fetch(x, x[0], x[0], torch.zeros(1, dtype=torch.int32), torch.device("cuda"), None, None,
                                  None, None, None, None)


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


def test_unbuilt_package_imports_and_every_call_raises(tmp_path):
    """A fresh checkout has no built extension: the package must still import
    (its build script lives inside it), and every API call must raise
    ImportError rather than fall back to anything else."""
    import os
    import shutil
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pkg = os.path.join(root, "invariantbitpacking_amd")
    dst = tmp_path / "invariantbitpacking_amd"
    shutil.copytree(pkg, dst, ignore=shutil.ignore_patterns("*.so", "_build", "__pycache__", "*.o"))
    code = ("import torch, invariantbitpacking_amd as m\n"
            "from invariantbitpacking_amd.csrc import build\n"
            "assert build.SOURCES\n"
            "try:\n"
            "    m.preprocess(torch.zeros(4, 4, dtype=torch.int64))\n"
            "except ImportError as e:\n"
            "    assert 'not built' in str(e), e\n"
            "else:\n"
            "    raise SystemExit('a call succeeded without the extension')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path), capture_output=True, text=True,
                       env=dict(os.environ, PYTHONPATH=str(tmp_path)))
    assert r.returncode == 0, r.stdout + r.stderr
