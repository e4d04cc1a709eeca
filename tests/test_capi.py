"""CPU tests of the C ABI boundary: libibp.so loads, exports every symbol that
include/ibp.h declares, its host-compiled bit arithmetic matches a brute-force
model, and argument errors come back as status codes (no GPU needed)."""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def capi():
    from invariantbitpacking_amd import _capi
    _capi.lib()
    return _capi


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ibp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ibp_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(capi):
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(capi.lib(), name), name
        assert name in capi.SIGNATURES, "no ctypes signature for " + name


def test_version_and_debug(capi):
    assert b"gfx950" in capi.lib().ibp_version()
    capi.lib().ibp_set_debug(1)
    capi.lib().ibp_set_debug(0)


def test_selftest_bitops_host(capi):
    assert capi.lib().ibp_selftest_bitops(50000, 12345) == 0, capi.last_error()


def test_selftest_rand_sequence_host(capi):
    """The private copy of glibc's rand() sequence (preprocess_kmeans with a
    seed) against rand() itself, seeds 0 (which glibc treats as 1), 1, a
    large one and 2^32 - 1."""
    for seed in (0, 1, 3, 12345, 2**32 - 1):
        assert capi.lib().ibp_selftest_rand(seed, 200000) == 0, capi.last_error()


def test_selftest_serve_admission_host(capi):
    """The server's admission bookkeeping with a mock launcher: the queue head
    is admitted by taking back later requests' pinned index space, and random
    submit / launch / complete / release sequences always drain."""
    assert capi.lib().ibp_selftest_serve_admission(300, 7) == 0, capi.last_error()


def test_selftest_index_copy_host(capi):
    """The server's split copy of request indices (the submitter and three
    helpers claiming 64 KiB parts), from three submitting threads at once,
    with idle gaps that put the helpers to sleep: every copy exact."""
    assert capi.lib().ibp_selftest_index_copy(150, 11) == 0, capi.last_error()


def test_invalid_arguments_are_status_codes(capi):
    L = capi.lib()
    st = L.ibp_decompress_fetch(3, None, None, 10, 4, None, None, None, None, 0, 10, 0, 0, 0, -1, 0, None)
    assert st == 1 and "1, 2, 4 or 8-byte" in capi.last_error()
    st = L.ibp_decompress_fetch(8, None, None, 10, 4, None, None, None, None, 0, 10, 0, 0, 0, 9, 0, None)
    assert st == 1 and "impl" in capi.last_error()
    st = L.ibp_decompress_fetch(8, None, None, 10, 4, None, None, None, None, 0, 11, 0, 0, 0, -1, 0, None)
    assert st == 1
    # n == 0 is a no-op
    assert L.ibp_decompress_fetch(8, None, None, 10, 4, None, None, None, None, 0, 0, 0, 0, 0, -1, 0, None) == 0
    buf = ctypes.create_string_buffer(64)
    st = L.ibp_compress(8, buf, buf, 2, 4, buf, buf, buf, 8, 2, buf, None)
    assert st == 1 and "alias" in capi.last_error()
    st = L.ibp_compress(3, buf, buf, 2, 4, buf, buf, None, 0, 2, buf, None)
    assert st == 1
    st = L.ibp_preprocess(3, buf, 2, 4, 0.9, buf, buf, None, None, 0, None)
    assert st == 1 and "1, 2, 4, or 8-byte" in capi.last_error()
    st = L.ibp_preprocess_kmeans(8, buf, 2, 4, 0, 0.9, 1, buf, buf, buf, None)
    assert st == 1 and "num_clusters" in capi.last_error()
    st = L.ibp_get_compress_size(8, buf, 2, 4, buf, buf, None, 0, 3, buf, None, None)
    assert st == 1
    # layouts: transfer baselines take none; the condensed stream must not
    # overlap its input; condensed offsets are the output of ibp_condensed_offsets
    cl = (ctypes.c_int32 * 2)(0, 0)
    lay = capi.Layout(None, 0, ctypes.addressof(cl), 1)
    st = L.ibp_decompress_fetch_ex(8, buf, buf, 2, 4, buf, buf, buf, None, 0, 2, 0, 0, 0, 2, 0, ctypes.byref(lay),
                                   None)
    assert st == 1 and "transfer baseline" in capi.last_error()
    offs = (ctypes.c_int64 * 2)(32, 64)
    lay = capi.Layout(ctypes.addressof(offs), 64, None, 1)
    st = L.ibp_compress_ex(8, buf, buf, 2, 4, buf, buf, None, 0, 2, None, ctypes.byref(lay), None)
    assert st == 1 and "overlap" in capi.last_error()
    st = L.ibp_condensed_offsets(8, buf, 2, 4, buf, buf, None, 0, 2, offs, ctypes.byref(lay), None)
    assert st == 1 and "output" in capi.last_error()
    lay = capi.Layout(None, 0, ctypes.addressof(cl), 0)
    st = L.ibp_get_compress_size_ex(8, buf, 2, 4, buf, buf, None, 0, 2, buf, None, ctypes.byref(lay), None)
    assert st == 1 and "num_clusters" in capi.last_error()
    plan = ctypes.c_void_p()
    assert L.ibp_plan_create(None, 0, ctypes.byref(plan)) == 1
    cfg = capi.ServerConfig(16, 1, 1, 1)
    srv = ctypes.c_void_p()
    assert L.ibp_server_create(ctypes.byref(cfg), ctypes.byref(srv)) == 1


def test_preprocess_workspace_size(capi):
    L = capi.lib()
    sweep = L.ibp_preprocess_workspace_size(8, 64, -1.0)
    single = L.ibp_preprocess_workspace_size(8, 64, 0.9)
    assert sweep > single >= 64 * 64 * 4


def test_python_api_rejects_unreachable_memory():
    """Plain (pageable) CPU tensors are not reachable by the GPU: the same
    RuntimeError as the reference's TORCH_CHECK, raised before any HIP call."""
    import invariantbitpacking_amd as ibp
    x = torch.zeros(4, 8, dtype=torch.int64)
    m = torch.zeros(8, dtype=torch.int64)
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp.decompress_fetch(x, m, m, torch.zeros(1, dtype=torch.int32), torch.device("cuda"))
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp.preprocess(x)
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp.compress_inplace(x, m, m)
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp.get_compress_size(x, m, m)
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp.compress(x, m, m)
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp.compress_condensed(x, m, m)
    with pytest.raises(RuntimeError, match="accessible by CUDA device"):
        ibp.decompress(x.view(-1), torch.zeros(4, dtype=torch.int64), m, m, torch.device("cuda"), 8)


def test_reference_api_surface():
    """Names and positional argument order of ibp/__init__.py:8-41."""
    import inspect

    import invariantbitpacking_amd as ibp
    want = {
        "print_debug": ["flag"],
        "preprocess": ["dataset", "threshold"],
        "preprocess_kmeans": ["dataset", "centroids", "threshold"],
        "get_compress_size": ["dataset", "mask", "bitval", "index_arr", "compress_total"],
        "get_single_compress_len": ["dataset", "mask", "bitval", "index_arr"],
        "compress_inplace": ["dataset", "mask", "bitval", "index_arr"],
        "compress": ["dataset", "mask", "bitval", "index_arr"],
        "decompress_fetch": ["comp_dataset", "mask", "bitval", "bitmask", "device", "comp_len", "index_arr",
                             "nblks", "nthds", "imp", "output_tensor"],
        # the extension module's further bindings (src/ibp_api.cu:131,133)
        "compress_condensed": ["dataset", "mask", "bitval", "index_arr"],
        "read_shared": ["filename", "shape", "dtype"],
    }
    for name, params in want.items():
        sig = inspect.signature(getattr(ibp, name))
        positional = [p.name for p in sig.parameters.values()
                      if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
        assert positional == params, name


def test_read_shared_loads_file_into_shared_memory(tmp_path):
    """read_shared (src/ibp_api.cu:52-122): the file's bytes in a shared-memory
    tensor of the given shape and dtype; no GPU involved."""
    import numpy as np

    import invariantbitpacking_amd as ibp
    rng = np.random.default_rng(0)
    data = rng.integers(-2**62, 2**62, size=(1000, 37), dtype=np.int64)
    f = tmp_path / "rows.bin"
    data.tofile(f)
    t = ibp.read_shared(str(f), [1000, 37], torch.int64)
    assert t.is_shared() and t.shape == (1000, 37) and t.dtype == torch.int64
    assert np.array_equal(t.numpy(), data)
    # several reader threads, pieces larger than one thread's share
    big = rng.integers(0, 255, size=(150 << 20) // 4, dtype=np.int32)
    g = tmp_path / "big.bin"
    big.tofile(g)
    t2 = ibp.read_shared(str(g), [big.size], torch.int32, threads=5)
    assert np.array_equal(t2.numpy(), big)
    # a shorter file leaves the rest zero; a longer one is an error
    t3 = ibp.read_shared(str(f), [1100, 37], torch.int64)
    assert np.array_equal(t3[:1000].numpy(), data) and not t3[1000:].any()
    with pytest.raises(RuntimeError, match="more than"):
        ibp.read_shared(str(f), [999, 37], torch.int64)
    with pytest.raises(RuntimeError, match="cannot open"):
        ibp.read_shared(str(tmp_path / "missing.bin"), [4], torch.int64)


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct the C ABI passes by pointer has the same size and field
    offsets in include/ibp.h (compiled here with gcc) and in the ctypes
    mirror the Python package uses."""
    import shutil
    import subprocess
    from invariantbitpacking_amd import _capi
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"ibp_fetch_task": _capi.FetchTask, "ibp_layout": _capi.Layout,
               "ibp_server_config": _capi.ServerConfig, "ibp_server_stats": _capi.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ibp.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (name, name))
        for field, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (name, field, name, field))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(got["%s sizeof" % name]) == ctypes.sizeof(cls), name
        for field, _ in cls._fields_:
            assert int(got["%s %s" % (name, field)]) == getattr(cls, field).offset, (name, field)


def _build(cmd, tmp_path):
    import subprocess
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(tmp_path))
    assert r.returncode == 0, "build failed:\n" + r.stdout + r.stderr


def test_index_copy_under_thread_sanitizer(tmp_path):
    """The server's split index copy (host code) built with ThreadSanitizer
    and driven from three submitting threads with idle gaps: no data race is
    reported and every copy is exact."""
    import os
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "copy_tsan"
    _build(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-pthread",
            "-I" + os.path.join(root, "invariantbitpacking_amd", "csrc"),
            os.path.join(root, "tests", "index_copy_tsan_driver.cpp"), "-o", str(exe)], tmp_path)
    r = subprocess.run([str(exe), "25"], capture_output=True, text=True, timeout=240,
                       env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0 and "WARNING: ThreadSanitizer" not in r.stderr, r.stdout + r.stderr
