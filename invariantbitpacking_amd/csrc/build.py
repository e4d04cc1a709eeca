"""Builds libibp.so (all HIP sources, gfx950) and the ibp_cuda torch extension
(_C, the host-side bindings over libibp's C ABI) next to this package.

Usage: python -m invariantbitpacking_amd.csrc.build [--jobs N]
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libibp.so")
EXT_OUT = os.path.join(PKG, "_C" + sysconfig.get_config_var("EXT_SUFFIX"))
OBJ_DIR = os.path.join(PKG, "_build")
SOURCES = ["ibp_core.hip", "ibp_fetch.hip", "ibp_fetch_u64.hip", "ibp_fetch_u32.hip", "ibp_fetch_narrow.hip",
           "ibp_pack.hip", "ibp_preproc.hip", "ibp_kmeans.hip", "ibp_batch.hip", "ibp_serve.hip", "ibp_io.hip"]
EXT_SOURCE = "ibp_torch.cpp"
ARCH = os.environ.get("IBP_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         # packed FP32 instructions are kept out of every HIP build
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fno-fast-math", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
         "-I" + os.path.join(ROOT, "include")]


def hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def rocm_path():
    return os.environ.get("ROCM_PATH", "/opt/rocm")


def _compile(src):
    obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
    cmd = [hipcc(), "-c", os.path.join(HERE, src), "-o", obj] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj, r.stderr


def _torch_dirs():
    """Include and library directories of the installed PyTorch (ROCm build),
    found without importing torch (the package's own import is the slow part)."""
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        raise RuntimeError("PyTorch is not installed")
    base = list(spec.submodule_search_locations)[0]
    inc = os.path.join(base, "include")
    return [inc, os.path.join(inc, "torch", "csrc", "api", "include")], os.path.join(base, "lib")


def _cxx11_abi():
    """_GLIBCXX_USE_CXX11_ABI of the installed PyTorch."""
    r = subprocess.run([sys.executable, "-c", "import torch; print(int(torch.compiled_with_cxx11_abi()))"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cannot query PyTorch's C++ ABI:\n" + r.stderr)
    return r.stdout.strip()


def _compile_ext(abi):
    """The torch extension's object: host C++ only (g++), no device code."""
    incs, _ = _torch_dirs()
    obj = os.path.join(OBJ_DIR, EXT_SOURCE.replace(".cpp", ".o"))
    cmd = (["g++", "-c", os.path.join(HERE, EXT_SOURCE), "-o", obj, "-O2", "-std=c++17", "-fPIC", "-Wall",
            "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-DTORCH_EXTENSION_NAME=_C",
            "-DTORCH_API_INCLUDE_EXTENSION_H", "-D_GLIBCXX_USE_CXX11_ABI=" + abi,
            "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(rocm_path(), "include"),
            "-I" + sysconfig.get_paths()["include"]] + ["-I" + d for d in incs])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for %s:\n%s%s" % (EXT_SOURCE, r.stdout, r.stderr))
    return obj, r.stderr


def _link_ext(obj):
    """Links _C against libibp.so (found next to it through $ORIGIN) and the
    PyTorch/HIP runtime libraries (already loaded by `import torch`)."""
    _, torch_lib = _torch_dirs()
    tmp = EXT_OUT + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp, obj, "-L" + PKG, "-l:libibp.so", "-Wl,-rpath,$ORIGIN",
           "-L" + torch_lib, "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python",
           "-L" + os.path.join(rocm_path(), "lib"), "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link of %s failed:\n%s%s" % (os.path.basename(EXT_OUT), r.stdout, r.stderr))
    os.replace(tmp, EXT_OUT)


def build(jobs=None, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    jobs = jobs or min(len(srcs) + 1, os.cpu_count() or 4)
    abi = _cxx11_abi()
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        ext_job = ex.submit(_compile_ext, abi)
        results = list(ex.map(_compile, srcs))
        ext_obj, ext_err = ext_job.result()
    if verbose:
        for _, err in results + [(ext_obj, ext_err)]:
            if err.strip():
                sys.stderr.write(err)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + [o for o, _ in results] + [
        "--offload-arch=" + ARCH, "-fPIC", "-L" + os.path.join(rocm_path(), "lib"), "-lrocprofiler-sdk-roctx"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    _link_ext(ext_obj)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
