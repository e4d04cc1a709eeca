"""Builds libibp.so (all HIP sources, gfx950) next to this package.

Usage: python -m invariantbitpacking_amd.csrc.build [--jobs N]
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libibp.so")
OBJ_DIR = os.path.join(PKG, "_build")
SOURCES = ["ibp_core.hip", "ibp_fetch.hip", "ibp_fetch_u64.hip", "ibp_fetch_u32.hip", "ibp_fetch_narrow.hip",
           "ibp_pack.hip", "ibp_preproc.hip", "ibp_kmeans.hip", "ibp_batch.hip", "ibp_serve.hip", "ibp_io.hip"]
ARCH = os.environ.get("IBP_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         # packed FP32 instructions are kept out of every HIP build
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fno-fast-math", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
         "-I" + os.path.join(os.path.dirname(PKG), "include")]


def hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _compile(src):
    obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
    cmd = [hipcc(), "-c", os.path.join(HERE, src), "-o", obj] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj, r.stderr


def build(jobs=None, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, [s for s in SOURCES if os.path.exists(os.path.join(HERE, s))]))
    if verbose:
        for obj, err in results:
            if err.strip():
                sys.stderr.write(err)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + [o for o, _ in results] + ["--offload-arch=" + ARCH, "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
