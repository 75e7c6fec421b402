"""Build liblkgpu.so (HIP kernels for gfx950 + C++ host runtime + C ABI).

Usage: python -m libkriging_amd.build   (or libkriging_amd.build.build())
Objects go to libkriging_amd/_build/, the library to libkriging_amd/liblkgpu.so.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

from ._srchash import source_hash

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblkgpu.so")

SOURCES = [
    "lk_kernels.hip",
    "lk_nested.hip",
    "lk_loo.hip",
    "lk_noalloc.hip",
    "lk_runtime.cpp",
    "lk_capi.cpp",
    "lbfgsb.cpp",
    "kriging.cpp",
    "kriging_capi.cpp",
    "kriging_io.cpp",
    "model_io.cpp",
    "nested.cpp",
    "nested_capi.cpp",
]

ARCH = os.environ.get("LKGPU_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-pass-failed",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src, defines=()):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, *defines, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


PROBE_SRC = os.path.join(CSRC, "probe", "lk_probe.hip")
PROBE_LIB = os.path.join(HERE, "liblkprobe.so")


def build_probe(verbose=True):
    """liblkprobe.so: the benchmark's copy and MFMA probes (not the library)"""
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC, "-x", "hip", *COMMON, "-shared", PROBE_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, PROBE_LIB)
    if verbose:
        print(f"built {PROBE_LIB}")
    return PROBE_LIB


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    # the C ABI translation unit carries the hash of the sources, which the
    # loader checks against the sources beside the library
    bid = [f'-DLKGPU_BUILD_ID="{source_hash()}"']
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, bid if s == "lk_capi.cpp" else ()), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_probe(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
