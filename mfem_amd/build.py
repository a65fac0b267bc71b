"""Build libh1pa.so (HIP for gfx950 + host C++) in-tree.

    python -m mfem_amd.build        # or mfem_amd.build.build()

Every translation unit is compiled with hipcc for gfx950 only; packed FP32
instructions are disabled (-packed-fp32-ops) for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libh1pa.so")
SERVE = os.path.join(HERE, "h1pa_serve")
ROOT = os.path.dirname(HERE)
# C++ callers of the C ABI that the tests run (each a main() over include/h1pa.h)
CAPI_TESTS = ["capi_mesh_arrays"]
CAPI_BIN = os.path.join(ROOT, "tests", "_bin")
CEILINGS = os.path.join(HERE, "h1pa_ceilings")

SOURCES = ["basis.cpp", "mesh.cpp", "capi.cpp", "restriction.hip", "geometry.hip",
           "apply.hip", "columns.hip", "pcg.hip", "batch.hip", "qinterp.hip", "transfer.hip",
           "multigrid.cpp", "devmem.cpp"]

ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
         "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # the serving driver: a native executable over the C ABI
    tmp = SERVE + ".tmp"
    cmd = [_hipcc(), "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "serve_main.cpp"),
           "-o", tmp, f"-L{HERE}", "-lh1pa", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"h1pa_serve build failed:\n{r.stderr}")
    os.replace(tmp, SERVE)
    # the roofline probes the kernels are measured against (standalone HIP)
    tmp = CEILINGS + ".tmp"
    cmd = [_hipcc(), *[f for f in FLAGS if f != "-fPIC"], os.path.join(CSRC, "ceilings.hip"),
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"h1pa_ceilings build failed:\n{r.stderr}")
    os.replace(tmp, CEILINGS)
    os.makedirs(CAPI_BIN, exist_ok=True)
    for name in CAPI_TESTS:
        exe = os.path.join(CAPI_BIN, name)
        cmd = [_hipcc(), "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
               os.path.join(ROOT, "tests", name + ".cpp"), "-o", exe + ".tmp", f"-L{HERE}",
               "-lh1pa", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../mfem_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name} build failed:\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {LIB} and {SERVE}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
