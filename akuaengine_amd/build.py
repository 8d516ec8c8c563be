"""Builds akuaengine_amd/libakua.so from csrc/ with hipcc for gfx950.

Every source is compiled with
  * -ffp-contract=off and correctly rounded FP32 division / square root, so the kernels round
    exactly like the CPU oracle (no fused multiply-add unless written explicitly);
  * packed-FP32 instructions disabled (-packed-fp32-ops).
Usage: python -m akuaengine_amd.build [-j N]
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libakua.so")
OBJ = os.path.join(HERE, "_build")

ARCH = os.environ.get("AKUA_OFFLOAD_ARCH", "gfx950")
SOURCES = ["radix_sort.hip", "pbf_kernels.hip", "mlp.hip", "mlp_noalloc.hip", "aos.hip",
           "stream_copy.hip", "selftest.hip", "solver.cpp", "boundary.cpp", "scene_io.cpp",
           "render.hip"]
RUNNER = os.path.join(HERE, "akua_run")

HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_runner()
    if verbose:
        print(f"built {LIB} and {RUNNER}")
    return LIB


def build_runner() -> str:
    """The headless driver akua_run: host C++ over the C ABI (include/akua.h, akua_io.h)."""
    cxx = os.environ.get("CXX") or shutil.which("g++") or shutil.which("c++")
    if not cxx:
        raise RuntimeError("no host C++ compiler (g++) found")
    tmp = RUNNER + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(CSRC, "akua_run.cpp"), "-o", tmp, f"-L{HERE}", "-lakua", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"akua_run build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, RUNNER)
    return RUNNER


if __name__ == "__main__":
    j = 4
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    build(j, verbose=True)
