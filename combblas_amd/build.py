"""Build recipe for libcbx.so (HIP kernels for gfx950 plus host C++).

Run ``python -m combblas_amd.build`` (or ``__graft_entry__.build()``).  HIP
sources go through hipcc for gfx950; host-only C++ (generators, OpenMP code)
goes through g++.  Objects land in ``combblas_amd/_build/`` and the shared
library in ``combblas_amd/libcbx.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB = PKG / "libcbx.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("CBX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _hipcc() -> str:
    p = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found; set ROCM_PATH")
    return p


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source and link libcbx.so.  Always rebuilds."""
    OUT_DIR.mkdir(exist_ok=True)
    hip_srcs = sorted(CSRC.glob("*.hip"))
    cpp_srcs = sorted(CSRC.glob("*.cpp"))
    hipcc = _hipcc()
    jobs = jobs or min(16, os.cpu_count() or 4)
    cmds = []
    objs = []
    for src in hip_srcs:
        obj = OUT_DIR / (src.stem + ".hip.o")
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    for src in cpp_srcs:
        obj = OUT_DIR / (src.stem + ".cpp.o")
        cmds.append(["g++", *GXX_FLAGS, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, c) for c in cmds]
        for f, c in zip(futs, cmds):
            if verbose:
                print(" ".join(c), flush=True)
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lgomp",
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
