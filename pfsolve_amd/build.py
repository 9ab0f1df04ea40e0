"""Build recipe of libpfsolve_gfx950.so (in-tree, next to this file).

C host sources are compiled with gcc, the HIP device source with hipcc for
gfx950 only, and the shared library is linked with hipcc.  Run as
``python -m pfsolve_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfsolve_gfx950.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

C_SOURCES = ["jw_plan.c", "pfsolve_api.c", "jw_batch.c", "jw_serve.c"]
HIP_SOURCES = ["jw_device.hip", "jw_main64.hip", "jw_long.hip", "jw_dd64.hip"]

C_FLAGS = [
    "-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-D__HIP_PLATFORM_AMD__", "-I" + INCLUDE, "-I" + CSRC, "-I" + os.path.join(ROCM, "include"),
    "-pthread",
]
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950", "-mcode-object-version=5",
    # packed FP32 forms are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # exact error-free transforms in the double-double path; FMAs are explicit
    "-ffp-contract=off",
    "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    # keep warnings visible, drop hipcc's note about the host-side feature flag
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        sys.stderr.write(out + "\n")


def build(verbose: bool = True) -> str:
    """Compile every source and link the library; returns the library path."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in C_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append(["gcc", *C_FLAGS, "-c", "-o", obj, path])
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-o", obj, path])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build()
