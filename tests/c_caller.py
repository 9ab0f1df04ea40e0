"""Build recipe of the C caller (tests/c_caller/jw_caller.c) against include/pfSolve.h and
libpfsolve_gfx950.so -- test helper."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_caller", "jw_caller.c")
LIBDIR = os.path.join(ROOT, "pfsolve_amd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
EXE = os.path.join(ROOT, "tests", "c_caller", "jw_caller")   # built in-tree by build()


def build(out_dir: str) -> str:
    exe = os.path.join(out_dir, "jw_caller")
    cmd = ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-O2",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROCM, "include"), SRC,
           "-L" + LIBDIR, "-lpfsolve_gfx950", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + LIBDIR, "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe
