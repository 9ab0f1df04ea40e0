"""The C caller (tests/c_caller/jw_caller.c, compiled against include/pfSolve.h) solves the
reference harness's 513-row tridiagonal system and seeded right-hand sides on the GPU;
its output matches the __float128 oracle (<= 1e-13) and rows outside the systems stay."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import c_caller
import jw_systems as J

pytestmark = pytest.mark.gpu


def test_c_caller_solves_reference_system(tmp_path):
    a, b, c, x0 = J.fixture_513()
    N, B, S = b.size, 16, b.size + 128       # the harness pads systems by 128 rows
    mat = J.tridiag_buffer(a, b, c, natural=False)
    x = O.uniform(77, -1, 1, B * S)
    x[:N] = x0
    exe = c_caller.EXE
    assert os.path.exists(exe), "tests/c_caller/jw_caller is built by __graft_entry__.build()"
    paths = [os.path.join(str(tmp_path), f) for f in ("m.bin", "x.bin", "y.bin")]
    mat.tofile(paths[0])
    x.tofile(paths[1])
    print(f"c caller N={N} B={B} stride={S}", flush=True)
    r = subprocess.run([exe, str(N), str(B), str(S), *paths], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(paths[2])
    p = O.problem(jw_type=13, upperBanded=2, N=N, B=B, mask=256, cfg_outStride=N, cfg_scaleC=1.0, run_outStride=S)
    want = O.jw_f128(p, mat, x)
    for k in range(B):
        assert O.relerr_norm(got[k * S:k * S + N], want[k * S:k * S + N]) <= 1e-13, k
        np.testing.assert_array_equal(got[k * S + N:(k + 1) * S], x[k * S + N:(k + 1) * S])
    assert J.tridiag_residual(a, b, c, got[:N], x0) <= 1e-14
