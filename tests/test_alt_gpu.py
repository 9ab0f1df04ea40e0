"""ALT variant (jw_type 110-113: rows couple with rows two apart, i.e. the even and the
odd rows are two interleaved systems) on the GPU through the C ABI, against the oracle's
ALT restatement (pinned by dense stride-2 operators and by the reference's reorder_i_alt /
get_ALT_stride compiled under oracle/_ref).  Matrices in the ALT storage: each diagonal a
block of get_ALT_stride entries in reorder_i_alt order, baked blocks at 0, S, 2S.  Bar: FP64
<= 1e-13, double-double <= 1e-30; rows outside the output keep-range and the gaps between
systems untouched."""
import numpy as np
import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

import _oracle as O
from jwgpu import run_plan

pytestmark = pytest.mark.gpu

MODES = [(0, 1), (0, 0), (1, 1), (1, 0), (2, 1), (2, 0), (3, 2), (3, 1), (3, 0)]
RT_IO = (_abi.RUNTIME_OFFSETSOLUTION | _abi.RUNTIME_OUTPUTBUFFERSTRIDE | _abi.RUNTIME_INPUTZEROPAD
         | _abi.RUNTIME_OUTPUTZEROPAD | _abi.RUNTIME_SCALEC)


def alt_buffer(jw, ub, N, seed):
    """junk everywhere, diagonal blocks with dominant diagonals where the mode reads them"""
    S = O.alt_stride(N)
    P = np.array([O.alt_pos(i, N) for i in range(N)], dtype=np.int64)
    buf = O.uniform(seed, -0.3, 0.3, 3 * S + 8)
    if ub == 2 or jw == 3:
        md = S if (ub == 2 or ub == 1) else 0
    elif jw in (0, 1):
        md = 0 if ub else S
    else:
        md = None
    if md is not None:
        buf[md + P] = O.uniform(seed + 1, 0.5, 1.5, N)
    return buf


@pytest.mark.parametrize("jw,ub", MODES)
@pytest.mark.parametrize("N", [1, 2, 3, 17, 64, 65, 130, 257, 1000, 2048])
@pytest.mark.parametrize("rt", [0, 1])
def test_alt_modes(jw, ub, N, rt):
    B = 5
    buf = alt_buffer(jw, ub, N, 300 + N + jw)
    S = N + 7 if rt else N
    offS = 3 if rt else 0
    izp = (1, N - 1) if rt else (0, 0)
    ozp = (2, N) if rt else (0, 0)
    x = O.uniform(11 + N, -1, 1, B * S + offS + 6)
    cfg = dict(size=[N, B, 1], jw_type=110 + jw, upperBanded=ub, doublePrecision=1, scaleC=1.0,
               jw_control_bitmask=RT_IO if rt else 0)
    launch = dict(offsetSolution=offS, outputBufferStride=S, inputZeropad=izp, outputZeropad=ozp,
                  scaleC=0.75) if rt else {}
    got = run_plan(cfg, buf, x, launch)
    p = O.problem(jw_type=110 + jw, upperBanded=ub, N=N, B=B, mask=RT_IO if rt else 0, cfg_outStride=N,
                  cfg_scaleC=1.0, run_offsetSolution=offS, run_izp=list(izp), run_ozp=list(ozp),
                  run_scaleC=0.75, run_outStride=S)
    want = O.jw_f128(p, buf, x)
    lo, hi = (ozp[0], min(ozp[1], N)) if rt else (0, N)
    keep = np.zeros(x.size, bool)
    for b in range(B):
        base = b * S + offS
        if hi > lo:
            assert O.relerr_norm(got[base + lo:base + hi], want[base + lo:base + hi]) <= 1e-13, b
        keep[base + lo:base + hi] = True
    np.testing.assert_array_equal(got[~keep], x[~keep])


@pytest.mark.parametrize("N", [5, 200, 1024])
@pytest.mark.parametrize("jw,ub", [(0, 1), (3, 2)])
def test_alt_double_double(N, jw, ub):
    B = 4
    hi = alt_buffer(jw, ub, N, 77 + N)
    lo = O.uniform(78 + N, -1, 1, hi.size) * 2.0 ** -60 * np.abs(hi)
    buf = np.ascontiguousarray(np.stack([hi, lo], -1))
    xh = O.uniform(79 + N, -1, 1, B * N)
    x = np.ascontiguousarray(np.stack([xh, O.uniform(80 + N, -1, 1, B * N) * 2.0 ** -60 * np.abs(xh)], -1))
    cfg = dict(size=[N, B, 1], jw_type=110 + jw, upperBanded=ub, quadDoubleDoublePrecision=1, scaleC=0.5)
    got = run_plan(cfg, buf.ravel(), x.ravel()).reshape(-1, 2)
    p = O.problem(jw_type=110 + jw, upperBanded=ub, N=N, B=B, cfg_outStride=N, cfg_scaleC=0.5)
    want = O.jw_f128_dd(p, buf, x)
    for b in range(B):
        assert O.dd_relerr_norm(got[b * N:(b + 1) * N], want[b * N:(b + 1) * N]) <= 1e-30, b


def test_alt_buffer_size_guard():
    """the matrix blocks of the ALT storage are get_ALT_stride long: a buffer that holds 3N
    but not the blocks is refused before any launch"""
    N, B = 100, 2
    S = O.alt_stride(N)
    assert S > N
    buf = alt_buffer(0, 1, N, 5)[:3 * N + 2]
    x = O.uniform(1, -1, 1, B * N)
    code = run_plan(dict(size=[N, B, 1], jw_type=110, upperBanded=1, doublePrecision=1), buf, x, {}, expect=1)
    assert code == _abi.PFSOLVE_ERROR_EMPTY_bufferSize
