"""Chained JW steps (numConsecutiveJWIterations = K > 1) on the GPU through the C ABI,
against the oracle's chained restatement (__float128, pinned by the dense composition
test): one launch keeps each pair of systems in registers through the K iterations,
iteration i has N -/+ i rows and its own matrices.  Bar: FP64 <= 1e-13 normwise relative
error per system; rows outside the last iteration's output keep-range and the gaps between
systems bit-exactly untouched."""
import numpy as np
import pytest

from pfsolve_amd import _abi
import pfsolve_amd as pf

import _oracle as O
import jw_chain as JC

pytestmark = pytest.mark.gpu
TOL64 = 1e-13
RT_IO = (_abi.RUNTIME_OFFSETSOLUTION | _abi.RUNTIME_OUTPUTBUFFERSTRIDE | _abi.RUNTIME_INPUTZEROPAD
         | _abi.RUNTIME_OUTPUTZEROPAD | _abi.RUNTIME_SCALEC)


def run_chain(N, K, jw, ub, rt, multi, B=6, natural=False, izp=None, ozp=None, seed=0):
    import torch
    offV = 7 + 2 * (N + K) if rt else None
    bufs, dense, Ns = JC.build(N, K, jw, ub, rt, seed=seed or 17 * N + K, natural=natural, offV=offV)
    S = max(N, Ns[-1]) + 6
    S += S & 1   # the chained kernel needs 16-byte aligned pairs of systems
    offS = 2
    x = O.uniform(N * 3 + K, -1, 1, B * S + offS + 5)
    izp = izp or (0, N)
    ozp = ozp or (0, Ns[-1])
    mask = RT_IO | ((_abi.RUNTIME_OFFSETM | _abi.RUNTIME_OFFSETV) if rt else 0)
    jt = (1000 if natural else 0) + 10 + jw
    if multi:
        dmats = [torch.from_numpy(b).cuda() for b in bufs]
    else:   # one buffer for every iteration: the first iteration's matrices
        dmats = [torch.from_numpy(bufs[0]).cuda()]
        bufs = [bufs[0]] * K
    drhs = torch.from_numpy(x.copy()).cuda()
    stream = torch.cuda.current_stream().cuda_stream
    print(f"chain N={N} K={K} jw=1{jw} ub={ub} rt={rt} multi={multi}", flush=True)
    with pf.Plan([m.data_ptr() for m in dmats], drhs.data_ptr(), stream=stream, size=[N, B, 1], jw_type=jt,
                 upperBanded=ub, doublePrecision=1, scaleC=1.0, numConsecutiveJWIterations=K,
                 useMultipleInputBuffers=K if multi else 0, jw_control_bitmask=mask) as plan:
        lp = dict(offsetSolution=offS, outputBufferStride=S, inputZeropad=izp, outputZeropad=ozp, scaleC=0.8)
        if rt:
            lp.update(offsetM=7, offsetV=offV)
        plan.append(**lp)
    torch.cuda.synchronize()
    got = drhs.cpu().numpy()
    p = O.problem(jw_type=jt, upperBanded=ub, N=N, B=B, mask=mask, cfg_outStride=N, cfg_scaleC=1.0,
                  run_offsetM=7, run_offsetV=offV or 0, run_offsetSolution=offS, run_outStride=S,
                  run_izp=list(izp), run_ozp=list(ozp), run_scaleC=0.8)
    want = O.jw_chain(p, K, bufs, x)
    lo, hi = max(0, ozp[0]), min(ozp[1], Ns[-1])
    for b in range(B):
        base = b * S + offS
        if hi > lo:
            assert O.relerr_norm(got[base + lo:base + hi], want[base + lo:base + hi]) <= TOL64, b
    keep = np.zeros(got.size, bool)
    for b in range(B):
        keep[b * S + offS + lo:b * S + offS + hi] = True
    np.testing.assert_array_equal(got[~keep], x[~keep])
    return got, want


@pytest.mark.parametrize("K", [2, 4, 8])
@pytest.mark.parametrize("ub", [1, 0])
@pytest.mark.parametrize("N", [64, 100, 512, 1000])
def test_chain_connection_modes(N, K, ub):
    """jw 10 (matvec + solve) with one buffer per iteration"""
    if not ub and N + K - 1 > 1024:
        pytest.skip("growing systems beyond 1024 rows")
    run_chain(N, K, 0, ub, rt=0, multi=True)


@pytest.mark.parametrize("jw", [0, 1, 2])
@pytest.mark.parametrize("ub", [1, 0])
@pytest.mark.parametrize("rt", [0, 1])
def test_chain_modes_offsets(jw, ub, rt):
    """matvec only, solve only, both; baked and runtime offsets; keep-ranges"""
    run_chain(200, 4, jw, ub, rt, multi=True, izp=(3, 190), ozp=(1, 195 if ub else 201))


@pytest.mark.parametrize("ub", [1, 0])
def test_chain_single_buffer_and_natural_order(ub):
    """without multiple input buffers every iteration reads buffer[0]; natural storage order"""
    run_chain(300, 3, 0, ub, rt=0, multi=False)
    run_chain(130, 5, 0, ub, rt=1, multi=True, natural=True)


def test_chain_one_iteration_equals_single_step():
    """K = 1 through the chained test harness: the plain single-step kernels"""
    run_chain(256, 1, 0, 1, rt=0, multi=True)


def test_chain_refuses_unaligned_systems():
    """an odd stride (systems not in 16-byte aligned pairs) is refused before any launch"""
    import torch
    N, K, B = 100, 3, 4
    bufs, _, _ = JC.build(N, K, 0, 1, 0, seed=5)
    dmats = [torch.from_numpy(b).cuda() for b in bufs]
    drhs = torch.zeros(B * (N + 1) + 8, dtype=torch.float64, device="cuda")
    with pf.Plan([m.data_ptr() for m in dmats], drhs.data_ptr(), size=[N, B, 1], jw_type=10, upperBanded=1,
                 doublePrecision=1, numConsecutiveJWIterations=K, useMultipleInputBuffers=K,
                 outputBufferStride=[N + 1]) as plan:
        with pytest.raises(pf.PfSolveError) as e:
            plan.append()
        assert e.value.code == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
