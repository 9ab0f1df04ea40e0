"""GPU parity of long systems (FP64 N > 2048, double-double N > 512): every system cut into
chunks of 64 R rows, the pivots continued from chunk to chunk, each chunk reduced to its
boundary values, the reduced recurrences solved per system and every chunk finished from
its carries.  Matrices in the reference's storage: Msplit0-row blocks of reorder_i0 beyond
3840 rows.  Checked against the __float128 oracle (FP64 <= 1e-13, DD <= 1e-30), with the
rows outside the output keep-range and the gaps between systems bit-exactly untouched."""
import numpy as np
import pytest

from pfsolve_amd import _abi

import _oracle as O
from jwgpu import run_plan, to_dev
from test_jw_gpu import MODES, RT_ALL, TOL64, check_systems, dd_inputs, fused_buffers, mode_buffer

pytestmark = pytest.mark.gpu

LONG_SIZES = [2049, 3000, 3841, 10000]


@pytest.mark.parametrize("jw_type,ub", MODES)
@pytest.mark.parametrize("N", LONG_SIZES)
@pytest.mark.parametrize("rt", [0, 1, 2])
def test_long_reference_modes(jw_type, ub, N, rt):
    """every reference mode; rt: runtime matrix / solution offsets, keep-ranges that cut
    chunks and scaleC, with an odd solution offset (rt = 1: the chunk kernel) or with even
    strides and offsets (rt = 2: 16-byte aligned systems, the pair kernel's chunks)"""
    print(f"long modes jw{jw_type} ub{ub} N={N} rt={rt}", flush=True)
    jw = jw_type % 10
    B = 5
    offM, offV = (13, 3 * N + 29) if rt else (0, 0)
    total = max(3 * N, offV + N, offM + 2 * N) + 8
    buf = mode_buffer(jw, ub, N, rt, offM, total, 3000 + N + jw)
    S = (N + 7 if rt == 1 else N + 8 - N % 2) if rt else N
    offS = (5 if rt == 1 else 6) if rt else 0
    izp = (N // 4 + 1, N - 3) if rt else (0, 0)
    ozp = (515, N - N // 8) if rt else (0, 0)
    x = O.uniform(91 + N, -1, 1, B * S + offS + 9)
    cfg = dict(size=[N, B, 1], jw_type=jw_type, upperBanded=ub, doublePrecision=1, scaleC=1.0,
               outputBufferStride=[N, 0, 0], jw_control_bitmask=RT_ALL if rt else 0)
    launch = dict(offsetM=offM, offsetV=offV, offsetSolution=offS, inputZeropad=izp, outputZeropad=ozp,
                  scaleC=0.75, outputBufferStride=S) if rt else {}
    got = run_plan(cfg, buf, x, launch)
    p = O.problem(jw_type=jw_type, upperBanded=ub, N=N, B=B, mask=RT_ALL if rt else 0, cfg_outStride=N,
                  cfg_scaleC=1.0, run_offsetM=offM, run_offsetV=offV, run_offsetSolution=offS,
                  run_izp=list(izp), run_ozp=list(ozp), run_scaleC=0.75, run_outStride=S)
    want = O.jw_f128(p, buf, x)
    lo, hi = (ozp[0], min(ozp[1], N)) if rt else (0, N)
    check_systems(got, want, x, B, S, offS, N, lo, max(lo, hi), TOL64)


@pytest.mark.parametrize("N", [2049, 3000, 3841, 4097, 10000, 100000])
@pytest.mark.parametrize("permuted", [1, 0])
def test_long_fused_step(N, permuted):
    """the fused step y = M^-1 T^-1 V x (pivots through every chunk), baked positions"""
    print(f"long fused N={N} permuted={permuted}", flush=True)
    B = 4 if N == 100000 else 7
    raw, buf, (offV, offT, offM) = fused_buffers(N, N + 1, permuted)
    x = O.uniform(17 + N, -1, 1, B * N)
    jt = 14 if permuted else 1014
    cfg = dict(size=[N, B, 1], jw_type=jt, doublePrecision=1, scaleC=0.5, KL=2, KU=2,
               offsetV=offV, offsetT=offT, offsetM=offM)
    got = run_plan(cfg, buf, x)
    p = O.problem(jw_type=jt, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=0.5,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, N, 0, N, 0, N, TOL64)
    # and against the raw-matrix __float128 evaluation (independent of storage) for one system
    if N <= 10000:
        y = np.zeros(N)
        O.lib().pfo_fused_raw_f128(N, 2, 2, *[O.P(a) for a in raw], O.P(np.ascontiguousarray(x[:N])), O.P(y))
        assert O.relerr_norm(got[:N], y * np.r_[0.5, np.ones(N - 1)]) <= TOL64   # row-0 scaleC


@pytest.mark.parametrize("N", [3000, 10000])
def test_long_chunk_kernel_on_aligned_systems(N):
    """kernelVariant 1 keeps 16-byte aligned FP64 systems on the chunk kernel (format-1 tables,
    separate launches); the default takes the pair path: both against the oracle, and the two
    against each other"""
    print(f"long chunk kernel aligned N={N}", flush=True)
    B = 6
    raw, buf, (offV, offT, offM) = fused_buffers(N, N + 2, 1)
    x = O.uniform(23 + N, -1, 1, B * N)
    outs = []
    for variant in (1, 0):
        cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=0.5, KL=2, KU=2, kernelVariant=variant,
                   offsetV=offV, offsetT=offT, offsetM=offM)
        outs.append(run_plan(cfg, buf, x))
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=0.5,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    for got in outs:
        check_systems(got, want, x, B, N, 0, N, 0, N, TOL64)
    assert O.relerr_norm(outs[0], outs[1]) <= 2 * TOL64


@pytest.mark.parametrize("pad", [3, 4])
def test_long_fused_runtime_keep_ranges_and_custom_split(pad):
    """caller-chosen Msplit0 (1000-row storage blocks), runtime keep-ranges whose ends fall
    inside chunks, runtime scaleC; odd stride (the chunk kernel) or even (the pair kernel)"""
    N, B = 5000, 5
    S = N + pad
    raw = O.raw_matrices(77, N)
    offV, offT, offM = 0, 5 * N + 3, 8 * N + 11
    V, Ta, Tb, Tc, Md, Me = raw
    buf = np.zeros(offM + 2 * N + 4)
    O.lib().pfo_pack_fused(N, 2, 2, O.P(V), O.P(Ta), O.P(Tb), O.P(Tc), O.P(Md), O.P(Me), 1, 1000,
                           offV, offT, offM, O.P(buf))
    x = O.uniform(78, -1, 1, B * S)
    mask = (_abi.RUNTIME_OUTPUTBUFFERSTRIDE | _abi.RUNTIME_SCALEC | _abi.RUNTIME_INPUTZEROPAD
            | _abi.RUNTIME_OUTPUTZEROPAD)
    cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, KL=2, KU=2, jw_control_bitmask=mask,
               offsetV=offV, offsetT=offT, offsetM=offM, Msplit=[1000, 0])
    launch = dict(outputBufferStride=S, scaleC=-1.25, inputZeropad=[511, N - 700], outputZeropad=[3, N - 513])
    got = run_plan(cfg, buf, x, launch)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, mask=mask, Msplit0=1000, cfg_outStride=N, run_outStride=S,
                  cfg_scaleC=1.0, run_scaleC=-1.25, run_izp=[511, N - 700], run_ozp=[3, N - 513],
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, S, 0, N, 3, N - 513, TOL64)


def test_long_step_replayed_from_a_graph():
    """the pair path's launches (fill kernel, front, reduced solve, finishing pass) captured
    once in a hipGraph and replayed on fresh right-hand sides, with eager copies and
    synchronizations between the replays: every replay bit-identical to the eager step"""
    import torch as t
    import pfsolve_amd as pf
    N, B = 10000, 6
    raw, buf, (offV, offT, offM) = fused_buffers(N, 31, 1)
    dmat = to_dev(buf)
    xs = [O.uniform(40 + i, -1, 1, B * N) for i in range(3)]
    drhs = to_dev(xs[0])
    stream = t.cuda.Stream()
    plan = pf.Plan(dmat.data_ptr(), drhs.data_ptr(), stream=stream.cuda_stream, size=[N, B, 1], jw_type=14,
                   doublePrecision=1, scaleC=0.5, KL=2, KU=2, offsetV=offV, offsetT=offT, offsetM=offM)
    try:
        want = []
        for x in xs:
            drhs.copy_(t.from_numpy(x).cuda())
            t.cuda.synchronize()
            plan.append()
            stream.synchronize()
            want.append(drhs.cpu().numpy())
        graph = t.cuda.CUDAGraph()
        t.cuda.synchronize()
        with t.cuda.graph(graph, stream=stream, capture_error_mode="relaxed"):
            plan.append()
        for i, x in enumerate(xs):
            drhs.copy_(t.from_numpy(x).cuda())
            t.cuda.synchronize()
            graph.replay()
            t.cuda.synchronize()
            got = drhs.cpu().numpy()
            assert np.array_equal(got, want[i]), f"replay {i} differs from the eager step"
        plan.sync()
    finally:
        plan.close()
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=0.5,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    check_systems(want[2], O.jw_f128(p, buf, xs[2]), xs[2], B, N, 0, N, 0, N, TOL64)


def test_long_slices_of_the_batch():
    """a launch with more systems than the plan was made for runs in slices of the plan's
    reduced data (batchSize > size[1]), odd count"""
    N, B = 6000, 9
    raw, buf, (offV, offT, offM) = fused_buffers(N, 5, 1)
    x = O.uniform(6, -1, 1, B * N)
    cfg = dict(size=[N, 2, 1], jw_type=14, doublePrecision=1, scaleC=1.0, KL=2, KU=2,
               offsetV=offV, offsetT=offT, offsetM=offM)
    got = run_plan(cfg, buf, x, dict(batchSize=B))
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, N, 0, N, 0, N, TOL64)


@pytest.mark.parametrize("N,variant,stride_pad", [(1500, 0, 1), (1025, 1, 0), (2048, 1, 3)])
def test_wave_range_in_chunks(N, variant, stride_pad):
    """1024 < N <= 2048: systems the wave-per-system kernel does not take (odd stride) or
    the general kernel asked for (kernelVariant 1) run in chunks"""
    B = 3
    S = N + stride_pad
    raw, buf, (offV, offT, offM) = fused_buffers(N, 3, 1)
    x = O.uniform(4, -1, 1, B * S)
    cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, KL=2, KU=2, offsetV=offV,
               offsetT=offT, offsetM=offM, kernelVariant=variant, jw_control_bitmask=_abi.RUNTIME_OUTPUTBUFFERSTRIDE)
    got = run_plan(cfg, buf, x, dict(outputBufferStride=S))
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, mask=_abi.RUNTIME_OUTPUTBUFFERSTRIDE, cfg_outStride=N,
                  run_outStride=S, cfg_scaleC=1.0, cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, S, 0, N, 0, N, TOL64)


@pytest.mark.parametrize("N", [513, 1000, 3000])
def test_long_double_double_fused(N):
    print(f"long DD fused N={N}", flush=True)
    B = 5
    buf, x, (offV, offT, offM) = dd_inputs(N, B, 41 + N)
    cfg = dict(size=[N, B, 1], jw_type=14, quadDoubleDoublePrecision=1, scaleC=1.0,
               offsetV=offV, offsetT=offT, offsetM=offM)
    got = run_plan(cfg, buf.ravel(), x.ravel()).reshape(-1, 2)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128_dd(p, buf, x)
    for b in range(B):
        err = O.dd_relerr_norm(got[b * N:(b + 1) * N], want[b * N:(b + 1) * N])
        assert err <= 1e-30, (b, err)


@pytest.mark.parametrize("jw_type,ub", [(10, 1), (12, 0), (13, 2)])
def test_long_double_double_modes(jw_type, ub):
    """DD reference modes past 512 rows (matrices from the fused packing's T block)"""
    N, B = 2000, 3
    buf, x, _ = dd_inputs(N, B, 9)
    tri = np.ascontiguousarray(buf[5 * N:8 * N])
    cfg = dict(size=[N, B, 1], jw_type=jw_type, upperBanded=ub, quadDoubleDoublePrecision=1, scaleC=2.0)
    got = run_plan(cfg, tri.ravel(), x.ravel()).reshape(-1, 2)
    p = O.problem(jw_type=jw_type, upperBanded=ub, N=N, B=B, cfg_outStride=N, cfg_scaleC=2.0)
    want = O.jw_f128_dd(p, tri, x)
    assert O.dd_relerr_norm(got, want) <= 1e-30


def test_long_double_double_fp64_memory():
    N, B = 1500, 4
    raw, buf, (offV, offT, offM) = fused_buffers(N, 8, 1)
    x = O.uniform(8, -1, 1, B * N)
    cfg = dict(size=[N, B, 1], jw_type=14, quadDoubleDoublePrecisionDoubleMemory=1, scaleC=1.0,
               offsetV=offV, offsetT=offT, offsetM=offM)
    got = run_plan(cfg, buf, x)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    assert O.relerr_norm(got, want) <= 2.3e-16


def test_long_repeat_launches_identical():
    """the tables and reduced data are rebuilt per launch: two launches on the same input give
    the same bits as one (the second on a fresh copy)"""
    N, B = 3000, 6
    raw, buf, (offV, offT, offM) = fused_buffers(N, 12, 1)
    x = O.uniform(13, -1, 1, B * N)
    cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, KL=2, KU=2,
               offsetV=offV, offsetT=offT, offsetM=offM)
    one = run_plan(cfg, buf, x)
    again = run_plan(cfg, buf, x)
    np.testing.assert_array_equal(one, again)
