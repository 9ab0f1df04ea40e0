"""GPU parity of the HIP path (libpfsolve_gfx950.so through the C ABI) against
the CPU oracle.

Bar: FP64 max normwise relative error <= 1e-13 against the __float128 oracle
(max_i |y_i - y*_i| / max_i |y*_i| per system); double-double <= 1e-30;
rows outside the output keep-range and the gaps between systems are
bit-exactly untouched; layout-only block plans are bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

import _oracle as O
from jwgpu import run_plan

pytestmark = pytest.mark.gpu

RT_ALL = (_abi.RUNTIME_OFFSETM | _abi.RUNTIME_OFFSETV | _abi.RUNTIME_OFFSETSOLUTION | _abi.RUNTIME_INPUTZEROPAD
          | _abi.RUNTIME_OUTPUTZEROPAD | _abi.RUNTIME_SCALEC | _abi.RUNTIME_OUTPUTBUFFERSTRIDE)
TOL64 = 1e-13


def md_base(jw, ub, N, rt, offM):
    if ub == 2:
        return N
    if jw in (0, 1):
        return (offM + N if rt else 0) if ub else (offM if rt else N)
    if jw == 3:
        return (offM if rt else N) if ub else (offM + N if rt else 0)
    return None


def mode_buffer(jw, ub, N, rt, offM, total, seed):
    buf = O.uniform(seed, -0.3, 0.3, total)
    mb = md_base(jw, ub, N, rt, offM)
    if mb is not None:
        buf[mb:mb + N] = O.uniform(seed + 1, 0.5, 1.5, N)
    return buf


def check_systems(got, want, x, nsys, stride, off, N, out_lo, out_hi, tol):
    for b in range(nsys):
        base = b * stride + off
        g = got[base:base + N]
        w = want[base:base + N]
        if out_hi > out_lo:
            err = O.relerr_norm(g[out_lo:out_hi], w[out_lo:out_hi])
            assert err <= tol, (b, err)
        np.testing.assert_array_equal(g[:out_lo], x[base:base + out_lo])
        np.testing.assert_array_equal(g[out_hi:], x[base + out_hi:base + N])
    # everything outside the systems is untouched
    mask = np.ones(got.size, bool)
    for b in range(nsys):
        mask[b * stride + off:b * stride + off + N] = False
    np.testing.assert_array_equal(got[mask], x[mask])


MODES = [(10, 1), (10, 0), (11, 1), (11, 0), (12, 1), (12, 0), (13, 2), (13, 1), (13, 0)]
SIZES = [1, 2, 3, 17, 64, 65, 127, 200, 512, 513, 1000, 1024]


@pytest.mark.parametrize("jw_type,ub", MODES)
@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("rt", [0, 1])
def test_reference_modes(jw_type, ub, N, rt):
    jw = jw_type % 10
    B = 5
    offM, offV = (13, 3 * N + 29) if rt else (0, 0)
    total = max(3 * N, offV + N, offM + 2 * N) + 8
    buf = mode_buffer(jw, ub, N, rt, offM, total, 1000 + N + jw)
    S = N + 7 if rt else N
    offS = 5 if rt else 0
    izp = (N // 4, N) if rt else (0, 0)
    ozp = (1, N - N // 8) if rt else (0, 0)
    x = O.uniform(77 + N, -1, 1, B * S + offS + 9)
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


def fused_buffers(N, seed, permuted, pad=0):
    offV, offT, offM = pad, pad + 5 * N + 3, pad + 8 * N + 11
    raw = O.raw_matrices(seed, N)
    buf = O.pack_fused(N, raw, permuted, offV, offT, offM, offM + 2 * N + 4)
    return raw, buf, (offV, offT, offM)


@pytest.mark.parametrize("N", [1, 2, 5, 16, 63, 64, 65, 100, 256, 511, 512, 777, 1024])
@pytest.mark.parametrize("permuted", [1, 0])
@pytest.mark.parametrize("variant", [1, 2])
def test_fused_step(N, permuted, variant):
    """variant 1: general kernel (table format 1); variant 2: wave-per-system kernel
    (format 2, used when the batch is 16-byte aligned: here B*N even needs N even)."""
    B = 7 if variant == 1 else 8
    raw, buf, (offV, offT, offM) = fused_buffers(N, N, permuted)
    x = O.uniform(5 + N, -1, 1, B * N)
    jt = 14 if permuted else 1014
    cfg = dict(size=[N, B, 1], jw_type=jt, doublePrecision=1, scaleC=1.0, KL=2, KU=2,
               offsetV=offV, offsetT=offT, offsetM=offM, kernelVariant=variant)
    got = run_plan(cfg, buf, x)
    p = O.problem(jw_type=jt, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, N, 0, N, 0, N, TOL64)
    # and against the raw-matrix __float128 evaluation (independent of storage)
    for b in (0, B - 1):
        y = np.zeros(N)
        O.lib().pfo_fused_raw_f128(N, 2, 2, *[O.P(a) for a in raw], O.P(np.ascontiguousarray(x[b * N:(b + 1) * N])),
                                   O.P(y))
        assert O.relerr_norm(got[b * N:(b + 1) * N], y) <= TOL64


def test_fused_config0_exact_workload():
    """N=64, 1024 real systems, seed 0, V/T/M/x from the stated distributions."""
    N, B = 64, 1024
    raw, buf, (offV, offT, offM) = fused_buffers(N, 0, 1)
    x = O.uniform(0, -1, 1, B * N)
    cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, offsetV=offV, offsetT=offT, offsetM=offM)
    got = run_plan(cfg, buf, x)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    want64 = O.jw_f64(p, buf, x)
    check_systems(got, want, x, B, N, 0, N, 0, N, TOL64)
    assert O.relerr_norm(got, want64) <= TOL64


def cfg1_like(nsys_complex, N=512):
    """complex XXII (2 real systems per complex one), runtime scaleC 0.75 and input keep-range [480,512)"""
    raw, buf, (offV, offT, offM) = fused_buffers(N, 1, 1)
    B = 2 * nsys_complex
    mask = _abi.RUNTIME_SCALEC | _abi.RUNTIME_INPUTZEROPAD
    cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, offsetV=offV, offsetT=offT,
               offsetM=offM, jw_control_bitmask=mask)
    launch = dict(scaleC=0.75, inputZeropad=[480, 512])
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, mask=mask, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM, run_scaleC=0.75, run_izp=[480, 512])
    return raw, buf, cfg, launch, p


def test_fused_config1_reduced_batch():
    raw, buf, cfg, launch, p = cfg1_like(256)
    N, B = 512, 512
    x = O.uniform(3, -1, 1, B * N)
    got = run_plan(cfg, buf, x, launch)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, N, 0, N, 0, N, TOL64)


def test_fused_config1_full_size_properties():
    """65536 complex systems (131072 real) at N=512: linearity in x, re/im independence,
    and a sample of systems against the oracle."""
    import torch
    raw, buf, cfg, launch, p = cfg1_like(65536)
    N, B = 512, 131072
    g = torch.Generator(device="cuda").manual_seed(0)
    x1 = torch.rand(B * N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    x2 = torch.rand(B * N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dmat = torch.from_numpy(buf).cuda()
    stream = torch.cuda.current_stream().cuda_stream

    def step(x):
        y = x.clone()
        with pf.Plan(dmat.data_ptr(), y.data_ptr(), stream=stream, **cfg) as plan:
            plan.append(**launch)
        torch.cuda.synchronize()
        return y

    y1, y2 = step(x1), step(x2)
    y12 = step(0.5 * x1 - 2.0 * x2)
    lin = (y12 - (0.5 * y1 - 2.0 * y2)).abs().max().item() / y12.abs().max().item()
    assert lin < 1e-13
    xs = x1.cpu().numpy()
    ys = y1.cpu().numpy()
    rng = np.random.default_rng(0)
    sample = np.unique(np.r_[0, 1, B - 2, B - 1, rng.integers(0, B, 60)])
    for b in sample:
        xb = xs[b * N:(b + 1) * N].copy()
        pb = O.problem(**{f: getattr(p, f) for f, _ in O.JwProblem._fields_ if f not in ("B",)}, B=1)
        want = O.jw_f128(pb, buf, xb)
        assert O.relerr_norm(ys[b * N:(b + 1) * N], want) <= TOL64, b


def test_fused_runtime_offsets_strides_and_buffers_swap():
    N, B = 300, 9
    raw, buf, (offV, offT, offM) = fused_buffers(N, 4, 1, pad=17)
    S, offS = N + 64, 12
    x = O.uniform(9, -1, 1, B * S + offS)
    mask = RT_ALL | _abi.RUNTIME_OFFSETT
    cfg = dict(size=[N, 3, 1], jw_type=14, doublePrecision=1, jw_control_bitmask=mask)
    launch = dict(offsetV=offV, offsetT=offT, offsetM=offM, offsetSolution=offS, outputBufferStride=S,
                  inputZeropad=[0, N], outputZeropad=[2, N - 3], scaleC=-1.5, batchSize=B)
    got = run_plan(cfg, buf, x, launch)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, mask=mask, run_offsetV=offV, run_offsetT=offT,
                  run_offsetM=offM, run_offsetSolution=offS, run_outStride=S, run_izp=[0, N],
                  run_ozp=[2, N - 3], run_scaleC=-1.5)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, S, offS, N, 2, N - 3, TOL64)


def test_odd_and_tiny_batches():
    N = 70
    raw, buf, (offV, offT, offM) = fused_buffers(N, 2, 1)
    for B in (1, 2, 3, 131):
        x = O.uniform(B, -1, 1, B * N)
        cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, offsetV=offV, offsetT=offT,
                   offsetM=offM)
        got = run_plan(cfg, buf, x)
        p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                      cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
        check_systems(got, O.jw_f128(p, buf, x), x, B, N, 0, N, 0, N, TOL64)


def test_size_guard_turns_bad_offsets_into_errors():
    N, B = 100, 4
    raw, buf, (offV, offT, offM) = fused_buffers(N, 2, 1)
    x = O.uniform(1, -1, 1, B * N)
    cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, offsetV=offV, offsetT=offT, offsetM=offM,
               jw_control_bitmask=_abi.RUNTIME_OFFSETSOLUTION)
    code = run_plan(cfg, buf, x, dict(offsetSolution=5), expect=1)
    assert code == _abi.PFSOLVE_ERROR_EMPTY_outputBufferSize
    cfg2 = dict(cfg, offsetM=buf.size)   # M beyond the matrix buffer
    cfg2["jw_control_bitmask"] = 0
    assert run_plan(cfg2, buf, x, {}, expect=1) == _abi.PFSOLVE_ERROR_EMPTY_bufferSize


# ---------------------------------------------------------------- double-double
def dd_inputs(N, B, seed):
    raw = O.raw_matrices(seed, N)
    rawdd = []
    for k, a in enumerate(raw):
        lo = O.uniform(seed * 7 + k, -1, 1, a.size) * 2.0 ** -60 * np.abs(a)
        rawdd.append(np.ascontiguousarray(np.stack([a, lo], -1)))
    offV, offT, offM = 0, 5 * N, 8 * N
    buf = np.zeros((10 * N, 2))
    O.lib().pfo_pack_fused_dd(N, 2, 2, *[O.PDD(a) for a in rawdd], 1, 0, offV, offT, offM, O.PDD(buf))
    xh = O.uniform(seed + 1, -1, 1, B * N)
    x = np.ascontiguousarray(np.stack([xh, O.uniform(seed + 2, -1, 1, B * N) * 2.0 ** -60 * np.abs(xh)], -1))
    return buf, x, (offV, offT, offM)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("N", [1, 7, 64, 100, 128, 129, 256, 257, 512])
def test_double_double_fused(N, variant):
    """variant 0: the wave-per-system double-double kernel (N <= 512); 1: the general kernel"""
    B = 64 if N == 256 else 6
    buf, x, (offV, offT, offM) = dd_inputs(N, B, 31 + N)
    cfg = dict(size=[N, B, 1], jw_type=14, quadDoubleDoublePrecision=1, scaleC=1.0,
               offsetV=offV, offsetT=offT, offsetM=offM, kernelVariant=variant)
    got = run_plan(cfg, buf.ravel(), x.ravel()).reshape(-1, 2)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128_dd(p, buf, x)
    for b in range(B):
        err = O.dd_relerr_norm(got[b * N:(b + 1) * N], want[b * N:(b + 1) * N])
        assert err <= 1e-30, (b, err)


@pytest.mark.parametrize("variant", [0, 1])
def test_double_double_reference_mode_tridiagonal(variant):
    N, B = 300, 4
    buf, x, _ = dd_inputs(N, B, 5)
    # jw 13 / ub 2 reads [ld | md | ud] at [0, N, 2N): reuse the T block of the fused packing
    tri = np.ascontiguousarray(buf[5 * N:8 * N])
    cfg = dict(size=[N, B, 1], jw_type=13, upperBanded=2, quadDoubleDoublePrecision=1, scaleC=1.0,
               kernelVariant=variant)
    got = run_plan(cfg, tri.ravel(), x.ravel()).reshape(-1, 2)
    p = O.problem(jw_type=13, upperBanded=2, N=N, B=B, cfg_outStride=N, cfg_scaleC=1.0)
    want = O.jw_f128_dd(p, tri, x)
    assert O.dd_relerr_norm(got, want) <= 1e-30


@pytest.mark.parametrize("variant", [0, 1])
def test_double_double_compute_fp64_memory(variant):
    N, B = 200, 8
    raw, buf, (offV, offT, offM) = fused_buffers(N, 6, 1)
    x = O.uniform(6, -1, 1, B * N)
    cfg = dict(size=[N, B, 1], jw_type=14, quadDoubleDoublePrecisionDoubleMemory=1, scaleC=1.0,
               offsetV=offV, offsetT=offT, offsetM=offM, kernelVariant=variant)
    got = run_plan(cfg, buf, x)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    # DD arithmetic, result rounded to FP64: within one rounding of the exact result
    assert O.relerr_norm(got, want) <= 2.3e-16


def promote_dd(a, seed):
    """FP64 values as double-double with a low word of U(-1,1) 2^-60 |hi|"""
    lo = O.uniform(seed, -1, 1, a.size) * 2.0 ** -60 * np.abs(a)
    return np.ascontiguousarray(np.stack([a, lo], -1))


def check_systems_dd(got, want, x, nsys, stride, off, N, out_lo, out_hi, tol):
    """DD arrays (n, 2): rows of the output keep-range within tol of the oracle (relative
    norm per system), every other element untouched"""
    keep = np.zeros(got.shape[0], bool)
    for b in range(nsys):
        base = b * stride + off
        if out_hi > out_lo:
            err = O.dd_relerr_norm(got[base + out_lo:base + out_hi], want[base + out_lo:base + out_hi])
            assert err <= tol, (b, err)
        keep[base + out_lo:base + out_hi] = True
    np.testing.assert_array_equal(got[~keep], x[~keep])


@pytest.mark.parametrize("jw_type,ub", MODES)
@pytest.mark.parametrize("N", [1, 2, 3, 17, 64, 65, 127, 200, 300, 512])
@pytest.mark.parametrize("variant", [0, 1])
def test_double_double_reference_modes(jw_type, ub, N, variant):
    """every reference mode in double-double with runtime offsets, keep ranges, output
    stride and scaleC, on the wave-per-system double-double kernel (0) and the general
    kernel (1)"""
    jw = jw_type % 10
    B = 5
    offM, offV = 13, 3 * N + 29
    total = max(3 * N, offV + N, offM + 2 * N) + 8
    buf = promote_dd(mode_buffer(jw, ub, N, 1, offM, total, 1000 + N + jw), 3000 + N + jw)
    S, offS = N + 7, 5
    izp, ozp = (N // 4, N), (1, N - N // 8)
    x = promote_dd(O.uniform(77 + N, -1, 1, B * S + offS + 9), 78 + N)
    cfg = dict(size=[N, B, 1], jw_type=jw_type, upperBanded=ub, quadDoubleDoublePrecision=1, scaleC=1.0,
               outputBufferStride=[N, 0, 0], jw_control_bitmask=RT_ALL, kernelVariant=variant)
    launch = dict(offsetM=offM, offsetV=offV, offsetSolution=offS, inputZeropad=izp, outputZeropad=ozp,
                  scaleC=0.75, outputBufferStride=S)
    got = run_plan(cfg, buf.ravel(), x.ravel(), launch).reshape(-1, 2)
    p = O.problem(jw_type=jw_type, upperBanded=ub, N=N, B=B, mask=RT_ALL, cfg_outStride=N,
                  cfg_scaleC=1.0, run_offsetM=offM, run_offsetV=offV, run_offsetSolution=offS,
                  run_izp=list(izp), run_ozp=list(ozp), run_scaleC=0.75, run_outStride=S)
    want = O.jw_f128_dd(p, buf, x)
    lo, hi = ozp[0], min(ozp[1], N)
    check_systems_dd(got, want, x, B, S, offS, N, lo, max(lo, hi), 1e-30)


@pytest.mark.parametrize("N", [7, 200])
def test_double_double_large_odd_batch(N):
    """more pairs of systems than one sweep of the grid holds (the wavefronts loop), odd
    batch: the last pass has one system"""
    B = 20001 if N == 7 else 4097
    buf, x, (offV, offT, offM) = dd_inputs(N, B, 900 + N)
    cfg = dict(size=[N, B, 1], jw_type=14, quadDoubleDoublePrecision=1, scaleC=1.0,
               offsetV=offV, offsetT=offT, offsetM=offM)
    got = run_plan(cfg, buf.ravel(), x.ravel()).reshape(-1, 2)
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128_dd(p, buf, x)
    errs = [O.dd_relerr_norm(got[b * N:(b + 1) * N], want[b * N:(b + 1) * N]) for b in range(B)]
    assert max(errs) <= 1e-30


def test_double_double_geometry_reports_the_wave_per_system_kernel():
    import torch
    for N, R in [(64, 2), (200, 4), (512, 8)]:
        buf, _, (offV, offT, offM) = dd_inputs(N, 1, 3)
        dmat = torch.from_numpy(buf.ravel()).cuda()
        x = torch.zeros(2 * N, dtype=torch.float64, device="cuda")
        with pf.Plan(dmat.data_ptr(), x.data_ptr(), size=[N, 1, 1], jw_type=14, quadDoubleDoublePrecision=1,
                     offsetV=offV, offsetT=offT, offsetM=offM) as plan:
            g = plan.geometry()
            assert g.fast_rows_per_lane == R and g.fast_waves == 8
        with pf.Plan(dmat.data_ptr(), x.data_ptr(), size=[N, 1, 1], jw_type=14, quadDoubleDoublePrecision=1,
                     offsetV=offV, offsetT=offT, offsetM=offM, kernelVariant=1) as plan:
            assert plan.geometry().fast_rows_per_lane == 0


# ---------------------------------------------------------------- block plan
@pytest.mark.parametrize("layout", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("op", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("acc", [0, 2])
def test_block_plan(layout, op, acc):
    N, B = 37, 6
    block = op * 100 + acc * 10 + layout
    in_c = layout in (2, 3, 4)
    out_c = layout in (5, 6, 7)
    inS = N + 3
    outS = N + 5
    offIn, offOut = 2, 4
    n_in = (2 if (in_c or layout == 7) else 1) * (B * inS + offIn + 2 * inS)
    n_out = (2 if (out_c or layout == 4) else 1) * (B * outS + offOut + 2 * outS)
    inb = O.uniform(1, -1, 1, n_in)
    outb = O.uniform(2, -1, 1, n_out)
    mask = (_abi.RUNTIME_SCALEC | _abi.RUNTIME_INPUTBUFFERSTRIDE | _abi.RUNTIME_OUTPUTBUFFERSTRIDE
            | _abi.RUNTIME_OFFSETM | _abi.RUNTIME_OFFSETSOLUTION | _abi.RUNTIME_INPUTZEROPAD | _abi.RUNTIME_OUTPUTZEROPAD)
    cfg = dict(size=[N, B, 1], block=block, lshift=3, doublePrecision=1, jw_control_bitmask=mask)
    launch = dict(scaleC=0.75, inputBufferStride=inS, outputBufferStride=outS, offsetM=offIn,
                  offsetSolution=offOut, inputZeropad=[3, N - 2], outputZeropad=[1, N])
    got = run_plan(cfg, inb, outb, launch)
    bp = O.BlockProblem(block=block, lshift=3, N=N, B=B, mask=mask, offsetM=offIn, offsetSolution=offOut,
                        izp_runtime=1, ozp_runtime=1, scaleC=0.75, scaleC_runtime=1, inStride=inS, outStride=outS)
    bp.izp[0], bp.izp[1] = 3, N - 2
    bp.ozp[0], bp.ozp[1] = 1, N
    want = outb.copy()
    assert O.lib().pfo_block_f64(C.byref(bp), O.P(inb), O.P(want)) == 0
    if op == 0:
        np.testing.assert_array_equal(got, want)
    else:
        np.testing.assert_allclose(got, want, rtol=4e-16, atol=0)


# ---------------------------------------------------------------- dgbmv plan
@pytest.mark.parametrize("KL,KU", [(2, 2), (1, 0), (0, 3), (3, 1), (0, 0)])
@pytest.mark.parametrize("N", [1, 7, 100, 1000])
def test_dgbmv_plan(KL, KU, N):
    B = 5
    LDA = KL + KU + 1
    offM = 3
    mat = O.uniform(N + KL, -1, 1, offM + LDA * N)
    S = N + 2
    x = O.uniform(4, -1, 1, B * S)
    mask = _abi.RUNTIME_OUTPUTBUFFERSTRIDE
    cfg = dict(size=[N, B, 1], LDA=LDA, KL=KL, KU=KU, doublePrecision=1, offsetM=offM, jw_control_bitmask=mask)
    got = run_plan(cfg, mat, x, dict(outputBufferStride=S))
    p = O.problem(N=N, B=B, mask=mask, cfg_offsetM=offM, run_outStride=S, cfg_outStride=N)
    want = x.copy()
    assert O.lib().pfo_dgbmv_f128(C.byref(p), LDA, KU, KL, O.P(mat), O.P(want)) == 0
    check_systems(got, want, x, B, S, 0, N, 0, N, 1e-14)


# ---------------------------------------------------------------- kernel selection
def test_wave_per_system_kernel_is_selected_for_aligned_fp64():
    import torch
    N, B = 512, 64
    raw, buf, (offV, offT, offM) = fused_buffers(N, 3, 1)
    dmat = torch.from_numpy(buf).cuda()
    x = torch.zeros(B * N, dtype=torch.float64, device="cuda")
    with pf.Plan(dmat.data_ptr(), x.data_ptr(), size=[N, B, 1], jw_type=14, doublePrecision=1,
                 offsetV=offV, offsetT=offT, offsetM=offM) as plan:
        g = plan.geometry()
        # compute / store split: computing and storing wavefronts
        assert g.fast_rows_per_lane == 8 and g.fast_waves >= 4 and g.fast_store_waves >= 1
        plan.append()
        assert pf.PfSolveSync(plan.app) == 0   # no bounded wait ran out
    with pf.Plan(dmat.data_ptr(), x.data_ptr(), size=[N, B, 1], jw_type=14, doublePrecision=1,
                 offsetV=offV, offsetT=offT, offsetM=offM, kernelVariant=1) as plan:
        assert plan.geometry().fast_rows_per_lane == 0


def test_split_kernel_time_out_is_reported_and_recovers():
    """The split kernel's bounded waits, made to give up at once from the host
    (pfsolve_set_wait_polls with a negative bound): every wavefront still leaves,
    PfSolveSync reports FAILED_TO_SYNCHRONIZE and clears the flag, and with the
    default bound the next launch gives the oracle's result again."""
    import torch
    N, B = 512, 64
    raw, buf, (offV, offT, offM) = fused_buffers(N, 3, 1)
    x = O.uniform(17, -1, 1, B * N)
    dmat = torch.from_numpy(buf).cuda()
    dx = torch.from_numpy(x.copy()).cuda()
    L = pf.lib()
    with pf.Plan(dmat.data_ptr(), dx.data_ptr(), size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0,
                 KL=2, KU=2, offsetV=offV, offsetT=offT, offsetM=offM) as plan:
        assert plan.geometry().fast_store_waves >= 1
        assert L.pfsolve_set_wait_polls(C.byref(plan.app), -1) == 0
        plan.append()
        assert pf.PfSolveSync(plan.app) == _abi.PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE
        assert pf.PfSolveSync(plan.app) == 0   # the flag was cleared
        assert L.pfsolve_set_wait_polls(C.byref(plan.app), 0) == 0
        dx.copy_(torch.from_numpy(x))
        plan.append()
        assert pf.PfSolveSync(plan.app) == 0
        got = dx.cpu().numpy()
    p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, N, 0, N, 0, N, TOL64)


def test_split_kernel_many_passes_per_wavefront():
    """enough pairs that every computing wavefront hands many outputs to its storer:
    N = 256 (R = 4) and N = 100 (R = 2), 40000 real systems, a sample of them against
    the oracle"""
    import torch
    for N in (256, 100):
        B = 40000
        raw, buf, (offV, offT, offM) = fused_buffers(N, 11, 1)
        x = O.uniform(17 + N, -1, 1, B * N)
        cfg = dict(size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, offsetV=offV, offsetT=offT,
                   offsetM=offM)
        got = run_plan(cfg, buf, x)
        idx = list(range(0, B, 997)) + [B - 2, B - 1]
        sub = np.concatenate([x[b * N:(b + 1) * N] for b in idx])
        p = O.problem(jw_type=14, N=N, B=len(idx), KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                      cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
        want = O.jw_f128(p, buf, sub)
        for j, b in enumerate(idx):
            assert O.relerr_norm(got[b * N:(b + 1) * N], want[j * N:(j + 1) * N]) <= TOL64, (N, b)


@pytest.mark.parametrize("N", [1025, 1300, 2047, 2048])
@pytest.mark.parametrize("permuted", [1, 0])
def test_fused_step_long_systems(N, permuted):
    """1024 < N <= 2048: one system per wavefront, table read from global memory;
    odd batch, runtime keep-ranges with odd ends, row-0 scaleC"""
    B = 5
    S = N + 4 + (N % 2)
    raw, buf, (offV, offT, offM) = fused_buffers(N, N + 7, permuted)
    x = O.uniform(50 + N, -1, 1, B * S)
    jt = 14 if permuted else 1014
    mask = (_abi.RUNTIME_OUTPUTBUFFERSTRIDE | _abi.RUNTIME_SCALEC | _abi.RUNTIME_INPUTZEROPAD
            | _abi.RUNTIME_OUTPUTZEROPAD)
    cfg = dict(size=[N, B, 1], jw_type=jt, doublePrecision=1, scaleC=1.0, KL=2, KU=2, jw_control_bitmask=mask,
               offsetV=offV, offsetT=offT, offsetM=offM)
    launch = dict(outputBufferStride=S, scaleC=0.5, inputZeropad=[3, N - 1], outputZeropad=[1, N - 2])
    got = run_plan(cfg, buf, x, launch)
    p = O.problem(jw_type=jt, N=N, B=B, KL=2, KU=2, mask=mask, cfg_outStride=N, run_outStride=S,
                  cfg_scaleC=1.0, run_scaleC=0.5, run_izp=[3, N - 1], run_ozp=[1, N - 2],
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, S, 0, N, 1, N - 2, TOL64)


@pytest.mark.parametrize("jw_type,ub", MODES)
@pytest.mark.parametrize("N", [2, 64, 130, 512, 1024, 1800])
def test_reference_modes_wave_per_system(jw_type, ub, N):
    """aligned batches of the reference modes run on the wave-per-system kernel"""
    jw = jw_type % 10
    B, S = 6, N + 64
    buf = mode_buffer(jw, ub, N, 0, 0, 3 * N + 8, 2000 + N)
    x = O.uniform(91 + N, -1, 1, B * S)
    mask = _abi.RUNTIME_OUTPUTBUFFERSTRIDE | _abi.RUNTIME_SCALEC | _abi.RUNTIME_INPUTZEROPAD
    cfg = dict(size=[N, B, 1], jw_type=jw_type, upperBanded=ub, doublePrecision=1, jw_control_bitmask=mask,
               kernelVariant=2)
    launch = dict(outputBufferStride=S, scaleC=1.25, inputZeropad=[1, N - 1])
    got = run_plan(cfg, buf, x, launch)
    p = O.problem(jw_type=jw_type, upperBanded=ub, N=N, B=B, mask=mask, cfg_outStride=N, run_outStride=S,
                  run_scaleC=1.25, run_izp=[1, N - 1])
    want = O.jw_f128(p, buf, x)
    check_systems(got, want, x, B, S, 0, N, 0, N, TOL64)


# ---------------------------------------------------------------- several streams
@pytest.mark.parametrize("N,variant", [(100, 1), (512, 0), (1000, 0)])
def test_round_robin_streams(N, variant):
    """num_streams = 3: launches go round robin over the caller's streams (the stream
    index advances per launch and PfSolveSync restarts it), each launch with its own
    right-hand sides and its own matrix set (runtime offsets), all in flight together;
    every result matches the oracle, as each slot has its own coefficient table."""
    import torch
    B, K = 6, 7
    sets = [fused_buffers(N, 300 + k, 1) for k in range(2)]
    # both matrix sets in one buffer, the second behind the first
    span = sets[0][1].size
    mat = np.concatenate([sets[0][1], sets[1][1]])
    offs = [sets[0][2], tuple(o + span for o in sets[1][2])]
    xs = [O.uniform(400 + k, -1, 1, B * N) for k in range(K)]
    dmat = torch.from_numpy(mat).cuda()
    drhs = [torch.from_numpy(x.copy()).cuda() for x in xs]
    streams = [torch.cuda.Stream() for _ in range(3)]
    torch.cuda.synchronize()
    rt = _abi.RUNTIME_OFFSETM | _abi.RUNTIME_OFFSETV | _abi.RUNTIME_OFFSETT
    with pf.Plan(dmat.data_ptr(), drhs[0].data_ptr(), stream=[s.cuda_stream for s in streams],
                 size=[N, B, 1], jw_type=14, doublePrecision=1, scaleC=1.0, KL=2, KU=2,
                 jw_control_bitmask=rt, kernelVariant=variant) as plan:
        seen = []
        for k in range(K):
            oV, oT, oM = offs[k % 2]
            seen.append(plan.app.configuration.streamID)
            plan.append(outputBuffer=drhs[k].data_ptr(), offsetV=oV, offsetT=oT, offsetM=oM)
        plan.sync()
        assert seen == [k % 3 for k in range(K)]
        assert plan.app.configuration.streamID == 0
    for k in range(K):
        oV, oT, oM = offs[k % 2]
        p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0, mask=rt,
                      run_offsetV=oV, run_offsetT=oT, run_offsetM=oM)
        want = O.jw_f128(p, mat, xs[k])
        check_systems(drhs[k].cpu().numpy(), want, xs[k], B, N, 0, N, 0, N, TOL64)
