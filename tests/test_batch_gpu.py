"""Uneven-size batches (pfsolve_batch_*): many matrix sets of different N solved in
one step, checked set by set against the __float128 oracle, against the plain
one-plan-per-set form (same kernels, bit for bit), across tunings, through
hipGraph replays and at the full 1024-degree workload."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

import _oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13

# (N, real systems, permuted storage, scaleC)
SMALL_SETS = [(2, 3, 1, 1.0), (3, 1, 0, 0.5), (17, 4, 1, 1.0), (64, 5, 1, 2.0), (65, 2, 0, 1.0),
              (100, 6, 1, 1.0), (128, 1, 1, 0.75), (129, 3, 1, 1.0), (256, 8, 0, 1.0), (257, 2, 1, 1.0),
              (300, 5, 1, -1.5), (512, 4, 1, 1.0), (513, 3, 0, 1.0), (700, 2, 1, 1.0), (1000, 6, 1, 0.25),
              (1024, 7, 1, 1.0), (511, 0, 1, 1.0), (1025, 3, 1, 1.0), (1600, 4, 0, 0.5), (2048, 5, 1, 1.0)]


def torch():
    import torch as t
    return t


def row_class(N):
    """row class c of a set: R = 2 << c rows per lane, N <= 64 R"""
    c = 0
    while c < 4 and (128 << c) < N:
        c += 1
    return c


class BatchData:
    """Host and device buffers of a batch: one matrix buffer and one RHS buffer
    holding every set at 16-byte aligned (even) element offsets."""

    def __init__(self, specs, seed=0, pad=6, off_sol=2):
        t = torch()
        self.specs = specs
        mats, rhs_parts = [], []
        self.sets, self.problems, self.rhs_off, self.mat_off = [], [], [], []
        moff = roff = 0
        for i, (N, nsys, perm, sc) in enumerate(specs):
            offV, offT, offM = 0, 5 * N + 2, 8 * N + 4
            raw = O.raw_matrices(seed * 7919 + i, N)
            buf = O.pack_fused(N, raw, perm, offV, offT, offM, offM + 2 * N + 2)
            buf = np.concatenate([buf, np.zeros((-buf.size) % 2)])
            stride = N + pad + (N % 2)
            n = nsys * stride + off_sol if nsys else 2
            x = O.uniform(seed * 104729 + 31 * i + 7, -1, 1, n + (n % 2))
            self.mat_off.append(moff)
            self.rhs_off.append(roff)
            mats.append(buf)
            rhs_parts.append(x)
            self.sets.append(dict(N=N, nsys=nsys, offsetV=moff + offV, offsetT=moff + offT, offsetM=moff + offM,
                                  stride=stride, offsetSolution=off_sol, naturalOrder=0 if perm else 1,
                                  KL=2, KU=2, scaleC=sc))
            self.problems.append(O.problem(jw_type=14 if perm else 1014, N=N, B=max(nsys, 1), KL=2, KU=2,
                                           cfg_outStride=stride, cfg_offsetSolution=off_sol, cfg_scaleC=sc,
                                           cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM))
            moff += buf.size
            roff += x.size
        self.mat = np.concatenate(mats)
        self.x = np.concatenate(rhs_parts)
        self.dmat = t.from_numpy(self.mat).cuda()
        self.drhs = t.from_numpy(self.x).cuda()
        base_m, base_r = self.dmat.data_ptr(), self.drhs.data_ptr()
        for s, r0, m0 in zip(self.sets, self.rhs_off, self.mat_off):
            s["matrices"] = base_m
            s["rhs"] = base_r + 8 * r0

    def reset(self):
        self.drhs.copy_(torch().from_numpy(self.x).cuda())
        torch().cuda.synchronize()

    def result(self):
        torch().cuda.synchronize()
        return self.drhs.cpu().numpy()

    def oracle(self, i):
        N, nsys, perm, sc = self.specs[i]
        s = self.sets[i]
        m0 = self.mat_off[i]
        r0 = self.rhs_off[i]
        n = (self.rhs_off[i + 1] if i + 1 < len(self.rhs_off) else self.x.size) - r0
        buf = self.mat[m0:m0 + (s["offsetM"] - m0) + 2 * N + 2]
        return O.jw_f128(self.problems[i], np.ascontiguousarray(buf), np.ascontiguousarray(self.x[r0:r0 + n]))

    def check(self, got, sets=None, tol=TOL64):
        for i in (range(len(self.specs)) if sets is None else sets):
            N, nsys, perm, sc = self.specs[i]
            s = self.sets[i]
            r0 = self.rhs_off[i]
            n = (self.rhs_off[i + 1] if i + 1 < len(self.rhs_off) else self.x.size) - r0
            g = got[r0:r0 + n]
            x = self.x[r0:r0 + n]
            if nsys == 0:
                np.testing.assert_array_equal(g, x)
                continue
            want = self.oracle(i)
            mask = np.ones(n, bool)
            for b in range(nsys):
                base = b * s["stride"] + s["offsetSolution"]
                err = O.relerr_norm(g[base:base + N], want[base:base + N])
                assert err <= tol, (i, N, b, err)
                mask[base:base + N] = False
            np.testing.assert_array_equal(g[mask], x[mask])


def run_batch(data, tuning=None, mode="run"):
    t = torch()
    stream = t.cuda.current_stream().cuda_stream
    data.reset()
    with pf.Batch(data.sets, tuning) as b:
        getattr(b, mode)(stream)
        b.sync(stream)
        return data.result()


def test_batch_matches_oracle():
    d = BatchData(SMALL_SETS)
    got = run_batch(d)
    d.check(got)


def test_batch_equals_plain_per_set_launches():
    d = BatchData(SMALL_SETS, seed=1)
    a = run_batch(d, mode="run")
    b = run_batch(d, mode="run_plain")
    np.testing.assert_array_equal(a, b)
    d.check(a)


def tuning(**kw):
    t = _abi.PfSolveBatchTuning()
    for k, v in kw.items():
        cur = getattr(t, k)
        if isinstance(cur, int):
            setattr(t, k, v)
        else:
            for c in range(_abi.BATCH_CLASSES):
                cur[c] = v[c] if isinstance(v, (list, tuple)) else v
    return t


@pytest.mark.parametrize("kw", [
    dict(unit_pairs=1),
    dict(unit_pairs=2, waves=1),
    dict(waves=2, grid_per_cu=1),
    dict(order=1, unit_pairs=3),
    dict(rows_per_lane_shift=[1, 1, 1, 0, 0]),
    dict(rows_per_lane_shift=[1, 0, 1, 0, 0], unit_pairs=1, waves=1),
    dict(kernel=1),
    dict(kernel=2, waves=1),
    dict(kernel=2, waves=2, unit_pairs=3),
    dict(kernel=[2, 2, 2, 2, 1], unit_pairs=1),
    dict(kernel=2, waves=64, order=1),
    dict(kernel=[1, 2, 1, 2, 1], rows_per_lane_shift=[1, 1, 0, 0, 0]),
    dict(kernel=[2, 2, 2, 2, 1], waves=5, unit_pairs=2),
    dict(kernel=3),
    dict(kernel=[3, 3, 3, 3, 1], waves=4, unit_pairs=2),
    dict(kernel=3, waves=2, order=1, unit_pairs=1),
    dict(kernel=[3, 1, 3, 2, 1], rows_per_lane_shift=[1, 0, 1, 0, 0]),
    dict(systems_per_wave=4),
    dict(systems_per_wave=8),
    dict(systems_per_wave=8, unit_pairs=1),
    dict(systems_per_wave=4, unit_pairs=3, order=1),
    dict(systems_per_wave=8, waves=2, kernel=[1, 2, 2, 2, 1]),
    dict(systems_per_wave=4, kernel=4),
    dict(kernel=4),
    dict(kernel=[4, 4, 4, 4, 1], waves=4, unit_pairs=2),
    dict(kernel=4, waves=2, order=1, unit_pairs=1),
    dict(kernel=[4, 1, 2, 4, 1], unit_pairs=3),
    dict(kernel=[4, 1, 4, 3, 1], rows_per_lane_shift=[1, 0, 1, 0, 0]),
])
def test_batch_tunings(kw):
    """every tuning gives the oracle's results; without a rows-per-lane change the
    arithmetic is the same as the default's, bit for bit"""
    d = BatchData(SMALL_SETS, seed=2)
    ref = run_batch(d)
    got = run_batch(d, tuning(**kw))
    d.check(got)
    if "rows_per_lane_shift" not in kw:
        np.testing.assert_array_equal(got, ref)


# small systems for the sub-wavefront groups: N around the 16- and 32-lane group limits,
# odd and even system counts, one set with more systems than a pass of every wavefront holds
GROUP_SETS = [(2, 7, 1, 1.0), (5, 1, 0, 0.5), (16, 8, 1, 1.0), (31, 9, 1, 2.0), (32, 301, 1, 1.0),
              (33, 15, 0, 1.0), (47, 16, 1, 0.75), (63, 17, 1, 1.0), (64, 33, 1, -1.0), (65, 3, 1, 1.0),
              (128, 5, 1, 1.0), (300, 2, 1, 1.0)]


@pytest.mark.parametrize("spw", [4, 8])
@pytest.mark.parametrize("unit_pairs", [0, 1, 5])
def test_batch_sub_wavefront_groups(spw, unit_pairs):
    """2 or 4 groups of lanes per wavefront (4 or 8 systems per pass) on the R = 2 class:
    the oracle's results, and bit for bit the pair kernel's"""
    d = BatchData(GROUP_SETS, seed=11)
    ref = run_batch(d, tuning(unit_pairs=unit_pairs))
    got = run_batch(d, tuning(systems_per_wave=spw, unit_pairs=unit_pairs))
    d.check(got)
    np.testing.assert_array_equal(got, ref)


def test_batch_refuses_bad_systems_per_wave_and_graph_steps():
    """systems_per_wave must be 0, 2, 4 or 8 and a captured graph holds 1..1024 steps; a refused
    call leaves the batch usable"""
    t = torch()
    d = BatchData(SMALL_SETS[:4], seed=12)
    stream = t.cuda.current_stream().cuda_stream
    L = pf.lib()
    with pf.Batch(d.sets) as b:
        for spw in (1, 3, 6, 16, -2):
            assert L.pfsolve_batch_set_tuning(b.handle, C.byref(tuning(systems_per_wave=spw))) == \
                _abi.PFSOLVE_ERROR_EMPTY_size
        s = t.cuda.Stream()
        for steps in (0, -1, 1025):
            assert L.pfsolve_batch_capture_graph_steps(b.handle, s.cuda_stream, 0, steps) == \
                _abi.PFSOLVE_ERROR_EMPTY_size
        d.reset()
        b.run(stream)
        b.sync(stream)
        d.check(d.result())


def test_batch_split_kernel_time_out_is_reported_and_recovers():
    """The split queue kernel's bounded waits, made to give up at once from the host
    (pfsolve_batch_set_wait_polls with a negative bound): the step ends,
    pfsolve_batch_sync reports FAILED_TO_SYNCHRONIZE once, and with the default bound
    the next step equals the plain per-set form, bit for bit."""
    t = torch()
    L = pf.lib()
    d = BatchData(SMALL_SETS, seed=8)
    ref = run_batch(d, mode="run_plain")
    stream = t.cuda.current_stream().cuda_stream
    d.reset()
    with pf.Batch(d.sets, tuning(kernel=[2, 2, 2, 2, 1])) as b:
        assert L.pfsolve_batch_set_wait_polls(b.handle, -1) == 0
        b.run(stream)
        assert L.pfsolve_batch_sync(b.handle, stream) == _abi.PFSOLVE_ERROR_FAILED_TO_SYNCHRONIZE
        assert L.pfsolve_batch_sync(b.handle, stream) == 0   # the flag was cleared
        assert L.pfsolve_batch_set_wait_polls(b.handle, 0) == 0
        d.reset()
        b.run(stream)
        b.sync(stream)
        got = d.result()
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kern", [0, 4])
def test_batch_graph_replays(kern):
    """the captured step replays as a full step every time (queue counters are reset
    inside the graph), with eager work and synchronisation between replays"""
    t = torch()
    d = BatchData(SMALL_SETS, seed=3)
    eager = run_batch(d)
    s = t.cuda.Stream()
    with pf.Batch(d.sets, tuning(kernel=kern) if kern else None) as b:
        b.capture_graph(s.cuda_stream)
        for rep in range(3):
            d.reset()
            junk = t.ones(1 << 20, device="cuda").sum()   # unrelated eager work
            b.launch_graph(s.cuda_stream)
            s.synchronize()
            got = d.result()
            np.testing.assert_array_equal(got, eager, err_msg=f"replay {rep}")
            assert junk.item() == float(1 << 20)


@pytest.mark.parametrize("solve_only", [False, True])
def test_batch_tuner_keeps_caller_data_and_results(solve_only):
    t = torch()
    d = BatchData(SMALL_SETS, seed=4)
    d.reset()
    stream = t.cuda.current_stream().cuda_stream
    with pf.Batch(d.sets) as b:
        best = b.tune(stream, repeats=1, solve_only=solve_only)
        # the tuner runs on scratch right-hand sides
        np.testing.assert_array_equal(d.result(), d.x)
        info = b.info()
        assert info.tuning.order == best.order
        b.run(stream)
        got = d.result()
    d.check(got)
    for c in range(_abi.BATCH_CLASSES):
        assert best.rows_per_lane_shift[c] in (0, 1)
        assert best.kernel[c] in (0, 1, 2, 3, 4)
        assert best.waves[c] >= 0


def test_batch_tuning_table_is_consulted(tmp_path):
    """a tuning saved by the tuner is reloaded and used by a batch created without one"""
    t = torch()
    L = pf.lib()
    L.pfsolve_tuning_clear()
    d = BatchData(SMALL_SETS[:6], seed=5)
    stream = t.cuda.current_stream().cuda_stream
    with pf.Batch(d.sets) as b:
        best = b.tune(stream, repeats=1)
    path = str(tmp_path / "tune.txt")
    assert L.pfsolve_tuning_save(path.encode()) == 0
    L.pfsolve_tuning_clear()
    assert L.pfsolve_tuning_load(path.encode()) == 0
    with pf.Batch(d.sets) as b:
        info = b.info()
    classes = {row_class(N) for N, nsys, *_ in SMALL_SETS[:6] if nsys > 0}
    assert classes
    for c in classes:
        assert info.tuning.unit_pairs[c] == best.unit_pairs[c]
        assert info.tuning.waves[c] == best.waves[c]
        assert info.tuning.grid_per_cu[c] == best.grid_per_cu[c]
        assert info.tuning.rows_per_lane_shift[c] == best.rows_per_lane_shift[c]
        assert info.tuning.kernel[c] == best.kernel[c]
    L.pfsolve_tuning_clear()


def test_batch_harmonic_degrees_full_size():
    """1024 degrees, N_l = 1024 - floor(l/2), 32 complex (64 real) systems each, in one
    step: sampled sets against the oracle, everything against the plain per-set form
    (bit for bit) and linearity of the whole step."""
    t = torch()
    L = 1024
    specs = [(1024 - l // 2, 64, 1, 1.0) for l in range(L)]
    d = BatchData(specs, seed=6, pad=0, off_sol=0)
    a = run_batch(d)
    assert np.isfinite(a).all()
    d.check(a, sets=[0, 1, 2, 511, 512, 1022, 1023])
    b = run_batch(d, mode="run_plain")
    np.testing.assert_array_equal(a, b)
    # linearity: step(x1 + 2 x2) = step(x1) + 2 step(x2)
    x1 = d.x.copy()
    x2 = O.uniform(99, -1, 1, d.x.size)
    d.x = x2
    y2 = run_batch(d)
    d.x = x1 + 2.0 * x2
    y12 = run_batch(d)
    assert O.relerr_norm(y12, a + 2.0 * y2) <= 1e-13


@pytest.mark.parametrize("kern", [0, 4])
@pytest.mark.parametrize("solve_only", [0, 1])
def test_batch_graph_of_several_steps(kern, solve_only):
    """three steps captured in one graph: one replay equals three eager steps, bit for bit,
    and a second replay continues from there"""
    t = torch()
    d = BatchData(SMALL_SETS, seed=9)
    stream = t.cuda.current_stream().cuda_stream
    tun = tuning(kernel=kern) if kern else None
    d.reset()
    with pf.Batch(d.sets, tun) as b:
        b.run(stream)
        b.sync(stream)
        one = d.result()
        b.run(stream)
        b.run(stream)
        b.sync(stream)
        three = d.result()
        for _ in range(3):
            b.run(stream)
        b.sync(stream)
        six = d.result()
        d.reset()
        if solve_only:
            b.prepare(stream)
        s = t.cuda.Stream()
        b.capture_graph(s.cuda_stream, solve_only=bool(solve_only), steps=3)
        b.launch_graph(s.cuda_stream)
        s.synchronize()
        np.testing.assert_array_equal(d.result(), three)
        b.launch_graph(s.cuda_stream)
        s.synchronize()
        np.testing.assert_array_equal(d.result(), six)
    assert not np.array_equal(one, three)


@pytest.mark.parametrize("kern", [0, 4, [4, 1, 4, 2, 1]])
def test_batch_prepare_then_solve_matches_run(kern):
    """tables prepared once and kept resident: every solve equals a full run (row classes
    on queue kernel 4 build their tables inside every solve)"""
    t = torch()
    d = BatchData(SMALL_SETS, seed=7)
    ref = run_batch(d)
    stream = t.cuda.current_stream().cuda_stream
    with pf.Batch(d.sets, tuning(kernel=kern) if kern else None) as b:
        b.prepare(stream)
        for _ in range(2):
            d.reset()
            b.solve(stream)
            np.testing.assert_array_equal(d.result(), ref)
        s = t.cuda.Stream()
        b.capture_graph(s.cuda_stream, solve_only=True)
        for rep in range(3):
            d.reset()
            b.launch_graph(s.cuda_stream)
            s.synchronize()
            np.testing.assert_array_equal(d.result(), ref, err_msg=f"replay {rep}")
