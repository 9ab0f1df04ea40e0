"""Host-side behaviour of the serving loop and the extension structs that needs
no GPU: struct layouts against the library, and argument validation, which
happens before any device call."""
from __future__ import annotations

import ctypes as C

import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi


@pytest.mark.parametrize("name,cls", [
    ("pfsolve_geometry", _abi.PfSolveGeometry),
    ("pfsolve_batch_set", _abi.PfSolveBatchSet),
    ("pfsolve_batch_tuning", _abi.PfSolveBatchTuning),
    ("pfsolve_batch_info", _abi.PfSolveBatchInfo),
    ("pfsolve_serve_config", _abi.PfSolveServeConfig),
    ("pfsolve_serve_request", _abi.PfSolveServeRequest),
    ("pfsolve_serve_stats", _abi.PfSolveServeStats),
])
def test_extension_struct_sizes_match_the_library(name, cls):
    assert pf.lib().pfsolve_sizeof_ext(name.encode()) == C.sizeof(cls)


def test_unknown_struct_name():
    assert pf.lib().pfsolve_sizeof_ext(b"nope") == 0


def req(**over):
    d = dict(N=100, nsys=4, matrix_key=1, matrices=0x1000, offsetV=0, offsetT=500, offsetM=800, rhs_in=0x2000,
             rhs_out=0x3000, arrival_s=0.0)
    d.update(over)
    return d


@pytest.mark.parametrize("reqs,cfg,code", [
    ([req()], dict(hbm_budget=1 << 10), "PFSOLVE_ERROR_EMPTY_size"),
    ([req()], dict(max_inflight=0), "PFSOLVE_ERROR_EMPTY_size"),
    ([req()], dict(streams=0), "PFSOLVE_ERROR_EMPTY_size"),
    ([req()], dict(streams=17), "PFSOLVE_ERROR_EMPTY_size"),
    ([req()], dict(max_batch=0), "PFSOLVE_ERROR_EMPTY_size"),
    ([req(N=1)], {}, "PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH"),
    ([req(N=2049)], {}, "PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH"),
    ([req(nsys=-1)], {}, "PFSOLVE_ERROR_EMPTY_size"),
    ([req(matrices=0)], {}, "PFSOLVE_ERROR_EMPTY_buffer"),
    ([req(offsetM=-2)], {}, "PFSOLVE_ERROR_EMPTY_bufferSize"),
    ([req(KL=3)], {}, "PFSOLVE_ERROR_UNSUPPORTED_RADIX"),
    ([req(rhs_out=0)], {}, "PFSOLVE_ERROR_EMPTY_outputBuffer"),
    ([req(arrival_s=1.0), req(arrival_s=0.5)], {}, "PFSOLVE_ERROR_EMPTY_size"),   # not sorted
])
def test_serve_rejects_bad_arguments_before_touching_the_device(reqs, cfg, code):
    c = dict(hbm_budget=1 << 30)
    c.update(cfg)
    with pytest.raises(pf.PfSolveError) as e:
        pf.serve(reqs, **c)
    assert e.value.code == getattr(_abi, code)


def test_serve_with_no_requests_is_a_no_op():
    st, out = pf.serve([], hbm_budget=1 << 30)
    assert st.requests == 0 and out == []


# ------------------------------------------------------------------ the policy on the host (dry run)
import numpy as np


class DryWorkload:
    """Requests over `keys` matrix sets of N rows (fused-step layout), each bundle with its
    own input and output; inputs hold distinct values so that a mixed-up copy shows."""

    def __init__(self, specs, seed=0):
        rng = np.random.default_rng(seed)
        self.mats, self.ins, self.outs, self.reqs = {}, [], [], []
        for i, (N, nsys, key) in enumerate(specs):
            if key not in self.mats:
                self.mats[key] = rng.uniform(-1, 1, 10 * N + 16)
            x = rng.uniform(-1, 1, nsys * N)
            y = np.full(nsys * N, np.nan)
            self.ins.append(x)
            self.outs.append(y)
            self.reqs.append(dict(N=N, nsys=nsys, matrix_key=key, matrices=self.mats[key].ctypes.data,
                                  offsetV=0, offsetT=5 * N, offsetM=8 * N, rhs_in=x.ctypes.data,
                                  rhs_out=y.ctypes.data, arrival_s=0.0))


def test_dry_run_serves_every_request():
    """without a device every result equals its input and every request completes"""
    w = DryWorkload([(64 + 37 * (i % 4), 3 + i % 5, i % 4) for i in range(40)], seed=1)
    st, out = pf.serve(w.reqs, hbm_budget=8 << 20, streams=3, max_batch=4, dry_run=3)
    assert st.requests == 40 and st.rejected == 0 and st.failed == 0 and st.integrity_failures == 0
    assert all(c == 0 for _, _, c in out)
    for x, y in zip(w.ins, w.outs):
        np.testing.assert_array_equal(x, y)
    assert st.resident_keys == 4 and st.evictions == 0


def test_eviction_never_takes_a_key_in_flight():
    """A budget that holds a few keys, launches that stay in flight for many polls while
    new keys keep arriving: keys are evicted least recently used first, but never one whose
    request is in flight (each completion checks its key's matrices are still the ones it
    copied).  More keys than the old fixed table held (4096)."""
    nkeys = 5000
    N = 64
    specs = [(N, 2, k) for k in range(nkeys)]
    # a request that re-uses an early key now and then
    specs += [(N, 2, k) for k in range(0, nkeys, 97)]
    w = DryWorkload(specs, seed=2)
    budget = 1 << 20          # about 60 keys of 64 rows (matrices + table) next to the bundles
    st, out = pf.serve(w.reqs, hbm_budget=budget, streams=4, max_batch=3, max_inflight=12, dry_run=25)
    assert st.requests == len(specs) and st.failed == 0 and st.rejected == 0
    assert st.integrity_failures == 0
    assert st.evictions > nkeys - 200
    assert st.peak_bytes <= budget
    assert all(c == 0 for _, _, c in out)
    for x, y in zip(w.ins, w.outs):
        np.testing.assert_array_equal(x, y)


def test_many_keys_stay_resident_when_they_fit():
    """the key table grows past 4096 entries; nothing is evicted when the budget holds all"""
    w = DryWorkload([(32, 1, k) for k in range(6000)], seed=3)
    st, out = pf.serve(w.reqs, hbm_budget=1 << 30, streams=2, max_batch=64, dry_run=2)
    assert st.requests == 6000 and st.evictions == 0
    assert st.resident_keys == 6000 and st.peak_keys == 6000


def test_injected_launch_failure_ends_only_that_launch():
    """the second launch fails before its kernels: its requests end with
    PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL, their memory and pins are returned, the key
    whose table that launch would have written is written by a later launch, and every
    other request completes (the device form: test_serve_gpu.py)"""
    dry = 4
    specs = [(100, 3, 0), (100, 2, 1), (200, 4, 2), (100, 1, 1), (200, 2, 2), (300, 5, 3), (100, 2, 0)]
    w = DryWorkload(specs, seed=4)
    st, out = pf.serve(w.reqs, hbm_budget=4 << 20, streams=1, max_batch=2, max_inflight=2, dry_run=dry,
                       fail_launch=2)
    codes = [c for _, _, c in out]
    failed = [i for i, c in enumerate(codes) if c != 0]
    assert failed == [2, 3]   # the second launch: the third and fourth requests
    assert all(codes[i] == _abi.PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL for i in failed)
    assert st.failed == 2 and st.requests == len(specs) - 2 and st.integrity_failures == 0
    for i, (x, y) in enumerate(zip(w.ins, w.outs)):
        if i in failed:
            continue
        np.testing.assert_array_equal(x, y)


def test_budget_exhaustion_queues_instead_of_failing():
    """bundles that fill the budget wait for memory rather than being rejected"""
    w = DryWorkload([(2048, 20, i % 2) for i in range(12)], seed=5)   # 330 KB each
    st, out = pf.serve(w.reqs, hbm_budget=1 << 20, streams=2, max_batch=1, max_inflight=16, dry_run=10)
    assert st.requests == 12 and st.rejected == 0 and st.peak_bytes <= 1 << 20


def test_key_reused_with_another_layout_is_refused():
    """a key's table is built for one band, set of matrix offsets and storage order: a later
    request of the key with another KL/KU, offsets or naturalOrder is refused, the others run"""
    w = DryWorkload([(100, 2, 7)] * 6, seed=6)
    w.reqs[1]["KL"] = 1
    w.reqs[2]["offsetT"] = w.reqs[2]["offsetT"] + 1
    w.reqs[3]["naturalOrder"] = 1
    w.reqs[4]["KU"] = 1
    st, out = pf.serve(w.reqs, hbm_budget=8 << 20, streams=2, max_batch=2, dry_run=2)
    codes = [c for _, _, c in out]
    assert codes[0] == 0 and codes[5] == 0
    assert codes[1:5] == [_abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX] * 4
    assert st.rejected == 4
