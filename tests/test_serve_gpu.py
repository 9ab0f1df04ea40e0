"""Serving loop (pfsolve_serve) on the GPU: every result against the __float128
oracle, each result the same alone and in the mix (bit for bit: the kernels
solve every system independently of what shares the launch), memory recycled
under a small budget, requests that cannot fit rejected without stalling the
others, and a real-time run with arrivals spread over time."""
from __future__ import annotations

import numpy as np
import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

import _oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


class Workload:
    """Requests with host matrices (one set per key) and host input/output."""

    def __init__(self, specs, seed=0, spacing_s=0.0):
        # specs: (N, nsys, key, permuted, scaleC)
        self.specs = specs
        self.mats, self.ins, self.outs, self.reqs, self.probs = {}, [], [], [], {}
        for i, (N, nsys, key, perm, sc) in enumerate(specs):
            if key not in self.mats:
                offV, offT, offM = 0, 5 * N + 2, 8 * N + 6
                raw = O.raw_matrices(1000 + key, N)
                self.mats[key] = (O.pack_fused(N, raw, perm, offV, offT, offM, offM + 2 * N + 2), (offV, offT, offM),
                                  perm)
            buf, (offV, offT, offM), perm_k = self.mats[key]
            x = O.uniform(seed * 7919 + i, -1, 1, max(nsys, 1) * N)
            y = np.full_like(x, np.nan)
            self.ins.append(x)
            self.outs.append(y)
            self.reqs.append(dict(N=N, nsys=nsys, matrix_key=key, matrices=buf.ctypes.data, offsetV=offV,
                                  offsetT=offT, offsetM=offM, naturalOrder=0 if perm_k else 1, scaleC=sc,
                                  rhs_in=x.ctypes.data, rhs_out=y.ctypes.data, arrival_s=i * spacing_s))

    def subset(self, idx):
        return [self.reqs[i] for i in idx]

    def oracle(self, i):
        N, nsys, key, perm, sc = self.specs[i]
        buf, (offV, offT, offM), perm_k = self.mats[key]
        p = O.problem(jw_type=14 if perm_k else 1014, N=N, B=nsys, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=sc,
                      cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
        return O.jw_f128(p, buf, self.ins[i])

    def check(self, i, got=None):
        N, nsys, *_ = self.specs[i]
        if nsys == 0:
            return
        got = self.outs[i] if got is None else got
        want = self.oracle(i)
        for b in range(nsys):
            err = O.relerr_norm(got[b * N:(b + 1) * N], want[b * N:(b + 1) * N])
            assert err <= TOL64, (i, N, b, err)


MIX = [(32, 40, 1, 1, 1.0), (2048, 7, 2, 1, 1.0), (100, 3, 3, 0, 0.5), (513, 64, 4, 1, 1.0), (1025, 5, 5, 1, 1.0),
       (64, 1, 6, 1, 2.0), (300, 17, 7, 1, 1.0), (32, 9, 1, 1, 1.0), (777, 33, 8, 0, 1.0), (2048, 2, 2, 1, 1.0),
       (129, 128, 9, 1, 1.0), (1024, 11, 10, 1, 0.75), (257, 0, 11, 1, 1.0), (513, 6, 4, 1, 1.0), (3, 5, 12, 1, 1.0)]


def test_serve_results_match_oracle_and_are_the_same_alone_and_in_the_mix():
    w = Workload(MIX, seed=1)
    st, out = pf.serve(w.reqs, hbm_budget=1 << 30, streams=4, max_batch=4)
    assert st.requests == len(MIX) and st.rejected == 0
    assert all(s == 0 for _, _, s in out)
    assert st.launches >= len(MIX) // 4
    for i in range(len(MIX)):
        w.check(i)
    mixed = [o.copy() for o in w.outs]
    # the same requests one at a time, each in a server of its own
    for i in (0, 1, 3, 4, 10, 13):
        w.outs[i][:] = np.nan
        st1, out1 = pf.serve(w.subset([i]), hbm_budget=1 << 30, streams=1, max_batch=1)
        assert st1.requests == 1 and out1[0][2] == 0
        np.testing.assert_array_equal(w.outs[i], mixed[i], err_msg=f"request {i} alone vs mixed")


def test_serve_recycles_memory_under_a_small_budget():
    specs = [(1000, 100, k % 3, 1, 1.0) for k in range(12)]   # ~0.8 MB each, ~10 MB in total
    w = Workload(specs, seed=2)
    budget = 3 << 20
    st, out = pf.serve(w.reqs, hbm_budget=budget, streams=3, max_batch=2, max_inflight=8)
    assert st.requests == 12 and st.rejected == 0
    assert st.peak_bytes <= budget
    assert st.resident_keys == 3
    for i in (0, 5, 11):
        w.check(i)


def test_serve_rejects_what_cannot_fit_and_finishes_the_rest():
    specs = [(64, 5, 1, 1, 1.0), (2048, 200, 2, 1, 1.0), (64, 3, 1, 1, 1.0), (128, 4, 1, 1, 1.0)]
    # (2048 x 200 x 8 B = 3.3 MB does not fit 2 MB; key 1 reused with another N is refused)
    w = Workload(specs, seed=3)
    st, out = pf.serve(w.reqs, hbm_budget=2 << 20, streams=2, max_batch=4)
    codes = [s for _, _, s in out]
    assert codes[0] == 0 and codes[2] == 0
    assert codes[1] == _abi.PFSOLVE_ERROR_FAILED_TO_ALLOCATE
    assert codes[3] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert st.rejected == 2 and st.requests == 2
    w.check(0)
    w.check(2)
    assert np.isnan(w.outs[1]).all()


def test_serve_real_time_arrivals():
    rng = np.random.default_rng(5)
    n = 80
    Ns = np.exp(rng.uniform(np.log(32), np.log(2048), n)).astype(int)
    sizes = np.exp(rng.uniform(np.log(4), np.log(64), n)).astype(int)
    specs = [(int(N), int(s), int(N), 1, 1.0) for N, s in zip(Ns, sizes)]
    w = Workload(specs, seed=4, spacing_s=0.0)
    t = np.cumsum(rng.exponential(1.0 / 800.0, n))   # Poisson, 800 requests/s
    for q, a in zip(w.reqs, t):
        q["arrival_s"] = float(a)
    st, out = pf.serve(w.reqs, hbm_budget=1 << 30, streams=4, max_batch=16, realtime=True)
    assert st.requests == n and st.rejected == 0
    assert 0 < st.p50_ms <= st.p99_ms <= st.max_ms
    for i, (adm, done, s) in enumerate(out):
        assert s == 0 and adm >= w.reqs[i]["arrival_s"] - 1e-6 and done >= adm
    for i in (0, n // 2, n - 1):
        w.check(i)
