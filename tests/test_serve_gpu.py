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


def test_serve_injected_launch_failure_on_the_device():
    """the second launch fails before its kernels (forced from the host): its two requests
    end with PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL, the key whose table it would have written
    is written by a later launch, and every other request matches the oracle"""
    specs = [(100, 3, 1, 1, 1.0), (513, 2, 2, 1, 1.0), (300, 4, 3, 1, 1.0), (513, 1, 2, 1, 1.0),
             (300, 2, 3, 1, 0.5), (64, 5, 4, 1, 1.0), (100, 2, 1, 1, 1.0)]
    w = Workload(specs, seed=6)
    st, out = pf.serve(w.reqs, hbm_budget=64 << 20, streams=1, max_batch=2, max_inflight=2, fail_launch=2)
    codes = [c for _, _, c in out]
    assert [i for i, c in enumerate(codes) if c != 0] == [2, 3]
    assert codes[2] == codes[3] == _abi.PFSOLVE_ERROR_FAILED_TO_LAUNCH_KERNEL
    assert st.failed == 2 and st.requests == len(specs) - 2
    for i in (0, 1, 4, 5, 6):
        w.check(i)


def test_serve_evicts_under_a_budget_with_more_keys_than_the_old_table():
    """4600 matrix sets through a budget that holds a few hundred: keys with no request in
    flight are evicted least recently used first, keys come back after eviction (uploaded
    and their table written again), and the results match the oracle"""
    nkeys = 4600
    specs = [(48 + (k % 3) * 8, 2, k, 1, 1.0) for k in range(nkeys)]
    # early keys come back after they were evicted
    specs += [(48 + (k % 3) * 8, 2, k, 1, 1.0) for k in range(0, 400, 7)]
    w = Workload(specs, seed=7)
    budget = 8 << 20
    st, out = pf.serve(w.reqs, hbm_budget=budget, streams=4, max_batch=8, max_inflight=64)
    assert st.requests == len(specs) and st.rejected == 0 and st.failed == 0
    assert st.evictions > nkeys - 400 and st.peak_bytes <= budget
    for i in list(range(0, nkeys, 211)) + list(range(nkeys, len(specs), 5)):
        w.check(i)


def ring_offsets(nbytes, ring):
    """sequential regions of a ring (a region that would cross the end starts again at 0):
    offsets, cycle numbers"""
    offs, cyc, pos, c = [], [], 0, 0
    for b in nbytes:
        if pos + b > ring:
            pos, c = 0, c + 1
        offs.append(pos)
        cyc.append(c)
        pos += b
    return np.array(offs), np.array(cyc)


def ring_overlaps(offs, cyc, nbytes):
    """for each request, the requests of the previous cycle whose regions it reuses"""
    out = [[] for _ in offs]
    byc = {}
    for i, c in enumerate(cyc):
        byc.setdefault(int(c), []).append(i)
    for c, js in byc.items():
        prev = byc.get(c - 1, [])
        k = 0
        for j in js:
            lo, hi = offs[j], offs[j] + nbytes[j]
            while k < len(prev) and offs[prev[k]] + nbytes[prev[k]] <= lo:
                k += 1
            m = k
            while m < len(prev) and offs[prev[m]] < hi:
                out[j].append(prev[m])
                m += 1
    return out


def test_serve_poisson_stream_of_complex_bundles():
    """A Poisson stream at 2000 requests/s for 2 s: each request a bundle of log-uniform
    [64, 4096] complex systems (XXII: the re/im halves of a system adjacent, 2 real systems
    each) of N log-uniform in [32, 2048], matrices resident per N.  Inputs are windows of one
    pinned pool; every request writes its own region of a pinned output ring (a region is
    reused only by a request admitted after the earlier one finished: checked from the
    admission and completion times).  Every request completes; sampled requests whose
    outputs are never reused match the oracle and give the same bits alone as in the stream."""
    import torch
    rng = np.random.default_rng(21)
    rate, seconds = 2000.0, 2.0
    n = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    Ns = np.exp(rng.uniform(np.log(32), np.log(2048), n)).astype(np.int64)
    ncx = np.exp(rng.uniform(np.log(64), np.log(4096), n)).astype(np.int64)
    mats = {}
    for N in np.unique(Ns):
        N = int(N)
        offV, offT, offM = 0, 5 * N + 2, 8 * N + 6
        raw = O.raw_matrices(5000 + N, N)
        mats[N] = (O.pack_fused(N, raw, 1, offV, offT, offM, offM + 2 * N + 2), (offV, offT, offM))
    pool_t = torch.empty(2 * 4096 * 2048 + 4096, dtype=torch.float64, pin_memory=True)
    pool = pool_t.numpy()
    pool[:] = O.uniform(22, -1, 1, pool.size)
    nbytes = [2 * int(k) * int(N) * 8 for k, N in zip(ncx, Ns)]
    ring = 6 << 30
    offs, cyc = ring_offsets(nbytes, ring)
    ring_t = torch.empty(ring // 8, dtype=torch.float64, pin_memory=True)
    out_np = ring_t.numpy()
    reqs, starts = [], []
    for i in range(n):
        N, nsys = int(Ns[i]), 2 * int(ncx[i])
        buf, (offV, offT, offM) = mats[N]
        start = (i * 7919) % (pool.size - nsys * N)
        starts.append(start)
        reqs.append(dict(N=N, nsys=nsys, matrix_key=N, matrices=buf.ctypes.data, offsetV=offV, offsetT=offT,
                         offsetM=offM, naturalOrder=0, rhs_in=pool.ctypes.data + 8 * start,
                         rhs_out=ring_t.data_ptr() + int(offs[i]), arrival_s=float(arrivals[i])))
    print(f"poisson stream: {n} requests, {sum(b // 8 for b in nbytes)} values, {cyc[-1] + 1} ring cycles",
          flush=True)
    st, out = pf.serve(reqs, hbm_budget=64 << 30, streams=8, max_batch=64, max_inflight=512, realtime=True)
    print(f"done: p50 {st.p50_ms:.2f} ms p99 {st.p99_ms:.2f} ms, {st.systems_per_s:.3g} systems/s, "
          f"{st.launches} launches, mean batch {st.mean_batch:.1f}", flush=True)
    assert st.requests == n and st.rejected == 0 and st.failed == 0
    assert all(c == 0 for _, _, c in out)
    for i, (adm, done, _) in enumerate(out):
        assert adm >= reqs[i]["arrival_s"] - 1e-6 and done >= adm
    # no output region was shared by requests in flight together
    for j, prev in enumerate(ring_overlaps(offs, cyc, nbytes)):
        for i in prev:
            assert out[i][1] <= out[j][0], (i, j)
    # samples among the requests of the last cycle (their outputs are still in the ring)
    last = [i for i in range(n) if cyc[i] == cyc[-1]]
    sample = [int(i) for i in rng.choice(last, 12, replace=False)]

    def got_of(i):
        o = int(offs[i]) // 8
        return out_np[o:o + reqs[i]["nsys"] * reqs[i]["N"]]

    for i in sample:
        N, nsys = reqs[i]["N"], reqs[i]["nsys"]
        buf, (offV, offT, offM) = mats[N]
        # one complex system (two real systems) at the start and at the end of the bundle
        for b0 in (0, nsys - 2):
            x = pool[starts[i] + b0 * N: starts[i] + (b0 + 2) * N].copy()
            p = O.problem(jw_type=14, N=N, B=2, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                          cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
            want = O.jw_f128(p, buf, x)
            for b in range(2):
                got = got_of(i)[(b0 + b) * N:(b0 + b + 1) * N]
                assert O.relerr_norm(got, want[b * N:(b + 1) * N]) <= TOL64, (i, N, b0 + b)
    # alone against the stream, bit for bit
    for i in sample[:4]:
        mixed = got_of(i).copy()
        got_of(i)[:] = np.nan
        st1, out1 = pf.serve([dict(reqs[i], arrival_s=0.0)], hbm_budget=8 << 30, streams=1, max_batch=1)
        assert st1.requests == 1 and out1[0][2] == 0
        np.testing.assert_array_equal(got_of(i), mixed, err_msg=f"request {i} alone vs in the stream")
