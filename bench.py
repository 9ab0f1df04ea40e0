#!/usr/bin/env python3
"""Benchmark of the MI355X Jones-Worland step (one GPU).

Default workload (the throughput configuration): the fused JW step
y = M^{-1} T^{-1} (V x) on N=512 rows, 65536 complex systems in strided XXII
layout (131072 real systems, re/im halves of one complex system adjacent),
runtime scaleC = 0.75 on solution row 0 and runtime input keep-range
[480, 512) (rows outside it read as zero, as the reference's zero-pad does),
FP64.  V: band KL=KU=2, U(-1,1); T: diag U(2,3), off U(-0.5,0.5); M: diag
U(1,2), super U(-0.5,0.5); x: U(-1,1).  A step is one PfSolveAppend (the
coefficient prologue plus the main kernel) over the whole batch, inputs
resident in HBM.

Prints one JSON line: throughput in complex systems/s, the dominant kernel's
HBM roofline fraction (its time from HIP events recorded by the library on
the launch stream), and the CPU baseline (the oracle's FP64 restatement of
the reference step, OpenMP over systems) timed on a bounded sample.

  python bench.py [--gpus 1] [--steps 50] [--warmup 10] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "JW-step systems/s + % HBM roofline; p50/p99 latency, batched vs plain"
HBM_PEAK = 8.0e12
N_ROWS = 512
N_COMPLEX = 65536
ZP = (480, 512)
SCALE_C = 0.75


def gen_fused_matrices(N: int, seed: int, reorder) -> np.ndarray:
    """V (band-major, KL=KU=2), T block [a/b | 1/b | c/b], M block [1/d | e/d] in the
    reference's permuted storage; returns the buffer and offsets."""
    rng = np.random.default_rng(seed)
    V = rng.uniform(-1, 1, (5, N))
    for j in range(5):
        for col in range(N):
            r = col - 2 + j
            if not (0 <= r < N):
                V[j, col] = 0.0
    a, b, c = rng.uniform(-0.5, 0.5, N), rng.uniform(2, 3, N), rng.uniform(-0.5, 0.5, N)
    d, e = rng.uniform(1, 2, N), rng.uniform(-0.5, 0.5, N)
    a[0] = 0.0
    c[-1] = 0.0
    e[-1] = 0.0
    P = np.array([reorder(i, N, N, 0) for i in range(N)], dtype=np.int64)
    offV, offT, offM = 0, 5 * N, 8 * N
    buf = np.zeros(10 * N)
    buf[offV:offV + 5 * N] = V.ravel()
    buf[offT + N + P] = 1.0 / b
    buf[offT + P] = a / b
    buf[offT + 2 * N + P] = c / b
    buf[offM + P] = 1.0 / d
    buf[offM + N + P] = e / d
    return buf, (offV, offT, offM)


def algorithmic_bytes(N: int, nreal: int, rows_read: int, rows_written: int, elt: int = 8) -> int:
    """bytes = elt*(rows_read + rows_written) per real system + the matrices once (10 N values)."""
    return nreal * elt * (rows_read + rows_written) + 10 * N * elt


def cpu_baseline(buf, offs, N, budget_s=10.0):
    """Oracle FP64 restatement of the reference step on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    offV, offT, offM = offs
    sample_complex = 4096
    nreal = 2 * sample_complex
    x = np.random.default_rng(1).uniform(-1, 1, nreal * N)
    p = O.problem(jw_type=14, N=N, B=nreal, KL=2, KU=2, mask=16 | 64, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM, run_scaleC=SCALE_C, run_izp=list(ZP))
    lib = O.lib()
    threads = lib.pfo_num_threads()
    out = {}
    for label, nt, budget in (("all", threads, budget_s), ("one", 1, budget_s / 2)):
        done = 0
        t0 = time.perf_counter()
        while True:
            y = x.copy()
            rc = lib.pfo_jw_step_f64_range(C.byref(p), O.P(buf), O.P(y), 0, nreal, nt)
            assert rc == 0
            done += sample_complex
            if time.perf_counter() - t0 >= budget:
                break
        out[label] = done / (time.perf_counter() - t0)
    return {
        "value": out["all"], "unit": "systems/s", "cores": threads, "kind": "port",
        "value_1core": out["one"],
        "sample": f"{sample_complex} complex systems (N={N}, XXII, keep-range {list(ZP)}, scaleC {SCALE_C}) "
                  f"repeated for >= {budget_s:.0f} s on {threads} threads and {budget_s / 2:.0f} s on 1 thread; "
                  "oracle/pfsolve_oracle.c FP64 dgbmv + Thomas + bidiagonal backsolve",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the JW batch shards across GPUs with no exchange, "
              "and this build runs and reports the single-GPU path only (use --gpus 1).", file=sys.stderr)
        return 2

    import torch
    import pfsolve_amd as pf

    if not torch.cuda.is_available():
        print("no GPU available", file=sys.stderr)
        return 1
    torch.cuda.init()
    lib = pf.lib()
    N, nreal = N_ROWS, 2 * N_COMPLEX
    buf, (offV, offT, offM) = gen_fused_matrices(N, 0, lib.pfsolve_reorder_i0)
    dmat = torch.from_numpy(buf).cuda()
    g = torch.Generator(device="cuda").manual_seed(0)
    x0 = torch.rand(nreal * N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    x = x0.clone()
    stream = torch.cuda.Stream()
    sptr = stream.cuda_stream
    mask = pf.RUNTIME_SCALEC | pf.RUNTIME_INPUTZEROPAD
    bsize = (C.c_uint64 * 1)(dmat.numel() * 8)
    osize = (C.c_uint64 * 1)(x.numel() * 8)
    plan = pf.Plan(dmat.data_ptr(), x.data_ptr(), stream=sptr, size=[N, nreal, 1], jw_type=14, doublePrecision=1,
                   scaleC=1.0, offsetV=offV, offsetT=offT, offsetM=offM, jw_control_bitmask=mask)
    plan.app.configuration.bufferSize = C.cast(bsize, pf._abi.u64p)
    plan.app.configuration.outputBufferSize = C.cast(osize, pf._abi.u64p)
    geo = plan.geometry()
    launch = dict(scaleC=SCALE_C, inputZeropad=list(ZP))

    for _ in range(args.warmup):
        plan.append(**launch)
    torch.cuda.synchronize()
    x.copy_(x0)
    torch.cuda.synchronize()

    K = args.steps
    step_ev = [torch.cuda.Event(enable_timing=True) for _ in range(K + 1)]
    kst = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    ken = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    # torch creates the underlying HIP event on first record; create them before handing
    # the raw handles to the library
    for e in kst + ken + step_ev:
        e.record(stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for i in range(K):
            step_ev[i].record(stream)
            lib.pfsolve_set_kernel_events(C.byref(plan.app), C.c_void_p(kst[i].cuda_event),
                                          C.c_void_p(ken[i].cuda_event))
            plan.append(**launch)
        step_ev[K].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lib.pfsolve_set_kernel_events(C.byref(plan.app), None, None)
    assert torch.isfinite(x).all().item(), "non-finite output"

    total_ms = step_ev[0].elapsed_time(step_ev[K])
    ms_step = total_ms / K
    steps_ms = np.array([step_ev[i].elapsed_time(step_ev[i + 1]) for i in range(K)])
    kern_ms = np.array([kst[i].elapsed_time(ken[i]) for i in range(K)])
    kavg = float(kern_ms.mean())
    rows_read = ZP[1] - ZP[0]
    alg = algorithmic_bytes(N, nreal, rows_read, N)
    achieved = alg / (kavg * 1e-3)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "jw_fused_n512_b65536c_traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    plan.close()

    result = {
        "metric": METRIC,
        "value": N_COMPLEX / (ms_step * 1e-3),
        "unit": "systems/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (numpy/torch RNG, seeded; distributions of the stated workload)",
        "config": {
            "workload": "fused JW step y=M^-1 T^-1 (V x), N=512, 65536 complex systems (131072 real, XXII), "
                        "KL=KU=2, scaleC=0.75 (runtime, row 0), inputZeropad=[480,512) (runtime keep-range), FP64",
            "N": N, "complex_systems": N_COMPLEX, "real_systems": nreal, "jw_type": 14,
            "lanes": geo.lanes, "rows_per_lane": geo.rows_per_lane, "waves_per_workgroup": geo.waves,
        },
        "p50_ms": float(np.percentile(steps_ms, 50)),
        "p99_ms": float(np.percentile(steps_ms, 99)),
        "wall_s": wall,
        "roofline": {
            "bound": "hbm",
            "kernel": "jw_main_kernel<double,8,64>",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "algorithmic_bytes": alg,
            "kernel_ms": kavg,
        },
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(buf, (offV, offT, offM), N)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
