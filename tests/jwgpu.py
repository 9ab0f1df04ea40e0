"""Helpers of the GPU tests: run a PfSolve plan of libpfsolve_gfx950.so on
device copies of numpy arrays (torch is used for device memory only)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import pfsolve_amd as pf


def torch():
    import torch as t
    return t


def to_dev(a: np.ndarray):
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run_plan(fields: dict, mat: np.ndarray, rhs: np.ndarray, launch: dict | None = None,
             launches: int = 1, expect: int = 0):
    """Creates the plan on device copies, runs `launches` appends, returns the RHS buffer
    (or the error code when `expect` != 0)."""
    t = torch()
    dmat = to_dev(mat)
    drhs = to_dev(rhs)
    stream = t.cuda.current_stream().cuda_stream
    f = dict(fields)
    bsize = (C.c_uint64 * 1)(dmat.numel() * dmat.element_size())
    osize = (C.c_uint64 * 1)(drhs.numel() * drhs.element_size())
    plan = pf.Plan(dmat.data_ptr(), drhs.data_ptr(), stream=stream, **f)
    plan.app.configuration.bufferSize = C.cast(bsize, pf._abi.u64p)
    plan.app.configuration.outputBufferSize = C.cast(osize, pf._abi.u64p)
    try:
        for _ in range(launches):
            lp = dict(launch or {})
            if expect:
                try:
                    plan.append(**lp)
                except pf.PfSolveError as e:
                    return e.code
                raise AssertionError("append succeeded, expected an error")
            plan.append(**lp)
        t.cuda.synchronize()
        return drhs.cpu().numpy()
    finally:
        plan.close()
