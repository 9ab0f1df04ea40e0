"""pfsolve_amd -- MI355X (gfx950) Jones-Worland step behind PfSolve's plan API.

The compute path is libpfsolve_gfx950.so (hand-written HIP kernels behind a C
ABI, include/pfSolve.h).  This module binds that library with ctypes and
mirrors the reference's plan interface: a PfSolveConfiguration describes the
plan, initializePfSolve creates it, PfSolveAppend enqueues one step with a
PfSolveLaunchParams, deletePfSolve frees it (reference:
PfSolve/pfSolve_AppManagement/pfSolve_InitializeApp.h:1480,
PfSolve/pfSolve_AppManagement/pfSolve_RunApp.h:80,
PfSolve/pfSolve_AppManagement/pfSolve_DeleteApp.h:28).

There is no CPU fallback: if the library is missing, importing the product
API raises.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _abi
from ._abi import *  # noqa: F401,F403  (structs and macros)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfsolve_gfx950.so")
_lib: C.CDLL | None = None


class PfSolveError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        name = "?"
        if _lib is not None:
            name = _lib.getPfSolveErrorString(code).decode()
            detail = _lib.pfsolve_last_error()
            if detail:
                name += " (" + detail.decode() + ")"
        super().__init__(f"{where}: {name} [{code}]")


def lib() -> C.CDLL:
    """The product library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m pfsolve_amd.build` "
                "(there is no CPU fallback for the JW step)")
        handle = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        _abi.declare(handle)
        _abi.declare_ext(handle)
        _lib = handle
    return _lib


def check(code: int, where: str) -> None:
    if code != 0:
        raise PfSolveError(code, where)


def initializePfSolve(app: _abi.PfSolveApplication, cfg: _abi.PfSolveConfiguration) -> int:
    return lib().initializePfSolve(C.byref(app), cfg)


def PfSolveAppend(app: _abi.PfSolveApplication, inverse: int, lp: _abi.PfSolveLaunchParams | None) -> int:
    return lib().PfSolveAppend(C.byref(app), inverse, C.byref(lp) if lp is not None else None)


def deletePfSolve(app: _abi.PfSolveApplication) -> None:
    lib().deletePfSolve(C.byref(app))


def PfSolveSync(app: _abi.PfSolveApplication) -> int:
    return lib().PfSolveSync(C.byref(app))


def getPfSolveErrorString(code: int) -> str:
    return lib().getPfSolveErrorString(code).decode()


def PfSolveGetVersion() -> int:
    return lib().PfSolveGetVersion()


class Plan:
    """One PfSolve application with its pointer arrays kept alive.

    ``fields`` are PfSolveConfiguration fields (size=[N, B, 1] etc.);
    ``matrices`` and ``rhs`` are device pointers (ints) of the plan's
    `buffer` and `outputBuffer` (``matrices`` a list of pointers for
    useMultipleInputBuffers: one buffer per chained iteration); ``stream`` an optional hipStream_t (int), or a
    list of them: launches then go round robin over the streams (num_streams).
    """

    def __init__(self, matrices: int, rhs: int, stream: int | None = None, **fields):
        self._keep = []
        cfg = _abi.PfSolveConfiguration()
        for k, v in fields.items():
            cur = getattr(cfg, k)
            if isinstance(cur, C.Array):
                arr = getattr(cfg, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(cfg, k, v)
        self._dev = C.c_int(0)
        cfg.device = C.pointer(self._dev)
        mats = list(matrices) if isinstance(matrices, (list, tuple)) else [matrices]
        self._buf = (C.c_void_p * len(mats))(*mats)
        self._out = (C.c_void_p * 1)(rhs)
        cfg.buffer = C.cast(self._buf, _abi.vpp)
        cfg.outputBuffer = C.cast(self._out, _abi.vpp)
        if stream is not None:
            streams = list(stream) if isinstance(stream, (list, tuple)) else [stream]
            self._stream = (C.c_void_p * len(streams))(*streams)
            cfg.stream = C.cast(self._stream, _abi.vpp)
            cfg.num_streams = len(streams)
        self.app = _abi.PfSolveApplication()
        check(initializePfSolve(self.app, cfg), "initializePfSolve")
        self.alive = True

    def append(self, **launch) -> None:
        lp = _abi.PfSolveLaunchParams()
        for k, v in launch.items():
            if k in ("buffer", "outputBuffer", "inputBuffer"):
                arr = (C.c_void_p * 1)(v)
                self._keep.append(arr)
                setattr(lp, k, C.cast(arr, _abi.vpp))
                continue
            cur = getattr(lp, k)
            if isinstance(cur, C.Array):
                for i, x in enumerate(v):
                    cur[i] = x
            else:
                setattr(lp, k, v)
        check(PfSolveAppend(self.app, 0, lp), "PfSolveAppend")
        del self._keep[:-4]

    def sync(self) -> None:
        """PfSolveSync: joins the plan's streams; raises if a kernel's bounded wait ran out"""
        check(PfSolveSync(self.app), "PfSolveSync")

    def geometry(self) -> _abi.PfSolveGeometry:
        g = _abi.PfSolveGeometry()
        check(lib().pfsolve_plan_geometry(C.byref(self.app), C.byref(g)), "pfsolve_plan_geometry")
        return g

    def close(self) -> None:
        if self.alive:
            deletePfSolve(self.app)
            self.alive = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """An uneven-size batch of fused JW steps (pfsolve_batch_* in include/pfsolve_amd.h).

    ``sets`` is a list of dicts with the pfsolve_batch_set fields (device
    pointers as ints): N, nsys, matrices, offsetV, offsetT, offsetM, rhs,
    stride, offsetSolution, naturalOrder, KL, KU, scaleC.
    """

    def __init__(self, sets, tuning: _abi.PfSolveBatchTuning | None = None):
        arr = (_abi.PfSolveBatchSet * len(sets))()
        for i, st in enumerate(sets):
            d = {"KL": 2, "KU": 2, "scaleC": 1.0}
            d.update(st)
            for k, v in d.items():
                setattr(arr[i], k, v)
        self._sets = arr
        self.handle = C.c_void_p()
        check(lib().pfsolve_batch_create(C.byref(self.handle), arr, len(sets),
                                         C.byref(tuning) if tuning is not None else None),
              "pfsolve_batch_create")

    def run(self, stream: int = 0) -> None:
        check(lib().pfsolve_batch_run(self.handle, stream), "pfsolve_batch_run")

    def prepare(self, stream: int = 0) -> None:
        check(lib().pfsolve_batch_prepare(self.handle, stream), "pfsolve_batch_prepare")

    def solve(self, stream: int = 0) -> None:
        check(lib().pfsolve_batch_solve(self.handle, stream), "pfsolve_batch_solve")

    def sync(self, stream: int = 0) -> None:
        """waits for `stream`; raises if a kernel's bounded wait ran out"""
        check(lib().pfsolve_batch_sync(self.handle, stream), "pfsolve_batch_sync")

    def run_plain(self, stream: int = 0) -> None:
        check(lib().pfsolve_batch_run_plain(self.handle, stream), "pfsolve_batch_run_plain")

    def capture_graph(self, stream: int, solve_only: bool = False, steps: int = 1) -> None:
        check(lib().pfsolve_batch_capture_graph_steps(self.handle, stream, int(solve_only), int(steps)),
              "pfsolve_batch_capture_graph")

    def launch_graph(self, stream: int) -> None:
        check(lib().pfsolve_batch_launch_graph(self.handle, stream), "pfsolve_batch_launch_graph")

    def tune(self, stream: int, repeats: int = 3, solve_only: bool = False) -> _abi.PfSolveBatchTuning:
        t = _abi.PfSolveBatchTuning()
        check(lib().pfsolve_batch_tune_mode(self.handle, stream, repeats, int(solve_only), C.byref(t)),
              "pfsolve_batch_tune")
        return t

    def set_tuning(self, t: _abi.PfSolveBatchTuning) -> None:
        check(lib().pfsolve_batch_set_tuning(self.handle, C.byref(t)), "pfsolve_batch_set_tuning")

    def info(self) -> _abi.PfSolveBatchInfo:
        i = _abi.PfSolveBatchInfo()
        check(lib().pfsolve_batch_get_info(self.handle, C.byref(i)), "pfsolve_batch_get_info")
        return i

    def close(self) -> None:
        if self.handle:
            lib().pfsolve_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(requests, hbm_budget: int, max_inflight: int = 256, streams: int = 4, max_batch: int = 64,
          realtime: bool = False, fail_launch: int = 0, dry_run: int = 0):
    """Runs the serving loop (pfsolve_serve) over `requests`, dicts with the
    pfsolve_serve_request fields (host pointers as ints), sorted by arrival_s.
    fail_launch = k > 0 makes the k-th launch fail; dry_run = n > 0 runs the policy
    on the host (no device, launches done after n polls).
    Returns (stats, per-request (admit_s, done_s, status) list)."""
    arr = (_abi.PfSolveServeRequest * max(1, len(requests)))()
    for i, q in enumerate(requests):
        d = {"KL": 2, "KU": 2, "scaleC": 1.0}
        d.update(q)
        for k, v in d.items():
            setattr(arr[i], k, v)
    cfg = _abi.PfSolveServeConfig(hbm_budget, max_inflight, streams, max_batch, int(realtime), int(fail_launch),
                                  int(dry_run))
    st = _abi.PfSolveServeStats()
    check(lib().pfsolve_serve(C.byref(cfg), arr, len(requests), C.byref(st)), "pfsolve_serve")
    return st, [(arr[i].admit_s, arr[i].done_s, arr[i].status) for i in range(len(requests))]
