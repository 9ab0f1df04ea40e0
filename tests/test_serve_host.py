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
