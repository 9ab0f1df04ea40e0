"""Host-side behaviour of the uneven-size batch API that needs no GPU: argument
validation (rejected before any device call) and the tuning table file."""
from __future__ import annotations

import ctypes as C

import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

def R_(name):
    return getattr(_abi, name)


def make_sets(**over):
    s = _abi.PfSolveBatchSet()
    d = dict(N=100, nsys=4, matrices=0x10000, offsetV=0, offsetT=500, offsetM=800, rhs=0x20000, stride=100,
             offsetSolution=0, naturalOrder=0, KL=2, KU=2, scaleC=1.0)
    d.update(over)
    for k, v in d.items():
        setattr(s, k, v)
    arr = (_abi.PfSolveBatchSet * 1)(s)
    return arr


@pytest.mark.parametrize("over,code", [
    (dict(N=1), "PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH"),
    (dict(N=2049), "PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH"),
    (dict(nsys=-1), "PFSOLVE_ERROR_EMPTY_size"),
    (dict(KL=3), "PFSOLVE_ERROR_UNSUPPORTED_RADIX"),
    (dict(KU=-1), "PFSOLVE_ERROR_UNSUPPORTED_RADIX"),
    (dict(matrices=0), "PFSOLVE_ERROR_EMPTY_buffer"),
    (dict(offsetT=-1), "PFSOLVE_ERROR_EMPTY_bufferSize"),
    (dict(rhs=0), "PFSOLVE_ERROR_EMPTY_outputBuffer"),
    (dict(rhs=0x20008), "PFSOLVE_ERROR_EMPTY_outputBufferSize"),
    (dict(stride=99), "PFSOLVE_ERROR_EMPTY_outputBufferSize"),
    (dict(stride=101), "PFSOLVE_ERROR_EMPTY_outputBufferSize"),
    (dict(offsetSolution=1), "PFSOLVE_ERROR_EMPTY_outputBufferSize"),
])
def test_batch_create_rejects(over, code):
    h = C.c_void_p()
    rc = pf.lib().pfsolve_batch_create(C.byref(h), make_sets(**over), 1, None)
    assert rc == R_(code)
    assert not h.value


def test_batch_create_rejects_empty_list():
    h = C.c_void_p()
    assert pf.lib().pfsolve_batch_create(C.byref(h), None, 0, None) == R_("PFSOLVE_ERROR_EMPTY_size")


def test_tuning_file_round_trip(tmp_path):
    L = pf.lib()
    L.pfsolve_tuning_clear()
    p = tmp_path / "t.txt"
    # the kernel and systems-per-wavefront columns (last two) are optional: 0 when absent
    p.write_text("# comment\n0 2 5 8 4 1 1 0\n3 16 10 5 0 1 0 1 2\n1 4 6 8 0 1 0 0 3\n2 8 7 12 0 1 0 0 4\n"
                 "0 2 9 16 0 1 0 0 1 8\n")
    assert L.pfsolve_tuning_load(str(p).encode()) == 0
    q = tmp_path / "u.txt"
    assert L.pfsolve_tuning_save(str(q).encode()) == 0
    lines = [l.split() for l in q.read_text().splitlines() if not l.startswith("#")]
    assert lines == [["0", "2", "5", "8", "4", "1", "1", "0", "0", "0"],
                     ["3", "16", "10", "5", "0", "1", "0", "1", "2", "0"],
                     ["1", "4", "6", "8", "0", "1", "0", "0", "3", "0"], ["2", "8", "7", "12", "0", "1", "0", "0", "4", "0"],
                     ["0", "2", "9", "16", "0", "1", "0", "0", "1", "8"]]
    L.pfsolve_tuning_clear()


def test_tuning_file_rejects_bad_lines(tmp_path):
    L = pf.lib()
    L.pfsolve_tuning_clear()
    p = tmp_path / "bad.txt"
    # wrong R for class 1; shift 2; kernel 5; 3 systems per wavefront
    p.write_text("1 8 3 4 4 1 0 0\n0 2 1 1 1 1 2 0\n2 8 1 1 1 1 0 0 5\n0 2 1 1 1 1 0 0 1 3\n")
    assert L.pfsolve_tuning_load(str(p).encode()) == R_("PFSOLVE_ERROR_EMPTY_FILE")
    assert L.pfsolve_tuning_load(str(tmp_path / "missing.txt").encode()) == R_("PFSOLVE_ERROR_EMPTY_FILE")
    L.pfsolve_tuning_clear()
