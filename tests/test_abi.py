"""CPU tests of the C-ABI library: it loads, exports every symbol the headers
declare, its struct layout matches the ctypes mirror, and plan validation,
geometry, layout maths and error strings behave like the reference."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("pfSolve.h", "pfsolve_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, re.M):
            if m.group(1) not in ("defined",):
                names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = pf.lib()
    names = declared_functions()
    assert "initializePfSolve" in names and "PfSolveAppend" in names and "deletePfSolve" in names
    out = subprocess.run(["nm", "-D", "--defined-only", pf.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


def _fields(struct):
    out = []
    for name, _t in struct._fields_:
        out.append((name, getattr(struct, name).offset))
    return out


def test_struct_layout_matches_c_header():
    lib = pf.lib()
    assert lib.pfsolve_sizeof_configuration() == C.sizeof(_abi.PfSolveConfiguration)
    assert lib.pfsolve_sizeof_launch_params() == C.sizeof(_abi.PfSolveLaunchParams)
    assert lib.pfsolve_sizeof_application() == C.sizeof(_abi.PfSolveApplication)
    checked = 0
    for name, off in _fields(_abi.PfSolveConfiguration):
        c_off = lib.pfsolve_offsetof_configuration(name.encode())
        if c_off >= 0:
            assert c_off == off, name
            checked += 1
    assert checked >= 40
    for name, off in _fields(_abi.PfSolveLaunchParams):
        c_off = lib.pfsolve_offsetof_launch_params(name.encode())
        if c_off >= 0:
            assert c_off == off, name


def test_version_and_error_strings_match_reference():
    assert pf.PfSolveGetVersion() == 0o102 == 66
    if not os.path.exists(O.REF_SO):
        pytest.skip("oracle/_ref not built")
    ref = C.CDLL(O.REF_SO)
    ref.ref_error_string.argtypes = [C.c_int]
    ref.ref_error_string.restype = C.c_char_p
    ref.ref_version.restype = C.c_int
    assert ref.ref_version() == pf.PfSolveGetVersion()
    for code in list(range(0, 12)) + list(range(1000, 1013)) + list(range(2000, 2016)) + \
            list(range(3000, 3007)) + list(range(4000, 4057)) + [9999]:
        assert pf.getPfSolveErrorString(code) == ref.ref_error_string(code).decode(), code


def test_reorder_i0_library_matches_oracle():
    lib = pf.lib()
    for N in [1, 63, 64, 65, 192, 193, 512, 777, 1024, 3840, 4000, 10000]:
        for i in range(0, N, max(1, N // 97)):
            assert lib.pfsolve_reorder_i0(i, N, N, 0) == O.lib().pfo_reorder_i0(i, N, N, 0)
    for i in range(0, 7777, 13):
        assert lib.pfsolve_reorder_i0(i, 7777, 7777, 333) == O.lib().pfo_reorder_i0(i, 7777, 7777, 333)


def cfg(**kw):
    c = _abi.PfSolveConfiguration()
    c.size[0] = kw.pop("N", 64)
    c.size[1] = kw.pop("B", 1)
    c.doublePrecision = kw.pop("doublePrecision", 1)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def validate(c):
    g = _abi.PfSolveGeometry()
    rc = pf.lib().pfsolve_validate_configuration(C.byref(c), C.byref(g))
    return rc, g


def test_validation_error_codes():
    assert validate(cfg(N=0, jw_type=13, upperBanded=2))[0] == _abi.PFSOLVE_ERROR_EMPTY_size
    assert validate(cfg(jw_type=0))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX          # no plan kind
    assert validate(cfg(jw_type=13, block=104))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(jw_type=113, upperBanded=2))[0] == 0   # ALT
    assert validate(cfg(jw_type=213, upperBanded=2))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(jw_type=10, upperBanded=2))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(jw_type=13, upperBanded=3))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(jw_type=13, upperBanded=2, doublePrecision=0))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(N=5000, jw_type=13, upperBanded=2))[0] == 0          # long system, in chunks
    assert validate(cfg(N=(1 << 22) + 1, jw_type=13, upperBanded=2))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert validate(cfg(jw_type=14, KL=3, KU=2))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(LDA=4, KL=2, KU=2))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(block=108))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    for jt, ub in [(10, 0), (10, 1), (11, 0), (12, 1), (13, 0), (13, 1), (13, 2), (14, 0), (1013, 2), (1014, 0)]:
        assert validate(cfg(jw_type=jt, upperBanded=ub))[0] == 0, (jt, ub)
    assert validate(cfg(block=104))[0] == 0
    assert validate(cfg(LDA=5, KL=2, KU=2))[0] == 0


def test_fused_block_codes_validation():
    """block operators fused around a JW step: layout 4 in (no accumulate), layout 7
    out (accumulate 0 or 2) or no output block, operators 0-4, FP64, N <= 512, only
    on a JW plan"""
    ok = dict(jw_type=14, inputBlock=204, outputBlock=327)
    assert validate(cfg(**ok))[0] == 0
    for ib in (4, 104, 204, 304, 404):
        for ob in (0, 7, 27, 107, 227, 427):
            assert validate(cfg(jw_type=14, inputBlock=ib, outputBlock=ob))[0] == 0, (ib, ob)
    assert validate(cfg(N=512, **ok))[0] == 0
    bad_codes = [(0, 327), (207, 327), (224, 327), (204, 317), (204, 324), (504, 7), (4, 507), (-4, 7), (4, -7),
                 (204, 4)]
    for ib, ob in bad_codes:
        assert validate(cfg(jw_type=14, inputBlock=ib, outputBlock=ob))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX, \
            (ib, ob)
    assert validate(cfg(N=513, **ok))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert validate(cfg(quadDoubleDoublePrecision=1, **ok))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert validate(cfg(block=104, inputBlock=4, outputBlock=7))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(LDA=5, KL=2, KU=2, inputBlock=4, outputBlock=7))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX


@pytest.mark.parametrize("N", [1, 8, 64, 65, 128, 512, 513, 1000, 1024])
def test_geometry_covers_every_row(N):
    rc, g = validate(cfg(N=N, jw_type=14))
    assert rc == 0
    assert g.lanes * g.rows_per_lane >= N
    assert g.lanes * (g.rows_per_lane - 1) < N or g.rows_per_lane == 1
    assert pf.lib().pfsolve_kernel_supported(g.lanes, g.rows_per_lane, 1)
    assert g.lds_bytes <= 160 * 1024


@pytest.mark.parametrize("N", [1025, 1500, 2048])
def test_long_systems_validate_for_the_wave_per_system_kernel(N):
    """1024 < N <= 2048 (FP64): accepted with no general-kernel geometry (the wave-per-system
    kernel, or chunks for unaligned systems and kernelVariant 1)"""
    rc, g = validate(cfg(N=N, jw_type=14))
    assert rc == 0 and g.lanes == 0 and g.table_bytes > 0
    assert validate(cfg(N=N, jw_type=14, kernelVariant=1))[0] == 0


@pytest.mark.parametrize("N,dd", [(2049, 0), (3841, 0), (100000, 0), (1 << 22, 0), (513, 1), (5000, 1)])
def test_chunked_geometry(N, dd):
    """beyond one wavefront (FP64 N > 2048, double-double N > 512): chunks of 64 R rows,
    one table per chunk of every 4-chunk build; the wave-per-system kernel cannot be forced;
    the ALT variant, chained steps and fused blocks keep their own limits"""
    prec = dict(doublePrecision=0, quadDoubleDoublePrecision=1) if dd else {}
    for jt, ub in [(10, 0), (11, 1), (12, 0), (13, 1), (13, 2), (14, 0), (1013, 2)]:
        rc, g = validate(cfg(N=N, jw_type=jt, upperBanded=ub, **prec))
        assert rc == 0, (jt, ub)
        R = 4 if dd else 8
        assert g.lanes == 64 and g.rows_per_lane == R and g.waves == 4
        nbuild = -(-N // (4 * 64 * R))
        assert g.table_bytes == nbuild * 4 * (10 * 64 * R + 24 * 64) * (16 if dd else 8)
        assert g.lds_bytes == (10 * 64 * R + 24 * 64) * (16 if dd else 8)
    assert validate(cfg(N=N, jw_type=14, kernelVariant=2, **prec))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert validate(cfg(N=N, jw_type=14, inputBlock=4, **prec))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert validate(cfg(N=N, jw_type=10, upperBanded=1, numConsecutiveJWIterations=2, **prec))[0] != 0
    assert validate(cfg(N=2 * N + 2100, jw_type=113, upperBanded=2, **prec))[0] == \
        _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH


def test_dd_geometry_and_limits():
    rc, g = validate(cfg(N=256, jw_type=14, doublePrecision=0, quadDoubleDoublePrecision=1))
    assert rc == 0 and g.precision == 2 and g.lanes * g.rows_per_lane >= 256
    rc, g = validate(cfg(N=1024, jw_type=14, doublePrecision=0, quadDoubleDoublePrecision=1))
    assert rc == 0 and g.lanes == 64 and g.rows_per_lane == 4   # chunks
    rc, _ = validate(cfg(N=(1 << 22) + 1, jw_type=14, doublePrecision=0, quadDoubleDoublePrecision=1))
    assert rc == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH


def test_initialize_requires_device_pointer_and_cleans_up():
    app = _abi.PfSolveApplication()
    c = cfg(jw_type=13, upperBanded=2)
    assert pf.initializePfSolve(app, c) == _abi.PFSOLVE_ERROR_INVALID_DEVICE
    assert app.localFFTPlan is None
    # invalid configuration with a device pointer: validation fails before any device call
    dev = C.c_int(0)
    c = cfg(N=0, jw_type=13, upperBanded=2)
    c.device = C.pointer(dev)
    assert pf.initializePfSolve(app, c) == _abi.PFSOLVE_ERROR_EMPTY_size
    assert app.localFFTPlan is None
    pf.deletePfSolve(app)           # idempotent


def test_append_without_plan():
    app = _abi.PfSolveApplication()
    assert pf.PfSolveAppend(app, 0, _abi.PfSolveLaunchParams()) == _abi.PFSOLVE_ERROR_PLAN_NOT_INITIALIZED
    assert pf.PfSolveSync(app) == 0


def test_wait_bound_setters_need_a_plan_or_batch():
    app = _abi.PfSolveApplication()
    lib = pf.lib()
    assert lib.pfsolve_set_wait_polls(C.byref(app), -1) == _abi.PFSOLVE_ERROR_PLAN_NOT_INITIALIZED
    assert lib.pfsolve_batch_set_wait_polls(None, -1) == _abi.PFSOLVE_ERROR_PLAN_NOT_INITIALIZED
    assert lib.pfsolve_batch_sync(None, None) == _abi.PFSOLVE_ERROR_PLAN_NOT_INITIALIZED


def _ref_fields(ref, which):
    ref.ref_field.restype = C.c_char_p
    ref.ref_field.argtypes = [C.c_int, C.c_long, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    ref.ref_struct_size.restype = C.c_long
    ref.ref_field_count.restype = C.c_long
    out = []
    for i in range(ref.ref_field_count(which)):
        o, s = C.c_long(), C.c_long()
        out.append((ref.ref_field(which, i, C.byref(o), C.byref(s)).decode(), o.value, s.value))
    return out, ref.ref_struct_size(which)


@pytest.mark.parametrize("which,struct,first_ext", [(0, "PfSolveConfiguration", "offsetT"),
                                                   (1, "PfSolveLaunchParams", "offsetT"),
                                                   (2, "PfSolveApplication", None)])
def test_struct_layout_matches_reference_hip_build(which, struct, first_ext):
    """Every member of the reference's HIP build of the struct (compiled from the reference
    headers under oracle/_ref) exists here under the same name, at the same offset and with the
    same size; the reference's members are listed completely (no gap of 8 or more bytes between
    them or before the end of the struct); the extension members start right after them (in the
    launch parameters, in the tail padding that the long double scaleC gives the struct)."""
    if not os.path.exists(O.REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ref = C.CDLL(O.REF_SO)
    fields, size = _ref_fields(ref, which)
    end = 0
    for name, off, sz in fields:
        assert off - end < 8, ("member missing from the reference list before", name)
        end = off + sz
    assert size - end < (16 if which == 1 else 8)
    mine = getattr(_abi, struct)
    # the application embeds the configuration, whose extension members shift what follows
    shift = C.sizeof(_abi.PfSolveConfiguration) - ref.ref_struct_size(0) if which == 2 else 0
    for name, off, sz in fields:
        assert hasattr(mine, name), name
        d = getattr(mine, name)
        if name == "configuration":
            assert (d.offset, d.size) == (off, sz + shift)
        else:
            assert (d.offset, d.size) == (off + shift if off else 0, sz), (name, d.offset, d.size, off, sz)
    if first_ext:
        assert getattr(mine, first_ext).offset == end
    else:
        assert C.sizeof(mine) == size + shift
    # the C header agrees with the mirror
    lib = pf.lib()
    sizes = {0: lib.pfsolve_sizeof_configuration, 1: lib.pfsolve_sizeof_launch_params,
             2: lib.pfsolve_sizeof_application}
    assert sizes[which]() == C.sizeof(mine)


def test_c_caller_compiles_against_the_header(tmp_path):
    """A C11 caller written the way the reference harness drives a plan (fields set by
    name, including ignored FFT-machinery fields) compiles warning-free against
    include/pfSolve.h and links libpfsolve_gfx950.so (run on the GPU in test_c_caller_gpu)."""
    import c_caller
    exe = c_caller.build(str(tmp_path))
    assert os.path.exists(exe)


def test_chained_iterations_validation():
    """numConsecutiveJWIterations K > 1: the bidiagonal modes (jw 10/11/12, upperBanded 0/1)
    in FP64 with every iteration's system within 1024 rows; useMultipleInputBuffers is 0 or K"""
    ok = dict(N=512, jw_type=10, upperBanded=1, numConsecutiveJWIterations=4)
    assert validate(cfg(**ok))[0] == 0
    assert validate(cfg(**dict(ok, useMultipleInputBuffers=4)))[0] == 0
    assert validate(cfg(**dict(ok, useMultipleInputBuffers=3)))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(**dict(ok, jw_type=1012, upperBanded=0)))[0] == 0
    assert validate(cfg(**dict(ok, jw_type=11, upperBanded=0)))[0] == 0
    assert validate(cfg(**dict(ok, jw_type=13, upperBanded=2)))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(**dict(ok, jw_type=14, upperBanded=0)))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(**dict(ok, doublePrecision=0, quadDoubleDoublePrecision=1)))[0] == \
        _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(**dict(ok, numConsecutiveJWIterations=17)))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    # growing systems: the last one has N + K - 1 rows
    assert validate(cfg(**dict(ok, N=1021, upperBanded=0)))[0] == 0
    assert validate(cfg(**dict(ok, N=1022, upperBanded=0)))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    # shrinking systems keep at least one row
    assert validate(cfg(**dict(ok, N=3, upperBanded=1)))[0] == _abi.PFSOLVE_ERROR_EMPTY_size


def test_alt_validation():
    """ALT (jw_type hundreds digit 1): jw 10-13 with baked matrix positions, one iteration,
    parallel (permuted) layout; each parity subsystem within the general kernel's sizes"""
    assert validate(cfg(N=2048, jw_type=110, upperBanded=1))[0] == 0
    assert validate(cfg(N=2049, jw_type=110, upperBanded=1))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert validate(cfg(N=1024, jw_type=113, upperBanded=2, doublePrecision=0,
                        quadDoubleDoublePrecision=1))[0] == 0
    assert validate(cfg(N=100, jw_type=114))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(N=100, jw_type=1110, upperBanded=1))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(N=100, jw_type=110, upperBanded=1,
                        jw_control_bitmask=_abi.RUNTIME_OFFSETM))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(N=100, jw_type=110, upperBanded=1, numConsecutiveJWIterations=2))[0] == \
        _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    assert validate(cfg(N=100, jw_type=210, upperBanded=1))[0] == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX


def test_alt_stride_library_matches_oracle():
    lib = pf.lib()
    lib.jw_alt_stride.argtypes = [C.c_int64, C.c_uint64]
    lib.jw_alt_stride.restype = C.c_int64
    for N in [1, 2, 63, 64, 65, 511, 512, 513, 1023, 1024, 1025, 2048, 3840, 5000]:
        assert lib.jw_alt_stride(N, 0) == O.alt_stride(N), N
