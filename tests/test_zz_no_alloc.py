"""The allocation-free entry point (pfsolve_append_noalloc, through the C ABI) against
PfSolveAppend of the same plan: bit-identical results on two calls in a row on one GPU,
the second with a smaller input, for each kernel family that keeps its table in the
device scratch (prologue + pair kernel, long-system kernel, general kernel) and for
the compute / store split kernel (table in LDS, error flag in the scratch)."""
import ctypes as C

import numpy as np
import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

import _oracle as O

pytestmark = pytest.mark.gpu


def _fused(N, seed):
    offV, offT, offM = 0, 5 * N + 3, 8 * N + 11
    raw = O.raw_matrices(seed, N)
    return O.pack_fused(N, raw, 1, offV, offT, offM, offM + 2 * N + 4), (offV, offT, offM)


def _config(N, B, offs, variant, mat_ptr, rhs_ptr, keep):
    cfg = _abi.PfSolveConfiguration()
    cfg.size[0], cfg.size[1], cfg.size[2] = N, B, 1
    cfg.jw_type = 14
    cfg.doublePrecision = 1
    cfg.scaleC = 1.0
    cfg.KL = cfg.KU = 2
    cfg.offsetV, cfg.offsetT, cfg.offsetM = offs
    cfg.kernelVariant = variant
    dev = C.c_int(0)
    bufs = (C.c_void_p * 1)(mat_ptr)
    outs = (C.c_void_p * 1)(rhs_ptr)
    keep += [dev, bufs, outs]
    cfg.device = C.pointer(dev)
    cfg.buffer = C.cast(bufs, _abi.vpp)
    cfg.outputBuffer = C.cast(outs, _abi.vpp)
    return cfg


# (N, B, variant) of the first call, then of the second (smaller) one
CASES = [
    ((512, 64, 0), (300, 16, 0)),      # split kernel: table in LDS, error flag in the scratch
    ((1000, 32, 0), (700, 8, 0)),      # prologue + pair kernel (R = 16): table in the scratch
    ((2048, 8, 0), (1500, 4, 0)),      # long-system kernel (R = 32): 140 KiB table in the scratch
    ((777, 9, 1), (100, 3, 1)),        # general kernel: format-1 table in the scratch
]


@pytest.mark.parametrize("first,second", CASES)
def test_noalloc_matches_append(first, second):
    import torch
    lib = pf.lib()
    assert lib.pfsolve_noalloc_scratch_bytes() == 144 * 1024 + 256
    for k, (N, B, variant) in enumerate((first, second)):
        print(f"noalloc call {k}: N={N} B={B} variant={variant}", flush=True)
        mat, offs = _fused(N, 100 + N)
        x = O.uniform(200 + N, -1, 1, B * N)
        dmat = torch.from_numpy(mat).cuda()
        want_dev = torch.from_numpy(x.copy()).cuda()
        got_dev = torch.from_numpy(x.copy()).cuda()
        stream = torch.cuda.current_stream().cuda_stream
        with pf.Plan(dmat.data_ptr(), want_dev.data_ptr(), stream=stream, size=[N, B, 1], jw_type=14,
                     doublePrecision=1, scaleC=1.0, KL=2, KU=2, offsetV=offs[0], offsetT=offs[1],
                     offsetM=offs[2], kernelVariant=variant) as plan:
            plan.append()
        torch.cuda.synchronize()
        keep = []
        cfg = _config(N, B, offs, variant, dmat.data_ptr(), got_dev.data_ptr(), keep)
        st = (C.c_void_p * 1)(stream)
        cfg.stream = C.cast(st, _abi.vpp)
        cfg.num_streams = 1
        lp = _abi.PfSolveLaunchParams()
        rc = lib.pfsolve_append_noalloc(C.byref(cfg), C.byref(lp))
        assert rc == 0, pf.getPfSolveErrorString(rc)
        torch.cuda.synchronize()
        got, want = got_dev.cpu().numpy(), want_dev.cpu().numpy()
        np.testing.assert_array_equal(got, want)
        p = O.problem(jw_type=14, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                      cfg_offsetV=offs[0], cfg_offsetT=offs[1], cfg_offsetM=offs[2])
        ref = O.jw_f128(p, mat, x)
        assert max(O.relerr_norm(got[b * N:(b + 1) * N], ref[b * N:(b + 1) * N]) for b in range(B)) <= 1e-13


def test_noalloc_refusals():
    """several streams and configurations the plan itself refuses come back as error codes,
    before anything is launched"""
    import torch
    lib = pf.lib()
    dmat = torch.zeros(16 * 512, dtype=torch.float64, device="cuda")
    drhs = torch.zeros(512 * 4, dtype=torch.float64, device="cuda")
    keep = []
    cfg = _config(512, 4, (0, 5 * 512, 8 * 512), 0, dmat.data_ptr(), drhs.data_ptr(), keep)
    s = [torch.cuda.Stream() for _ in range(2)]
    st = (C.c_void_p * 2)(*[x.cuda_stream for x in s])
    cfg.stream = C.cast(st, _abi.vpp)
    cfg.num_streams = 2
    assert lib.pfsolve_append_noalloc(C.byref(cfg), None) == _abi.PFSOLVE_ERROR_INVALID_QUEUE
    cfg.num_streams = 1
    cfg.size[0] = 1501   # odd stride: the chunk path, whose tables do not fit the scratch
    assert lib.pfsolve_append_noalloc(C.byref(cfg), None) == _abi.PFSOLVE_ERROR_INSUFFICIENT_TEMP_BUFFER
    cfg.size[0] = 5000   # a long system: its chunk tables and reduced data exceed the scratch
    assert lib.pfsolve_append_noalloc(C.byref(cfg), None) == _abi.PFSOLVE_ERROR_INSUFFICIENT_TEMP_BUFFER
    cfg.size[0] = (1 << 22) + 1
    assert lib.pfsolve_append_noalloc(C.byref(cfg), None) == _abi.PFSOLVE_ERROR_UNSUPPORTED_FFT_LENGTH
    assert lib.pfsolve_append_noalloc(None, None) == _abi.PFSOLVE_ERROR_PLAN_NOT_INITIALIZED
