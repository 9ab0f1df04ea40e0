"""Block operators fused around the JW step (one plan, one launch): complex systems
read through block layout 4 and written through block layout 7 (with its
operator and optional accumulate) equal, bit for bit, the unfused chain of three
plans (block 4 -> JW step on the packed buffer -> block 7), and the composition
of the oracle's block and __float128 JW functions to the FP64 bar.  A baked
ScaleC of 0 on the input does not read the input (NaNs there stay out of the
result), rows outside the output systems are untouched, and misconfigurations
are refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import pfsolve_amd as pf
from pfsolve_amd import _abi

import _oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


def torch():
    import torch as t
    return t


class Case:
    """complex systems y = out_op(JW(in_op(x))) [+ prev]: host data and geometry"""

    def __init__(self, N, ncx, in_op, out_op, acc, seed=0, permuted=1, jw_scale=1.0, in_nan=False):
        self.N, self.ncx = N, ncx
        self.in_op, self.out_op, self.acc = in_op, out_op, acc
        self.inS, self.offIn = N + 3, 2          # complex elements
        self.outS, self.offOut = N + 1, 1
        self.S = N + (N & 1) + 2                 # packed (real) stride of the unfused chain
        self.in_scale = 0.0 if in_op == 1 and in_nan else 0.8
        self.out_scale = -1.25
        self.in_ls, self.out_ls = 1, 2
        self.jw_scale = jw_scale
        self.jt = 14 if permuted else 1014
        self.offV, self.offT, self.offM = 0, 5 * N + 2, 8 * N + 4
        raw = O.raw_matrices(seed + 17, N)
        self.mat = O.pack_fused(N, raw, permuted, self.offV, self.offT, self.offM, self.offM + 2 * N + 2)
        nin = 2 * ((ncx - 1) * self.inS + self.offIn + N + 3)
        nout = 2 * ((ncx - 1) * self.outS + self.offOut + N + 2)
        self.x = O.uniform(seed * 31 + 1, -1, 1, nin)
        if in_nan:
            self.x[:] = np.nan
        self.prev = O.uniform(seed * 31 + 2, -1, 1, nout)

    def jw_fields(self):
        return dict(size=[self.N, 2 * self.ncx, 1], jw_type=self.jt, doublePrecision=1, scaleC=self.jw_scale,
                    KL=2, KU=2, offsetV=self.offV, offsetT=self.offT, offsetM=self.offM)

    def fused_fields(self):
        f = self.jw_fields()
        f.update(outputBufferStride=[self.outS], offsetSolution=self.offOut, inputBufferStride=[self.inS],
                 inputBlock=self.in_op * 100 + 4, outputBlock=self.out_op * 100 + self.acc * 10 + 7,
                 inputBlockLshift=self.in_ls, outputBlockLshift=self.out_ls, inputBlockScale=self.in_scale,
                 outputBlockScale=self.out_scale, inputBlockOffset=self.offIn)
        return f

    def run_fused(self):
        t = torch()
        dmat = t.from_numpy(self.mat).cuda()
        din = t.from_numpy(self.x).cuda()
        dout = t.from_numpy(self.prev).cuda()
        stream = t.cuda.current_stream().cuda_stream
        with pf.Plan(dmat.data_ptr(), dout.data_ptr(), stream=stream, **self.fused_fields()) as p:
            p.append(inputBuffer=din.data_ptr())
            t.cuda.synchronize()
        return dout.cpu().numpy()

    def run_chain(self):
        t = torch()
        N, ncx, S = self.N, self.ncx, self.S
        dmat = t.from_numpy(self.mat).cuda()
        din = t.from_numpy(self.x).cuda()
        dout = t.from_numpy(self.prev).cuda()
        tmp = t.full((2 * ncx * S,), float("nan"), dtype=t.float64, device="cuda")
        stream = t.cuda.current_stream().cuda_stream
        with pf.Plan(din.data_ptr(), tmp.data_ptr(), stream=stream, size=[N, ncx, 1], doublePrecision=1,
                     block=self.in_op * 100 + 4, lshift=self.in_ls, scaleC=self.in_scale,
                     inputBufferStride=[self.inS], outputBufferStride=[S], offsetM=self.offIn,
                     offsetSolution=0) as p:
            p.append()
        jw = self.jw_fields()
        jw.update(outputBufferStride=[S])
        with pf.Plan(dmat.data_ptr(), tmp.data_ptr(), stream=stream, **jw) as p:
            p.append()
        with pf.Plan(tmp.data_ptr(), dout.data_ptr(), stream=stream, size=[N, ncx, 1], doublePrecision=1,
                     block=self.out_op * 100 + self.acc * 10 + 7, lshift=self.out_ls, scaleC=self.out_scale,
                     inputBufferStride=[S], outputBufferStride=[self.outS], offsetM=0,
                     offsetSolution=self.offOut) as p:
            p.append()
        t.cuda.synchronize()
        return dout.cpu().numpy()

    def oracle(self, systems):
        """block 4 (f64) -> JW step (__float128) -> block 7 (f64) on the CPU, all systems
        (the JW oracle runs only the listed real-system pairs' rows afterwards)"""
        N, ncx, S = self.N, self.ncx, self.S
        tmp = np.zeros(2 * ncx * S)
        b4 = O.BlockProblem(block=self.in_op * 100 + 4, lshift=self.in_ls, N=N, B=ncx, offsetM=self.offIn,
                            offsetSolution=0, scaleC=self.in_scale, inStride=self.inS, outStride=S)
        assert O.lib().pfo_block_f64(C.byref(b4), O.P(self.x), O.P(tmp)) == 0
        p = O.problem(jw_type=self.jt, N=N, B=2 * ncx, KL=2, KU=2, cfg_outStride=S, cfg_scaleC=self.jw_scale,
                      cfg_offsetV=self.offV, cfg_offsetT=self.offT, cfg_offsetM=self.offM)
        y = O.jw_f128(p, self.mat, tmp)
        out = self.prev.copy()
        b7 = O.BlockProblem(block=self.out_op * 100 + self.acc * 10 + 7, lshift=self.out_ls, N=N, B=ncx,
                            offsetM=0, offsetSolution=self.offOut, scaleC=self.out_scale, inStride=S,
                            outStride=self.outS)
        assert O.lib().pfo_block_f64(C.byref(b7), O.P(y), O.P(out)) == 0
        return out

    def system(self, a, q):
        """complex system q of an output buffer: (re, im)"""
        base = 2 * (q * self.outS + self.offOut)
        v = a[base:base + 2 * self.N]
        return v[0::2], v[1::2]

    def outside_mask(self):
        m = np.ones(self.prev.size, bool)
        for q in range(self.ncx):
            base = 2 * (q * self.outS + self.offOut)
            m[base:base + 2 * self.N] = False
        return m


OPS = [(0, 0, 0), (1, 1, 2), (2, 3, 0), (3, 4, 2), (4, 2, 2), (0, 1, 2)]


@pytest.mark.parametrize("N", [2, 3, 63, 100, 128, 129, 257, 511, 512])
@pytest.mark.parametrize("ops", OPS)
def test_fused_blocks_equal_unfused_chain_and_oracle(N, ops):
    c = Case(N, 37, *ops, seed=N, permuted=N % 2, jw_scale=0.5 if N % 3 else 1.0)
    got = c.run_fused()
    chain = c.run_chain()
    np.testing.assert_array_equal(got, chain)
    np.testing.assert_array_equal(got[c.outside_mask()], c.prev[c.outside_mask()])
    want = c.oracle(range(c.ncx))
    for q in (0, 1, c.ncx // 2, c.ncx - 1):
        for g, w in zip(c.system(got, q), c.system(want, q)):
            assert O.relerr_norm(g, w) <= TOL64, (N, ops, q)


@pytest.mark.parametrize("acc", [0, 2])
def test_fused_blocks_baked_zero_scale_does_not_read_input(acc):
    """input operator ScaleC with a baked scale of 0: zeros in, input never read (all
    NaN here), so the step gives 0 and the output is 0 or the previous content"""
    c = Case(100, 9, 1, 0, acc, seed=3, in_nan=True)
    got = c.run_fused()
    inside = ~c.outside_mask()
    if acc:
        np.testing.assert_array_equal(got, c.prev)
    else:
        assert (got[inside] == 0).all()
        np.testing.assert_array_equal(got[~inside], c.prev[~inside])


def test_fused_blocks_full_size_equals_chain():
    """N = 512, 32768 complex systems (65536 real): bit for bit against the unfused
    chain, sampled systems against the oracle composition"""
    c = Case(512, 32768, 2, 3, 2, seed=5)
    got = c.run_fused()
    chain = c.run_chain()
    np.testing.assert_array_equal(got, chain)
    small = Case(512, 4, 2, 3, 2, seed=5)
    small.x = c.x[:small.x.size].copy()
    small.prev = c.prev[:small.prev.size].copy()
    want = small.oracle(range(4))
    for q in range(4):
        for g, w in zip(c.system(got, q), small.system(want, q)):
            assert O.relerr_norm(g, w) <= TOL64, q


def append_code(fields, launch):
    t = torch()
    mat = t.zeros(16 * 512, dtype=t.float64, device="cuda")
    out = t.zeros(1 << 16, dtype=t.float64, device="cuda")
    stream = t.cuda.current_stream().cuda_stream
    with pf.Plan(mat.data_ptr(), out.data_ptr(), stream=stream, **fields) as p:
        try:
            p.append(**launch)
        except pf.PfSolveError as e:
            return e.code
        t.cuda.synchronize()
        return 0


def test_fused_blocks_refusals():
    t = torch()
    inp = t.zeros(1 << 16, dtype=t.float64, device="cuda")
    c = Case(64, 4, 2, 2, 0)
    f = c.fused_fields()
    assert append_code(f, dict(inputBuffer=inp.data_ptr())) == 0
    # no input buffer of its own
    assert append_code(f, {}) == _abi.PFSOLVE_ERROR_EMPTY_inputBuffer
    # an odd number of real systems (re/im halves come in pairs)
    g = dict(f, size=[64, 7, 1])
    assert append_code(g, dict(inputBuffer=inp.data_ptr())) == _abi.PFSOLVE_ERROR_EMPTY_size
    # partial keep-ranges are not fused
    g = dict(f, outputZeropad=[1, 60])
    assert append_code(g, dict(inputBuffer=inp.data_ptr())) == _abi.PFSOLVE_ERROR_UNSUPPORTED_RADIX
    # unaligned complex input
    assert append_code(f, dict(inputBuffer=inp.data_ptr() + 8)) != 0
