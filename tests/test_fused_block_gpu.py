"""Block operators fused around the JW step (one plan, one launch): complex systems
read through block layout 4 and written through block layout 7 (with its
operator and optional accumulate), or written as the step's real systems, equal
bit for bit the unfused chain of plans (block 4 -> JW step on the packed buffer
-> block 7), and the composition of the oracle's block and __float128 JW
functions to the FP64 bar.  The input keep-range selects the rows read (the
chain's block 4 zero pad), the output keep-range the rows written (block 7's).
A baked ScaleC of 0 on the input does not read the input (NaNs there stay out
of the result), rows outside the written rows are untouched, and
misconfigurations are refused."""
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

    def __init__(self, N, ncx, in_op, out_op, acc, seed=0, permuted=1, jw_scale=1.0, in_nan=False, izp=(0, 0),
                 ozp=(0, 0), real_out=False, in_scale=0.8):
        self.N, self.ncx = N, ncx
        self.in_op, self.out_op, self.acc = in_op, out_op, acc
        self.izp, self.ozp = list(izp), list(ozp)   # baked zero pads (0, 0: every row)
        self.real_out = real_out                    # no output block: the step's real systems
        self.inS, self.offIn = N + 3, 2          # complex elements
        self.outS, self.offOut = N + 1, 1
        self.S = N + (N & 1) + 2                 # packed (real) stride of the unfused chain
        self.in_scale = 0.0 if in_op == 1 and in_nan else in_scale
        self.out_scale = -1.25
        self.in_ls, self.out_ls = 1, 2
        self.jw_scale = jw_scale
        self.jt = 14 if permuted else 1014
        self.offV, self.offT, self.offM = 0, 5 * N + 2, 8 * N + 4
        raw = O.raw_matrices(seed + 17, N)
        self.mat = O.pack_fused(N, raw, permuted, self.offV, self.offT, self.offM, self.offM + 2 * N + 2)
        nin = 2 * ((ncx - 1) * self.inS + self.offIn + N + 3)
        nout = 2 * ((ncx - 1) * self.outS + self.offOut + N + 2)
        if real_out:   # real systems 2q, 2q+1 at stride S, offset 2 (16-byte aligned pairs)
            nout = 2 * ncx * self.S + 4
        self.x = O.uniform(seed * 31 + 1, -1, 1, nin)
        if in_nan:
            self.x[:] = np.nan
        self.prev = O.uniform(seed * 31 + 2, -1, 1, nout)

    def jw_fields(self):
        return dict(size=[self.N, 2 * self.ncx, 1], jw_type=self.jt, doublePrecision=1, scaleC=self.jw_scale,
                    KL=2, KU=2, offsetV=self.offV, offsetT=self.offT, offsetM=self.offM)

    def fused_fields(self):
        f = self.jw_fields()
        f.update(inputBufferStride=[self.inS], inputBlock=self.in_op * 100 + 4, inputBlockLshift=self.in_ls,
                 inputBlockScale=self.in_scale, inputBlockOffset=self.offIn, inputZeropad=self.izp,
                 outputZeropad=self.ozp)
        if self.real_out:
            f.update(outputBufferStride=[self.S], offsetSolution=2)
        else:
            f.update(outputBufferStride=[self.outS], offsetSolution=self.offOut,
                     outputBlock=self.out_op * 100 + self.acc * 10 + 7, outputBlockLshift=self.out_ls,
                     outputBlockScale=self.out_scale)
        return f

    def run_fused(self, runtime_izp=False):
        """runtime_izp: the input keep-range comes from the launch parameters
        (RUNTIME_INPUTZEROPAD, as the benchmark passes it), the plan's own is unset"""
        t = torch()
        dmat = t.from_numpy(self.mat).cuda()
        din = t.from_numpy(self.x).cuda()
        dout = t.from_numpy(self.prev).cuda()
        stream = t.cuda.current_stream().cuda_stream
        fields = self.fused_fields()
        launch = dict(inputBuffer=din.data_ptr())
        if runtime_izp:
            fields.update(jw_control_bitmask=_abi.RUNTIME_INPUTZEROPAD, inputZeropad=[0, 0])
            launch.update(inputZeropad=self.izp)
        with pf.Plan(dmat.data_ptr(), dout.data_ptr(), stream=stream, **fields) as p:
            p.append(**launch)
            t.cuda.synchronize()
        return dout.cpu().numpy()

    def run_chain(self):
        t = torch()
        N, ncx, S = self.N, self.ncx, self.S
        dmat = t.from_numpy(self.mat).cuda()
        din = t.from_numpy(self.x).cuda()
        dout = t.from_numpy(self.prev).cuda()
        tmp = t.full((2 * ncx * S,), float("nan"), dtype=t.float64, device="cuda")
        # real output: block 4 packs straight into the output buffer (offset 1 complex = 2 real)
        packed, poff = (dout, 1) if self.real_out else (tmp, 0)
        stream = t.cuda.current_stream().cuda_stream
        with pf.Plan(din.data_ptr(), packed.data_ptr(), stream=stream, size=[N, ncx, 1], doublePrecision=1,
                     block=self.in_op * 100 + 4, lshift=self.in_ls, scaleC=self.in_scale,
                     inputBufferStride=[self.inS], outputBufferStride=[S], offsetM=self.offIn,
                     offsetSolution=poff, inputZeropad=self.izp) as p:
            p.append()
        jw = self.jw_fields()
        jw.update(outputBufferStride=[S], offsetSolution=2 * poff)
        if self.real_out:
            jw.update(outputZeropad=self.ozp)
        with pf.Plan(dmat.data_ptr(), packed.data_ptr(), stream=stream, **jw) as p:
            p.append()
        if not self.real_out:
            with pf.Plan(tmp.data_ptr(), dout.data_ptr(), stream=stream, size=[N, ncx, 1], doublePrecision=1,
                         block=self.out_op * 100 + self.acc * 10 + 7, lshift=self.out_ls, scaleC=self.out_scale,
                         inputBufferStride=[S], outputBufferStride=[self.outS], offsetM=0,
                         offsetSolution=self.offOut, outputZeropad=self.ozp) as p:
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
        b4.izp[0], b4.izp[1] = self.izp
        assert O.lib().pfo_block_f64(C.byref(b4), O.P(self.x), O.P(tmp)) == 0
        p = O.problem(jw_type=self.jt, N=N, B=2 * ncx, KL=2, KU=2, cfg_outStride=S, cfg_scaleC=self.jw_scale,
                      cfg_offsetV=self.offV, cfg_offsetT=self.offT, cfg_offsetM=self.offM)
        y = O.jw_f128(p, self.mat, tmp)
        out = self.prev.copy()
        if self.real_out:
            for b in range(2 * ncx):
                for r in range(N):
                    if O.lib().pfo_zeropad_keep(r, self.ozp[0], self.ozp[1], 0):
                        out[2 + b * S + r] = y[b * S + r]
            return out
        b7 = O.BlockProblem(block=self.out_op * 100 + self.acc * 10 + 7, lshift=self.out_ls, N=N, B=ncx,
                            offsetM=0, offsetSolution=self.offOut, scaleC=self.out_scale, inStride=S,
                            outStride=self.outS)
        b7.ozp[0], b7.ozp[1] = self.ozp
        assert O.lib().pfo_block_f64(C.byref(b7), O.P(y), O.P(out)) == 0
        return out

    def rows(self):
        """rows the step writes (output keep-range), as an index array"""
        return np.array([r for r in range(self.N) if O.lib().pfo_zeropad_keep(r, self.ozp[0], self.ozp[1], 0)],
                        dtype=np.int64)

    def system(self, a, q):
        """complex system q of an output buffer, written rows only: (re, im)"""
        rows = self.rows()
        if self.real_out:
            return a[2 + 2 * q * self.S + rows], a[2 + (2 * q + 1) * self.S + rows]
        base = 2 * (q * self.outS + self.offOut)
        return a[base + 2 * rows], a[base + 2 * rows + 1]

    def outside_mask(self):
        """elements of the output buffer the step must leave untouched"""
        m = np.ones(self.prev.size, bool)
        rows = self.rows()
        for q in range(self.ncx):
            if self.real_out:
                m[2 + 2 * q * self.S + rows] = False
                m[2 + (2 * q + 1) * self.S + rows] = False
            else:
                base = 2 * (q * self.outS + self.offOut)
                m[base + 2 * rows] = False
                m[base + 2 * rows + 1] = False
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
    for q in sorted({0, 1, c.ncx // 2, c.ncx - 1} & set(range(c.ncx))):
        for g, w in zip(c.system(got, q), c.system(want, q)):
            assert O.relerr_norm(g, w) <= TOL64, (N, ops, q)


def check_against_chain_and_oracle(c, label):
    got = c.run_fused()
    chain = c.run_chain()
    out = c.outside_mask()
    np.testing.assert_array_equal(got[out], c.prev[out])
    # the chain's block 4 writes every row of a real output, the fused step only the
    # output keep-range: compare the rows the fused step writes
    np.testing.assert_array_equal(got[~out], chain[~out])
    if c.rows().size == 0:
        return
    want = c.oracle(range(c.ncx))
    for q in sorted({0, 1, c.ncx // 2, c.ncx - 1} & set(range(c.ncx))):
        for g, w in zip(c.system(got, q), c.system(want, q)):
            assert O.relerr_norm(g, w) <= TOL64, (label, q)


@pytest.mark.parametrize("N", [3, 100, 257, 512])
@pytest.mark.parametrize("izp,ozp", [((480, 512), (0, 0)), ((1, 50), (3, 0)), ((0, 0), (5, 61)), ((7, 8), (2, 3))])
def test_fused_blocks_keep_ranges(N, izp, ozp):
    """input keep-range = rows read (block 4's zero pad), output keep-range = rows
    written (block 7's); odd range ends included"""
    c = Case(N, 21, 2, 4, 2, seed=N + 1, izp=izp, ozp=ozp)
    check_against_chain_and_oracle(c, (N, izp, ozp))


@pytest.mark.parametrize("N", [2, 63, 128, 300, 512])
@pytest.mark.parametrize("in_op", [0, 1, 3])
def test_fused_blocks_real_output(N, in_op):
    """no output block: complex input through block 4, the step's real systems out
    (the chain block 4 -> JW step), with and without keep-ranges"""
    for izp, ozp in (((0, 0), (0, 0)), ((N // 3, N), (1, N - 1))):
        c = Case(N, 19, in_op, 0, 0, seed=2 * N + in_op, izp=izp, ozp=ozp, real_out=True)
        check_against_chain_and_oracle(c, (N, in_op, izp, ozp))


@pytest.mark.parametrize("ncx", [1, 2, 9])
@pytest.mark.parametrize("real_out", [False, True])
def test_fused_blocks_few_systems(ncx, real_out):
    """fewer complex systems than computing wavefronts: idle wavefronts and storers
    leave cleanly, every system right"""
    c = Case(100, ncx, 2, 1, 2, seed=40 + ncx, real_out=real_out)
    check_against_chain_and_oracle(c, (ncx, real_out))


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


def test_fused_blocks_headline_workload():
    """N = 512, 65536 complex systems read through block 104 (ScaleC 0.75), input
    keep-range [480, 512) baked and as a runtime launch parameter, packed real systems
    out: bit for bit against the chain (block 104 -> JW step), sampled systems against
    the oracle composition"""
    c = Case(512, 65536, 1, 0, 0, seed=9, izp=(480, 512), real_out=True, in_scale=0.75)
    got = c.run_fused()
    chain = c.run_chain()
    out = c.outside_mask()
    np.testing.assert_array_equal(got[~out], chain[~out])
    np.testing.assert_array_equal(got[out], c.prev[out])
    # the benchmark's form: the same keep-range as a runtime launch parameter
    np.testing.assert_array_equal(c.run_fused(runtime_izp=True), got)
    small = Case(512, 3, 1, 0, 0, seed=9, izp=(480, 512), real_out=True, in_scale=0.75)
    small.x = c.x[:small.x.size].copy()
    small.prev = c.prev[:small.prev.size].copy()
    want = small.oracle(range(3))
    for q in range(3):
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
    # unaligned complex input
    assert append_code(f, dict(inputBuffer=inp.data_ptr() + 8)) == _abi.PFSOLVE_ERROR_EMPTY_inputBufferSize
    # real output systems must be 16-byte aligned pairs (even stride and offset)
    r = Case(64, 4, 2, 0, 0, real_out=True).fused_fields()
    assert append_code(r, dict(inputBuffer=inp.data_ptr())) == 0
    assert append_code(dict(r, offsetSolution=1), dict(inputBuffer=inp.data_ptr())) == \
        _abi.PFSOLVE_ERROR_EMPTY_outputBufferSize
    assert append_code(dict(r, outputBufferStride=[65]), dict(inputBuffer=inp.data_ptr())) == \
        _abi.PFSOLVE_ERROR_EMPTY_outputBufferSize
    # matrix blocks before the start of the matrix buffer: refused on the host, nothing launches
    # (a baked negative offset, and runtime offsets whose unsigned value wraps to a negative one)
    assert append_code(dict(f, offsetM=-8), dict(inputBuffer=inp.data_ptr())) == _abi.PFSOLVE_ERROR_EMPTY_bufferSize
    assert append_code(dict(f, offsetV=-1), dict(inputBuffer=inp.data_ptr())) == _abi.PFSOLVE_ERROR_EMPTY_bufferSize
    rt = dict(f, jw_control_bitmask=_abi.RUNTIME_OFFSETT | _abi.RUNTIME_OFFSETM)
    for off in (dict(offsetT=-64, offsetM=0), dict(offsetT=0, offsetM=(1 << 64) - 16)):
        assert append_code(rt, dict(inputBuffer=inp.data_ptr(), **off)) == _abi.PFSOLVE_ERROR_EMPTY_bufferSize
