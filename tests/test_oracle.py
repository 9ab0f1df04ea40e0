"""Pins the CPU oracle (test infrastructure) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle fails it:
  * reorder_i0 against the reference's own header compiled under oracle/_ref;
  * each JW mode against a dense __float128 Gaussian elimination built from the
    mode's operator equation (a wrong diagonal, sign or index changes the result);
  * the tridiagonal solve against the closed-form inverse of tridiag(-1,2,-1)
    and against manufactured solutions;
  * the fused step (stored, pre-normalised matrices) against an independent
    __float128 evaluation on the raw matrices;
  * double-double primitives against exact error-free-transform identities;
  * block scale operators against closed forms in long double, layouts against
    round-trip identities.
The reference's tests hold no expected outputs (its 513-row a0/b0/c0/x0
arrays are inputs only); they are used through tests/golden as a realistic
system whose residual is checked in __float128.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import _oracle as O

RT = dict(M=2, V=4, SOL=8, IZP=16, OZP=32, SC=64, OSTR=256, T=512)


def lib():
    return O.lib()


# ---------------------------------------------------------------- layout maths
def test_reorder_i0_matches_reference_build():
    if not os.path.exists(O.REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ref = C.CDLL(O.REF_SO)
    ref.ref_reorder_i0.argtypes = [C.c_int] * 4
    for N in list(range(1, 200)) + [255, 256, 257, 511, 512, 513, 777, 1024, 2048, 3840, 3841, 5000, 7040, 10000]:
        for i in range(N):
            assert lib().pfo_reorder_i0(i, N, N, 0) == ref.ref_reorder_i0(i, N, N, 0), (N, i)
    # blocks of a caller-chosen Msplit0 beyond 3840 rows (odd block rows, a short last block)
    for N, ms in [(4000, 512), (5000, 1000), (7777, 333)]:
        for i in range(N):
            assert lib().pfo_reorder_i0(i, N, N, ms) == ref.ref_reorder_i0(i, N, N, ms), (N, ms, i)


def test_reorder_i0_is_permutation_and_identity_for_short_systems():
    for N in [1, 2, 7, 64, 65, 100, 191, 192, 193, 512, 1000, 3840]:
        pos = [lib().pfo_reorder_i0(i, N, N, 0) for i in range(N)]
        assert sorted(pos) == list(range(N))
        if N <= 64:
            assert pos == list(range(N))


def test_reorder_i0_long_systems_permute_within_blocks():
    """beyond 3840 rows every Msplit0-row block (704 by default) is a permutation of its
    own rows, laid out as a 704-row system (11 rows per lane, 64 lanes): the first lane's
    rows of a block sit at its first positions one register apart"""
    for N, ms in [(3841, 0), (7040, 0), (10000, 0), (5000, 1000)]:
        blk = ms or 704
        pos = [lib().pfo_reorder_i0(i, N, N, ms) for i in range(N)]
        assert sorted(pos) == list(range(N))
        for b0 in range(0, N, blk):
            assert sorted(pos[b0:b0 + blk]) == list(range(b0, min(N, b0 + blk)))
    pos = [lib().pfo_reorder_i0(i, 7040, 7040, 0) for i in range(704)]
    assert pos[0] == 0 and pos[1] == 64 and pos[11] == 1


def test_reorder_i0_row_major_chunks():
    # logical row t*R + k is stored at sum_{j<k} A_j + t, A_j = lanes owning register j
    N = 512
    R = lib().pfo_jw_registers_per_thread(N, 1, 1)
    assert R == 9
    A = [sum(1 for t in range(64) if t * R + j < N) for j in range(R)]
    for t in range(64):
        for k in range(R):
            r = t * R + k
            if r < N:
                assert lib().pfo_reorder_i0(r, N, N, 0) == sum(A[:k]) + t


def test_zeropad_keep_semantics():
    keep = lambda r, a, b, rt: lib().pfo_zeropad_keep(r, a, b, rt)
    # baked: 0/0 keeps everything, an upper bound of 0 means unbounded
    assert all(keep(r, 0, 0, 0) for r in range(10))
    assert [keep(r, 3, 0, 0) for r in range(6)] == [0, 0, 0, 1, 1, 1]
    assert [keep(r, 0, 2, 0) for r in range(4)] == [1, 1, 0, 0]
    assert [keep(r, 1, 3, 0) for r in range(5)] == [0, 1, 1, 0, 0]
    # runtime bounds are literal: [0,0) keeps nothing
    assert not any(keep(r, 0, 0, 1) for r in range(10))
    assert [keep(r, 1, 3, 1) for r in range(5)] == [0, 1, 1, 0, 0]


# ---------------------------------------------------------------- JW modes vs dense f128
def _perm(N, jw_type):
    if jw_type > 1000 or N <= 64:
        return np.arange(N)
    return np.array([lib().pfo_reorder_i0(i, N, N, 0) for i in range(N)])


def _dense_mode(jw, ub, N, buf, P, offM, offV, rtM, rtV):
    """Dense A (matvec) and S (solve) of a reference mode, from the operator equations
    y = (I+L_S)^{-1}(D_A+U_A)x (jw 10, ub 1), y = (I+U_S)^{-1}(D_A+L_A)x (jw 10, ub 0),
    y = (I+L+U)^{-1} D x (jw 13, ub 2), y = (I+L)^{-1} D x (jw 13, ub 1),
    y = (I+U)^{-1} D x (jw 13, ub 0); positions per pfSolve_JonesWorland.h:117-348."""
    g = lambda base, r: buf[base + P[r]]
    A = np.eye(N)
    S = np.eye(N)
    if ub == 2:
        A = np.diag([g(N, r) for r in range(N)])
        for r in range(1, N):
            S[r, r - 1] = g(0, r)
        for r in range(N - 1):
            S[r, r + 1] = g(2 * N, r)
        return A, S
    if jw in (0, 1):
        if ub:
            ud, md = (offM, offM + N) if rtM else (N, 0)
            A = np.diag([g(md, r) for r in range(N)])
            for r in range(N - 1):
                A[r, r + 1] = g(ud, r)
        else:
            md, ld = (offM, offM + N) if rtM else (N, 0)
            A = np.diag([g(md, r) for r in range(N)])
            for r in range(1, N):
                A[r, r - 1] = g(ld, r)
    if jw in (0, 2):
        base = offV if rtV else 2 * N
        if ub:
            for r in range(1, N):
                S[r, r - 1] = g(base, r)
        else:
            for r in range(N - 1):
                S[r, r + 1] = g(base, r)
    if jw == 3:
        if ub:
            md, ld = (offM, offM + N) if rtM else (N, 0)
            A = np.diag([g(md, r) for r in range(N)])
            for r in range(1, N):
                S[r, r - 1] = g(ld, r)
        else:
            ud, md = (offM, offM + N) if rtM else (N, 0)
            A = np.diag([g(md, r) for r in range(N)])
            for r in range(N - 1):
                S[r, r + 1] = g(ud, r)
    return A, S


MODES = [(10, 1), (10, 0), (11, 1), (11, 0), (12, 1), (12, 0), (13, 2), (13, 1), (13, 0)]


@pytest.mark.parametrize("jw_type,ub", MODES)
@pytest.mark.parametrize("N", [1, 5, 33, 64, 70])
@pytest.mark.parametrize("rt", [0, 1])
def test_modes_match_dense_f128(jw_type, ub, N, rt):
    jw = jw_type % 10
    offM, offV = (7, 3 * N + 11) if rt else (0, 0)
    mask = (RT["M"] | RT["V"]) if rt else 0
    total = max(3 * N, offV + N, offM + 2 * N) + 16
    buf = O.uniform(11 + N, -0.3, 0.3, total)
    # diagonal-type slots get values away from 0 so that D is well scaled
    buf = np.where(np.arange(total) % 3 == 0, buf + 1.0, buf)
    B, S = 3, N + 5
    x = O.uniform(5, -1, 1, B * S)
    p = O.problem(jw_type=jw_type, upperBanded=ub, N=N, B=B, mask=mask, cfg_outStride=S,
                  cfg_scaleC=1.0, run_offsetM=offM, run_offsetV=offV)
    got = O.jw_f64(p, buf, x)
    got_q = O.jw_f128(p, buf, x)
    P = _perm(N, jw_type)
    A, Sm = _dense_mode(jw, ub, N, buf, P, offM, offV, rt, rt)
    for b in range(B):
        xb = x[b * S:b * S + N]
        u = A @ xb if jw != 2 else xb.copy()
        want = np.zeros(N)
        if jw == 1:
            want = A @ xb
        else:
            lib().pfo_dense_solve_f128(N, O.P(np.ascontiguousarray(Sm)), O.P(np.ascontiguousarray(u)), O.P(want))
        assert O.relerr_norm(got_q[b * S:b * S + N], want) < 1e-15
        assert O.relerr_norm(got[b * S:b * S + N], want) < 1e-13
        # rows between systems untouched
        np.testing.assert_array_equal(got[b * S + N:(b + 1) * S], x[b * S + N:(b + 1) * S])


def test_tridiagonal_closed_form_inverse():
    # (I + L + U) with ld = ud = -1/2 and md = 1/2 is tridiag(-1,2,-1)/2; its inverse is
    # (T^{-1})_{ij} = min(i,j) (N+1-max(i,j)) / (N+1) in 1-based indices.
    for N in [1, 2, 3, 10, 64, 100]:
        buf = np.zeros(3 * N)
        P = _perm(N, 13)
        for r in range(N):
            buf[P[r]] = -0.5 if r > 0 else 0.0
            buf[N + P[r]] = 0.5
            buf[2 * N + P[r]] = -0.5 if r < N - 1 else 0.0
        x = O.uniform(3, -1, 1, N)
        p = O.problem(jw_type=13, upperBanded=2, N=N, B=1, cfg_outStride=N, cfg_scaleC=1.0)
        got = O.jw_f128(p, buf, x)
        i = np.arange(1, N + 1)
        Tinv = np.minimum.outer(i, i) * (N + 1 - np.maximum.outer(i, i)) / (N + 1)
        want = Tinv @ x
        assert O.relerr_norm(got, want) < 1e-14


def test_tridiagonal_manufactured_solution():
    N = 300
    rng = np.random.default_rng(1)
    ld = rng.uniform(-0.2, 0.2, N); ld[0] = 0
    ud = rng.uniform(-0.2, 0.2, N); ud[-1] = 0
    ystar = rng.uniform(-1, 1, N)
    # rhs = (I+L+U) y* computed exactly enough in long double, then rounded
    rhs = (ystar.astype(np.longdouble) + ld.astype(np.longdouble) * np.r_[0, ystar[:-1]]
           + ud.astype(np.longdouble) * np.r_[ystar[1:], 0]).astype(np.float64)
    P = _perm(N, 13)
    buf = np.zeros(3 * N)
    buf[P] = ld
    buf[N + P] = 1.0
    buf[2 * N + P] = ud
    p = O.problem(jw_type=13, upperBanded=2, N=N, B=1, cfg_outStride=N, cfg_scaleC=1.0)
    assert O.relerr_norm(O.jw_f64(p, buf, rhs), ystar) < 1e-14


def test_diagonal_only_reduces_to_scaling():
    N = 50
    buf = np.zeros(3 * N)
    buf[N:2 * N] = O.uniform(4, 1, 2, N)
    x = O.uniform(9, -1, 1, N)
    p = O.problem(jw_type=13, upperBanded=2, N=N, B=1, cfg_outStride=N, cfg_scaleC=1.0)
    np.testing.assert_array_equal(O.jw_f64(p, buf, x), buf[N:2 * N] * x)


def test_scalec_zeropad_and_offsets():
    N, B, S = 40, 4, 64
    buf = np.zeros(3 * N)
    buf[N:2 * N] = 1.0
    x = O.uniform(2, -1, 1, B * S + 10)
    # baked scaleC 0.75 multiplies row 0 only; runtime zero-pads are literal keep-ranges
    p = O.problem(jw_type=13, upperBanded=2, N=N, B=B, mask=RT["IZP"] | RT["OZP"] | RT["SOL"],
                  cfg_outStride=S, cfg_scaleC=0.75, run_izp=[5, 30], run_ozp=[0, 20], run_offsetSolution=3)
    got = O.jw_f64(p, buf, x)
    for b in range(B):
        base = b * S + 3
        want = x[base:base + N].copy()
        want[:5] = 0
        want[30:] = 0
        want[0] *= 0.75
        np.testing.assert_array_equal(got[base:base + 20], want[:20])
        np.testing.assert_array_equal(got[base + 20:base + N], x[base + 20:base + N])
    # baked scaleC = 0 zeroes row 0 (zero-initialised configuration)
    p0 = O.problem(jw_type=13, upperBanded=2, N=N, B=1, cfg_outStride=N, cfg_scaleC=0.0)
    assert O.jw_f64(p0, buf, x[:N])[0] == 0.0


def test_matvec_only_mode_skips_scalec():
    N = 20
    buf = O.uniform(1, 0.5, 1.5, 3 * N)
    x = O.uniform(2, -1, 1, N)
    p = O.problem(jw_type=11, upperBanded=1, N=N, B=1, cfg_outStride=N, cfg_scaleC=0.5)
    got = O.jw_f64(p, buf, x)
    want = buf[0:N] * x + np.r_[buf[N:2 * N - 1] * x[1:], 0]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


# ---------------------------------------------------------------- fused step
@pytest.mark.parametrize("N,permuted", [(1, 0), (2, 1), (5, 1), (64, 1), (65, 1), (512, 1), (200, 0)])
def test_fused_stored_vs_raw(N, permuted):
    raw = O.raw_matrices(0, N)
    offV, offT, offM = 0, 5 * N, 8 * N
    buf = O.pack_fused(N, raw, permuted, offV, offT, offM, 10 * N)
    x = O.uniform(123, -1, 1, 2 * N)
    jt = 14 if permuted else 1014
    p = O.problem(jw_type=jt, upperBanded=0, N=N, B=2, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    got64 = O.jw_f64(p, buf, x)
    gotq = O.jw_f128(p, buf, x)
    for b in range(2):
        want = np.zeros(N)
        lib().pfo_fused_raw_f128(N, 2, 2, *[O.P(a) for a in raw], O.P(np.ascontiguousarray(x[b * N:(b + 1) * N])),
                                 O.P(want))
        assert O.relerr_norm(gotq[b * N:(b + 1) * N], want) < 1e-14
        assert O.relerr_norm(got64[b * N:(b + 1) * N], want) < 1e-13


def test_fused_dense_bruteforce_tiny():
    N = 9
    raw = O.raw_matrices(7, N)
    V, Ta, Tb, Tc, Md, Me = raw
    Vd = np.zeros((N, N))
    for j in range(5):
        for col in range(N):
            r = col - 2 + j
            if 0 <= r < N:
                Vd[r, col] = V[j * N + col]
    T = np.diag(Tb) + np.diag(Ta[1:], -1) + np.diag(Tc[:-1], 1)
    M = np.diag(Md) + np.diag(Me[:-1], 1)
    x = O.uniform(1, -1, 1, N)
    want = np.linalg.solve(M, np.linalg.solve(T, Vd @ x))
    y = np.zeros(N)
    lib().pfo_fused_raw_f128(N, 2, 2, *[O.P(a) for a in raw], O.P(x), O.P(y))
    assert O.relerr_norm(y, want) < 1e-13


# ---------------------------------------------------------------- double-double
def _dd(h, l=0.0):
    return O.DD(h, l)


def test_error_free_transforms_are_exact():
    from fractions import Fraction as F
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a, b = rng.standard_normal(2) * 10.0 ** rng.integers(-20, 20, 2)
        s = lib().pfo_two_sum(a, b)
        assert F(s.hi) + F(s.lo) == F(a) + F(b)
        p = lib().pfo_two_prod(a, b)
        assert F(p.hi) + F(p.lo) == F(a) * F(b)


def test_dd_ops_against_float128():
    rng = np.random.default_rng(1)
    u2 = 2.0 ** -106
    worst = {"add": 0.0, "mul": 0.0, "div": 0.0}
    for _ in range(5000):
        ah, bh = rng.uniform(-2, 2, 2)
        a = lib().pfo_two_sum(ah, rng.uniform(-1, 1) * abs(ah) * 2 ** -55)
        b = lib().pfo_two_sum(bh, rng.uniform(-1, 1) * abs(bh) * 2 ** -55)
        if abs(b.hi) < 1e-3:
            continue
        worst["mul"] = max(worst["mul"], lib().pfo_dd_relerr(lib().pfo_dd_mul(a, b), lib().pfo_q_mul(a, b)))
        worst["div"] = max(worst["div"], lib().pfo_dd_relerr(lib().pfo_dd_div(a, b), lib().pfo_q_div(a, b)))
        if ah * bh > 0:  # no cancellation: relative bound holds
            worst["add"] = max(worst["add"], lib().pfo_dd_relerr(lib().pfo_dd_add(a, b), lib().pfo_q_add(a, b)))
    # Joldes-Muller-Popescu bounds: add 3u^2, mul 4u^2, div 15u^2 (u = 2^-53)
    assert worst["add"] <= 3 * u2 * 1.01
    assert worst["mul"] <= 4 * u2 * 1.01
    assert worst["div"] <= 15 * u2 * 1.01


def test_dd_jw_step_against_float128():
    N, B = 256, 4
    raw = O.raw_matrices(3, N)
    # DD inputs: FP64 draws promoted with a low word of U(-1,1) * 2^-60 * |hi|
    rawdd = []
    for k, a in enumerate(raw):
        lo = O.uniform(100 + k, -1, 1, a.size) * 2.0 ** -60 * np.abs(a)
        rawdd.append(np.ascontiguousarray(np.stack([a, lo], -1)))
    offV, offT, offM = 0, 5 * N, 8 * N
    buf = np.zeros((10 * N, 2))
    lib().pfo_pack_fused_dd(N, 2, 2, *[O.PDD(a) for a in rawdd], 1, 0, offV, offT, offM, O.PDD(buf))
    xh = O.uniform(7, -1, 1, B * N)
    x = np.ascontiguousarray(np.stack([xh, O.uniform(8, -1, 1, B * N) * 2.0 ** -60 * np.abs(xh)], -1))
    p = O.problem(jw_type=14, upperBanded=0, N=N, B=B, KL=2, KU=2, cfg_outStride=N, cfg_scaleC=1.0,
                  cfg_offsetV=offV, cfg_offsetT=offT, cfg_offsetM=offM)
    got = O.jw_dd(p, buf, x)
    want = O.jw_f128_dd(p, buf, x)
    assert O.dd_relerr_norm(got, want) < 1e-30


# ---------------------------------------------------------------- block operators
def _block(block, N, B, inbuf, outbuf, lshift=0, scale=1.0, rt_scale=1, izp=(0, 0), ozp=(0, 0),
           inStride=None, outStride=None):
    bp = O.BlockProblem(block=block, lshift=lshift, N=N, B=B, mask=0, offsetM=0, offsetSolution=0,
                        izp_runtime=0, ozp_runtime=0, scaleC=scale, scaleC_runtime=rt_scale,
                        inStride=inStride or N, outStride=outStride or N)
    bp.izp[0], bp.izp[1] = izp
    bp.ozp[0], bp.ozp[1] = ozp
    out = outbuf.copy()
    assert lib().pfo_block_f64(C.byref(bp), O.P(inbuf), O.P(out)) == 0
    return out


@pytest.mark.parametrize("op", [1, 2, 3, 4])
def test_block_scale_factor_closed_forms(op):
    import decimal
    for n in [0, 1, 5, 100, 1000]:
        for l in [0, 3, 17]:
            for s in [0.0, 0.75, 2.5]:
                f = lib().pfo_block_scale_factor(op, n, l, s)
                decimal.getcontext().prec = 40
                D = decimal.Decimal
                x, L, S = D(n), D(l), D(s)
                if op == 1:
                    want = S
                elif op == 2:
                    want = 2 * (x + L + 1) * ((x + 1) / (x + S + L + D("1.5"))).sqrt()
                elif op == 3:
                    want = 4 * (x + L + 2) * (x + L + 3) * ((x + 1) * (x + 2) / ((x + S + L + D("2.5")) * (x + S + L + D("3.5")))).sqrt()
                else:
                    want = (x + L + 1) * (4 * L + 6) * ((x + 1) / (x + S + L + D("1.5"))).sqrt()
                assert abs(D(f) - want) <= abs(want) * D(4e-16), (op, n, l, s)


def test_block_layout_round_trip():
    N, B = 33, 5
    z = O.uniform(1, -1, 1, 2 * N * B)            # complex packed (XIXI)
    xxii = _block(4, N, B, z, np.zeros(2 * N * B))
    # XXII: system y holds re at [2yN, 2yN+N) and im at [2yN+N, 2yN+2N)
    zc = z.reshape(B, N, 2)
    np.testing.assert_array_equal(xxii.reshape(B, 2, N)[:, 0], zc[:, :, 0])
    np.testing.assert_array_equal(xxii.reshape(B, 2, N)[:, 1], zc[:, :, 1])
    back = _block(7, N, B, xxii, np.zeros(2 * N * B))
    np.testing.assert_array_equal(back, z)
    re = _block(2, N, B, z, np.zeros(N * B))
    im = _block(3, N, B, z, np.zeros(N * B))
    np.testing.assert_array_equal(re, zc[:, :, 0].ravel())
    np.testing.assert_array_equal(im, zc[:, :, 1].ravel())
    r5 = _block(5, N, B, re, np.zeros(2 * N * B)).reshape(B, N, 2)
    np.testing.assert_array_equal(r5[..., 0], zc[:, :, 0])
    assert not r5[..., 1].any()
    acc = _block(1 + 20, N, B, re, re)
    np.testing.assert_array_equal(acc, 2 * re)


def test_block_scalec_zero_baked_writes_zeros_without_reading():
    N, B = 8, 2
    inb = np.full(N * B, np.nan)
    out = _block(101, N, B, inb, np.ones(N * B), scale=0.0, rt_scale=0)
    assert not out.any()


# ---------------------------------------------------------------- reference fixture
def test_reference_tridiagonal_fixture_residual():
    path = os.path.join(os.path.dirname(__file__), "golden", "jw_tridiag_513.json")
    d = json.load(open(path))
    a, b, c, x = (np.array(d[k]) for k in ("a", "b", "c", "x"))
    N = len(b)
    # pre-normalise as the reference harness does (sample_1_benchmark_VkFFT_double.cpp:2332-2337)
    P = _perm(N, 13)
    buf = np.zeros(3 * N)
    buf[N + P] = 1.0 / b
    buf[P] = a * buf[N + P]
    buf[2 * N + P] = c * buf[N + P]
    p = O.problem(jw_type=13, upperBanded=2, N=N, B=1, cfg_outStride=N, cfg_scaleC=1.0)
    y = O.jw_f128(p, buf, x)
    # residual of the raw system T y = x (a[0] and c[N-1] are outside the matrix)
    aa = a.copy(); aa[0] = 0
    cc = c.copy(); cc[-1] = 0
    res = b * y + aa * np.r_[0, y[:-1]] + cc * np.r_[y[1:], 0] - x
    assert np.max(np.abs(res)) / np.max(np.abs(x)) < 1e-14
    y64 = O.jw_f64(p, buf, x)
    assert O.relerr_norm(y64, y) < 1e-14


# ---------------------------------------------------------------- dgbmv against BLAS
@pytest.mark.parametrize("KL,KU", [(2, 2), (1, 0), (0, 3), (3, 1), (0, 0), (4, 2)])
@pytest.mark.parametrize("N", [1, 2, 5, 17, 64])
def test_dgbmv_matches_blas_dgbmv(KL, KU, N):
    """The band matvec of the dgbmv plan (PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel1/
    pfSolve_MatVecMul.h:383-446: y_r = sum_j AB[j*N + c] x_c, c = r + KU - j) is LAPACK's
    dgbmv on a band array stored row by row (AB[j*N + c] = A(c - KU + j, c)): checked against
    the BLAS routine itself (scipy.linalg.blas.dgbmv), with the plan's keep-ranges, offsets and
    strides applied around it.  A band read in the wrong direction (r - KL + j) or a swapped
    KL/KU fails for every KL != KU case."""
    from scipy.linalg.blas import dgbmv
    LDA = KL + KU + 1
    B, S, offS, offM = 3, N + 5, 2, 7
    buf = O.uniform(40 + N, -1, 1, offM + LDA * N + 3)
    ab = buf[offM:offM + LDA * N].reshape(LDA, N)
    x = O.uniform(41 + N, -1, 1, B * S + offS + 4)
    izp, ozp = (N // 3, N), (1 if N > 1 else 0, N - N // 4)
    p = O.problem(jw_type=10, N=N, B=B, mask=RT["M"] | RT["SOL"] | RT["OSTR"] | RT["IZP"] | RT["OZP"],
                  cfg_outStride=N, run_offsetM=offM, run_offsetSolution=offS, run_outStride=S,
                  run_izp=list(izp), run_ozp=list(ozp))
    for fn, tol in (("pfo_dgbmv_f64", 4e-16), ("pfo_dgbmv_f128", 4e-16)):
        out = x.copy()
        assert getattr(lib(), fn)(C.byref(p), LDA, KU, KL, O.P(buf), O.P(out)) == 0
        for b in range(B):
            base = b * S + offS
            xin = x[base:base + N].copy()
            xin[:izp[0]] = 0
            xin[izp[1]:] = 0
            # scipy's wrapper wants at least LDA rows; below that the dense definition stands in
            want = dgbmv(N, N, KL, KU, 1.0, ab, xin) if N >= LDA else _band_dense(ab.ravel(), N, KL, KU) @ xin
            np.testing.assert_allclose(want, _band_dense(ab.ravel(), N, KL, KU) @ xin, rtol=0, atol=1e-15 * LDA)
            keep = np.zeros(N, bool)
            keep[ozp[0]:ozp[1]] = True
            got = out[base:base + N]
            np.testing.assert_allclose(got[keep], want[keep], rtol=0, atol=tol * LDA)
            np.testing.assert_array_equal(got[~keep], x[base:base + N][~keep])
        mask = np.ones(x.size, bool)
        for b in range(B):
            mask[b * S + offS:b * S + offS + N] = False
        np.testing.assert_array_equal(out[mask], x[mask])


def _band_dense(ab, N, KL, KU):
    """Dense A from LAPACK's band definition A(i, c) = AB(KU + i - c, c) for
    max(0, c - KU) <= i <= min(N - 1, c + KL)."""
    A = np.zeros((N, N))
    for c in range(N):
        for i in range(max(0, c - KU), min(N - 1, c + KL) + 1):
            A[i, c] = ab[(KU + i - c) * N + c]
    return A


def test_dgbmv_dense_band_definition():
    """Same band, dense brute force from LAPACK's definition, against the __float128 oracle."""
    N, KL, KU = 9, 3, 1
    LDA = KL + KU + 1
    ab = O.uniform(11, -1, 1, LDA * N)
    A = _band_dense(ab, N, KL, KU)
    x = O.uniform(12, -1, 1, N)
    p = O.problem(jw_type=10, N=N, B=1, cfg_outStride=N)
    out = x.copy()
    assert lib().pfo_dgbmv_f128(C.byref(p), LDA, KU, KL, O.P(ab), O.P(out)) == 0
    np.testing.assert_allclose(out, A @ x, rtol=0, atol=1e-15)
    assert np.count_nonzero(A) == sum(min(N - 1, c + KL) - max(0, c - KU) + 1 for c in range(N))


def test_block_layout_6_real_to_imaginary():
    """Layout 6 reads a real system into the imaginary part of a complex element (re = 0)
    (pfSolve_ReadWrite.h:2342-2344 read side, :2330-2331 write side): layout 5 of the real
    parts plus accumulated layout 6 of the imaginary parts rebuilds the complex data, and
    layouts 2 / 3 take it apart again.  A swapped re/im fails both checks."""
    N, B = 21, 4
    z = O.uniform(3, -1, 1, 2 * N * B)
    zc = z.reshape(B, N, 2)
    re = np.ascontiguousarray(zc[:, :, 0].ravel())
    im = np.ascontiguousarray(zc[:, :, 1].ravel())
    only_im = _block(6, N, B, im, np.full(2 * N * B, 9.0)).reshape(B, N, 2)
    assert not only_im[..., 0].any()
    np.testing.assert_array_equal(only_im[..., 1], zc[:, :, 1])
    both = _block(20 + 6, N, B, im, _block(5, N, B, re, np.zeros(2 * N * B)))
    np.testing.assert_array_equal(both, z)
    np.testing.assert_array_equal(_block(3, N, B, both, np.zeros(N * B)), im)
    # with the ScaleC operator on the way in: the imaginary part is scaled, re stays 0
    sc = _block(106, N, B, im, np.zeros(2 * N * B), scale=0.75).reshape(B, N, 2)
    np.testing.assert_array_equal(sc[..., 1], 0.75 * zc[:, :, 1])
    assert not sc[..., 0].any()


# ---------------------------------------------------------------- the reference's own systems
@pytest.mark.parametrize("jw,ub", [(0, 1), (0, 0), (2, 1), (2, 0)])
@pytest.mark.parametrize("N,l", [(64, 128), (200, 400), (200, 1), (65, 3)])
@pytest.mark.parametrize("rt", [0, 1])
def test_connection_matrix_modes_against_dense(jw, ub, N, l, rt):
    """Jones-Worland connection matrices (mu/nu closed forms of the reference harness) in the
    matvec + solve (jw 10) and solve-only (jw 12) modes: the oracle's __float128 step against a
    dense LAPACK solve of the same operator, and its FP64 step against both."""
    import jw_systems as J
    offM, offV = (11, 3 * N + 5) if rt else (0, 0)
    buf, A, S = J.mode_buffer(jw, ub, N, bool(rt), offM, offV, l, natural=False)
    x = O.uniform(9 + N, -1, 1, N)
    p = O.problem(jw_type=10 + jw, upperBanded=ub, N=N, B=1, mask=(RT["M"] | RT["V"]) if rt else 0,
                  cfg_outStride=N, cfg_scaleC=1.0, run_offsetM=offM, run_offsetV=offV)
    want = np.linalg.solve(S, A @ x)
    assert O.relerr_norm(O.jw_f128(p, buf, x), want) < 1e-14
    assert O.relerr_norm(O.jw_f64(p, buf, x), want) < 1e-14


@pytest.mark.parametrize("N,l", [(64, 128), (300, 1), (513, 8)])
def test_gram_tridiagonal_oracle_residual(N, l):
    """T = C^T C of a connection matrix (critical pivot recurrence for small l): the oracle's
    __float128 step solves the raw system to a residual at FP64 rounding of the stored
    coefficients, and its FP64 step agrees with it to the condition number's allowance."""
    import jw_systems as J
    a, b, c = J.gram_tridiagonal(N, l)
    buf = J.tridiag_buffer(a, b, c, natural=False)
    x = O.uniform(3 + N, -1, 1, N)
    p = O.problem(jw_type=13, upperBanded=2, N=N, B=1, cfg_outStride=N, cfg_scaleC=1.0)
    yq = O.jw_f128(p, buf, x)
    assert J.tridiag_residual(a, b, c, yq, x) < 1e-14 * max(1.0, np.max(np.abs(yq)) * np.max(np.abs(b)))
    cond = np.linalg.cond(np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1), np.inf)
    assert O.relerr_norm(O.jw_f64(p, buf, x), yq) < 4 * cond * 2.0 ** -53


# ---------------------------------------------------------------- chained iterations
@pytest.mark.parametrize("ub", [1, 0])
@pytest.mark.parametrize("jw", [0, 1, 2])
@pytest.mark.parametrize("N,K", [(9, 1), (9, 2), (40, 4), (64, 8), (100, 3), (200, 5)])
@pytest.mark.parametrize("rt", [0, 1])
def test_chain_against_dense_composition(ub, jw, N, K, rt):
    """The oracle's chained step (numConsecutiveJWIterations = K) against the composition of
    dense rectangular operators solved by LAPACK: iteration i maps N_{i-1} values to
    N_i = N -/+ i, its matvec diagonal is 1 except in the first (shrinking) or last (growing)
    iteration, the solve's row-0 scaleC applies after every solve; input keep-range on the
    first read, output keep-range on the last write; junk everywhere the layout leaves free.
    K = 1 is also checked against the single-step oracle."""
    import jw_chain as JC
    B, S = 3, N + K + 4
    bufs, dense, Ns = JC.build(N, K, jw, ub, rt, seed=10 * N + K + jw)
    x = O.uniform(N + 7 * K, -1, 1, B * S + 3)
    izp, ozp = (1, N), (0, max(1, Ns[-1] - 1))
    mask = RT["SOL"] | RT["OSTR"] | RT["IZP"] | RT["OZP"] | RT["SC"] | ((RT["M"] | RT["V"]) if rt else 0)
    p = O.problem(jw_type=10 + jw, upperBanded=ub, N=N, B=B, mask=mask, cfg_outStride=N, cfg_scaleC=1.0,
                  run_offsetM=7, run_offsetV=0, run_offsetSolution=2, run_outStride=S,
                  run_izp=list(izp), run_ozp=list(ozp), run_scaleC=0.8)
    if rt:   # runtime offsetV is one value for every iteration: the solve block sits behind the rest
        bufs, dense, Ns = JC.build(N, K, jw, ub, rt, seed=10 * N + K + jw, offV=7 + 2 * (N + K))
        p.run_offsetV = 7 + 2 * (N + K)
    got = O.jw_chain(p, K, bufs, x)
    got64 = O.jw_chain(p, K, bufs, x, f128=False)
    for b in range(B):
        base = b * S + 2
        x0 = x[base:base + N].copy()
        x0[:izp[0]] = 0
        want = JC.apply_dense(dense, Ns, x0, 0.8)
        keep = np.zeros(Ns[-1], bool)
        keep[ozp[0]:ozp[1]] = True
        g = got[base:base + Ns[-1]]
        assert O.relerr_norm(g[keep], want[keep]) < 1e-14
        assert O.relerr_norm(got64[base:base + Ns[-1]][keep], want[keep]) < 1e-13
        np.testing.assert_array_equal(g[~keep], x[base:base + Ns[-1]][~keep])
    mask_out = np.ones(x.size, bool)
    for b in range(B):
        mask_out[b * S + 2 + ozp[0]:b * S + 2 + ozp[1]] = False
    np.testing.assert_array_equal(got[mask_out], x[mask_out])
    if K == 1:
        single = O.jw_f128(p, bufs[0], x)
        np.testing.assert_array_equal(got, single)


# ---------------------------------------------------------------- ALT variant
def test_alt_layout_matches_reference_build():
    """reorder_i_alt and get_ALT_stride of the reference headers (compiled under oracle/_ref)"""
    if not os.path.exists(O.REF_SO):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ref = C.CDLL(O.REF_SO)
    ref.ref_reorder_i_alt.argtypes = [C.c_int] * 4
    ref.ref_get_ALT_stride.argtypes = [C.c_int] * 2
    for N in list(range(1, 140)) + [255, 256, 511, 512, 513, 777, 1023, 1024, 1025, 2048, 3000]:
        assert O.alt_stride(N) == ref.ref_get_ALT_stride(N, 0), N
        pos = [O.alt_pos(i, N) for i in range(N)]
        assert pos == [ref.ref_reorder_i_alt(i, N, N, 0) for i in range(N)], N
        assert len(set(pos)) == N and max(pos) < O.alt_stride(N)


@pytest.mark.parametrize("jw,ub", [(0, 1), (0, 0), (1, 1), (2, 0), (3, 2), (3, 1), (3, 0)])
@pytest.mark.parametrize("N", [1, 2, 3, 8, 63, 64, 130, 257])
def test_alt_modes_against_dense(jw, ub, N):
    """ALT (jw_type 110-113): rows couple with rows +-2; baked positions are multiples of the
    ALT block stride, entries in reorder_i_alt order; row-0 scaleC on row 0 only.  Checked
    against a dense LAPACK solve of the stride-2 operator built from the same values."""
    S = O.alt_stride(N)
    P = np.array([O.alt_pos(i, N) for i in range(N)], dtype=np.int64)
    rng = np.random.default_rng(N + 10 * jw + ub)
    buf = rng.uniform(-3, 3, 3 * S + 4)                   # junk outside the matrix
    A = np.eye(N)
    Sm = np.eye(N)
    if jw in (0, 1):
        d, e = rng.uniform(0.5, 1.5, N), rng.uniform(-0.5, 0.5, N)
        if ub:   # [D @0 | U @S]
            buf[P], buf[S + P] = d, e
            A = np.diag(d) + np.diag(e[:-2], 2) if N > 2 else np.diag(d)
        else:    # [L @0 | D @S]
            buf[S + P], buf[P] = d, e
            A = np.diag(d) + (np.diag(e[2:], -2) if N > 2 else 0)
    if jw in (0, 2):
        s = rng.uniform(-0.45, 0.45, N)
        buf[2 * S + P] = s
        if N > 2:
            Sm = np.eye(N) + (np.diag(s[2:], -2) if ub else np.diag(s[:-2], 2))
    if jw == 3:
        d = rng.uniform(0.5, 1.5, N)
        lo, up = rng.uniform(-0.3, 0.3, N), rng.uniform(-0.3, 0.3, N)
        A = np.diag(d)
        if ub == 2:   # [L @0 | D @S | U @2S]
            buf[P], buf[S + P], buf[2 * S + P] = lo, d, up
            if N > 2:
                Sm = np.eye(N) + np.diag(lo[2:], -2) + np.diag(up[:-2], 2)
        elif ub == 1:   # [D @S | L @0]
            buf[S + P], buf[P] = d, lo
            if N > 2:
                Sm = np.eye(N) + np.diag(lo[2:], -2)
        else:           # [U @S | D @0]
            buf[S + P], buf[P] = up, d
            if N > 2:
                Sm = np.eye(N) + np.diag(up[:-2], 2)
    x = O.uniform(N + 3, -1, 1, N)
    p = O.problem(jw_type=110 + jw, upperBanded=ub, N=N, B=1, cfg_outStride=N, cfg_scaleC=0.5)
    want = np.linalg.solve(Sm, A @ x)
    if jw != 1:
        want[0] *= 0.5
    assert O.relerr_norm(O.jw_f128(p, buf, x), want) < 1e-14
    assert O.relerr_norm(O.jw_f64(p, buf, x), want) < 1e-14
