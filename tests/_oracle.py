"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; it is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpfsolve_ref_layout.so")

i64 = C.c_int64
dptr = C.POINTER(C.c_double)


class DD(C.Structure):
    _fields_ = [("hi", C.c_double), ("lo", C.c_double)]


ddptr = C.POINTER(DD)


class JwProblem(C.Structure):
    _fields_ = [
        ("jw_type", C.c_int), ("upperBanded", C.c_int), ("N", i64), ("B", i64), ("mask", i64),
        ("Msplit0", C.c_int), ("KL", C.c_int), ("KU", C.c_int),
        ("cfg_offsetM", i64), ("cfg_offsetV", i64), ("cfg_offsetSolution", i64), ("cfg_offsetT", i64),
        ("cfg_izp", i64 * 2), ("cfg_ozp", i64 * 2), ("cfg_scaleC", C.c_double), ("cfg_outStride", i64),
        ("run_offsetM", i64), ("run_offsetV", i64), ("run_offsetSolution", i64), ("run_offsetT", i64),
        ("run_izp", i64 * 2), ("run_ozp", i64 * 2), ("run_scaleC", C.c_double), ("run_outStride", i64),
    ]


class BlockProblem(C.Structure):
    _fields_ = [
        ("block", C.c_int), ("lshift", C.c_int), ("N", i64), ("B", i64), ("mask", i64),
        ("offsetM", i64), ("offsetSolution", i64), ("izp", i64 * 2), ("ozp", i64 * 2),
        ("izp_runtime", C.c_int), ("ozp_runtime", C.c_int), ("scaleC", C.c_double),
        ("scaleC_runtime", C.c_int), ("inStride", i64), ("outStride", i64),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = C.CDLL(ORACLE_SO)
        L.pfo_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.pfo_splitmix64.restype = C.c_uint64
        L.pfo_fill_uniform.argtypes = [C.c_uint64, C.c_double, C.c_double, dptr, i64]
        L.pfo_reorder_i0.argtypes = [i64, i64, i64, i64]
        L.pfo_reorder_i0.restype = i64
        L.pfo_jw_registers_per_thread.argtypes = [i64, C.c_int, C.c_int]
        L.pfo_jw_uses_parallel_thomas.argtypes = [i64, C.c_int, C.c_int]
        L.pfo_zeropad_keep.argtypes = [i64, i64, i64, C.c_int]
        for f in ("pfo_jw_step_f64",):
            getattr(L, f).argtypes = [C.POINTER(JwProblem), dptr, dptr]
        L.pfo_jw_step_f128.argtypes = [C.POINTER(JwProblem), dptr, dptr, dptr]
        L.pfo_jw_step_dd.argtypes = [C.POINTER(JwProblem), ddptr, ddptr]
        L.pfo_jw_step_f128_dd.argtypes = [C.POINTER(JwProblem), ddptr, ddptr]
        L.pfo_jw_step_f64_range.argtypes = [C.POINTER(JwProblem), dptr, dptr, i64, i64, C.c_int]
        L.pfo_jw_step_dd_range.argtypes = [C.POINTER(JwProblem), ddptr, ddptr, i64, i64, C.c_int]
        L.pfo_num_threads.restype = C.c_int
        L.pfo_dgbmv_f64.argtypes = [C.POINTER(JwProblem), C.c_int, C.c_int, C.c_int, dptr, dptr]
        L.pfo_dgbmv_f128.argtypes = [C.POINTER(JwProblem), C.c_int, C.c_int, C.c_int, dptr, dptr]
        L.pfo_block_f64.argtypes = [C.POINTER(BlockProblem), dptr, dptr]
        L.pfo_block_scale_factor.argtypes = [C.c_int, i64, C.c_int, C.c_double]
        L.pfo_block_scale_factor.restype = C.c_double
        L.pfo_block_scale_factor_ld.argtypes = [C.c_int, i64, C.c_int, C.c_longdouble]
        L.pfo_block_scale_factor_ld.restype = C.c_longdouble
        for f in ("pfo_dd_add", "pfo_dd_sub", "pfo_dd_mul", "pfo_dd_div", "pfo_q_add", "pfo_q_mul", "pfo_q_div"):
            getattr(L, f).argtypes = [DD, DD]
            getattr(L, f).restype = DD
        for f in ("pfo_two_sum", "pfo_two_prod"):
            getattr(L, f).argtypes = [C.c_double, C.c_double]
            getattr(L, f).restype = DD
        L.pfo_dd_relerr.argtypes = [DD, DD]
        L.pfo_dd_relerr.restype = C.c_double
        L.pfo_gen_raw_matrices.argtypes = [C.c_uint64, i64, C.c_int, C.c_int] + [dptr] * 6
        L.pfo_pack_fused.argtypes = [i64, C.c_int, C.c_int] + [dptr] * 6 + [C.c_int, C.c_int, i64, i64, i64, dptr]
        L.pfo_pack_fused_dd.argtypes = [i64, C.c_int, C.c_int] + [ddptr] * 6 + [C.c_int, C.c_int, i64, i64, i64, ddptr]
        L.pfo_fused_raw_f128.argtypes = [i64, C.c_int, C.c_int] + [dptr] * 8
        L.pfo_dense_solve_f128.argtypes = [i64, dptr, dptr, dptr]
        _lib = L
    return _lib


def P(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dptr)


def PDD(a: np.ndarray):
    """a: float64 array of shape (..., 2) holding (hi, lo) pairs."""
    assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape[-1] == 2
    return a.ctypes.data_as(ddptr)


def uniform(seed: int, a: float, b: float, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().pfo_fill_uniform(seed, a, b, P(out), n)
    return out


def problem(**kw) -> JwProblem:
    p = JwProblem()
    p.Msplit0 = 0
    for k, v in kw.items():
        cur = getattr(p, k)
        if isinstance(cur, C.Array):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(p, k, v)
    return p


def jw_f64(p: JwProblem, buf: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    out = rhs.copy()
    rc = lib().pfo_jw_step_f64(C.byref(p), P(buf), P(out))
    assert rc == 0, rc
    return out


def jw_f128(p: JwProblem, buf: np.ndarray, rhs: np.ndarray, with_lo: bool = False):
    out = rhs.copy()
    lo = np.zeros_like(rhs)
    rc = lib().pfo_jw_step_f128(C.byref(p), P(buf), P(out), P(lo))
    assert rc == 0, rc
    return (out, lo) if with_lo else out


def jw_dd(p: JwProblem, buf: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    out = rhs.copy()
    rc = lib().pfo_jw_step_dd(C.byref(p), PDD(buf), PDD(out))
    assert rc == 0, rc
    return out


def jw_f128_dd(p: JwProblem, buf: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    out = rhs.copy()
    rc = lib().pfo_jw_step_f128_dd(C.byref(p), PDD(buf), PDD(out))
    assert rc == 0, rc
    return out


def raw_matrices(seed: int, N: int, KL: int = 2, KU: int = 2):
    LDA = KL + KU + 1
    V = np.zeros(LDA * N)
    Ta, Tb, Tc, Md, Me = (np.zeros(N) for _ in range(5))
    lib().pfo_gen_raw_matrices(seed, N, KL, KU, P(V), P(Ta), P(Tb), P(Tc), P(Md), P(Me))
    return V, Ta, Tb, Tc, Md, Me


def pack_fused(N, raw, permuted, offV, offT, offM, total, KL=2, KU=2):
    buf = np.zeros(total)
    V, Ta, Tb, Tc, Md, Me = raw
    lib().pfo_pack_fused(N, KL, KU, P(V), P(Ta), P(Tb), P(Tc), P(Md), P(Me), int(permuted), 0,
                         offV, offT, offM, P(buf))
    return buf


def dd_relerr_norm(a: np.ndarray, b: np.ndarray) -> float:
    """max_i |a_i - b_i| / max_i |b_i| for DD arrays, differences taken exactly in DD."""
    a = a.reshape(-1, 2)
    b = b.reshape(-1, 2)
    num = np.abs((a[:, 0] - b[:, 0]) + (a[:, 1] - b[:, 1]))
    den = np.max(np.abs(b[:, 0]))
    return float(np.max(num) / den) if den > 0 else float(np.max(num))


def relerr_norm(a: np.ndarray, b: np.ndarray) -> float:
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    return num / den if den > 0 else num


def _chain_fn(name):
    f = getattr(lib(), name)
    f.argtypes = [C.POINTER(JwProblem), C.c_int, C.POINTER(dptr), dptr]
    return f


def jw_chain(p: JwProblem, K: int, bufs, rhs: np.ndarray, f128: bool = True) -> np.ndarray:
    """chained step (numConsecutiveJWIterations = K) of the oracle: bufs[i] are the
    matrices of iteration i (one array K times without multiple input buffers)"""
    out = rhs.copy()
    arr = (dptr * K)(*[P(b) for b in bufs])
    rc = _chain_fn("pfo_jw_chain_f128" if f128 else "pfo_jw_chain_f64")(C.byref(p), K, arr, P(out))
    assert rc == 0, rc
    return out


def alt_pos(i, N, Nalign=0):
    L = lib()
    L.pfo_reorder_i_alt.argtypes = [i64, i64, i64, i64]
    L.pfo_reorder_i_alt.restype = i64
    return L.pfo_reorder_i_alt(i, N, Nalign or N, 0)


def alt_stride(N):
    L = lib()
    L.pfo_alt_stride.argtypes = [i64, i64]
    L.pfo_alt_stride.restype = i64
    return L.pfo_alt_stride(N, 0)
