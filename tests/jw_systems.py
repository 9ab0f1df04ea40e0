"""Matrices of the systems the reference itself solves, in the storage its JW plans read.

* The reference's embedded tridiagonal system (a0/b0/c0/x0, 513 rows,
  benchmark_scripts/vkFFT_scripts/src/sample_1_benchmark_VkFFT_double.cpp:49-2106), stored
  pre-normalised as that harness stores it (:2332-2337: md = 1/b, ld = a * md, ud = c * md).
* Jones-Worland connection matrices from the harness's closed forms
  (sample_1_benchmark_VkFFT_double.cpp:2109-2114, the same mu/nu in
  sample_1_benchmark_VkFFT_double_JW.cpp:45-50), normalised per row as
  sample_1_benchmark_VkFFT_double_JW.cpp:161-180 does: the upper bidiagonal connection matrix
  C has C[r, r] = mu(r) and C[r, r+1] = nu(r+1) (its back substitution is the harness's CPU
  check, :236-243); stored, the diagonal is 1/mu(r) and the off-diagonal nu/mu.  The harness uses
  alpha = -1/2, beta = l + 1/2 with l = 2N; small l makes the off-diagonal ratio approach 1.

Test infrastructure: the positions of each mode's diagonals follow the reference's kernel
(PfSolve/pfSolve_CodeGen/pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:117-348).
"""
from __future__ import annotations

import json
import os

import numpy as np

import _oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture_513():
    d = json.load(open(os.path.join(HERE, "golden", "jw_tridiag_513.json")))
    return tuple(np.array(d[k], dtype=np.float64) for k in ("a", "b", "c", "x"))


def perm(N: int, natural: bool) -> np.ndarray:
    if natural or N <= 64:
        return np.arange(N)
    return np.array([O.lib().pfo_reorder_i0(i, N, N, 0) for i in range(N)], dtype=np.int64)


def tridiag_buffer(a, b, c, natural: bool) -> np.ndarray:
    """jw 13 / upperBanded 2 storage of the tridiagonal (a sub, b diag, c super):
    [ld | md | ud] at [0, N, 2N), each permuted by reorder_i0 unless natural."""
    N = b.size
    P = perm(N, natural)
    buf = np.zeros(3 * N)
    md = 1.0 / b
    buf[N + P] = md
    buf[P] = a * md
    buf[2 * N + P] = c * md
    return buf


def tridiag_residual(a, b, c, y, x) -> float:
    """max |T y - x| / max |x| for the raw tridiagonal (a[0], c[N-1] outside the matrix)."""
    r = b * y - x
    r[1:] += a[1:] * y[:-1]
    r[:-1] += c[:-1] * y[1:]
    return float(np.max(np.abs(r)) / np.max(np.abs(x)))


def mu(n, alpha, beta):
    n = np.asarray(n, dtype=np.float64)
    return np.sqrt(2 * (n + beta) * (n + alpha + beta) / (2 * n + alpha + beta) / (2 * n + alpha + beta + 1))


def nu(n, alpha, beta):
    n = np.asarray(n, dtype=np.float64)
    return np.sqrt(2 * (n + 1) * (n + alpha + 1) / (2 * n + alpha + beta + 1) / (2 * n + alpha + beta + 2))


def connection(N: int, l: float, shift: float = 0.5):
    """Upper bidiagonal connection matrix of degree l: (diag mu(r), super nu(r+1)) with
    alpha = -1/2, beta = l + shift; super[N-1] is outside the matrix (0)."""
    r = np.arange(N, dtype=np.float64)
    d = mu(r, -0.5, l + shift)
    e = np.zeros(N)
    e[:-1] = nu(r[1:], -0.5, l + shift)
    return d, e


def mode_buffer(jw: int, ub: int, N: int, rt: bool, offM: int, offV: int, l: float, natural: bool,
                junk_seed: int = 5):
    """Storage of the reference mode jw (0 = matvec + solve, 1 = matvec, 2 = solve) with
    upperBanded ub for connection matrices of degree l.  Matvec A = C_A (ub 1: upper, diag at
    D_A and super at U_A) or C_A^T (ub 0: lower); solve S = rows of C_S (ub 0) or C_S^T (ub 1)
    divided by their diagonal, C_A of beta = l + 3/2, C_S of beta = l + 1/2.  Positions
    (pfSolve_JonesWorland.h:145-198, 281-326): baked upper [D_A @0, U_A @N], lower
    [L_A @0, D_A @N], solve @2N; runtime upper [U_A @offM, D_A @offM+N], lower
    [D_A @offM, L_A @offM+N], solve @offV.  Entries outside the matrix hold junk (the kernels
    must not read them).  Returns (buffer, dense A, dense S) with A, S as the solve sees them."""
    P = perm(N, natural)
    total = max(3 * N, offM + 2 * N, offV + N) + 8
    buf = O.uniform(junk_seed, -3, 3, total)
    A = np.eye(N)
    S = np.eye(N)
    if jw in (0, 1):
        d, e = connection(N, l, 1.5)
        if ub:
            ud, md = (offM, offM + N) if rt else (N, 0)
            buf[md + P] = d
            buf[ud + P[:-1]] = e[:-1]
            A = np.diag(d) + np.diag(e[:-1], 1)
        else:
            md, ld = (offM, offM + N) if rt else (N, 0)
            buf[md + P] = d
            buf[ld + P[1:]] = e[:-1]          # lower: L[r, r-1] = nu(r)
            A = np.diag(d) + np.diag(e[:-1], -1)
    if jw in (0, 2):
        d, e = connection(N, l, 0.5)
        inv = 1.0 / d
        base = offV if rt else 2 * N
        if ub:
            lo = e[:-1] * inv[1:]             # row r of C^T: nu(r) y_{r-1} + mu(r) y_r
            buf[base + P[1:]] = lo
            S = np.eye(N) + np.diag(lo, -1)
        else:
            up = e[:-1] * inv[:-1]            # row r of C: mu(r) y_r + nu(r+1) y_{r+1}
            buf[base + P[:-1]] = up
            S = np.eye(N) + np.diag(up, 1)
    return buf, A, S


def gram_tridiagonal(N: int, l: float):
    """T = C^T C of the connection matrix C of degree l: symmetric positive definite and
    tridiagonal; for small l its off-diagonal / diagonal ratio reaches 0.5 (the pivot
    recurrence sits at its critical point, condition number ~ N^2).  Returns (a, b, c)."""
    d, e = connection(N, l, 0.5)
    b = d * d
    b[1:] += e[:-1] ** 2
    off = d[:-1] * e[:-1]
    a = np.zeros(N)
    c = np.zeros(N)
    a[1:] = off
    c[:-1] = off
    return a, b, c
