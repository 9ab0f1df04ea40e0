"""Chained JW steps (numConsecutiveJWIterations = K) -- test helper: logical matrices of
each iteration placed in the reference's storage (PfSolve/pfSolve_CodeGen/
pfSolve_KernelsLevel2/pfSolve_JonesWorland.h:108-330: iteration i has N_i = N -/+ i rows,
baked positions change with i, runtime ones use N_i, the matvec diagonal is stored only for
the first (upperBanded 1) or last (upperBanded 0) iteration; rows stored in reorder_i0 order
of N_i rows with the register layout of the largest system), and the operator as dense
matrices for an independent evaluation."""
from __future__ import annotations

import numpy as np

import _oracle as O


def sizes(N, K, ub):
    return [N - i if ub else N + i for i in range(K)]


def perm(Ni, Nalign, natural):
    if natural:
        return np.arange(Ni)
    return np.array([O.lib().pfo_reorder_i0(r, Ni, Nalign, 0) for r in range(Ni)], dtype=np.int64)


def build(N, K, jw, ub, rt, seed, natural=False, offM=7, offV=None, junk=True):
    """Returns (bufs, dense) where bufs[i] is iteration i's matrix buffer and
    dense[i] = (A_i or None, S_i or None) with A_i of shape (N_i, N_{i-1})."""
    Ns = sizes(N, K, ub)
    Nalign = N if ub else N + K - 1
    bufs, dense = [], []
    rng = np.random.default_rng(seed)
    for i, Ni in enumerate(Ns):
        Nprev = N if i == 0 else Ns[i - 1]
        P = perm(Ni, Nalign, natural)
        ov = offV if offV is not None else 3 * Ni + 5
        total = max(3 * Ni, offM + 2 * Ni, ov + Ni) + 8
        buf = rng.uniform(-3, 3, total) if junk else np.zeros(total)
        diag_here = (i == 0) if ub else (i == K - 1)
        A = S = None
        if jw in (0, 1):
            d = rng.uniform(0.5, 1.5, Ni)
            e = rng.uniform(-0.5, 0.5, Ni)
            A = np.zeros((Ni, Nprev))
            for r in range(Ni):
                if r < Nprev:
                    A[r, r] = d[r] if diag_here else 1.0
                if ub and r + 1 < Nprev:
                    A[r, r + 1] = e[r]
                if not ub and r >= 1 and r - 1 < Nprev:
                    A[r, r - 1] = e[r]
            if ub:
                ud, md = (offM, offM + Ni) if rt else (Ni if i == 0 else 0, 0)
                buf[ud + P] = e
                if diag_here:
                    buf[md + P] = d
            else:
                md, ld = (offM, offM + Ni) if rt else (Ni, 0)
                buf[ld + P] = e
                if diag_here:
                    buf[md + P] = d
        if jw in (0, 2):
            s = rng.uniform(-0.45, 0.45, Ni)
            S = np.eye(Ni)
            if ub:
                base = ov if rt else (2 * Ni if i == 0 else Ni)
                S[np.arange(1, Ni), np.arange(Ni - 1)] = s[1:]
            else:
                base = ov if rt else (2 * Ni if i == K - 1 else Ni)
                S[np.arange(Ni - 1), np.arange(1, Ni)] = s[:-1]
            buf[base + P] = s
        bufs.append(np.ascontiguousarray(buf))
        dense.append((A, S))
    return bufs, dense, Ns


def apply_dense(dense, Ns, x0, scale):
    """y = prod_i S_i^{-1} A_i x with the solve's row-0 scaleC after every solve"""
    x = x0.copy()
    for (A, S), Ni in zip(dense, Ns):
        if A is not None:
            v = A @ x
        else:
            v = np.zeros(Ni)
            m = min(Ni, x.size)
            v[:m] = x[:m]
        if S is not None:
            v = np.linalg.solve(S, v)
            v[0] *= scale
        x = v
    return x
