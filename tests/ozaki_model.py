"""numpy model of the library's Ozaki fp64 form (test infrastructure).

Follows the algorithm documented in ssblasex_amd/csrc/ozaki.hip step by step
with exactly the same floating-point operations, so that for alpha a power
of two the library's result must match it bit for bit:
  e_i = frexp exponent of max_k |A(i,k)|; slices by r *= 128, a = trunc(r),
  r -= a; D_t = sum_{p+q=t} a_p b_q (exact int64 here); fold
  D = fl(D + D_t 2^(-7t)) for t = s+1 .. 2; C = fl(ldexp(D, e_i+f_j)*alpha
  + fl(beta*C0)).
"""
from __future__ import annotations

import numpy as np


def exponents(X: np.ndarray, axis: int) -> np.ndarray:
    mx = np.max(np.abs(X), axis=axis)
    _, e = np.frexp(mx)
    return np.where(mx == 0, 0, e).astype(np.int64)


def slices(X: np.ndarray, e: np.ndarray, axis: int, s: int) -> list[np.ndarray]:
    shape = [1, 1]
    shape[1 - axis] = -1
    r = np.ldexp(X, -e.reshape(shape))
    out = []
    for _ in range(s):
        r = r * 128.0
        t = np.trunc(r)
        r = r - t
        out.append(t.astype(np.int64))
    return out


def ozaki_gemm(A: np.ndarray, B: np.ndarray, s: int, alpha: float = 1.0, beta: float = 0.0,
               C0: np.ndarray | None = None) -> np.ndarray:
    """A: op(A) m x k, B: op(B) k x n (float64)."""
    ea = exponents(A, axis=1)  # per row of op(A)
    fb = exponents(B, axis=0)  # per column of op(B)
    As = slices(A, ea, 1, s)
    Bs = slices(B, fb, 0, s)
    D = np.zeros((A.shape[0], B.shape[1]), dtype=np.float64)
    for t in range(s + 1, 1, -1):
        Dt = np.zeros_like(D, dtype=np.int64)
        for p in range(max(1, t - s), min(s, t - 1) + 1):
            Dt += As[p - 1] @ Bs[t - p - 1]
        D = D + np.ldexp(Dt.astype(np.float64), -7 * t)
    val = np.ldexp(D, ea[:, None] + fb[None, :])
    t0 = np.zeros_like(val) if beta == 0 else C0 * beta
    return val * alpha + t0
