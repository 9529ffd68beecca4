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


U = 2.0 ** -53
TARGET = 2.0 ** -55


def digit_depth(X: np.ndarray, e: np.ndarray, axis: int) -> int:
    """Most base-128 digits any value of X (scaled by 2^-e along `axis`)
    needs; 10 when some value needs more than nine."""
    shape = [1, 1]
    shape[1 - axis] = -1
    r = np.abs(np.ldexp(X, -e.reshape(shape)))
    depth = np.zeros(r.shape, dtype=np.int64)
    for q in range(1, 10):
        r = r * 128.0
        t = np.trunc(r)
        r = r - t
        depth = np.where(t != 0, q, depth)
    depth = np.where(r != 0, 10, depth)
    return int(depth.max()) if depth.size else 0


def _bound(wa, wb, x, y, a2, b2, s):
    q = (wa[:, None] * wb[None, :] * (x[:, None] + y[None, :]) ** 2).sum()
    return float(2.0 ** (-7 * s) * np.sqrt(q / ((wa * a2).sum() * (wb * b2).sum())))


def error_bound(A: np.ndarray, B: np.ndarray, s: int) -> float:
    """The library's a-priori normwise bound on ||op(A)op(B) - Ozaki result||_F
    / (||op(A)||_F ||op(B)||_F) with s slices (truncation, dropped products and
    the folds before the last one), evaluated here in float64:

        |dP_ij| <= 2^(e_i+f_j) 2^-7s (x_i + y_j),
        x_i = S_i(s)/128 + c_s a1_i,  y_j = (1 + c_s) b1_j,  c_s = 2^(7s-53)/127,

    S_i(s) = sum over row i and digits p <= s of |digit|, a1_i = sum |A'_ik|,
    b1_j = sum |B'_kj|; with exactly representable operands (Da + Db - 1 <= s
    digits) x_i = c_s a1_i and y_j = c_s b1_j."""
    ea = exponents(A, axis=1)
    fb = exponents(B, axis=0)
    Ap = np.ldexp(np.abs(A), -ea[:, None])
    Bp = np.ldexp(np.abs(B), -fb[None, :])
    a1, b1 = Ap.sum(axis=1), Bp.sum(axis=0)
    if not np.isfinite(A).all() or not np.isfinite(B).all():
        return float("inf")
    if (a1 == 0).all() or (b1 == 0).all():
        return 0.0
    S = sum(np.abs(d) for d in slices(A, ea, 1, s)).sum(axis=1).astype(np.float64)
    c = 2.0 ** (7 * s - 53) / 127
    exact = digit_depth(A, ea, 1) + digit_depth(B, fb, 0) - 1 <= s
    x = c * a1 + (0.0 if exact else S / 128)
    y = (c if exact else 1 + c) * b1
    wa = np.where(a1 > 0, np.exp2(2.0 * (ea - ea[a1 > 0].max())), 0.0)
    wb = np.where(b1 > 0, np.exp2(2.0 * (fb - fb[b1 > 0].max())), 0.0)
    return _bound(wa, wb, x, y, (Ap ** 2).sum(axis=1), (Bp ** 2).sum(axis=0), s)


def error_bound_fused(A: np.ndarray, B: np.ndarray, s: int) -> float:
    """The same bound as the fused one-wave kernel (m, n, k <= 32) evaluates
    it: the row and column sums come from the leading digits (d_p the p-th
    digit magnitude of a value, each row or column held as two halves of 16
    values) --
        a1 = sum (d1 2^-7 + d2 2^-14 + d3 2^-21) + 32 * 2^-21  (>= sum |v|),
        a2 = sum (d1 2^-7 + d2 2^-14)^2                      (<= sum v^2) --
    and no exactness refinement."""
    if not np.isfinite(A).all() or not np.isfinite(B).all():
        return float("inf")
    ea = exponents(A, axis=1)
    fb = exponents(B, axis=0)
    da = [np.abs(d).astype(np.float64) for d in slices(A, ea, 1, max(s, 3))]
    db = [np.abs(d).astype(np.float64) for d in slices(B, fb, 0, max(s, 3))]

    def sums(d, axis):
        nz = d[0].sum(axis=axis) > 0
        l1 = (d[0] / 128 + d[1] / 2 ** 14 + d[2] / 2 ** 21).sum(axis=axis) + 32 * 2.0 ** -21
        l2 = ((d[0] / 128 + d[1] / 2 ** 14) ** 2).sum(axis=axis)
        return np.where(nz, l1, 0.0), l2, nz

    a1, a2, nza = sums(da, 1)
    b1, b2, nzb = sums(db, 0)
    if not nza.any() or not nzb.any():
        return 0.0
    S = sum(da[:s]).sum(axis=1)
    c = 2.0 ** (7 * s - 53) / 127
    x = S / 128 + c * a1
    y = (1 + c) * b1
    wa = np.where(nza, np.exp2(2.0 * (ea - ea[nza].max())), 0.0)
    wb = np.where(nzb, np.exp2(2.0 * (fb - fb[nzb].max())), 0.0)
    return _bound(wa, wb, x, y, a2, b2, s)


def auto_slices(A: np.ndarray, B: np.ndarray, smax: int = 9) -> int:
    """The automatic slice count: the smallest s (down from smax while the
    bound stays within TARGET), smax if the bound at smax misses it."""
    if error_bound(A, B, smax) > TARGET:
        return smax
    s = smax
    for t in range(smax - 1, 2, -1):
        if not error_bound(A, B, t) <= TARGET:
            break
        s = t
    return s
