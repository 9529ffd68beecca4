"""Pins the CPU oracle (oracle/ssblas_oracle.c) before anything trusts it.

Parity anchors (the reference ships no golden vectors or tests, so each
oracle function is pinned by what the mathematics fixes):

* oracle_gemm_batched_ex, int8 path -- bit-exact against numpy int64
  matmul reduced mod 2^32 (a library routine), on every op pair, with
  leading-dimension padding, wrap-around scalars, and brute force on tiny
  shapes; plus the hand-derived fixtures in tests/golden.
* oracle_gemm_batched_ex, fp paths -- (a) integer-valued inputs where all
  arithmetic is exact must equal exact integer arithmetic; (b) the
  hand-derived rounding fixtures pin the 512-wide K blocks, the separate
  beta pre-multiply and the fused K chain; (c) random data stays within the
  standard error bound of the double-double oracle.
* oracle_gemm_dd -- against exact rational arithmetic (fractions) on tiny
  random inputs.
* oracle_validate -- the status table, case by case, with the reference
  file:line each row comes from.
* oracle_naive_baseline_i8 -- equal to the int8 restatement.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

import oracle
from helpers import (Batch, TYPES, UNIT_ROUNDOFF, load_golden, make_batch, normwise, padding_untouched,
                     run_oracle)

OPS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def numpy_exact(b: Batch, i: int):
    """alpha*op(A)op(B) + beta*C0 in int64 / exact arithmetic (integer inputs)."""
    ra, ca = b.dims_a()
    rb, cb = b.dims_b()
    A = oracle.colmajor(b.A[i], ra, ca, b.lda).astype(np.int64)
    B = oracle.colmajor(b.B[i], rb, cb, b.ldb).astype(np.int64)
    A = A if b.transa == 0 else A.T
    B = B if b.transb == 0 else B.T
    C0 = oracle.colmajor(b.C0[i], b.m, b.n, b.ldc).astype(np.int64)
    S = A @ B
    if b.type == "i8":
        r = (np.int64(b.alpha) * S) & 0xFFFFFFFF
        if b.beta != 0:
            r = (r + ((np.int64(b.beta) * C0) & 0xFFFFFFFF)) & 0xFFFFFFFF
        return r.astype(np.uint32).view(np.int32)
    return b.alpha * S + (b.beta * C0 if b.beta != 0 else 0)


# --------------------------------------------------------------------------
# golden fixtures
# --------------------------------------------------------------------------
@pytest.mark.parametrize("case", load_golden(), ids=lambda c: c[0])
def test_golden_fixture(case):
    name, applies, b, expect = case
    assert "oracle" in applies
    st, C = run_oracle(b)
    assert st == 0
    for got, exp in zip(C, expect):
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (name, got, exp)


# --------------------------------------------------------------------------
# int8: bit exact vs numpy int64 mod 2^32
# --------------------------------------------------------------------------
@pytest.mark.parametrize("ta,tb", OPS)
@pytest.mark.parametrize("alpha,beta", [(1, 0), (1, 1), (-3, 7), (1 << 20, -(1 << 30))])
def test_int8_exact_vs_numpy(ta, tb, alpha, beta):
    b = make_batch("i8", ta, tb, 19, 13, 37, 3, alpha=alpha, beta=beta, pad=(3, 1, 2), seed=ta * 2 + tb)
    st, C = run_oracle(b)
    assert st == 0
    for i in range(b.batch):
        assert np.array_equal(oracle.colmajor(C[i], b.m, b.n, b.ldc), numpy_exact(b, i))
    assert padding_untouched(b, C)


def test_int8_brute_force_tiny():
    for m in (1, 2, 5):
        for n in (1, 3):
            for k in (1, 4, 7):
                for ta, tb in OPS:
                    for beta in (0, 1, -3):
                        b = make_batch("i8", ta, tb, m, n, k, 2, alpha=2, beta=beta, seed=m * 100 + n * 10 + k)
                        st, C = run_oracle(b)
                        assert st == 0
                        for i in range(2):
                            A = oracle.colmajor(b.A[i], *b.dims_a(), b.lda)
                            B = oracle.colmajor(b.B[i], *b.dims_b(), b.ldb)
                            C0 = oracle.colmajor(b.C0[i], m, n, b.ldc)
                            for r in range(m):
                                for c in range(n):
                                    s = 0
                                    for kk in range(k):
                                        a = int(A[r, kk] if ta == 0 else A[kk, r])
                                        x = int(B[kk, c] if tb == 0 else B[c, kk])
                                        s += a * x
                                    v = 2 * s + (beta * int(C0[r, c]) if beta else 0)
                                    v &= 0xFFFFFFFF
                                    v = v - (1 << 32) if v >= (1 << 31) else v
                                    assert int(oracle.colmajor(C[i], m, n, b.ldc)[r, c]) == v


# --------------------------------------------------------------------------
# fp: exact on integer inputs, bounded error on random inputs
# --------------------------------------------------------------------------
@pytest.mark.parametrize("t", ["f32", "f64"])
@pytest.mark.parametrize("ta,tb", OPS)
def test_fp_exact_integer_inputs(t, ta, tb):
    rng = np.random.default_rng(5)
    b = make_batch(t, ta, tb, 17, 9, 600, 2, alpha=3.0, beta=-2.0, pad=(1, 2, 3), seed=1)
    # overwrite with small integers: |sum| < 2^24, exact in fp32
    for arrs in (b.A, b.B, b.C0):
        for x in arrs:
            x[:] = rng.integers(-4, 5, size=x.shape)
    st, C = run_oracle(b)
    assert st == 0
    for i in range(b.batch):
        assert np.array_equal(oracle.colmajor(C[i], b.m, b.n, b.ldc), numpy_exact(b, i).astype(TYPES[t][3]))
    assert padding_untouched(b, C)


@pytest.mark.parametrize("t", ["f32", "f64"])
@pytest.mark.parametrize("ta,tb", OPS)
def test_fp_error_bound(t, ta, tb):
    """Deterministic bound for a blocked fused-dot product: |err| <=
    gamma_k * |alpha| |A||B| + small terms; normwise it is far below k*u."""
    k = 700  # two K blocks
    b = make_batch(t, ta, tb, 23, 11, k, 2, alpha=1.5, beta=-0.5, seed=3)
    st, C = run_oracle(b)
    assert st == 0
    err = normwise(b, C)
    u = UNIT_ROUNDOFF[t]
    assert err <= (k + 2) * u
    assert err > 0.0  # the restatement rounds: it is not the dd oracle


def test_dd_oracle_vs_fractions():
    rng = np.random.default_rng(11)
    for t in ("f32", "f64"):
        for ta, tb in OPS:
            b = make_batch(t, ta, tb, 3, 4, 9, 1, alpha=1.25, beta=-0.75, seed=int(rng.integers(1 << 30)))
            hi, lo = oracle.gemm_dd(b.transa, b.transb, b.m, b.n, b.k, b.alpha, b.A[0], b.ab_code, b.lda, b.B[0],
                                    b.ldb, b.beta, b.C0[0], b.ldc)
            A = oracle.colmajor(b.A[0], *b.dims_a(), b.lda)
            B = oracle.colmajor(b.B[0], *b.dims_b(), b.ldb)
            C0 = oracle.colmajor(b.C0[0], b.m, b.n, b.ldc)
            for r in range(b.m):
                for c in range(b.n):
                    s = Fraction(0)
                    for kk in range(b.k):
                        a = A[r, kk] if ta == 0 else A[kk, r]
                        x = B[kk, c] if tb == 0 else B[c, kk]
                        s += Fraction(float(a)) * Fraction(float(x))
                    exact = Fraction(b.alpha) * s + Fraction(b.beta) * Fraction(float(C0[r, c]))
                    got = Fraction(hi[r, c]) + Fraction(lo[r, c])
                    assert abs(got - exact) <= abs(exact) * Fraction(2) ** -100 + Fraction(2) ** -1000


def test_dd_oracle_int8_exact():
    b = make_batch("i8", 1, 1, 8, 6, 64, 1, alpha=3, beta=-2, seed=4)
    hi, lo = oracle.gemm_dd(b.transa, b.transb, b.m, b.n, b.k, b.alpha, b.A[0], b.ab_code, b.lda, b.B[0], b.ldb,
                            b.beta, b.C0[0], b.ldc)
    ex = (3 * (oracle.colmajor(b.A[0], *b.dims_a(), b.lda).astype(np.int64).T @
               oracle.colmajor(b.B[0], *b.dims_b(), b.ldb).astype(np.int64).T)
          - 2 * oracle.colmajor(b.C0[0], 8, 6, b.ldc).astype(np.int64))
    assert np.array_equal(hi + lo, ex.astype(np.float64))


# --------------------------------------------------------------------------
# status table
# --------------------------------------------------------------------------
R8, R32I, R32F, R64F, R16F, R8U = 0, 3, 4, 5, 2, 1
STATUS_CASES = [
    # (transa, transb, m, n, k, Atype, lda, Btype, ldb, Ctype, ldc, batch, expected, source)
    (0, 0, 4, 4, 4, R32F, 4, R32F, 4, R32F, 4, 2, 0, "supported fp32"),
    (1, 1, 4, 4, 4, R64F, 4, R64F, 4, R64F, 4, 2, 0, "supported fp64"),
    (1, 0, 4, 4, 4, R8, 4, R8, 4, R32I, 4, 2, 0, "supported int8"),
    (0, 0, 4, 4, 4, R32F, 4, R32F, 4, R64F, 4, 2, 2, "Selector.hpp:105-107 mixed triple"),
    (0, 0, 4, 4, 4, R16F, 4, R16F, 4, R16F, 4, 2, 2, "Selector.hpp:105-107 R_16F"),
    (0, 0, 4, 4, 4, R8U, 4, R8U, 4, R32I, 4, 2, 2, "Selector.hpp:105-107 R_8U"),
    (0, 0, 4, 4, 4, R8, 4, R8, 4, R8, 4, 2, 2, "Selector.hpp:105-107 int8->int8"),
    (2, 0, 4, 4, 4, R32F, 4, R32F, 4, R32F, 4, 2, 2, "ALLDATASIZESAME.hpp:85-88 OP_C fp"),
    (0, 2, 4, 4, 4, R64F, 4, R64F, 4, R64F, 4, 2, 2, "ALLDATASIZESAME.hpp:117-120 OP_C fp"),
    (2, 0, 4, 4, 4, R8, 4, R8, 4, R32I, 4, 2, 2, "int8 OP_C undefined in reference -> rejected"),
    (0, 0, 4, 0, 4, R32F, 4, R32F, 4, R32F, 4, 2, 3, "launcher.hpp:174-176 n == 0"),
    (0, 0, 4, 4, 4, R32F, 4, R32F, 4, R32F, 4, 0, 3, "launcher.hpp:174-176 batch == 0"),
    (0, 0, 0, 4, 4, R32F, 1, R32F, 4, R32F, 1, 2, 3, "m == 0 (reference divides by zero) -> INTERNAL_ERROR"),
    (0, 0, 4, 4, 0, R32F, 4, R32F, 1, R32F, 4, 2, 0, "k == 0 computes nothing, SUCCESS"),
    (0, 0, -1, 4, 4, R32F, 4, R32F, 4, R32F, 4, 2, 2, "negative m rejected (unchecked in reference)"),
    (0, 0, 4, 4, -2, R32F, 4, R32F, 4, R32F, 4, 2, 2, "negative k rejected"),
    (0, 0, 4, 4, 4, R32F, 4, R32F, 4, R32F, 4, -1, 2, "negative batch rejected"),
    (0, 0, 4, 4, 4, R32F, 3, R32F, 4, R32F, 4, 2, 2, "lda < m rejected"),
    (1, 0, 4, 4, 5, R32F, 4, R32F, 5, R32F, 4, 2, 2, "lda < k (T) rejected"),
    (0, 1, 4, 6, 4, R32F, 4, R32F, 5, R32F, 4, 2, 2, "ldb < n (T) rejected"),
    (0, 0, 4, 4, 4, R32F, 4, R32F, 4, R32F, 3, 2, 2, "ldc < m rejected"),
]


@pytest.mark.parametrize("case", STATUS_CASES, ids=lambda c: c[-1])
def test_status_table(case):
    *args, expected, _src = case
    assert oracle.validate(*args) == expected


def test_naive_baseline_equals_restatement():
    for ta, tb in OPS:
        b = make_batch("i8", ta, tb, 33, 21, 70, 3, alpha=5, beta=-1, seed=9)
        st, C = run_oracle(b)
        C2 = [c.copy() for c in b.C0]
        oracle.naive_baseline_i8(ta, tb, b.m, b.n, b.k, b.alpha, b.A, b.lda, b.B, b.ldb, b.beta, C2, b.ldc, 2)
        for x, y in zip(C, C2):
            assert np.array_equal(x, y)


def test_planted_mistake_is_caught():
    """A transposed operand in the oracle input must break exactness
    (guards against a check that cannot fail)."""
    b = make_batch("i8", 0, 0, 6, 6, 6, 1, seed=2)
    st, C = run_oracle(b)
    bt = make_batch("i8", 0, 0, 6, 6, 6, 1, seed=2)
    bt.transb = 1  # plant: B read transposed
    st2, C2 = run_oracle(bt)
    assert not np.array_equal(C[0], C2[0])
    assert np.array_equal(oracle.colmajor(C[0], 6, 6, 6), numpy_exact(b, 0))


@pytest.mark.parametrize("t", ["f32", "f64", "i8"])
@pytest.mark.parametrize("beta", [0.0, -0.75])
def test_dd_batched_equals_dd(t, beta):
    """oracle_gemm_dd_batched is oracle_gemm_dd per instance, bit for bit
    (the batched form only changes the parallel loop)."""
    for ta, tb in OPS:
        b = make_batch(t, ta, tb, 9, 7, 13, 5, alpha=1.25 if t != "i8" else 3, beta=beta if t != "i8" else int(beta * 4),
                       pad=(1, 2, 3), seed=77)
        hi, lo = oracle.gemm_dd_batched(ta, tb, b.m, b.n, b.k, b.alpha, b.A, b.ab_code, b.lda, b.B, b.ldb, b.beta,
                                        b.C0, b.ldc, nthreads=3)
        for i in range(b.batch):
            h1, l1 = oracle.gemm_dd(ta, tb, b.m, b.n, b.k, b.alpha, b.A[i], b.ab_code, b.lda, b.B[i], b.ldb, b.beta,
                                    b.C0[i], b.ldc)
            assert np.array_equal(hi[i], h1) and np.array_equal(lo[i], l1)


@pytest.mark.skipif(oracle.find_openblas() is None, reason="scipy's OpenBLAS not installed")
@pytest.mark.parametrize("t", ["f32", "f64"])
def test_blas_baseline_matches_numpy(t):
    """The example's cblas loop (scipy's OpenBLAS, opened at run time):
    trans flags, leading dimensions, alpha and beta map as in the call, so
    on integer-valued inputs (every sum exact) it equals exact arithmetic;
    a transposed operand planted in the call must not."""
    lib = oracle.find_openblas()
    for ta, tb in OPS:
        b = make_batch(t, ta, tb, 11, 6, 9, 3, alpha=2.0, beta=-3.0, pad=(2, 1, 3), seed=91)
        for arr in b.A + b.B + b.C0:
            arr[:] = np.round(arr * 8)
        C = [c.copy() for c in b.C0]
        assert oracle.blas_baseline(lib, b.ab_code, ta, tb, b.m, b.n, b.k, b.alpha, b.A, b.lda, b.B, b.ldb, b.beta,
                                    C, b.ldc, 2) == 0
        assert padding_untouched(b, C)
        for i in range(b.batch):
            assert np.array_equal(oracle.colmajor(C[i], b.m, b.n, b.ldc), numpy_exact(b, i))
    b = make_batch(t, 0, 0, 6, 6, 6, 1, alpha=1.0, beta=0.0, seed=92)
    for arr in b.A + b.B:
        arr[:] = np.round(arr * 8)
    C = [c.copy() for c in b.C0]
    oracle.blas_baseline(lib, b.ab_code, 0, 1, 6, 6, 6, 1.0, b.A, b.lda, b.B, b.ldb, 0.0, C, b.ldc, 1)
    assert not np.array_equal(oracle.colmajor(C[0], 6, 6, 6), numpy_exact(b, 0))
