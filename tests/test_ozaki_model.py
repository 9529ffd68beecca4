"""Pins the numpy model of the Ozaki form's a-priori error bound
(tests/ozaki_model.py: error_bound, auto_slices) on the CPU, before the GPU
report is compared with it.  The bound must hold -- the model's own Ozaki
result, at every slice count, stays within it (plus the final fp64
rounding) against the double-double oracle -- on random data of several
distributions and on adversarial data whose digits are all 127 (the dropped
products at their largest), and it must not be loose by more than a few
binary orders on random data.  Exactly representable data must be computed
exactly at the slice count the automatic choice takes."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from ozaki_model import TARGET, U, auto_slices, digit_depth, error_bound, error_bound_fused, exponents, ozaki_gemm


def dd_error(A, B, C):
    m, k = A.shape
    n = B.shape[1]
    Af = np.asfortranarray(A).ravel(order="F")
    Bf = np.asfortranarray(B).ravel(order="F")
    hi, lo = oracle.gemm_dd(0, 0, m, n, k, 1.0, Af, oracle.R_64F, m, Bf, k, 0.0, None, m)
    return oracle.normwise_error(C, hi, lo, np.linalg.norm(A) * np.linalg.norm(B))


def draw(rng, dist, shape):
    if dist == "wide":
        return rng.uniform(-1, 1, shape) * np.exp2(rng.integers(-8, 9, shape))
    if dist == "normal":
        return rng.standard_normal(shape)
    if dist.startswith("all127"):  # every digit 127 (values just below 1); random or equal signs
        v = np.where(rng.random(shape) < 0.5, 1 - 2.0 ** -53, 1 - 2.0 ** -49)
        return v * (rng.choice([-1.0, 1.0], shape) if dist == "all127" else 1.0)
    return rng.uniform(-1, 1, shape)


@pytest.mark.parametrize("dist", ["uniform", "wide", "normal", "all127", "all127pos"])
@pytest.mark.parametrize("shape", [(16, 24, 32), (40, 33, 7)], ids=str)
def test_bound_holds_and_is_not_vacuous(dist, shape):
    rng = np.random.default_rng(hash((dist, shape)) % 2 ** 32)
    m, n, k = shape
    for _ in range(3):
        A, B = draw(rng, dist, (m, k)), draw(rng, dist, (k, n))
        for s in range(3, 10):
            bound = error_bound(A, B, s)
            err = dd_error(A, B, ozaki_gemm(A, B, s))
            assert err <= bound + U, (s, err, bound)
            if max(shape) <= 32:  # the fused kernel's form of it holds too, and is no tighter
                bf = error_bound_fused(A, B, s)
                assert err <= bf + U and bf >= bound * (1 - 1e-9), (s, err, bound, bf)
            if not dist.startswith("all127") and s < 9:  # truncation-dominated: within 2^8 of the actual error
                assert bound <= 256 * err, (s, err, bound)


def test_bound_falls_with_slices():
    rng = np.random.default_rng(5)
    A, B = draw(rng, "wide", (32, 32)), draw(rng, "wide", (32, 32))
    b = [error_bound(A, B, s) for s in range(3, 10)]
    assert all(y < x / 64 for x, y in zip(b, b[1:])), b


def test_exact_data():
    rng = np.random.default_rng(6)
    A = rng.integers(-100, 101, (30, 40)).astype(np.float64)
    B = rng.integers(-100, 101, (40, 20)).astype(np.float64)
    assert digit_depth(A, exponents(A, 1), 1) <= 2 and digit_depth(B, exponents(B, 0), 0) <= 2
    s = auto_slices(A, B)
    assert s == 3
    assert error_bound(A, B, s) <= TARGET
    assert np.array_equal(ozaki_gemm(A, B, s), A @ B)  # small integers: every sum exact in fp64


def test_full_precision_data_takes_nine_slices():
    rng = np.random.default_rng(7)
    A, B = draw(rng, "uniform", (32, 32)), draw(rng, "uniform", (32, 32))
    assert auto_slices(A, B) == 9
    assert error_bound(A, B, 9) <= TARGET < error_bound(A, B, 8)


def test_zero_and_nonfinite():
    A = np.zeros((4, 5))
    B = np.ones((5, 3))
    assert error_bound(A, B, 4) == 0.0
    A[1, 2] = np.inf
    assert error_bound(A, B, 9) == float("inf")


def test_planted_mistake_is_caught():
    """A bound that forgets the dropped products of the triangle (only the
    digit truncation kept) fails on the all-127 data of equal signs (the
    dropped products add up): the checks above can catch a missing term."""
    rng = np.random.default_rng(8)
    A, B = draw(rng, "all127pos", (16, 32)), draw(rng, "all127pos", (32, 16))
    s = 5
    ea, fb = exponents(A, 1), exponents(B, 0)
    Ap, Bp = np.ldexp(np.abs(A), -ea[:, None]), np.ldexp(np.abs(B), -fb[None, :])
    # truncation of both operands only: 2^-7s (b1_j + a1_i) per element
    dP = 2.0 ** (-7 * s) * (Ap.sum(axis=1)[:, None] + Bp.sum(axis=0)[None, :])
    wrong = np.sqrt((np.exp2(2.0 * (ea[:, None] + fb[None, :])) * dP ** 2).sum()) / (np.linalg.norm(A) * np.linalg.norm(B))
    err = dd_error(A, B, ozaki_gemm(A, B, s))
    assert err > wrong + U
    assert err <= error_bound(A, B, s) + U
