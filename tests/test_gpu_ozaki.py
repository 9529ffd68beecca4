"""GPU checks of the Ozaki fp64 form (SSBLAS_COMPUTE_64F_EMULATED_OZAKI*).

The form has no counterpart in the reference; its parity bar is the
north-star one: normwise error against the FP64 (double-double) oracle,
reported per slice count, and with the automatic slice count no worse than
2 x the reference-order fp64 restatement's error + 1 unit roundoff.  In
addition the library must reproduce tests/ozaki_model.py (the same
algorithm in numpy) bit for bit when alpha is a power of two, at the slice
count each instance used, and its per-instance error report
(ssblasGetOzakiReport) must equal the model's a-priori bound and hold
against the double-double oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import UNIT_ROUNDOFF, dd_reference, make_batch, normwise, padding_untouched, run_gpu, run_oracle
from ozaki_model import U, auto_slices, error_bound, error_bound_fused, ozaki_gemm

pytestmark = pytest.mark.gpu
OPS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def op_mats(b, i):
    A = oracle.colmajor(b.A[i], *b.dims_a(), b.lda).astype(np.float64)
    B = oracle.colmajor(b.B[i], *b.dims_b(), b.ldb).astype(np.float64)
    return (A if b.transa == 0 else A.T), (B if b.transb == 0 else B.T)


@pytest.mark.parametrize("ta,tb", OPS)
@pytest.mark.parametrize("s", [3, 5, 8])
@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 32, 17), (16, 16, 32), (32, 32, 32), (37, 29, 70), (64, 64, 64),
                                   (100, 65, 129)], ids=str)
def test_ozaki_bit_exact_vs_model(ta, tb, s, shape):
    m, n, k = shape
    b = make_batch("f64", ta, tb, m, n, k, 2, alpha=2.0, beta=-0.5, pad=(1, 2, 3), seed=s * 7 + m, dist="wide")
    st, C = run_gpu(b, compute_type=0x100 + s)
    assert st == 0
    assert padding_untouched(b, C)
    for i in range(b.batch):
        A, B = op_mats(b, i)
        C0 = oracle.colmajor(b.C0[i], m, n, b.ldc)
        want = ozaki_gemm(A, B, s, alpha=2.0, beta=-0.5, C0=C0)
        got = oracle.colmajor(C[i], m, n, b.ldc)
        assert np.array_equal(got, want), (i, np.max(np.abs(got - want)))


@pytest.mark.parametrize("ta,tb", OPS)
def test_ozaki_error_decreases_and_meets_native(ta, tb):
    b = make_batch("f64", ta, tb, 48, 40, 96, 3, alpha=1.0, beta=0.0, seed=3, dist="wide")
    errs = {}
    for s in range(3, 10):
        st, C = run_gpu(b, compute_type=0x100 + s)
        assert st == 0
        errs[s] = normwise(b, C)
    for s in range(3, 8):
        assert errs[s + 1] < errs[s] or errs[s + 1] <= UNIT_ROUNDOFF["f64"], errs
    _, R = run_oracle(b)
    e_ref = normwise(b, R)
    st, C = run_gpu(b, compute_type=6)  # automatic slice count
    assert st == 0
    assert normwise(b, C) <= 2 * e_ref + UNIT_ROUNDOFF["f64"], (normwise(b, C), e_ref, errs)


def test_ozaki_nonfinite_rows_become_nan():
    b = make_batch("f64", 0, 0, 20, 12, 30, 1, alpha=1.0, beta=0.0, seed=4)
    oracle.colmajor(b.A[0], 20, 30, b.lda)[3, 7] = np.inf
    st, C = run_gpu(b, compute_type=6)
    assert st == 0
    out = oracle.colmajor(C[0], 20, 12, b.ldc)
    assert np.isnan(out[3]).all()
    assert np.isfinite(np.delete(out, 3, axis=0)).all()


@pytest.mark.parametrize("ta,tb", [(0, 1), (1, 0)])
def test_ozaki_fused_subnormal_zero_and_nonfinite_rows(ta, tb, monkeypatch):
    """Whole 32^3 instances whose rows of op(A) / columns of op(B) have a
    subnormal maximum (the row exponent then comes from frexp, not from the
    biased exponent field), are all zero, mix subnormal and normal values,
    or hold an infinity: the fused kernel gives the general path's result
    bit for bit (NaN where an infinity reaches), and both give the numpy
    model's, whose row exponents come from numpy's frexp."""
    b = make_batch("f64", ta, tb, 32, 32, 32, 6, alpha=1.0, beta=0.0, seed=61, dist="wide")
    for i in range(6):
        A = oracle.colmajor(b.A[i], *oracle.op_dims(ta, 32, 32), b.lda)
        B = oracle.colmajor(b.B[i], *oracle.op_dims(tb, 32, 32), b.ldb)
        opA = A.T if ta else A      # views: writes land in the stored matrix
        opB = B.T if tb else B
        opA[2 * i] *= 2.0 ** -1060          # a row of subnormals
        opA[2 * i + 1] = 0.0                # an all-zero row
        opA[2 * i + 3, :5] *= 2.0 ** -1065  # a row mixing subnormal and normal values
        opB[:, 3 * i] *= 2.0 ** -1040       # a column of subnormals
        opB[:, 3 * i + 1] = 0.0
        if i == 4:
            opA[7, 9] = np.inf
    st, fused = run_gpu(b, compute_type=0x109)
    assert st == 0
    monkeypatch.setenv("SSBLAS_OZAKI_NO_FUSED", "1")
    st, general = run_gpu(b, compute_type=0x109)
    assert st == 0
    for i, (x, y) in enumerate(zip(fused, general)):
        assert np.array_equal(x, y, equal_nan=True), f"instance {i}"
        if i != 4:  # and the numpy model (its exponents from numpy's frexp)
            A, B = op_mats(b, i)
            want = ozaki_gemm(A, B, 9)
            assert np.array_equal(oracle.colmajor(y, 32, 32, b.ldc), want), f"instance {i} vs model"
    out = oracle.colmajor(fused[4], 32, 32, b.ldc)
    assert np.isnan(out[7]).all()


def test_ozaki_long_k_chunked():
    """K beyond one int32-exact chunk (8192) is folded chunk by chunk."""
    b = make_batch("f64", 1, 0, 8, 8, 9000, 1, alpha=1.0, beta=0.0, seed=5)
    st, C = run_gpu(b, compute_type=0x108)
    assert st == 0
    _, R = run_oracle(b)
    assert normwise(b, C) <= 2 * normwise(b, R) + UNIT_ROUNDOFF["f64"]


def test_ozaki_config_fp64_32cubed_sample():
    """fp64 NT m=n=k=32, entries U[-1,1] * 2^U{-8..8}: error per slice count
    on 256 instances (the benchmark configuration's shape)."""
    b = make_batch("f64", 0, 1, 32, 32, 32, 256, alpha=1.0, beta=0.0, seed=0x55B1A5 + 3, dist="wide")
    st, C = run_gpu(b)
    e_native = normwise(b, C)
    _, R = run_oracle(b, 8)
    e_ref = normwise(b, R)
    assert e_native <= 2 * e_ref + UNIT_ROUNDOFF["f64"]
    res = {}
    for s in range(4, 9):
        st, C = run_gpu(b, compute_type=0x100 + s)
        assert st == 0
        res[s] = normwise(b, C)
    print("ozaki normwise error by slice count", res, "native", e_native, "reference order", e_ref)
    assert res[8] <= 2 * e_ref + UNIT_ROUNDOFF["f64"]


@pytest.mark.parametrize("ta,tb", OPS)
@pytest.mark.parametrize("s", [4, 9])
@pytest.mark.parametrize("shape,pad,grid", [((32, 29, 31), (0, 3, 1), None), ((32, 29, 31), (0, 3, 1), "3"),
                                            ((32, 32, 32), (0, 0, 2), "3"), ((28, 32, 32), (2, 0, 4), None)])
def test_ozaki_fused_small_equals_general(ta, tb, s, shape, pad, grid, monkeypatch):
    """Instances with m, n, k <= 32 run the fused one-wave kernel; it must
    give the general (workspace) path's result bit for bit, also when every
    wave walks several instances (grid capped: the next instance's operands
    are loaded while the current one is computed) and when C is written as
    16-byte pairs (even ldc) or element by element (odd ldc)."""
    m, n, k = shape
    if grid is not None:
        monkeypatch.setenv("SSBLAS_MAX_GRID", grid)
    b = make_batch("f64", ta, tb, m, n, k, 40, alpha=-1.5, beta=0.25, pad=pad, seed=11 + s, dist="wide")
    st, fused = run_gpu(b, compute_type=0x100 + s)
    assert st == 0
    monkeypatch.setenv("SSBLAS_OZAKI_NO_FUSED", "1")
    st, general = run_gpu(b, compute_type=0x100 + s)
    assert st == 0
    for x, y in zip(fused, general):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("ta,tb", OPS)
@pytest.mark.parametrize("pad", [(0, 0, 0), (1, 0, 0), (0, 2, 1)], ids=str)
def test_ozaki_fused_full_blocks(ta, tb, pad, monkeypatch):
    """Whole 32 x 32 x 32 instances (the fused kernel's unguarded loads):
    even and odd leading dimensions, one instance 8 bytes off 16-byte
    alignment, every wave walking several instances (grid capped).
    Bit-identical to the general path."""
    import torch

    from helpers import DeviceBatch
    from ssblasex_amd import pointer_array

    monkeypatch.setenv("SSBLAS_MAX_GRID", "2")
    b = make_batch("f64", ta, tb, 32, 32, 32, 24, alpha=0.5, beta=-1.0, pad=pad, seed=97, dist="wide")
    d = DeviceBatch(b)
    x = d.A[5]
    big = torch.empty(x.numel() + 1, dtype=x.dtype, device=x.device)
    big[1:].copy_(x)
    d.A[5] = big[1:]
    d.pA = pointer_array(d.A)
    st, fused = run_gpu(b, compute_type=0x109, dev=d)
    assert st == 0
    monkeypatch.setenv("SSBLAS_OZAKI_NO_FUSED", "1")
    d.reset_c()
    st, general = run_gpu(b, compute_type=0x109, dev=d)
    assert st == 0
    for i, (x, y) in enumerate(zip(fused, general)):
        assert np.array_equal(x, y), f"instance {i}"



def report():
    import torch

    import ssblasex_amd as S

    st, rep, bounds, slices = S.ozaki_report(torch.cuda.current_stream())
    assert st == S.Status.SUCCESS, st
    return rep, bounds, slices


REPORT_CASES = [((32, 32, 32), 0x109), ((32, 32, 32), 6), ((20, 31, 17), 0x105), ((32, 32, 32), 0x107),
                ((48, 40, 96), 0x109), ((48, 40, 96), 6), ((70, 33, 65), 0x104), ((64, 64, 64), 0x108)]


@pytest.mark.parametrize("ta,tb", OPS)
@pytest.mark.parametrize("shape,ct", REPORT_CASES, ids=[f"{c[0]}-{c[1]:#x}" for c in REPORT_CASES])
def test_ozaki_report_bound_holds(ta, tb, shape, ct):
    """Per instance: the reported slice count is the named one (or the
    model's automatic choice), the reported bound equals the model's bound
    (fp32 evaluation: 1e-3 relative), the result equals the model at that
    slice count bit for bit, and the error against the double-double oracle
    stays within the bound plus the final fp64 rounding (one unit roundoff
    of the normwise scale).  Fused (m, n, k <= 32) and general paths."""
    m, n, k = shape
    b = make_batch("f64", ta, tb, m, n, k, 6, alpha=1.0, beta=0.0, pad=(1, 0, 2), seed=ct + m, dist="wide")
    # rows and columns of very different scales, a zero row and a zero column
    oracle.colmajor(b.A[1], *b.dims_a(), b.lda)[:] *= 2.0 ** 40
    A2 = oracle.colmajor(b.A[2], *b.dims_a(), b.lda)
    if b.transa == 0:
        A2[0, :] = 0.0
    else:
        A2[:, 0] = 0.0
    B3 = oracle.colmajor(b.B[3], *b.dims_b(), b.ldb)
    if b.transb == 0:
        B3[:, 0] = 0.0
    else:
        B3[0, :] = 0.0
    st, C = run_gpu(b, compute_type=ct)
    assert st == 0
    rep, bounds, slices = report()
    assert rep.instances == b.batch
    fused = max(m, n, k) <= 32  # the fused one-wave kernel: its automatic choice is always 9 slices
    for i in range(b.batch):
        A, B = op_mats(b, i)
        s = (9 if fused else auto_slices(A, B)) if ct == 6 else ct - 0x100
        assert slices[i] == s, (i, slices[i], s)
        want_bound = (error_bound_fused if fused else error_bound)(A, B, s)
        assert abs(bounds[i] - want_bound) <= 1e-3 * want_bound + 1e-30, (i, bounds[i], want_bound)
        got = oracle.colmajor(C[i], m, n, b.ldc)
        assert np.array_equal(got, ozaki_gemm(A, B, s)), i
        hi, lo = dd_reference(b, i)
        err = oracle.normwise_error(got, hi, lo, np.linalg.norm(A) * np.linalg.norm(B))
        assert err <= bounds[i] * (1 + 1e-3) + U, (i, s, err, bounds[i])
    assert rep.bound_max == pytest.approx(float(bounds.max()))
    assert rep.slices_min == slices.min() and rep.slices_max == slices.max()


@pytest.mark.parametrize("shape", [(32, 32, 32), (50, 40, 60)], ids=str)
def test_ozaki_auto_exact_data_takes_few_slices(shape):
    """Integer-valued operands need one or two digits: the general path's
    automatic choice takes 3 slices (the fused kernel always takes 9), the
    general path's reported bound is the fold term only, and the result is
    the exact integer product."""
    m, n, k = shape
    b = make_batch("f64", 0, 1, m, n, k, 4, alpha=1.0, beta=0.0, seed=21)
    rng = np.random.default_rng(22)
    for x in b.A + b.B:
        x[:] = rng.integers(-100, 101, x.shape)
    st, C = run_gpu(b, compute_type=6)
    assert st == 0
    rep, bounds, slices = report()
    assert (slices == (9 if max(shape) <= 32 else 3)).all(), slices
    if max(shape) > 32:
        assert (bounds <= 2.0 ** -55).all(), bounds
    for i in range(b.batch):
        A, B = op_mats(b, i)
        exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
        assert np.array_equal(oracle.colmajor(C[i], m, n, b.ldc), exact)


def test_ozaki_report_nonfinite_and_absent():
    import torch

    import ssblasex_amd as S

    s2 = torch.cuda.Stream()
    st, rep, _, _ = S.ozaki_report(s2)
    assert st == S.Status.NOT_COMPUTED_ERROR
    b = make_batch("f64", 0, 0, 20, 12, 30, 2, alpha=1.0, beta=0.0, seed=4)
    oracle.colmajor(b.A[1], 20, 30, b.lda)[3, 7] = np.nan
    st, C = run_gpu(b, compute_type=6)
    assert st == 0
    rep, bounds, slices = report()
    assert np.isfinite(bounds[0]) and np.isinf(bounds[1])
    assert rep.over_target == 1 and np.isinf(rep.bound_max)
