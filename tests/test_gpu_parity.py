"""GPU parity of ssblasGemmBatchedEx (HIP library, called through the C ABI
via ctypes) against the CPU oracle on the same seeded inputs.

Bars (stated per form):
* int8 -> int32: bit-exact against the oracle restatement.
* native fp32 / fp64: normwise error ||C - C_exact||_F / (|alpha| ||A|| ||B||
  + |beta| ||C0||) against the double-double oracle at most
  2 x (the same error of the reference-order restatement) + 1 unit roundoff,
  and C outside the m x n block untouched.
* BF16X6 (fp32 emulated): same bar as native fp32 but 4 x instead of 2 x.
* BF16X3 (fp32 fast): normwise error <= 2^-14 (16-bit operands).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from helpers import (DeviceBatch, UNIT_ROUNDOFF, load_golden, make_batch, normwise, padding_untouched, run_gpu,
                     run_oracle)

pytestmark = pytest.mark.gpu

OPS = [(0, 0), (0, 1), (1, 0), (1, 1)]
NTHREADS = min(16, os.cpu_count() or 4)
SHAPES = [(1, 1, 1), (7, 5, 3), (16, 16, 64), (33, 17, 65), (64, 64, 64), (65, 63, 130), (128, 128, 128),
          (130, 257, 200)]
CONFIGS = ["direct16", "direct32", "tile64", "tile128", "tile128x64", "tile64x128", "tile128k2", "tile128w8",
           "tile256x128", "tile64s2", "tile128s2", "tile128k2s2", "tile128w8s2", "tile256x128s2", "dma64",
           "dma128", "dma128k2", "dma256x128", "dma128s3", "kc256x128", "kc128", "dma256", "dma256w16", "dma256k2",
           "gdma128", "gdma64", "gdma256x128"]


@pytest.fixture(autouse=True)
def _gpu():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    yield
    os.environ.pop("SSBLAS_TILE_CONFIG", None)


def fp_bar(t, ref_err, factor=2.0):
    return factor * ref_err + UNIT_ROUNDOFF[t]


def check(b, compute_type=0, entry="sync", factor=2.0):
    st, C = run_gpu(b, compute_type=compute_type, entry=entry)
    assert st == 0, st
    assert padding_untouched(b, C)
    if b.type == "i8":
        st2, R = run_oracle(b, NTHREADS)
        assert st2 == 0
        for i, (x, y) in enumerate(zip(C, R)):
            if not np.array_equal(x, y):
                bad = np.argwhere(oracle.colmajor(x, b.m, b.n, b.ldc) != oracle.colmajor(y, b.m, b.n, b.ldc))
                pytest.fail(f"instance {i}: {len(bad)} mismatches, first at {bad[:4].tolist()}")
        return 0.0, 0.0
    st2, R = run_oracle(b, NTHREADS)
    e_gpu = normwise(b, C)
    e_ref = normwise(b, R)
    assert e_gpu <= fp_bar(b.type, e_ref, factor), (e_gpu, e_ref)
    return e_gpu, e_ref


# --------------------------------------------------------------------------
@pytest.mark.parametrize("case", [c for c in load_golden() if "gpu" in c[1]], ids=lambda c: c[0])
def test_golden_fixture_gpu(case):
    name, _, b, expect = case
    st, C = run_gpu(b)
    assert st == 0
    for got, exp in zip(C, expect):
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (name, got, exp)


@pytest.mark.parametrize("shape", SHAPES, ids=str)
@pytest.mark.parametrize("ta,tb", OPS)
@pytest.mark.parametrize("beta", [0, -3])
def test_int8_bit_exact(shape, ta, tb, beta):
    m, n, k = shape
    check(make_batch("i8", ta, tb, m, n, k, 3, alpha=3, beta=beta, pad=(1, 2, 3), seed=m + n + k))


@pytest.mark.parametrize("t", ["f32", "f64"])
@pytest.mark.parametrize("shape", SHAPES, ids=str)
@pytest.mark.parametrize("ta,tb", OPS)
def test_fp_normwise(t, shape, ta, tb):
    m, n, k = shape
    check(make_batch(t, ta, tb, m, n, k, 3, alpha=1.5, beta=-0.5, pad=(2, 1, 0), seed=k))


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("t", ["i8", "f32", "f64"])
@pytest.mark.parametrize("ta,tb", OPS)
def test_every_tile_config(cfg, t, ta, tb):
    """Each kernel variant, forced, on a ragged shape with a K tail."""
    os.environ["SSBLAS_TILE_CONFIG"] = cfg
    alpha, beta = (2, 1) if t == "i8" else (0.75, 1.25)
    check(make_batch(t, ta, tb, 150, 97, 203, 2, alpha=alpha, beta=beta, pad=(3, 0, 5), seed=17))
    check(make_batch(t, ta, tb, 128, 128, 128, 2, alpha=alpha, beta=0, seed=18))


@pytest.mark.parametrize("cfg", ["dma64", "dma128", "dma128k2", "dma256x128", "dma128s3"])
@pytest.mark.parametrize("ta,tb", OPS)
def test_int8_dma_tiles(cfg, ta, tb):
    """Direct-to-LDS int8 tiles on whole tiles (DMA path) and on a ragged
    shape (guarded path), beta 0 and 1."""
    os.environ["SSBLAS_TILE_CONFIG"] = cfg
    check(make_batch("i8", ta, tb, 256, 256, 256, 3, alpha=1, beta=0, seed=61))
    check(make_batch("i8", ta, tb, 256, 256, 384, 2, alpha=-2, beta=1, seed=62))
    check(make_batch("i8", ta, tb, 200, 136, 100, 2, alpha=3, beta=-1, pad=(1, 0, 2), seed=63))
    # whole and partial tiles in one batch
    check(make_batch("i8", ta, tb, 320, 256, 128, 2, alpha=1, beta=0, seed=64))
    check(make_batch("i8", ta, tb, 256, 320, 64, 3, alpha=-1, beta=2, seed=65))


@pytest.mark.parametrize("cfg", ["dma64", "dma128", "dma128s3", "dma256x128", "dma256", "dma256k2"])
@pytest.mark.parametrize("grid", ["3", "8"])
@pytest.mark.parametrize("ta,tb", OPS)
def test_int8_stream_several_tiles_per_workgroup(cfg, grid, ta, tb):
    """The streaming int8 kernels with the grid capped so that every
    workgroup walks several whole tiles (the next tile's k-blocks are in
    flight while a tile's C is written), beta 0 and 1, one instance whose A
    is not 16-byte aligned in the middle of the batch (the stream stops in
    front of it, runs it on the guarded path and restarts) and one whose C
    is not."""
    import torch

    from ssblasex_amd import pointer_array

    os.environ["SSBLAS_TILE_CONFIG"] = cfg
    os.environ["SSBLAS_MAX_GRID"] = grid
    try:
        check(make_batch("i8", ta, tb, 256, 384, 128, 5, alpha=3, beta=0, seed=66))
        check(make_batch("i8", ta, tb, 256, 256, 192, 4, alpha=-1, beta=1, seed=67))
        # four 128-k blocks per tile: the next tile's blocks are requested
        # before a tile's C is written (beta 0: stores only; beta 1: loads
        # of the old C too), and waited for past that epilogue
        check(make_batch("i8", ta, tb, 256, 512, 512, 3, alpha=1, beta=0, seed=69))
        check(make_batch("i8", ta, tb, 512, 256, 512, 3, alpha=-3, beta=1, seed=70))
        b = make_batch("i8", ta, tb, 256, 256, 256, 5, alpha=2, beta=0, seed=68)
        d = DeviceBatch(b)
        x = d.A[2]
        big = torch.empty(x.numel() + 1, dtype=x.dtype, device=x.device)
        big[1:].copy_(x)
        d.A[2] = big[1:]
        d.pA = pointer_array(d.A)
        # and one C 4 bytes off 16-byte alignment (written without vectors)
        y = d.C[3]
        bigc = torch.empty(y.numel() + 1, dtype=y.dtype, device=y.device)
        bigc[1:].copy_(y)
        d.C[3] = bigc[1:]
        d.pC = pointer_array(d.C)
        st, C = run_gpu(b, dev=d)
        assert st == 0
        _, R = run_oracle(b, NTHREADS)
        for i, (x, y) in enumerate(zip(C, R)):
            assert np.array_equal(x, y), f"instance {i} differs"
    finally:
        os.environ.pop("SSBLAS_MAX_GRID", None)


KC_FORMS = [("f32", 0, 2.0), ("f64", 0, 2.0), ("bf16x6", 5, 4.0), ("bf16x3", 4, None)]


@pytest.mark.parametrize("cfg", ["kc256x128", "kc128"])
@pytest.mark.parametrize("form,ct,factor", KC_FORMS, ids=[f[0] for f in KC_FORMS])
@pytest.mark.parametrize("grid", ["3", "0"])
def test_kc_dma_tiles(cfg, form, ct, factor, grid):
    """Direct-to-LDS tiles of K-contiguous fp operands (op(A) = A^T, op(B) =
    B) in every form: whole tiles streamed through capped and uncapped grids,
    beta 0 and nonzero, and one instance whose A is one element off alignment (its
    tiles run as register-staged sub-tiles inside the same launch)."""
    import torch

    from ssblasex_amd import pointer_array

    t = "f64" if form == "f64" else "f32"
    bar = (lambda e, r: e <= 2.0 ** -14) if factor is None else (lambda e, r: e <= fp_bar(t, r, factor))
    os.environ["SSBLAS_TILE_CONFIG"] = cfg
    if grid != "0":
        os.environ["SSBLAS_MAX_GRID"] = grid
    try:
        for (m, n, k, batch, alpha, beta, seed) in [(256, 384, 256, 3, 1.5, -0.5, 71), (256, 256, 128, 2, 1.0, 0.0, 72)]:
            b = make_batch(t, 1, 0, m, n, k, batch, alpha=alpha, beta=beta, seed=seed, dist="normal")
            st, C = run_gpu(b, compute_type=ct)
            assert st == 0
            _, R = run_oracle(b, NTHREADS)
            assert bar(normwise(b, C), normwise(b, R)), (m, n, k, normwise(b, C), normwise(b, R))
        b = make_batch(t, 1, 0, 256, 256, 256, 4, alpha=0.75, beta=0.5, seed=73, dist="normal")
        d = DeviceBatch(b)
        x = d.A[1]
        big = torch.empty(x.numel() + 1, dtype=x.dtype, device=x.device)
        big[1:].copy_(x)
        d.A[1] = big[1:]
        d.pA = pointer_array(d.A)
        st, C = run_gpu(b, compute_type=ct, dev=d)
        assert st == 0
        _, R = run_oracle(b, NTHREADS)
        # normwise() is the worst instance, the misaligned one included
        assert bar(normwise(b, C), normwise(b, R)), (normwise(b, C), normwise(b, R))
    finally:
        os.environ.pop("SSBLAS_MAX_GRID", None)


SPLIT_FORMS = [("bf16x6", 5, 4.0), ("bf16x3", 4, None)]


@pytest.mark.parametrize("cfg", ["tile256x128", "tile256x128s2", "tile128w8", "tile128", "tile64", "gdma128",
                                 "gdma64", "gdma256x128"])
@pytest.mark.parametrize("form,ct,factor", SPLIT_FORMS, ids=[f[0] for f in SPLIT_FORMS])
@pytest.mark.parametrize("ta,tb", OPS)
def test_split_forms_forced_tiles(cfg, form, ct, factor, ta, tb):
    """The BF16 split forms on the register-staged tiles the tuned table
    sends large fp32 batches to (256x128 with its swizzled unpadded bf16 LDS
    rows), forced: whole tiles at beta 0, ragged tiles with a K tail and
    padded leading dimensions at beta != 0, and a grid capped so that every
    workgroup walks several tiles."""
    os.environ["SSBLAS_TILE_CONFIG"] = cfg
    bar = (lambda e, r: e <= 2.0 ** -14) if factor is None else (lambda e, r: e <= fp_bar("f32", r, factor))
    cases = [dict(m=256, n=256, k=256, batch=2, alpha=1.0, beta=0.0, pad=(0, 0, 0), seed=81),
             dict(m=300, n=200, k=150, batch=2, alpha=1.5, beta=-0.5, pad=(3, 1, 2), seed=82),
             dict(m=512, n=384, k=96, batch=3, alpha=-0.75, beta=2.0, pad=(0, 0, 0), seed=83)]
    try:
        for i, c in enumerate(cases):
            if i == 2:
                os.environ["SSBLAS_MAX_GRID"] = "5"
            b = make_batch("f32", ta, tb, c["m"], c["n"], c["k"], c["batch"], alpha=c["alpha"], beta=c["beta"],
                           pad=c["pad"], seed=c["seed"], dist="normal")
            st, C = run_gpu(b, compute_type=ct)
            assert st == 0
            assert padding_untouched(b, C)
            _, R = run_oracle(b, NTHREADS)
            e, r = normwise(b, C), normwise(b, R)
            assert bar(e, r), (c, e, r)
    finally:
        os.environ.pop("SSBLAS_MAX_GRID", None)


@pytest.mark.parametrize("t", ["i8", "f32", "f64"])
@pytest.mark.parametrize("ta,tb", OPS)
def test_unaligned_pointers_and_odd_ld(t, ta, tb):
    """Instances that start one element past an allocation and odd leading
    dimensions take the guarded load paths."""
    import torch

    alpha, beta = (1, -2) if t == "i8" else (1.0, 0.5)
    b = make_batch(t, ta, tb, 70, 45, 99, 2, alpha=alpha, beta=beta, pad=(1, 3, 1), seed=23)
    d = DeviceBatch(b)
    # re-home every buffer one element into a larger allocation
    for lst in (d.A, d.B, d.C):
        for i, x in enumerate(lst):
            big = torch.empty(x.numel() + 1, dtype=x.dtype, device=x.device)
            big[1:].copy_(x)
            lst[i] = big[1:]
    from ssblasex_amd import pointer_array

    d.pA, d.pB, d.pC = pointer_array(d.A), pointer_array(d.B), pointer_array(d.C)
    for cfg in ("direct32", "tile64", "tile128"):
        os.environ["SSBLAS_TILE_CONFIG"] = cfg
        for x, c0 in zip(d.C, b.C0):
            x.copy_(torch.from_numpy(c0))
        st, C = run_gpu(b, dev=d)
        assert st == 0
        assert padding_untouched(b, C)
        if t == "i8":
            _, R = run_oracle(b)
            assert all(np.array_equal(x, y) for x, y in zip(C, R))
        else:
            _, R = run_oracle(b)
            assert normwise(b, C) <= fp_bar(t, normwise(b, R))


@pytest.mark.parametrize("entry", ["64", "async"])
def test_other_entry_points(entry):
    check(make_batch("i8", 1, 1, 96, 80, 72, 4, alpha=2, beta=1, seed=31), entry=entry)
    check(make_batch("f64", 0, 1, 40, 33, 20, 4, alpha=1.0, beta=0.25, seed=32), entry=entry)


@pytest.mark.parametrize("form,ct,factor", [("bf16x6", 5, 4.0)])
@pytest.mark.parametrize("ta,tb", OPS)
def test_bf16x6_matches_fp32_accuracy(form, ct, factor, ta, tb):
    for shape in [(16, 16, 32), (64, 64, 64), (150, 97, 203)]:
        b = make_batch("f32", ta, tb, *shape, 2, alpha=1.5, beta=-0.5, seed=41, dist="normal")
        check(b, compute_type=ct, factor=factor)


@pytest.mark.parametrize("ta,tb", OPS)
def test_bf16x3_error_bound(ta, tb):
    for shape in [(16, 16, 32), (64, 64, 64), (150, 97, 203)]:
        b = make_batch("f32", ta, tb, *shape, 2, alpha=1.5, beta=-0.5, seed=43, dist="normal")
        st, C = run_gpu(b, compute_type=4)
        assert st == 0 and padding_untouched(b, C)
        assert normwise(b, C) <= 2.0 ** -14


def test_beta_zero_ignores_nan_in_c():
    for t in ("f32", "f64"):
        b = make_batch(t, 0, 0, 64, 64, 64, 2, alpha=1.0, beta=0.0, seed=5)
        for c in b.C0:
            oracle.colmajor(c, 64, 64, b.ldc)[:] = np.nan
        st, C = run_gpu(b)
        assert st == 0
        assert all(np.isfinite(oracle.colmajor(c, 64, 64, b.ldc)).all() for c in C)


# --------------------------------------------------------------------------
# benchmark configurations at full size (properties) and in samples (oracle)
# --------------------------------------------------------------------------
def test_config_int8_512cubed_x64_bit_exact():
    """int8 TT m=n=k=512 batch 64, beta = 0 then beta = 1: bit-exact on all
    64 instances."""
    b0 = make_batch("i8", 1, 1, 512, 512, 512, 64, alpha=1, beta=0, seed=0x55B1A5 + 2)
    check(b0)
    b1 = make_batch("i8", 1, 1, 512, 512, 512, 64, alpha=1, beta=1, seed=0x55B1A5 + 2)
    check(b1)


def test_config_fp32_64cubed_x128():
    check(make_batch("f32", 0, 0, 64, 64, 64, 128, alpha=1.0, beta=0.5, seed=0x55B1A5 + 1))
