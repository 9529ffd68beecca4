"""ssblasGemmBatchedExNoAlloc (no counterpart in the reference): the
asynchronous batched call that allocates no device memory, its Ozaki
workspace taken from a fixed per-GPU scratch array.  Through the C ABI its
results must equal ssblasGemmBatchedExAsync's bit for bit (the same kernels
on the same inputs), on two calls in a row, the second smaller, so that the
second reads a scratch array still holding the first call's data."""
from __future__ import annotations

import numpy as np
import pytest

import ssblasex_amd as S
from helpers import DeviceBatch, make_batch

pytestmark = pytest.mark.gpu


def run(entry, b, ct):
    import torch

    d = DeviceBatch(b)
    fn = S.ssblasGemmBatchedExNoAlloc if entry == "noalloc" else S.ssblasGemmBatchedExAsync
    st = fn(b.transa, b.transb, b.m, b.n, b.k, b.alpha, d.pA, b.ab_code, b.lda, d.pB, b.ab_code, b.ldb, b.beta, d.pC,
            b.c_code, b.ldc, b.batch, ct, stream=torch.cuda.current_stream())
    return st, d.results()


CASES = [
    # (type, compute type, first call (ta, tb, m, n, k, batch), second smaller call)
    ("f64", 0x109, (0, 0, 96, 80, 70, 6), (1, 0, 64, 48, 40, 3)),   # general Ozaki path: slices in scratch
    ("f64", 0x107, (1, 1, 130, 66, 200, 4), (0, 1, 33, 35, 37, 2)),
    ("f64", 6, (0, 1, 100, 100, 100, 5), (0, 1, 20, 30, 10, 7)),     # then the fused form (no scratch)
    ("i8", 0, (1, 1, 256, 256, 256, 4), (0, 0, 70, 45, 99, 2)),      # plain forms: no scratch at all
    ("f32", 5, (1, 0, 256, 256, 128, 3), (0, 0, 64, 64, 64, 2)),
]


@pytest.mark.parametrize("t,ct,first,second", CASES, ids=[f"{c[0]}-{c[1]:#x}" for c in CASES])
def test_noalloc_matches_async(t, ct, first, second):
    for i, (ta, tb, m, n, k, batch) in enumerate((first, second)):
        alpha, beta = (2, -1) if t == "i8" else (1.5, -0.5)
        b = make_batch(t, ta, tb, m, n, k, batch, alpha=alpha, beta=beta, pad=(1, 2, 3), seed=100 + i,
                       dist="wide" if t == "f64" else "normal")
        st1, C1 = run("async", b, ct)
        st2, C2 = run("noalloc", b, ct)
        assert st1 == S.Status.SUCCESS and st2 == S.Status.SUCCESS
        for x, y in zip(C1, C2):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (t, ct, i)


def test_noalloc_refuses_beyond_scratch():
    import torch

    # 64 instances of fp64 256^3 at 9 slices need 72 MiB of slices: more than the scratch
    b = make_batch("f64", 0, 0, 256, 256, 256, 1, seed=5)
    d = DeviceBatch(b)
    big = torch.zeros(64, dtype=torch.int64, device="cuda") + int(d.pA[0].item())
    bigb = torch.zeros(64, dtype=torch.int64, device="cuda") + int(d.pB[0].item())
    bigc = torch.zeros(64, dtype=torch.int64, device="cuda") + int(d.pC[0].item())
    st = S.ssblasGemmBatchedExNoAlloc(0, 0, 256, 256, 256, 1.0, big, S.DataType.R_64F, 256, bigb, S.DataType.R_64F,
                                      256, 0.0, bigc, S.DataType.R_64F, 256, 64, 0x109)
    assert st == S.Status.ALLOC_FAILED
    torch.cuda.synchronize()
