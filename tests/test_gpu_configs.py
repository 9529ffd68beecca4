"""The benchmark configurations at their full batch sizes, called exactly as
the benchmark calls them (one ssblasGemmBatchedEx over the whole batch, so
the tuned kernel for that batch size is the one that runs), every instance
checked.

* fp64 NT m=n=k=32, batchCount 65536, entries U[-1,1]*2^U{-8..8}: native
  and the Ozaki form (3..9 slices and automatic), every instance against the
  double-double oracle, and the reference-order restatement (the oracle) on
  every instance for the bar.
* fp32 TN m=n=k=256, batchCount 4096, N(0,1), alpha 1.5, beta -0.5, with
  lda=ldb=ldc=256 and 264: native, BF16x3 and BF16x6, every instance against
  an fp64 product computed on the GPU by torch (itself checked against the
  double-double oracle on a seeded sample of 16 instances, where the forms
  are also checked directly against the double-double oracle) and the
  reference-order restatement on every instance for the bar; C outside the
  m x n block untouched.

Bars (normwise error ||C - C_exact||_F / (|alpha| ||A|| ||B|| + |beta| ||C0||),
worst instance of the batch):
* native fp32 / fp64: <= 2 x the reference-order restatement's + 1 unit roundoff
* BF16x6: <= 4 x the reference-order restatement's + 1 unit roundoff
* BF16x3: <= 2^-14
* Ozaki, automatic slice count: <= the reference-order restatement's + 1 unit
  roundoff; more slices never lose accuracy.

$SSBLAS_ACCURACY_OUT (a JSON path) receives the errors and kernel names.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import oracle
from helpers import UNIT_ROUNDOFF, FlatBatch, FlatDeviceBatch, normwise

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 4)
RESULTS = {}


def record(cfg, key, value):
    RESULTS.setdefault(cfg, {})[key] = value
    out = os.environ.get("SSBLAS_ACCURACY_OUT")
    if out:
        prev = {}
        if os.path.exists(out):
            with open(out) as f:
                prev = json.load(f)
        prev.update(RESULTS)
        with open(out, "w") as f:
            json.dump(prev, f, indent=1)


def kernel_of(fb, ct):
    import ssblasex_amd as S

    return S.kernel_config(fb.transa, fb.transb, fb.m, fb.n, fb.k, fb.ab_code, fb.ab_code, fb.c_code, fb.lda, fb.ldb,
                           fb.ldc, fb.batch, ct)


def reference_order(fb):
    """The oracle's restatement of the reference on every instance -> (batch, ldc*n)."""
    C = fb.C0flat.copy()
    st = oracle.gemm_batched_ex(fb.transa, fb.transb, fb.m, fb.n, fb.k, fb.alpha, fb.A, fb.ab_code, fb.lda, fb.B,
                                fb.ab_code, fb.ldb, fb.beta, list(C), fb.c_code, fb.ldc, nthreads=NTHREADS)
    assert st == 0
    return C


# --------------------------------------------------------------------------
# fp64 NT 32^3 x 65536
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3():
    import torch

    fb = FlatBatch("f64", 0, 1, 32, 32, 32, 65536, alpha=1.0, beta=0.0, seed=0x55B1A5 + 3, dist="wide")
    hi, lo = oracle.gemm_dd_batched(0, 1, 32, 32, 32, 1.0, fb.A, fb.ab_code, fb.lda, fb.B, fb.ldb, 0.0, fb.C0,
                                    fb.ldc, nthreads=NTHREADS)
    e_ref = fb.errors(reference_order(fb), hi, lo)
    d = FlatDeviceBatch(fb)
    yield fb, d, hi, lo, e_ref
    del d
    torch.cuda.empty_cache()


def test_c3_fp64_native_full_batch(c3):
    fb, d, hi, lo, e_ref = c3
    cfg = "fp64 NT 32^3 x65536 (all instances)"
    st, C = d.run(0)
    assert st == 0
    e = fb.errors(C, hi, lo)
    record(cfg, "reference_order", {"max": float(e_ref.max()), "mean": float(e_ref.mean())})
    record(cfg, "native", {"kernel": kernel_of(fb, 0), "max": float(e.max()), "mean": float(e.mean()),
                           "ratio_max": float(e.max() / e_ref.max())})
    assert np.all(np.isfinite(e))
    assert e.max() <= 2 * e_ref.max() + UNIT_ROUNDOFF["f64"], (e.max(), e_ref.max())


def test_c3_fp64_ozaki_full_batch(c3):
    fb, d, hi, lo, e_ref = c3
    cfg = "fp64 NT 32^3 x65536 (all instances)"
    worst = {}
    for s in range(3, 10):
        st, C = d.run(0x100 + s)
        assert st == 0
        e = fb.errors(C, hi, lo)
        assert np.all(np.isfinite(e)), s
        worst[s] = float(e.max())
        record(cfg, f"ozaki_s{s}", {"kernel": kernel_of(fb, 0x100 + s), "max": worst[s], "mean": float(e.mean()),
                                    "ratio_max": worst[s] / float(e_ref.max())})
    # more slices never lose accuracy (down to the fp64 rounding floor)
    for s in range(3, 9):
        assert worst[s + 1] <= max(worst[s], 2.0 ** -52), (s, worst)
    st, C = d.run(6)
    assert st == 0
    e = fb.errors(C, hi, lo)
    record(cfg, "ozaki_auto", {"kernel": kernel_of(fb, 6), "max": float(e.max()), "mean": float(e.mean()),
                               "ratio_max": float(e.max() / e_ref.max())})
    assert e.max() <= e_ref.max() + UNIT_ROUNDOFF["f64"], (e.max(), e_ref.max())


# --------------------------------------------------------------------------
# fp32 TN 256^3 x 4096, ld 256 and 264
# --------------------------------------------------------------------------
def fp64_product(fb, d):
    """alpha * op(A) op(B) + beta * C0 for every instance in fp64 on the GPU
    (torch) -> (batch, m, n) float64 device tensor."""
    import torch

    ra, ca = fb.dims_a()
    rb, cb = fb.dims_b()
    out = torch.empty((fb.batch, fb.m, fb.n), dtype=torch.float64, device="cuda")
    C0 = torch.from_numpy(fb.C0flat).to("cuda")
    step = 512
    for i0 in range(0, fb.batch, step):
        i1 = min(fb.batch, i0 + step)
        Av = d.A[i0:i1].view(i1 - i0, ca, fb.lda)[:, :, :ra].double()
        Bv = d.B[i0:i1].view(i1 - i0, cb, fb.ldb)[:, :, :rb].double()
        opA = (Av if fb.transa else Av.transpose(1, 2)).contiguous()
        opB = (Bv.transpose(1, 2) if fb.transb == 0 else Bv).contiguous()
        r = torch.bmm(opA, opB) * fb.alpha
        if fb.beta != 0:
            r += fb.beta * C0[i0:i1].view(i1 - i0, fb.n, fb.ldc)[:, :, :fb.m].double().transpose(1, 2)
        out[i0:i1] = r
    return out


def gpu_errors(fb, Cflat, R, scale):
    """Per-instance ||C - R||_F / scale with C a host (batch, ldc*n) array."""
    import torch

    C = torch.from_numpy(Cflat).to("cuda").view(fb.batch, fb.n, fb.ldc)[:, :, :fb.m].double().transpose(1, 2)
    return (torch.linalg.vector_norm(C - R, dim=(1, 2)).cpu().numpy() / scale)


C4_FORMS = [("native", 0, lambda e, r: e <= 2 * r + UNIT_ROUNDOFF["f32"]),
            ("bf16x6", 5, lambda e, r: e <= 4 * r + UNIT_ROUNDOFF["f32"]),
            ("bf16x3", 4, lambda e, r: e <= 2.0 ** -14)]


@pytest.mark.parametrize("pad", [0, 8], ids=["ld256", "ld264"])
def test_c4_fp32_tn_full_batch(pad):
    import torch

    cfg = f"fp32 TN 256^3 x4096 alpha=1.5 beta=-0.5 ld={256 + pad} (all instances)"
    fb = FlatBatch("f32", 1, 0, 256, 256, 256, 4096, alpha=1.5, beta=-0.5, pad=pad, seed=0x55B1A5 + 4,
                   dist="normal")
    d = FlatDeviceBatch(fb)
    scale = fb.norms()
    R = fp64_product(fb, d)
    # the fp64 product against the double-double oracle on a seeded sample
    idx = np.random.default_rng(7).choice(fb.batch, 16, replace=False)
    sb = fb.sample(idx)
    Rs = R[torch.as_tensor(idx, device="cuda")].cpu().numpy()
    for j, i in enumerate(idx):
        hi, lo = oracle.gemm_dd(1, 0, 256, 256, 256, fb.alpha, sb.A[j], fb.ab_code, fb.lda, sb.B[j], fb.ldb, fb.beta,
                                sb.C0[j], fb.ldc)
        assert np.sqrt((((Rs[j] - hi) - lo) ** 2).sum()) / scale[i] <= 1e-13
    e_ref = gpu_errors(fb, reference_order(fb), R, scale)
    record(cfg, "reference_order", {"max": float(e_ref.max()), "mean": float(e_ref.mean())})
    # the sample's reference-order error against the double-double oracle
    Cref_s = [c.copy() for c in sb.C0]
    oracle.gemm_batched_ex(1, 0, 256, 256, 256, fb.alpha, sb.A, fb.ab_code, fb.lda, sb.B, fb.ab_code, fb.ldb, fb.beta,
                           Cref_s, fb.c_code, fb.ldc, nthreads=NTHREADS)
    ref_sample = normwise(sb, Cref_s)
    for name, ct, bar in C4_FORMS:
        st, C = d.run(ct)
        assert st == 0
        assert fb.padding_untouched(C), name
        e = gpu_errors(fb, C, R, scale)
        e_sample = normwise(sb, [C[i] for i in idx])
        record(cfg, name, {"kernel": kernel_of(fb, ct), "max": float(e.max()), "mean": float(e.mean()),
                           "ratio_max": float(e.max() / e_ref.max()), "sample16_vs_dd": e_sample,
                           "sample16_reference_order_vs_dd": ref_sample})
        assert np.all(np.isfinite(e)), name
        assert bar(e.max(), e_ref.max()), (name, e.max(), e_ref.max())
        assert bar(e_sample, ref_sample), (name, e_sample, ref_sample)
    del d, R
    torch.cuda.empty_cache()
