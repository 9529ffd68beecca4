"""Accuracy of every arithmetic form on the benchmark configurations, against
the FP64 (double-double) oracle, on seeded samples of the batches.

Writes the table to $SSBLAS_ACCURACY_OUT (JSON) when that is set, so the
error of each form can be reported beside its throughput.  Bars:
* native fp32 / fp64 <= 2 x reference-order error + 1 unit roundoff
* BF16X6 <= 4 x reference-order fp32 error + 1 unit roundoff
* BF16X3 <= 2^-14
* Ozaki, automatic slice count <= reference-order fp64 error + 1 unit
  roundoff ("no worse than the plain path" with the same rounding slack)
"""
from __future__ import annotations

import json
import os

import pytest

from helpers import UNIT_ROUNDOFF, make_batch, normwise, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

RESULTS = {}


def record(cfg, form, err, ref):
    RESULTS.setdefault(cfg, {})[form] = {"normwise_error": err, "ratio_to_reference_order": (err / ref) if ref else None}
    out = os.environ.get("SSBLAS_ACCURACY_OUT")
    if out:
        with open(out, "w") as f:
            json.dump(RESULTS, f, indent=1)


def test_fp32_64cubed_nn():
    b = make_batch("f32", 0, 0, 64, 64, 64, 32, alpha=1.0, beta=0.5, seed=0x55B1A5 + 1)
    _, R = run_oracle(b, 8)
    ref = normwise(b, R)
    record("fp32 NN 64^3 (32 of 128)", "reference_order", ref, ref)
    st, C = run_gpu(b)
    e = normwise(b, C)
    record("fp32 NN 64^3 (32 of 128)", "native", e, ref)
    assert e <= 2 * ref + UNIT_ROUNDOFF["f32"]


def test_fp64_32cubed_nt_native_and_ozaki():
    cfg = "fp64 NT 32^3 wide range (512 of 65536)"
    b = make_batch("f64", 0, 1, 32, 32, 32, 512, alpha=1.0, beta=0.0, seed=0x55B1A5 + 3, dist="wide")
    _, R = run_oracle(b, 8)
    ref = normwise(b, R)
    record(cfg, "reference_order", ref, ref)
    st, C = run_gpu(b)
    e = normwise(b, C)
    record(cfg, "native", e, ref)
    assert e <= 2 * ref + UNIT_ROUNDOFF["f64"]
    for s in range(4, 10):
        st, C = run_gpu(b, compute_type=0x100 + s)
        assert st == 0
        record(cfg, f"ozaki_s{s}", normwise(b, C), ref)
    st, C = run_gpu(b, compute_type=6)
    e_auto = normwise(b, C)
    record(cfg, "ozaki_auto", e_auto, ref)
    assert e_auto <= ref + UNIT_ROUNDOFF["f64"]


def test_fp64_256cubed_nn_ozaki():
    cfg = "fp64 NN 256^3 wide range (8 of 512)"
    b = make_batch("f64", 0, 0, 256, 256, 256, 8, alpha=1.0, beta=0.0, seed=0x55B1A5 + 5, dist="wide")
    _, R = run_oracle(b, 8)
    ref = normwise(b, R)
    record(cfg, "reference_order", ref, ref)
    st, C = run_gpu(b)
    record(cfg, "native", normwise(b, C), ref)
    for s in (7, 8, 9):
        st, C = run_gpu(b, compute_type=0x100 + s)
        record(cfg, f"ozaki_s{s}", normwise(b, C), ref)
    st, C = run_gpu(b, compute_type=6)
    e_auto = normwise(b, C)
    record(cfg, "ozaki_auto", e_auto, ref)
    assert e_auto <= ref + UNIT_ROUNDOFF["f64"]


def test_fp32_256cubed_tn_bf16_forms():
    cfg = "fp32 TN 256^3 N(0,1) alpha=1.5 beta=-0.5 (8 of 4096)"
    b = make_batch("f32", 1, 0, 256, 256, 256, 8, alpha=1.5, beta=-0.5, seed=0x55B1A5 + 4, dist="normal")
    _, R = run_oracle(b, 8)
    ref = normwise(b, R)
    record(cfg, "reference_order", ref, ref)
    st, C = run_gpu(b)
    e = normwise(b, C)
    record(cfg, "native", e, ref)
    assert e <= 2 * ref + UNIT_ROUNDOFF["f32"]
    st, C = run_gpu(b, compute_type=5)
    e6 = normwise(b, C)
    record(cfg, "bf16x6", e6, ref)
    assert e6 <= 4 * ref + UNIT_ROUNDOFF["f32"]
    st, C = run_gpu(b, compute_type=4)
    e3 = normwise(b, C)
    record(cfg, "bf16x3", e3, ref)
    assert e3 <= 2.0 ** -14
