"""SSBLAS_STATUS_ALLOC_FAILED.  The reference never returns it (its scratch
allocation is unchecked, src/include/Utils/ssblas_malloc.hpp:8-27); here the
library-owned device workspace (the Ozaki form's slices and exponents, a
grouped plan's device image) is checked against SSBLAS_WORKSPACE_LIMIT
before any device work, so the status is decided on the host (no GPU)."""
from __future__ import annotations

import ctypes

import pytest

import ssblasex_amd as S

R_8I, R_32I, R_32F, R_64F = 0, 3, 4, 5
DUMMY = ctypes.create_string_buffer(64)  # a non-null address; never dereferenced by these calls
ADDR = ctypes.addressof(DUMMY)


@pytest.fixture
def limit(monkeypatch):
    def set_(v):
        monkeypatch.setenv("SSBLAS_WORKSPACE_LIMIT", v)
    yield set_


def ozaki_call(m=256, batch=64, ct=0x109):
    return S.ssblasGemmBatchedEx(0, 0, m, m, m, 1.0, ADDR, R_64F, m, ADDR, R_64F, m, 0.0, ADDR, R_64F, m, batch,
                                 computeType=ct)


def test_ozaki_workspace_over_limit(limit):
    # 64 instances of 256^3 at 9 slices need 9 * 2 * 256^2 * 64 bytes of slices (72 MiB)
    limit("64M")
    assert ozaki_call() == S.Status.ALLOC_FAILED
    assert ozaki_call(ct=6) == S.Status.ALLOC_FAILED  # the automatic slice count sizes for the most slices
    limit("1G")
    # within the cap the call goes on to its next check: host pointer arrays are rejected
    assert ozaki_call() == S.Status.INVALID_VALUE


def test_limit_suffixes(limit):
    limit(str(70 * 2 ** 20))  # bytes, below the need
    assert ozaki_call() == S.Status.ALLOC_FAILED
    limit("80M")
    assert ozaki_call() == S.Status.INVALID_VALUE
    limit("1g")
    assert ozaki_call() == S.Status.INVALID_VALUE


def test_plain_forms_need_no_workspace(limit):
    limit("0")
    for ct, t in ((0, R_64F), (0, R_32F), (4, R_32F), (5, R_32F)):
        st = S.ssblasGemmBatchedEx(0, 0, 256, 256, 256, 1.0, ADDR, t, 256, ADDR, t, 256, 0.0, ADDR, t, 256, 64,
                                   computeType=ct)
        assert st == S.Status.INVALID_VALUE, ct
    st = S.ssblasGemmBatchedEx(0, 0, 256, 256, 256, 1, ADDR, R_8I, 256, ADDR, R_8I, 256, 0, ADDR, R_32I, 256, 64)
    assert st == S.Status.INVALID_VALUE


def test_fused_small_ozaki_needs_only_its_report(limit):
    # m, n, k <= 32: slices stay in registers; the workspace is the
    # per-instance report (4-byte bound + 1-byte slice count, 256-byte aligned)
    limit(str(65536 * 5 + 256))
    assert ozaki_call(m=32, batch=65536) == S.Status.INVALID_VALUE
    limit(str(65536 * 5 - 1))
    assert ozaki_call(m=32, batch=65536) == S.Status.ALLOC_FAILED


def test_grouped_plan_over_limit(limit):
    sizes = [(64, 64, 64)] * 100
    args = ([100], [0], [0], [R_32F], [R_32F], [R_32F], [1.0], [0.0], [s[0] for s in sizes], [s[1] for s in sizes],
            [s[2] for s in sizes], ADDR, [64] * 100, ADDR, [64] * 100, ADDR, [64] * 100)
    limit("1K")
    st, plan = S.GroupedPlan.create(*args)
    assert plan is None and st == S.Status.ALLOC_FAILED
    limit("1G")
    st, plan = S.GroupedPlan.create(*args)
    assert plan is None and st == S.Status.INVALID_VALUE


def test_noalloc_entry_refuses_beyond_scratch(limit):
    """ssblasGemmBatchedExNoAlloc takes the Ozaki workspace from a fixed
    per-GPU scratch array of SSBLAS_NOALLOC_SCRATCH_BYTES (64 MiB): a call
    needing more is refused on the host, whatever SSBLAS_WORKSPACE_LIMIT says."""
    limit("1G")

    def call(batch, ct=0x109):
        return S.ssblasGemmBatchedExNoAlloc(0, 0, 256, 256, 256, 1.0, ADDR, R_64F, 256, ADDR, R_64F, 256, 0.0, ADDR,
                                            R_64F, 256, batch, ct)

    assert call(64) == S.Status.ALLOC_FAILED   # 72 MiB of slices
    assert call(48) == S.Status.INVALID_VALUE  # 54 MiB: fits, then the host pointer arrays are refused
    assert call(64, ct=0) == S.Status.INVALID_VALUE  # native fp64 needs no scratch
