"""The C ABI boundary, checked without a GPU: the built library loads and
exports every function include/ssblasBatch.h declares, the enum values are
the reference's, and every status that is decided before any device work
matches the oracle's status table (no compute call is made)."""
from __future__ import annotations

import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

import oracle
import ssblasex_amd as S
from ssblasex_amd import _lib
from test_oracle import STATUS_CASES

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = []
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:ssblasStatus_t|int)\s+(ssblas\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return names


def test_header_declares_reference_entry_points():
    names = declared_functions()
    # reference include/ssblasBatch.h:15-57
    assert "ssblasGemmBatchedEx" in names and "ssblasGemmBatchedEx_64" in names
    assert len(names) == len(set(names))


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = _lib.load()
    assert hasattr(lib, name)
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.library_path())], capture_output=True,
                         text=True, check=True).stdout
    assert re.search(rf"\bT {name}$", out, flags=re.M), name


def test_python_mirror_covers_every_export():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def enum_values(header_text, enum_name):
    m = re.search(r"typedef enum\s*\{([^}]*)\}\s*" + enum_name + r"\s*;", header_text)
    assert m, enum_name
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    vals, nxt = {}, 0
    for item in body.split(","):
        item = item.strip()
        if not item:
            continue
        if "=" in item:
            k, v = (s.strip() for s in item.split("="))
            nxt = int(v, 0)
        else:
            k = item
        vals[k] = nxt
        nxt += 1
    return vals


def test_enum_values_match_reference():
    text = (ROOT / "include" / "ssblasBatch_type.h").read_text()
    # values of the reference's include/ssblasBatch_type.h:14-79
    assert enum_values(text, "ssblasStatus_t") == {
        "SSBLAS_STATUS_SUCCESS": 0, "SSBLAS_STATUS_ALLOC_FAILED": 1, "SSBLAS_STATUS_INVALID_VALUE": 2,
        "SSBLAS_STATUS_INTERNAL_ERROR": 3, "SSBLAS_STATUS_NOTIMPLEMENTED_ERROR": 4,
        "SSBLAS_STATUS_NOT_COMPUTED_ERROR": 5}
    assert enum_values(text, "ssblasOperation_t") == {"SSBLAS_OP_N": 0, "SSBLAS_OP_T": 1, "SSBLAS_OP_C": 2}
    assert enum_values(text, "ssblasDataType_t") == {
        "SSBLAS_R_8I": 0, "SSBLAS_R_8U": 1, "SSBLAS_R_16F": 2, "SSBLAS_R_32I": 3, "SSBLAS_R_32F": 4,
        "SSBLAS_R_64F": 5}
    ct = enum_values(text, "ssblasComputeType_t")
    assert (ct["SSBLAS_COMPUTE_DEFAULT_TYPE"], ct["SSBLAS_COMPUTE_32I"], ct["SSBLAS_COMPUTE_32F"]) == (0, 1, 2)
    assert min(v for k, v in ct.items() if k not in (
        "SSBLAS_COMPUTE_DEFAULT_TYPE", "SSBLAS_COMPUTE_32I", "SSBLAS_COMPUTE_32F")) > 2  # appended only
    assert enum_values(text, "ssblasGemmAlgo_t") == {"SSBLAS_COMPUTE_DEFAULT": 0}
    assert "typedef long int SSINT;" in text and "typedef long int SSLONG;" in text
    for name, v in ct.items():
        short = name.replace("SSBLAS_COMPUTE_", "")
        py = {"DEFAULT_TYPE": "DEFAULT_TYPE", "32I": "C32I", "32F": "C32F", "64F": "C64F"}.get(short, short.replace(
            "32F_", "F32_").replace("64F_", "F64_"))
        assert S.ComputeType[py] == v, name


def _call(entry, case):
    ta, tb, m, n, k, At, lda, Bt, ldb, Ct, ldc, batch = case
    fn = getattr(S, entry)
    return fn(ta, tb, m, n, k, 1.0 if At != 0 else 1, 0, At, lda, 0, Bt, ldb, 0.0 if At != 0 else 0, 0, Ct,
              ldc, batch)


@pytest.mark.parametrize("entry", ["ssblasGemmBatchedEx", "ssblasGemmBatchedEx_64", "ssblasGemmBatchedExAsync"])
@pytest.mark.parametrize("case", [c for c in STATUS_CASES if c[-2] != 0 or c[4] == 0], ids=lambda c: c[-1])
def test_status_before_device_work(entry, case):
    """Statuses decided on the host equal the oracle's table."""
    *args, expected, _src = case
    assert int(_call(entry, tuple(args))) == expected == oracle.validate(*args)


@pytest.mark.parametrize("ct,types", [
    (S.ComputeType.F32_FAST_BF16X3, (5, 5, 5)),
    (S.ComputeType.F32_EMULATED_BF16X6, (0, 0, 3)),
    (S.ComputeType.F64_EMULATED_OZAKI, (4, 4, 4)),
    (S.ComputeType.F64_EMULATED_OZAKI_S5, (0, 0, 3)),
    (S.ComputeType.C64F, (4, 4, 4)),
    (77, (4, 4, 4)),
])
def test_compute_type_must_fit_triple(ct, types):
    At, Bt, Ct = types
    st = S.ssblasGemmBatchedEx(0, 0, 4, 4, 4, 1.0 if At else 1, 0, At, 4, 0, Bt, 4, 0.0 if At else 0, 0, Ct, 4, 1,
                               computeType=ct)
    assert st == S.Status.INVALID_VALUE


def test_reference_compute_types_are_ignored_for_validation():
    # DEFAULT/32I/32F are accepted for every triple (reference ignores them);
    # with NULL arrays the call then stops at the pointer check.
    for ct in (0, 1, 2):
        for At, Ct in ((4, 4), (5, 5), (0, 3)):
            st = S.ssblasGemmBatchedEx(0, 0, 4, 4, 4, 1.0 if At else 1, 0, At, 4, 0, At, 4, 0.0 if At else 0, 0, Ct,
                                       4, 1, computeType=ct)
            assert st == S.Status.INVALID_VALUE  # NULL pointer arrays


def test_debug_dump_format():
    """SSBLAS_GEMMBATCHEDEX_DEBUG prints the argument block of the reference
    (Impl/GemmBatchedEx/ssblasGemmBatchedEx.cpp:48-93)."""
    code = ("import ssblasex_amd as S; "
            "S.ssblasGemmBatchedEx(1, 1, 1024, 1024, 1024, 1, 0, 0, 1024, 0, 0, 1024, 0, 0, 3, 1024, 3)")
    env = dict(os.environ, SSBLAS_GEMMBATCHEDEX_DEBUG="1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, cwd=ROOT,
                         check=True).stdout
    lines = ["ssblasGemmBatchedExShowdebug: START", "transa: SSBLAS_OP_T", "transb: SSBLAS_OP_T", "M: 1024",
             "N: 1024", "K: 1024", "alpha: ", "A: ", "Atype: SSBLAS_R_8I", "lda: 1024", "B: ",
             "Btype: SSBLAS_R_8I", "ldb: 1024", "beta: ", "C: ", "Ctype: SSBLAS_R_32I", "ldc: 1024",
             "batchCount: 3", "computeType: SSBLAS_COMPUTE_DEFAULT_TYPE", "algo: SSBLAS_COMPUTE_DEFAULT",
             "ssblasGemmBatchedExShowdebug: END"]
    expected = "\n".join(lines) + "\n"
    assert out == expected


def test_no_oracle_in_product():
    """The shipped package never imports the oracle."""
    for p in (ROOT / "ssblasex_amd").rglob("*"):
        if p.suffix in (".py", ".cpp", ".hip", ".hpp", ".h"):
            text = p.read_text()
            assert "import oracle" not in text and "from oracle" not in text and "libssblas_oracle" not in text, p
