"""CPU oracle for ssblasGemmBatchedEx -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
use this package, and only as the checker or the timed CPU baseline.  The
library (ssblasex_amd) never imports it.

ctypes binding of oracle/libssblas_oracle.so (see ssblas_oracle.c for the
reference file:line each function restates) plus numpy helpers:

* gemm_batched_ex   restatement of the reference's ssblasGemmBatchedEx
                    (status table + exact fp operation order), host arrays
* gemm_dd           double-double FP64 oracle, returns (hi, lo)
* gemm_dd_batched   the same over a batch, parallel over instances
* naive_baseline_i8 the example's OpenMP int8 baseline loop
* blas_baseline     the example's per-instance cblas_?gemm loop (scipy's OpenBLAS)
* colmajor / normwise_error  helpers for column-major buffers

Parity status of each function is recorded in tests/test_oracle.py, which
pins them against closed forms, exact rational arithmetic and numpy.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_SO = _DIR / "libssblas_oracle.so"
_lib = None

R_8I, R_32I, R_32F, R_64F = 0, 3, 4, 5
_NP = {R_8I: np.int8, R_32I: np.int32, R_32F: np.float32, R_64F: np.float64}
_C_OF = {R_8I: R_32I, R_32F: R_32F, R_64F: R_64F}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_DIR)], check=True)
    return _SO


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _SO.exists():
            build()
        lib = ctypes.CDLL(str(_SO))
        L, I, V = ctypes.c_long, ctypes.c_int, ctypes.c_void_p
        lib.oracle_validate.restype = I
        lib.oracle_validate.argtypes = [I, I, L, L, L, I, L, I, L, I, L, L]
        lib.oracle_gemm_batched_ex.restype = I
        lib.oracle_gemm_batched_ex.argtypes = [I, I, L, L, L, V, V, I, L, V, I, L, V, V, I, L, L, I]
        lib.oracle_gemm_dd.restype = None
        lib.oracle_gemm_dd.argtypes = [I, I, L, L, L, ctypes.c_double, V, I, L, V, L, ctypes.c_double, V, L, V, V]
        lib.oracle_naive_baseline_i8.restype = None
        lib.oracle_naive_baseline_i8.argtypes = [I, I, L, L, L, ctypes.c_int32, V, L, V, L, ctypes.c_int32, V, L, L, I]
        lib.oracle_gemm_dd_batched.restype = None
        lib.oracle_gemm_dd_batched.argtypes = [I, I, L, L, L, ctypes.c_double, V, I, L, V, L, ctypes.c_double, V, L,
                                               V, V, L, I]
        lib.oracle_blas_baseline.restype = I
        lib.oracle_blas_baseline.argtypes = [ctypes.c_char_p, I, I, I, L, L, L, ctypes.c_double, V, L, V, L,
                                             ctypes.c_double, V, L, L, I]
        lib.oracle_max_threads.restype = I
        _lib = lib
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _ptr_array(arrs) -> ctypes.Array:
    return (ctypes.c_void_p * len(arrs))(*[_ptr(a) for a in arrs])


def scale_ctype(Atype):
    return {R_32F: ctypes.c_float, R_64F: ctypes.c_double, R_8I: ctypes.c_int32}[Atype]


def validate(transa, transb, m, n, k, Atype, lda, Btype, ldb, Ctype, ldc, batch) -> int:
    return load().oracle_validate(transa, transb, m, n, k, Atype, lda, Btype, ldb, Ctype, ldc, batch)


def gemm_batched_ex(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype, ldc,
                    batch=None, nthreads=1) -> int:
    """Restated reference call on lists of host numpy buffers (C updated in place)."""
    batch = len(C) if batch is None else batch
    st = scale_ctype(Atype) if Atype in (R_32F, R_64F, R_8I) else ctypes.c_double
    a, b = st(alpha), st(beta)
    pa, pb, pc = _ptr_array(A), _ptr_array(B), _ptr_array(C)
    return load().oracle_gemm_batched_ex(transa, transb, m, n, k, ctypes.addressof(a), ctypes.addressof(pa),
                                         Atype, lda, ctypes.addressof(pb), Btype, ldb, ctypes.addressof(b),
                                         ctypes.addressof(pc), Ctype, ldc, batch, nthreads)


def gemm_dd(transa, transb, m, n, k, alpha, A, Atype, lda, B, ldb, beta, C0, ldc):
    """Double-double alpha*op(A)op(B) + beta*C0 for one instance -> (hi, lo),
    each a dense (m, n) float64 array."""
    hi = np.empty(m * n, dtype=np.float64)
    lo = np.empty(m * n, dtype=np.float64)
    c0 = C0 if C0 is not None else np.zeros(1, dtype=_NP[_C_OF[Atype]])
    load().oracle_gemm_dd(transa, transb, m, n, k, float(alpha), _ptr(A), Atype, lda, _ptr(B), ldb,
                          float(beta), _ptr(c0), ldc, _ptr(hi), _ptr(lo))
    return hi.reshape(n, m).T, lo.reshape(n, m).T


def gemm_dd_batched(transa, transb, m, n, k, alpha, A, Atype, lda, B, ldb, beta, C0, ldc, nthreads=1):
    """oracle_gemm_dd over lists of host buffers, parallel over instances ->
    (hi, lo), each (batch, m, n) float64."""
    batch = len(A)
    hi = np.empty(batch * m * n, dtype=np.float64)
    lo = np.empty(batch * m * n, dtype=np.float64)
    pa, pb = _ptr_array(A), _ptr_array(B)
    pc = _ptr_array(C0 if beta != 0 else A)  # C0 unread when beta == 0
    load().oracle_gemm_dd_batched(transa, transb, m, n, k, float(alpha), ctypes.addressof(pa), Atype, lda,
                                  ctypes.addressof(pb), ldb, float(beta), ctypes.addressof(pc), ldc, _ptr(hi),
                                  _ptr(lo), batch, nthreads)
    return hi.reshape(batch, n, m).transpose(0, 2, 1), lo.reshape(batch, n, m).transpose(0, 2, 1)


def find_openblas():
    """Path of the OpenBLAS shipped with scipy (scipy.libs/libscipy_openblas*.so), or None."""
    import glob
    import importlib.util

    spec = importlib.util.find_spec("scipy")
    if spec is None or spec.origin is None:
        return None
    site = Path(spec.origin).resolve().parent.parent
    hits = sorted(glob.glob(str(site / "scipy.libs" / "libscipy_openblas*.so")))
    return hits[0] if hits else None


def blas_baseline(libpath, Atype, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nthreads) -> int:
    """The example's per-instance cblas_?gemm loop (C updated in place); -1
    when the library or its symbols are missing."""
    pa, pb, pc = _ptr_array(A), _ptr_array(B), _ptr_array(C)
    return load().oracle_blas_baseline(str(libpath).encode(), Atype, transa, transb, m, n, k, float(alpha),
                                       ctypes.addressof(pa), lda, ctypes.addressof(pb), ldb, float(beta),
                                       ctypes.addressof(pc), ldc, len(C), nthreads)


def naive_baseline_i8(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nthreads):
    pa, pb, pc = _ptr_array(A), _ptr_array(B), _ptr_array(C)
    load().oracle_naive_baseline_i8(transa, transb, m, n, k, int(alpha), ctypes.addressof(pa), lda,
                                    ctypes.addressof(pb), ldb, int(beta), ctypes.addressof(pc), ldc, len(C),
                                    nthreads)


def max_threads() -> int:
    return load().oracle_max_threads()


def colmajor(buf: np.ndarray, rows: int, cols: int, ld: int) -> np.ndarray:
    """(rows, cols) view of a column-major buffer with leading dimension ld."""
    return buf[: ld * cols].reshape(cols, ld)[:, :rows].T


def op_dims(trans: int, rows_of_op: int, cols_of_op: int):
    """(stored rows, stored cols) of X for op(X) of the given shape."""
    return (rows_of_op, cols_of_op) if trans == 0 else (cols_of_op, rows_of_op)


def normwise_error(C: np.ndarray, hi: np.ndarray, lo: np.ndarray, scale: float) -> float:
    """||C - (hi + lo)||_F / scale, the difference taken in double-double
    order ((C - hi) - lo) so the oracle's own rounding does not enter."""
    d = (C.astype(np.float64) - hi) - lo
    return float(np.sqrt(np.sum(d * d)) / scale) if scale > 0 else float(np.sqrt(np.sum(d * d)))
