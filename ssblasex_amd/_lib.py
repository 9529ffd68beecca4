"""ctypes binding of libssblasGemmBatchedEx.so (the C ABI in include/ssblasBatch.h).

The library is loaded from ssblasex_amd/lib/.  There is no fallback: if the
built library is missing, importing raises with the build command to run.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libssblasGemmBatchedEx.so"

_c_long = ctypes.c_long
_c_int = ctypes.c_int
_vp = ctypes.c_void_p

# (name, restype, argtypes) for every function include/ssblasBatch.h declares
_GEMM_ARGS = [
    _c_int, _c_int,                 # transa, transb
    _c_long, _c_long, _c_long,      # m, n, k
    _vp,                            # alpha (host pointer)
    _vp, _c_int, _c_long,           # A, Atype, lda
    _vp, _c_int, _c_long,           # B, Btype, ldb
    _vp,                            # beta
    _vp, _c_int, _c_long,           # C, Ctype, ldc
    _c_long, _c_int, _c_int,        # batchCount, computeType, algo
]
_GROUPED_ARGS = [
    _c_long, _vp,                   # group_count, group_size
    _vp, _vp,                       # transa_array, transb_array
    _vp, _vp, _vp,                  # Atype_array, Btype_array, Ctype_array
    _vp, _vp,                       # alpha_array, beta_array
    _vp, _vp, _vp,                  # m, n, k arrays
    _vp, _vp,                       # A, lda_array
    _vp, _vp,                       # B, ldb_array
    _vp, _vp,                       # C, ldc_array
    _c_int, _c_int,                 # computeType, algo
]
SIGNATURES = {
    "ssblasGemmBatchedEx": (_c_int, _GEMM_ARGS),
    "ssblasGemmBatchedEx_64": (_c_int, _GEMM_ARGS),
    "ssblasGemmBatchedExAsync": (_c_int, _GEMM_ARGS + [_vp]),
    "ssblasGemmBatchedExNoAlloc": (_c_int, _GEMM_ARGS + [_vp]),
    "ssblasGroupedPlanCreate": (_c_int, [_vp] + _GROUPED_ARGS),
    "ssblasGroupedPlanExecute": (_c_int, [_vp, _vp]),
    "ssblasGroupedPlanGraphLaunch": (_c_int, [_vp, _vp]),
    "ssblasGroupedPlanDestroy": (_c_int, [_vp]),
    "ssblasGemmGroupedBatchedEx": (_c_int, _GROUPED_ARGS),
    "ssblasGetKernelConfig": (_c_int, [_c_int, _c_int, _c_long, _c_long, _c_long, _c_int, _c_int, _c_int, _c_long,
                                       _c_long, _c_long, _c_long, _c_int, ctypes.c_char_p, _c_long]),
    "ssblasGetOzakiReport": (_c_int, [_vp, _vp, _vp, _vp, _c_long]),
    "ssblasGetVersion": (_c_int, []),
}

_lib = None


def library_path() -> Path:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load the built library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(
            f"{_LIB_PATH} is missing: build it with `python -m ssblasex_amd.build` "
            "(the HIP library is the only implementation; there is no fallback)")
    lib = ctypes.CDLL(str(_LIB_PATH), mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
