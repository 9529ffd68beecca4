"""ssblasex_amd -- MI355X (gfx950) implementation of ssblasGemmBatchedEx.

The product is the C library ssblasex_amd/lib/libssblasGemmBatchedEx.so
(C ABI in include/ssblasBatch.h).  This package is its Python mirror:

* the reference's enums with the same names and values
  (include/ssblasBatch_type.h of hirokitokura/ssblasEx, lines 14-79);
* ``ssblasGemmBatchedEx`` / ``ssblasGemmBatchedEx_64`` /
  ``ssblasGemmBatchedExAsync`` taking the same 19 arguments as the C entry
  points (scalars as Python numbers, pointer arrays as device addresses) and
  returning the ``Status``;
* ``pointer_array`` to build the device array of device pointers from torch
  tensors, and the grouped-plan wrapper ``GroupedPlan``.

torch is only used here for device memory; every computation runs in the
HIP kernels of the library.
"""
from __future__ import annotations

import ctypes
import enum
from typing import Sequence

from . import _lib


class Status(enum.IntEnum):
    SUCCESS = 0
    ALLOC_FAILED = 1
    INVALID_VALUE = 2
    INTERNAL_ERROR = 3
    NOTIMPLEMENTED_ERROR = 4
    NOT_COMPUTED_ERROR = 5


class Operation(enum.IntEnum):
    N = 0
    T = 1
    C = 2


class DataType(enum.IntEnum):
    R_8I = 0
    R_8U = 1
    R_16F = 2
    R_32I = 3
    R_32F = 4
    R_64F = 5


class ComputeType(enum.IntEnum):
    DEFAULT_TYPE = 0
    C32I = 1
    C32F = 2
    C64F = 3
    F32_FAST_BF16X3 = 4
    F32_EMULATED_BF16X6 = 5
    F64_EMULATED_OZAKI = 6
    F64_EMULATED_OZAKI_S3 = 0x103
    F64_EMULATED_OZAKI_S4 = 0x104
    F64_EMULATED_OZAKI_S5 = 0x105
    F64_EMULATED_OZAKI_S6 = 0x106
    F64_EMULATED_OZAKI_S7 = 0x107
    F64_EMULATED_OZAKI_S8 = 0x108
    F64_EMULATED_OZAKI_S9 = 0x109

    @staticmethod
    def ozaki(slices: int) -> "ComputeType":
        return ComputeType(0x100 + slices)


class GemmAlgo(enum.IntEnum):
    DEFAULT = 0


# reference names (include/ssblasBatch_type.h)
SSBLAS_STATUS_SUCCESS = Status.SUCCESS
SSBLAS_STATUS_ALLOC_FAILED = Status.ALLOC_FAILED
SSBLAS_STATUS_INVALID_VALUE = Status.INVALID_VALUE
SSBLAS_STATUS_INTERNAL_ERROR = Status.INTERNAL_ERROR
SSBLAS_STATUS_NOTIMPLEMENTED_ERROR = Status.NOTIMPLEMENTED_ERROR
SSBLAS_STATUS_NOT_COMPUTED_ERROR = Status.NOT_COMPUTED_ERROR
SSBLAS_OP_N, SSBLAS_OP_T, SSBLAS_OP_C = Operation.N, Operation.T, Operation.C
SSBLAS_R_8I, SSBLAS_R_8U, SSBLAS_R_16F = DataType.R_8I, DataType.R_8U, DataType.R_16F
SSBLAS_R_32I, SSBLAS_R_32F, SSBLAS_R_64F = DataType.R_32I, DataType.R_32F, DataType.R_64F
SSBLAS_COMPUTE_DEFAULT_TYPE = ComputeType.DEFAULT_TYPE
SSBLAS_COMPUTE_32I = ComputeType.C32I
SSBLAS_COMPUTE_32F = ComputeType.C32F
SSBLAS_COMPUTE_DEFAULT = GemmAlgo.DEFAULT

_SCALE_CTYPE = {
    (DataType.R_32F, DataType.R_32F, DataType.R_32F): ctypes.c_float,
    (DataType.R_64F, DataType.R_64F, DataType.R_64F): ctypes.c_double,
    (DataType.R_8I, DataType.R_8I, DataType.R_32I): ctypes.c_int32,
}


def scale_ctype(Atype, Btype, Ctype):
    """ctypes type alpha/beta are read as for a type triple (int32 for int8:
    the reference code reads int32, its README table says R_8I)."""
    return _SCALE_CTYPE.get((DataType(Atype), DataType(Btype), DataType(Ctype)), ctypes.c_double)


def _scalar(value, ctype):
    if isinstance(value, ctypes._SimpleCData):
        return value
    return ctype(value)


def _addr(x) -> int:
    """Device address of a pointer array: int, ctypes pointer or torch tensor."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    if isinstance(x, ctypes.c_void_p):
        return int(x.value or 0)
    return int(ctypes.cast(x, ctypes.c_void_p).value or 0)


def _gemm_args(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype, ldc,
               batchCount, computeType, algo):
    st = scale_ctype(Atype, Btype, Ctype)
    a = _scalar(alpha, st)
    b = _scalar(beta, st)
    args = (int(transa), int(transb), int(m), int(n), int(k), ctypes.cast(ctypes.pointer(a), ctypes.c_void_p),
            _addr(A), int(Atype), int(lda), _addr(B), int(Btype), int(ldb),
            ctypes.cast(ctypes.pointer(b), ctypes.c_void_p), _addr(C), int(Ctype), int(ldc),
            int(batchCount), int(computeType), int(algo))
    return args, (a, b)


def ssblasGemmBatchedEx(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype,
                        ldc, batchCount, computeType=ComputeType.DEFAULT_TYPE,
                        algo=GemmAlgo.DEFAULT) -> Status:
    """C[i] = alpha*op(A[i])*op(B[i]) + beta*C[i]; synchronous.  Mirrors
    include/ssblasBatch.h:15-35 of the reference."""
    args, keep = _gemm_args(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype,
                            ldc, batchCount, computeType, algo)
    return Status(_lib.load().ssblasGemmBatchedEx(*args))


def ssblasGemmBatchedEx_64(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype,
                           ldc, batchCount, computeType=ComputeType.DEFAULT_TYPE,
                           algo=GemmAlgo.DEFAULT) -> Status:
    args, keep = _gemm_args(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype,
                            ldc, batchCount, computeType, algo)
    return Status(_lib.load().ssblasGemmBatchedEx_64(*args))


def ssblasGemmBatchedExAsync(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C,
                             Ctype, ldc, batchCount, computeType=ComputeType.DEFAULT_TYPE,
                             algo=GemmAlgo.DEFAULT, stream=0) -> Status:
    """Enqueue on `stream` (a hipStream_t address or a torch.cuda.Stream)."""
    args, keep = _gemm_args(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype,
                            ldc, batchCount, computeType, algo)
    s = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
    return Status(_lib.load().ssblasGemmBatchedExAsync(*args, s))


def ssblasGemmBatchedExNoAlloc(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C,
                               Ctype, ldc, batchCount, computeType=ComputeType.DEFAULT_TYPE,
                               algo=GemmAlgo.DEFAULT, stream=0) -> Status:
    """ssblasGemmBatchedExAsync that allocates no device memory (the Ozaki
    workspace comes from a fixed per-GPU scratch array)."""
    args, keep = _gemm_args(transa, transb, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype,
                            ldc, batchCount, computeType, algo)
    s = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
    return Status(_lib.load().ssblasGemmBatchedExNoAlloc(*args, s))


class OzakiReport(ctypes.Structure):
    """ssblasOzakiReport_t of include/ssblasBatch.h."""
    _fields_ = [("instances", ctypes.c_long), ("slices_min", ctypes.c_int), ("slices_max", ctypes.c_int),
                ("bound_max", ctypes.c_double), ("target", ctypes.c_double), ("over_target", ctypes.c_long)]


def ozaki_report(stream=0, per_instance=True):
    """(Status, OzakiReport, bounds, slices) of the last Ozaki call on
    `stream`; bounds / slices are numpy arrays (per instance) or None."""
    import numpy as np

    rep = OzakiReport()
    s = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)
    lib = _lib.load()
    st = Status(lib.ssblasGetOzakiReport(s, ctypes.byref(rep), None, None, 0))
    if st != Status.SUCCESS or not per_instance:
        return st, rep, None, None
    b = np.empty(rep.instances, dtype=np.float32)
    sl = np.empty(rep.instances, dtype=np.int8)
    st = Status(lib.ssblasGetOzakiReport(s, ctypes.byref(rep), b.ctypes.data, sl.ctypes.data, rep.instances))
    return st, rep, b, sl


NOALLOC_SCRATCH_BYTES = 64 << 20  # SSBLAS_NOALLOC_SCRATCH_BYTES of include/ssblasBatch.h


def kernel_config(transa, transb, m, n, k, Atype, Btype, Ctype, lda, ldb, ldc, batchCount,
                  computeType=0) -> str:
    """Name of the kernel (tile config or Ozaki family) the call would run."""
    buf = ctypes.create_string_buffer(64)
    st = _lib.load().ssblasGetKernelConfig(int(transa), int(transb), m, n, k, int(Atype), int(Btype), int(Ctype),
                                           lda, ldb, ldc, batchCount, int(computeType), buf, 64)
    if st != 0:
        raise ValueError(f"ssblasGetKernelConfig returned {Status(st).name}")
    return buf.value.decode()


def version() -> int:
    return int(_lib.load().ssblasGetVersion())


def pointer_array(tensors: Sequence, device=None):
    """Device int64 tensor holding the data pointers of `tensors` (the
    A[] / B[] / C[] argument)."""
    import torch

    ptrs = [int(t.data_ptr()) for t in tensors]
    dev = device if device is not None else (tensors[0].device if len(tensors) else "cuda")
    return torch.tensor(ptrs, dtype=torch.int64, device=dev)


def pointer_array_strided(base, count: int, stride_elems: int, offset_elems: int = 0):
    """Device int64 tensor of `count` pointers into one allocation: element
    offset_elems + i * stride_elems of the tensor `base` for instance i (a
    batch held in one buffer, instance i at a fixed stride)."""
    import torch

    step = base.element_size()
    start = int(base.data_ptr()) + offset_elems * step
    return torch.arange(count, dtype=torch.int64, device=base.device) * (stride_elems * step) + start


from .grouped import GroupedPlan, ssblasGemmGroupedBatchedEx  # noqa: E402,F401

__all__ = [
    "Status", "Operation", "DataType", "ComputeType", "GemmAlgo", "scale_ctype",
    "ssblasGemmBatchedEx", "ssblasGemmBatchedEx_64", "ssblasGemmBatchedExAsync", "ssblasGemmBatchedExNoAlloc",
    "ssblasGemmGroupedBatchedEx", "GroupedPlan", "pointer_array", "pointer_array_strided", "kernel_config", "OzakiReport", "ozaki_report", "version",
]
