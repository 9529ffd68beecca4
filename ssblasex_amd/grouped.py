"""Python mirror of the grouped entry points (include/ssblasBatch.h).

Per GROUP (sequences of length group_count): group_size, transa, transb,
Atype, Btype, Ctype, alpha, beta.  Per INSTANCE (sequences of length
sum(group_size), instances numbered group by group): m, n, k, lda, ldb, ldc.
A, B, C: device arrays of device pointers (torch int64 tensors or
addresses), as for the batched call.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

from . import _lib


def _enum_array(vals):
    return (ctypes.c_int * len(vals))(*[int(v) for v in vals])


def _long_array(vals):
    return (ctypes.c_long * len(vals))(*[int(v) for v in vals])


def _grouped_args(group_size, transa, transb, Atype, Btype, Ctype, alpha, beta, m, n, k, A, lda, B, ldb, C, ldc,
                  computeType, algo):
    from . import _addr, _scalar, scale_ctype

    G = len(group_size)
    for name, seq in (("transa", transa), ("transb", transb), ("Atype", Atype), ("Btype", Btype),
                      ("Ctype", Ctype), ("alpha", alpha), ("beta", beta)):
        if len(seq) != G:
            raise ValueError(f"{name} has {len(seq)} entries for {G} groups")
    total = int(sum(int(s) for s in group_size)) if G else 0
    for name, seq in (("m", m), ("n", n), ("k", k), ("lda", lda), ("ldb", ldb), ("ldc", ldc)):
        if len(seq) != max(total, 0):
            raise ValueError(f"{name} has {len(seq)} entries for {total} instances")
    scal_a = [_scalar(alpha[g], scale_ctype(Atype[g], Btype[g], Ctype[g])) for g in range(G)]
    scal_b = [_scalar(beta[g], scale_ctype(Atype[g], Btype[g], Ctype[g])) for g in range(G)]
    pa = (ctypes.c_void_p * max(G, 1))(*[ctypes.addressof(x) for x in scal_a])
    pb = (ctypes.c_void_p * max(G, 1))(*[ctypes.addressof(x) for x in scal_b])
    keep = [scal_a, scal_b, pa, pb]

    def arr(x, mk):
        a = mk(x) if len(x) else None
        keep.append(a)
        return ctypes.cast(a, ctypes.c_void_p) if a is not None else None

    args = (G, arr(group_size, _long_array), arr(transa, _enum_array), arr(transb, _enum_array),
            arr(Atype, _enum_array), arr(Btype, _enum_array), arr(Ctype, _enum_array),
            ctypes.cast(pa, ctypes.c_void_p), ctypes.cast(pb, ctypes.c_void_p),
            arr(m, _long_array), arr(n, _long_array), arr(k, _long_array),
            _addr(A), arr(lda, _long_array), _addr(B), arr(ldb, _long_array), _addr(C), arr(ldc, _long_array),
            int(computeType), int(algo))
    return args, keep


def _stream_handle(stream) -> int:
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0)


class GroupedPlan:
    """A grouped plan (ssblasGroupedPlanCreate).  Raises ValueError carrying
    the status when creation fails; use `GroupedPlan.create` to get the
    status instead."""

    def __init__(self, group_size: Sequence[int], transa, transb, Atype, Btype, Ctype, alpha, beta, m, n, k,
                 A, lda, B, ldb, C, ldc, computeType=0, algo=0):
        from . import Status

        st, handle = self._create(group_size, transa, transb, Atype, Btype, Ctype, alpha, beta, m, n, k, A, lda,
                                  B, ldb, C, ldc, computeType, algo)
        if st != Status.SUCCESS:
            raise ValueError(f"ssblasGroupedPlanCreate failed: {st.name}", st)
        self._handle = handle

    @staticmethod
    def _create(*a):
        from . import Status

        args, keep = _grouped_args(*a)
        h = ctypes.c_void_p(0)
        st = Status(_lib.load().ssblasGroupedPlanCreate(ctypes.byref(h), *args))
        return st, h

    @classmethod
    def create(cls, *a, **kw):
        """(Status, plan or None)."""
        from . import Status

        names = ["group_size", "transa", "transb", "Atype", "Btype", "Ctype", "alpha", "beta", "m", "n", "k", "A",
                 "lda", "B", "ldb", "C", "ldc", "computeType", "algo"]
        vals = list(a) + [kw[nm] for nm in names[len(a):] if nm in kw]
        if len(vals) < 17:
            raise TypeError("missing arguments")
        if len(vals) == 17:
            vals += [kw.get("computeType", 0), kw.get("algo", 0)]
        elif len(vals) == 18:
            vals += [kw.get("algo", 0)]
        st, h = cls._create(*vals)
        if st != Status.SUCCESS:
            return st, None
        plan = cls.__new__(cls)
        plan._handle = h
        return st, plan

    def execute(self, stream=0):
        from . import Status

        return Status(_lib.load().ssblasGroupedPlanExecute(self._handle, _stream_handle(stream)))

    def graph_launch(self, stream=0):
        from . import Status

        return Status(_lib.load().ssblasGroupedPlanGraphLaunch(self._handle, _stream_handle(stream)))

    def close(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            _lib.load().ssblasGroupedPlanDestroy(h)
        self._handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ssblasGemmGroupedBatchedEx(group_size, transa, transb, Atype, Btype, Ctype, alpha, beta, m, n, k, A, lda, B,
                               ldb, C, ldc, computeType=0, algo=0):
    """One-shot synchronous grouped call; returns the Status."""
    from . import Status

    args, keep = _grouped_args(group_size, transa, transb, Atype, Btype, Ctype, alpha, beta, m, n, k, A, lda, B,
                               ldb, C, ldc, computeType, algo)
    return Status(_lib.load().ssblasGemmGroupedBatchedEx(*args))
