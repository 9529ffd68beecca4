"""Shared test helpers: seeded batch generation, fixture decoding, running a
batch through the HIP library and through the oracle."""
from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden" / "closed_forms.json"

TYPES = {
    "f32": (oracle.R_32F, oracle.R_32F, np.float32, np.float32),
    "f64": (oracle.R_64F, oracle.R_64F, np.float64, np.float64),
    "i8": (oracle.R_8I, oracle.R_32I, np.int8, np.int32),
}
UNIT_ROUNDOFF = {"f32": 2.0 ** -24, "f64": 2.0 ** -53}


@dataclass
class Batch:
    """One ssblasGemmBatchedEx call on host buffers (column-major, flat)."""
    type: str
    transa: int
    transb: int
    m: int
    n: int
    k: int
    lda: int
    ldb: int
    ldc: int
    alpha: float
    beta: float
    A: list = field(default_factory=list)
    B: list = field(default_factory=list)
    C0: list = field(default_factory=list)

    @property
    def batch(self):
        return len(self.C0)

    @property
    def ab_code(self):
        return TYPES[self.type][0]

    @property
    def c_code(self):
        return TYPES[self.type][1]

    def dims_a(self):
        return oracle.op_dims(self.transa, self.m, self.k)

    def dims_b(self):
        return oracle.op_dims(self.transb, self.k, self.n)


def _fill(rng, shape_elems, kind, dtype):
    if kind == "i8":
        return rng.integers(-128, 128, size=shape_elems, dtype=np.int64).astype(np.int8)
    if kind == "i32c":
        return rng.integers(-1000, 1001, size=shape_elems, dtype=np.int64).astype(np.int32)
    if kind == "normal":
        return rng.standard_normal(shape_elems).astype(dtype)
    if kind == "wide":  # U[-1,1] * 2^U{-8..8}
        return (rng.uniform(-1, 1, shape_elems) * np.exp2(rng.integers(-8, 9, shape_elems))).astype(dtype)
    return rng.uniform(-1, 1, shape_elems).astype(dtype)


def make_batch(type_, transa, transb, m, n, k, batch, *, alpha=1.0, beta=0.0, pad=(0, 0, 0), seed=0,
               dist="uniform") -> Batch:
    """Seeded synthetic batch (numpy PCG64, platform independent).  Padding
    rows of every buffer hold a sentinel so stray reads/writes show up."""
    rng = np.random.default_rng(seed)
    _, _, abt, ct = TYPES[type_]
    ra, ca = oracle.op_dims(transa, m, k)
    rb, cb = oracle.op_dims(transb, k, n)
    lda, ldb, ldc = max(1, ra) + pad[0], max(1, rb) + pad[1], max(1, m) + pad[2]
    if type_ == "i8":
        alpha, beta = int(alpha), int(beta)
    b = Batch(type_, transa, transb, m, n, k, lda, ldb, ldc, alpha, beta)
    for _ in range(batch):
        A = np.full(lda * ca, 7, dtype=abt)
        B = np.full(ldb * cb, 7, dtype=abt)
        C = np.full(ldc * n, 3, dtype=ct)
        if type_ == "i8":
            oracle.colmajor(A, ra, ca, lda)[:] = _fill(rng, (ra, ca), "i8", abt)
            oracle.colmajor(B, rb, cb, ldb)[:] = _fill(rng, (rb, cb), "i8", abt)
            oracle.colmajor(C, m, n, ldc)[:] = _fill(rng, (m, n), "i32c", ct)
        else:
            oracle.colmajor(A, ra, ca, lda)[:] = _fill(rng, (ra, ca), dist, abt)
            oracle.colmajor(B, rb, cb, ldb)[:] = _fill(rng, (rb, cb), dist, abt)
            oracle.colmajor(C, m, n, ldc)[:] = _fill(rng, (m, n), dist, ct)
        b.A.append(A)
        b.B.append(B)
        b.C0.append(C)
    return b


def _dec(v, type_):
    if type_ == "i8":
        return int(v)
    if v == "nan":
        return float("nan")
    return float.fromhex(v)


def load_golden():
    data = json.loads(GOLDEN.read_text())
    out = []
    for c in data["cases"]:
        t = c["type"]
        _, _, abt, ct = TYPES[t]
        b = Batch(t, c["transa"], c["transb"], c["m"], c["n"], c["k"], c["lda"], c["ldb"], c["ldc"],
                  _dec(c["alpha"], t), _dec(c["beta"], t))
        b.A = [np.array([_dec(v, t) for v in a], dtype=abt) for a in c["A"]]
        b.B = [np.array([_dec(v, t) for v in x], dtype=abt) for x in c["B"]]
        b.C0 = [np.array([_dec(v, t) for v in x], dtype=ct) for x in c["C0"]]
        expect = [np.array([_dec(v, t) for v in x], dtype=ct) for x in c["expect"]]
        out.append((c["name"], c["applies"], b, expect))
    return out


def run_oracle(b: Batch, nthreads=1):
    C = [c.copy() for c in b.C0]
    st = oracle.gemm_batched_ex(b.transa, b.transb, b.m, b.n, b.k, b.alpha, b.A, b.ab_code, b.lda, b.B,
                                b.ab_code, b.ldb, b.beta, C, b.c_code, b.ldc, nthreads=nthreads)
    return st, C


def dd_reference(b: Batch, i: int):
    return oracle.gemm_dd(b.transa, b.transb, b.m, b.n, b.k, b.alpha, b.A[i], b.ab_code, b.lda, b.B[i], b.ldb,
                          b.beta, b.C0[i], b.ldc)


def normwise(b: Batch, C_list) -> float:
    """max over instances of ||C - C_exact||_F / (|alpha| ||op(A)||_F ||op(B)||_F + |beta| ||C0||_F)."""
    worst = 0.0
    ra, ca = b.dims_a()
    rb, cb = b.dims_b()
    for i, C in enumerate(C_list):
        hi, lo = dd_reference(b, i)
        na = np.linalg.norm(oracle.colmajor(b.A[i], ra, ca, b.lda).astype(np.float64))
        nb = np.linalg.norm(oracle.colmajor(b.B[i], rb, cb, b.ldb).astype(np.float64))
        nc = np.linalg.norm(oracle.colmajor(b.C0[i], b.m, b.n, b.ldc).astype(np.float64)) if b.beta != 0 else 0.0
        scale = abs(b.alpha) * na * nb + abs(b.beta) * nc
        worst = max(worst, oracle.normwise_error(oracle.colmajor(C, b.m, b.n, b.ldc), hi, lo, scale))
    return worst


def padding_untouched(b: Batch, C_list) -> bool:
    """Entries of C outside the m x n block equal their initial value."""
    for C, C0 in zip(C_list, b.C0):
        mask = np.ones_like(C, dtype=bool)
        oracle.colmajor(mask, b.m, b.n, b.ldc)[:] = False
        if not np.array_equal(C[mask], C0[mask]):
            return False
    return True


# ---------------------------------------------------------------------------
# GPU side (only called from tests marked gpu)
# ---------------------------------------------------------------------------
class DeviceBatch:
    """Device copies of a Batch plus the device pointer arrays."""

    def __init__(self, b: Batch, device="cuda"):
        import torch

        self.b = b
        self.A = [torch.from_numpy(a).to(device) for a in b.A]
        self.B = [torch.from_numpy(x).to(device) for x in b.B]
        self.C = [torch.from_numpy(c.copy()).to(device) for c in b.C0]
        from ssblasex_amd import pointer_array

        self.pA = pointer_array(self.A)
        self.pB = pointer_array(self.B)
        self.pC = pointer_array(self.C)

    def reset_c(self):
        for dst, src in zip(self.C, self.b.C0):
            dst.copy_(__import__("torch").from_numpy(src))

    def results(self):
        import torch

        torch.cuda.synchronize()
        return [c.cpu().numpy() for c in self.C]


def run_gpu(b: Batch, compute_type=0, entry="sync", dev: DeviceBatch | None = None):
    import ssblasex_amd as S

    d = dev or DeviceBatch(b)
    args = (b.transa, b.transb, b.m, b.n, b.k, b.alpha, d.pA, b.ab_code, b.lda, d.pB, b.ab_code, b.ldb, b.beta,
            d.pC, b.c_code, b.ldc, b.batch, compute_type)
    if entry == "sync":
        st = S.ssblasGemmBatchedEx(*args)
    elif entry == "64":
        st = S.ssblasGemmBatchedEx_64(*args)
    else:
        import torch

        st = S.ssblasGemmBatchedExAsync(*args, stream=torch.cuda.current_stream())
    return st, d.results()
