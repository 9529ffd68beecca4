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


# ---------------------------------------------------------------------------
# benchmark-sized batches: one buffer per operand, instance i at a fixed stride
# ---------------------------------------------------------------------------
class FlatBatch:
    """A batch of `batch` instances held in one (batch, elems) host array per
    operand, filled by vectorised numpy draws (seeded).  Padding rows of C
    hold a sentinel (3) so that writes outside the m x n block show up.
    Instance lists (A, B, C0) are row views, usable wherever a Batch is."""

    def __init__(self, type_, transa, transb, m, n, k, batch, *, alpha=1.0, beta=0.0, pad=0, seed=0,
                 dist="uniform"):
        rng = np.random.default_rng(seed)
        _, _, abt, ct = TYPES[type_]
        self.type, self.transa, self.transb, self.m, self.n, self.k = type_, transa, transb, m, n, k
        ra, ca = oracle.op_dims(transa, m, k)
        rb, cb = oracle.op_dims(transb, k, n)
        self.lda, self.ldb, self.ldc = ra + pad, rb + pad, m + pad
        if type_ == "i8":
            alpha, beta = int(alpha), int(beta)
        self.alpha, self.beta = alpha, beta

        def draw(rows, cols, ld, kind, dtype):
            x = np.full((batch, cols, ld), 7 if dtype != ct else 3, dtype=dtype)
            if kind == "i8":
                x[:, :, :rows] = rng.integers(-128, 128, size=(batch, cols, rows), dtype=np.int8)
            elif kind == "i32c":
                x[:, :, :rows] = rng.integers(-1000, 1001, size=(batch, cols, rows), dtype=np.int32)
            elif kind == "normal":
                x[:, :, :rows] = rng.standard_normal((batch, cols, rows), dtype=np.float64 if dtype == np.float64
                                                     else np.float32)
            elif kind == "wide":
                x[:, :, :rows] = rng.uniform(-1, 1, (batch, cols, rows)) * np.exp2(
                    rng.integers(-8, 9, (batch, cols, rows)))
            else:
                x[:, :, :rows] = rng.uniform(-1, 1, (batch, cols, rows))
            return x.reshape(batch, cols * ld)

        kind_ab = "i8" if type_ == "i8" else dist
        self.Aflat = draw(ra, ca, self.lda, kind_ab, abt)
        self.Bflat = draw(rb, cb, self.ldb, kind_ab, abt)
        self.C0flat = draw(m, n, self.ldc, "i32c" if type_ == "i8" else dist, ct)
        self.A = list(self.Aflat)
        self.B = list(self.Bflat)
        self.C0 = list(self.C0flat)

    batch = property(lambda self: len(self.C0))
    ab_code = property(lambda self: TYPES[self.type][0])
    c_code = property(lambda self: TYPES[self.type][1])

    def dims_a(self):
        return oracle.op_dims(self.transa, self.m, self.k)

    def dims_b(self):
        return oracle.op_dims(self.transb, self.k, self.n)

    def sample(self, idx) -> Batch:
        """The instances `idx` as an ordinary Batch (copies)."""
        b = Batch(self.type, self.transa, self.transb, self.m, self.n, self.k, self.lda, self.ldb, self.ldc,
                  self.alpha, self.beta)
        b.A = [self.Aflat[i].copy() for i in idx]
        b.B = [self.Bflat[i].copy() for i in idx]
        b.C0 = [self.C0flat[i].copy() for i in idx]
        return b

    def cview(self, flat):
        """(batch, m, n) view of the m x n blocks of a (batch, ldc*n) array."""
        return flat.reshape(self.batch, self.n, self.ldc)[:, :, :self.m].transpose(0, 2, 1)

    def norms(self):
        """Per-instance normwise-error scale |alpha| ||op(A)||_F ||op(B)||_F + |beta| ||C0||_F."""
        ra, ca = self.dims_a()
        rb, cb = self.dims_b()
        na = np.sqrt((self.Aflat.reshape(self.batch, ca, self.lda)[:, :, :ra].astype(np.float64) ** 2).sum(axis=(1, 2)))
        nb = np.sqrt((self.Bflat.reshape(self.batch, cb, self.ldb)[:, :, :rb].astype(np.float64) ** 2).sum(axis=(1, 2)))
        s = abs(self.alpha) * na * nb
        if self.beta != 0:
            s = s + abs(self.beta) * np.sqrt((self.cview(self.C0flat).astype(np.float64) ** 2).sum(axis=(1, 2)))
        return s

    def errors(self, Cflat, hi, lo):
        """Per-instance ||C - (hi + lo)||_F / scale, the difference taken in
        double-double order; hi, lo are (batch, m, n)."""
        d = (self.cview(Cflat).astype(np.float64) - hi) - lo
        return np.sqrt((d * d).sum(axis=(1, 2))) / self.norms()

    def padding_untouched(self, Cflat):
        x = Cflat.reshape(self.batch, self.n, self.ldc)[:, :, self.m:]
        y = self.C0flat.reshape(self.batch, self.n, self.ldc)[:, :, self.m:]
        return np.array_equal(x, y)


class FlatDeviceBatch:
    """Device copies of a FlatBatch: one allocation per operand, pointer
    arrays at the instance stride (three copies in all, however large the
    batch)."""

    def __init__(self, fb: FlatBatch, device="cuda"):
        import torch

        from ssblasex_amd import pointer_array_strided

        self.fb = fb
        self.A = torch.from_numpy(fb.Aflat).to(device)
        self.B = torch.from_numpy(fb.Bflat).to(device)
        self.C = torch.from_numpy(fb.C0flat).to(device)
        self.pA = pointer_array_strided(self.A, fb.batch, self.A.shape[1])
        self.pB = pointer_array_strided(self.B, fb.batch, self.B.shape[1])
        self.pC = pointer_array_strided(self.C, fb.batch, self.C.shape[1])

    def reset_c(self):
        import torch

        self.C.copy_(torch.from_numpy(self.fb.C0flat))

    def run(self, compute_type=0):
        """One synchronous ssblasGemmBatchedEx over the whole batch from the
        initial C; returns (status, host copy of C as (batch, ldc*n))."""
        import torch

        import ssblasex_amd as S

        fb = self.fb
        self.reset_c()
        torch.cuda.synchronize()
        st = S.ssblasGemmBatchedEx(fb.transa, fb.transb, fb.m, fb.n, fb.k, fb.alpha, self.pA, fb.ab_code, fb.lda,
                                   self.pB, fb.ab_code, fb.ldb, fb.beta, self.pC, fb.c_code, fb.ldc, fb.batch,
                                   compute_type)
        torch.cuda.synchronize()
        return st, self.C.cpu().numpy()
