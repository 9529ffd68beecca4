"""Grouped workloads for the tests: seeded uneven instances (log-uniform
m, n, k, padded leading dimensions, random N/T per group, mixed types),
each instance a one-element Batch so the oracle helpers apply unchanged."""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

import oracle
from helpers import Batch, TYPES, _fill

COMPUTE_DEFAULT = 0


@dataclass
class Grouped:
    group_size: list = field(default_factory=list)
    gtype: list = field(default_factory=list)     # "f32" / "i8" / "f64" per group
    transa: list = field(default_factory=list)
    transb: list = field(default_factory=list)
    alpha: list = field(default_factory=list)
    beta: list = field(default_factory=list)
    inst: list = field(default_factory=list)      # one Batch (batch = 1) per instance

    def group_of(self):
        out = []
        for g, s in enumerate(self.group_size):
            out += [g] * s
        return out

    def c_args(self):
        """Per-group and per-instance argument lists of the grouped call."""
        ab = [TYPES[t][0] for t in self.gtype]
        cc = [TYPES[t][1] for t in self.gtype]
        return dict(group_size=self.group_size, transa=self.transa, transb=self.transb, Atype=ab, Btype=ab,
                    Ctype=cc, alpha=self.alpha, beta=self.beta, m=[b.m for b in self.inst],
                    n=[b.n for b in self.inst], k=[b.k for b in self.inst], lda=[b.lda for b in self.inst],
                    ldb=[b.ldb for b in self.inst], ldc=[b.ldc for b in self.inst])


def log_uniform_dim(rng, lo, hi):
    return int(math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi + 1)))))


def make_grouped(count, seed, *, lo=8, hi=512, group=16, types=("f32", "i8"), beta=0.0, alpha=1.0,
                 pad=16, dims=None) -> Grouped:
    """`count` instances in groups of `group`; group g has type types[g % len(types)]
    and random transa/transb.  dims: optional list of (m, n, k) overriding the draw."""
    rng = np.random.default_rng(seed)
    G = Grouped()
    left = count
    g = 0
    idx = 0
    while left > 0:
        s = min(group, left)
        t = types[g % len(types)]
        G.group_size.append(s)
        G.gtype.append(t)
        G.transa.append(int(rng.integers(0, 2)))
        G.transb.append(int(rng.integers(0, 2)))
        G.alpha.append(int(alpha) if t == "i8" else float(alpha))
        G.beta.append(int(beta) if t == "i8" else float(beta))
        _, _, abt, ct = TYPES[t]
        for _ in range(s):
            if dims is not None:
                m, n, k = dims[idx]
            else:
                m, n, k = (log_uniform_dim(rng, lo, hi) for _ in range(3))
            ta, tb = G.transa[-1], G.transb[-1]
            ra, ca = oracle.op_dims(ta, m, k)
            rb, cb = oracle.op_dims(tb, k, n)
            lda = max(1, ra) + int(rng.integers(0, pad + 1))
            ldb = max(1, rb) + int(rng.integers(0, pad + 1))
            ldc = max(1, m) + int(rng.integers(0, pad + 1))
            b = Batch(t, ta, tb, m, n, k, lda, ldb, ldc, G.alpha[-1], G.beta[-1])
            A = np.full(lda * ca, 7, dtype=abt)
            B = np.full(ldb * cb, 7, dtype=abt)
            C = np.full(ldc * n, 3, dtype=ct)
            kind = "i8" if t == "i8" else ("normal" if t == "f32" else "wide")
            oracle.colmajor(A, ra, ca, lda)[:] = _fill(rng, (ra, ca), kind, abt)
            oracle.colmajor(B, rb, cb, ldb)[:] = _fill(rng, (rb, cb), kind, abt)
            oracle.colmajor(C, m, n, ldc)[:] = _fill(rng, (m, n), "i32c" if t == "i8" else kind, ct)
            b.A.append(A)
            b.B.append(B)
            b.C0.append(C)
            G.inst.append(b)
            idx += 1
        left -= s
        g += 1
    return G


class DeviceGrouped:
    """Device buffers of a Grouped workload, allocated in a shuffled order,
    and the three device pointer arrays."""

    def __init__(self, G: Grouped, seed=1, device="cuda"):
        import torch

        from ssblasex_amd import pointer_array

        self.G = G
        n = len(G.inst)
        order = np.random.default_rng(seed).permutation(3 * n)
        bufs = [None] * (3 * n)
        for j in order:
            i, w = divmod(int(j), 3)
            src = (G.inst[i].A[0], G.inst[i].B[0], G.inst[i].C0[0])[w]
            bufs[j] = torch.from_numpy(src.copy()).to(device)
        self.A = [bufs[3 * i] for i in range(n)]
        self.B = [bufs[3 * i + 1] for i in range(n)]
        self.C = [bufs[3 * i + 2] for i in range(n)]
        self.pA = pointer_array(self.A)
        self.pB = pointer_array(self.B)
        self.pC = pointer_array(self.C)

    def reset_c(self):
        import torch

        for dst, b in zip(self.C, self.G.inst):
            dst.copy_(torch.from_numpy(b.C0[0]))

    def results(self):
        import torch

        torch.cuda.synchronize()
        return [c.cpu().numpy() for c in self.C]
