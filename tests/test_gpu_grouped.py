"""Grouped (variable-size, mixed-type) plans on the GPU against the oracle.

Every instance is checked the way the batched tests check a batch: int8
bit-exact against the oracle, fp32 / fp64 normwise against the FP64
(double-double) oracle with the plain path's bar (2 x the reference-order
error + 1 unit roundoff), entries of C outside m x n untouched.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from grouped_helpers import DeviceGrouped, make_grouped
from helpers import UNIT_ROUNDOFF, normwise, padding_untouched, run_oracle

pytestmark = pytest.mark.gpu

BF16X3 = 4


def plan_for(dg, compute_type=0):
    import ssblasex_amd as S

    a = dg.G.c_args()
    return S.GroupedPlan(a["group_size"], a["transa"], a["transb"], a["Atype"], a["Btype"], a["Ctype"], a["alpha"],
                         a["beta"], a["m"], a["n"], a["k"], dg.pA, a["lda"], dg.pB, a["ldb"], dg.pC, a["ldc"],
                         computeType=compute_type)


def check(G, got, *, fp32_form="plain"):
    """Compare every instance with the oracle; returns the worst fp ratio."""
    worst = 0.0
    for i, b in enumerate(G.inst):
        C = [got[i]]
        assert padding_untouched(b, C), f"instance {i}: padding of C written"
        st, R = run_oracle(b)
        assert st == 0
        if b.type == "i8":
            assert np.array_equal(C[0], R[0]), f"instance {i} ({b.m}x{b.n}x{b.k}, ta={b.transa} tb={b.transb}) differs"
            continue
        ref = normwise(b, R)
        e = normwise(b, C)
        if b.type == "f32" and fp32_form == "bf16x3":
            assert e <= 2.0 ** -14, f"instance {i}: bf16x3 error {e}"
        else:
            bar = 2 * ref + UNIT_ROUNDOFF[b.type]
            assert e <= bar, f"instance {i} ({b.m}x{b.n}x{b.k}, {b.type}): error {e} > {bar}"
            worst = max(worst, e / bar)
    return worst


def test_grouped_mixed_small():
    G = make_grouped(160, seed=0x55B1A5 + 50, hi=160)
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.execute() == 0
        check(G, dg.results())


def test_grouped_alpha_beta():
    G = make_grouped(96, seed=0x55B1A5 + 51, hi=130, alpha=-2, beta=3)
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.execute() == 0
        check(G, dg.results())


def test_grouped_beta_zero_ignores_nan():
    G = make_grouped(32, seed=0x55B1A5 + 52, hi=70, types=("f32",))
    for b in G.inst:
        oracle.colmajor(b.C0[0], b.m, b.n, b.ldc)[:] = np.nan
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.execute() == 0
        got = dg.results()
    for b, C in zip(G.inst, got):
        assert np.isfinite(oracle.colmajor(C, b.m, b.n, b.ldc)).all()
    check(G, got)


@pytest.mark.parametrize("cfg", ["g32", "g64", "g128", "w32", "w32:g64", "w32:g128", "g32:w32", "g64/whole",
                                 "g128/whole"])
def test_grouped_every_tile_shape(cfg, monkeypatch):
    """Each tile shape forced (w32: one wave per 32x32 block, taken by the
    waves one at a time after the workgroup tiles), and queues that mix
    wave items and workgroup items ("small:rest"); "/whole": the shape over
    the whole instance, partial tiles at its edges instead of strips of a
    smaller shape."""
    if cfg.endswith("/whole"):
        cfg = cfg[:-6]
        monkeypatch.setenv("SSBLAS_GROUPED_EDGES", "0")
    monkeypatch.setenv("SSBLAS_GROUPED_CFG", cfg)
    G = make_grouped(48, seed=0x55B1A5 + 53, hi=300, group=8)
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.execute() == 0
        check(G, dg.results())


@pytest.mark.parametrize("cfg", [None, "w32"])
def test_grouped_edges(cfg, monkeypatch):
    if cfg is not None:
        monkeypatch.setenv("SSBLAS_GROUPED_CFG", cfg)
    dims = [(1, 1, 1), (1, 9, 3), (9, 1, 3), (7, 7, 0), (5, 3, 1), (33, 31, 65), (64, 64, 64), (65, 129, 127),
            (128, 128, 128), (129, 2, 300), (2, 129, 300), (512, 512, 512), (3, 500, 8), (500, 3, 8),
            (17, 17, 1000), (1, 1, 1000)]
    G = make_grouped(len(dims) * 2, seed=0x55B1A5 + 54, group=len(dims), dims=dims + dims)
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.execute() == 0
        got = dg.results()
    for b, C in zip(G.inst, got):
        if b.k == 0:
            assert np.array_equal(C, b.C0[0]), "k == 0 must leave C unchanged"
    check(G, got)


@pytest.mark.parametrize("cfg", [None, "g32", "g64", "g128", "w32", "g64/whole", "g128/whole"])
def test_grouped_matrix_ends(cfg, monkeypatch):
    """Leading dimensions equal to the stored rows (every matrix ends exactly
    at its last element, no padding after it) with tiny k, one-row and
    one-column operands: the word that holds an int8 matrix's last bytes
    falls in an earlier row tile or an earlier k-block than the last one,
    and k tails of K-contiguous operands read the next row's data."""
    if cfg is not None and cfg.endswith("/whole"):
        cfg = cfg[:-6]
        monkeypatch.setenv("SSBLAS_GROUPED_EDGES", "0")
    if cfg is not None:
        monkeypatch.setenv("SSBLAS_GROUPED_CFG", cfg)
    dims = [(33, 5, 1), (33, 33, 2), (1, 3, 65), (1, 1, 129), (2, 70, 65), (3, 3, 3), (31, 2, 17), (65, 1, 1),
            (129, 130, 1), (2, 2, 64), (3, 1, 128), (70, 3, 2), (1, 65, 63), (5, 5, 129)]
    dims = dims * 4
    G = make_grouped(len(dims), seed=0x55B1A5 + 57, group=1, dims=dims, pad=0, types=("i8", "f32"))
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.execute() == 0
        check(G, dg.results())


@pytest.mark.parametrize("cfg", [None, "g128", "g64", "w32", "g128/whole"])
def test_grouped_nan_padding(cfg, monkeypatch):
    """fp32 operands whose storage outside the logical matrix (rows between
    the stored rows and the leading dimension) holds NaN, every layout pair,
    k tails of every length: the elements past K that a K-contiguous k-block
    reads, and the rows and columns outside the instance, must not reach a
    stored result."""
    if cfg is not None and cfg.endswith("/whole"):
        cfg = cfg[:-6]
        monkeypatch.setenv("SSBLAS_GROUPED_EDGES", "0")
    if cfg is not None:
        monkeypatch.setenv("SSBLAS_GROUPED_CFG", cfg)
    dims = [(130, 129, 17), (128, 128, 1), (200, 140, 33), (129, 3, 15), (3, 129, 47), (140, 200, 100),
            (257, 130, 130), (128, 128, 16), (160, 150, 63), (131, 133, 65)]
    dims = dims * 4
    G = make_grouped(len(dims), seed=0x55B1A5 + 58, group=1, dims=dims, types=("f32",), pad=16)
    for b in G.inst:
        ra, ca = oracle.op_dims(b.transa, b.m, b.k)
        rb, cb = oracle.op_dims(b.transb, b.k, b.n)
        for X, r, c, ld in ((b.A[0], ra, ca, b.lda), (b.B[0], rb, cb, b.ldb)):
            mask = np.ones_like(X, dtype=bool)
            oracle.colmajor(mask, r, c, ld)[:] = False
            X[mask] = np.nan
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.execute() == 0
        check(G, dg.results())


@pytest.mark.parametrize("cfg", [None, "w32", "w32:g64"])
def test_grouped_fp64_and_bf16x3_families(cfg, monkeypatch):
    if cfg is not None:
        monkeypatch.setenv("SSBLAS_GROUPED_CFG", cfg)
    G = make_grouped(60, seed=0x55B1A5 + 55, hi=140, group=10, types=("f64", "f32", "i8"))
    dg = DeviceGrouped(G)
    with plan_for(dg, compute_type=BF16X3) as p:
        assert p.execute() == 0
        check(G, dg.results(), fp32_form="bf16x3")


@pytest.mark.parametrize("cfg", [None, "w32:g64"])
def test_grouped_graph_replays(cfg, monkeypatch):
    """Captured graph replayed with eager work between replays: each replay
    matches the oracle on the data it saw (a launch keeps no state between
    replays)."""
    import torch

    if cfg is not None:
        monkeypatch.setenv("SSBLAS_GROUPED_CFG", cfg)
    G = make_grouped(64, seed=0x55B1A5 + 56, hi=200)
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        for rep in range(3):
            dg.reset_c()
            if rep == 2:  # new operand values behind the same pointers
                for b, A in zip(G.inst, dg.A):
                    b.A[0] = (-b.A[0]).astype(b.A[0].dtype) if b.type != "i8" else np.where(
                        b.A[0] == -128, 127, -b.A[0]).astype(np.int8)
                    A.copy_(torch.from_numpy(b.A[0]))
            x = torch.randn(256, 256, device="cuda")
            (x @ x).sum().item()  # unrelated eager work and a sync between replays
            assert p.graph_launch(torch.cuda.current_stream()) == 0
            check(G, dg.results())


@pytest.mark.parametrize("grid", ["1", "3", "13"])
def test_grouped_capped_grid(grid, monkeypatch):
    """A grid of 1, 3 or 13 workgroups: queues without a workgroup of their
    own are drained by workgroups of other XCDs, every workgroup claims many
    entries (tiles and bundles of one-wave items, both forms); the plan is
    executed twice in a row and then replayed from its graph, each launch
    starting from the counters the previous one re-armed."""
    import torch

    monkeypatch.setenv("SSBLAS_MAX_GRID", grid)
    dims = [(300, 200, 150), (40, 300, 70), (20, 20, 500), (129, 257, 33), (8, 9, 10), (64, 64, 64),
            (200, 31, 300), (128, 128, 1)]
    G = make_grouped(len(dims) * 6, seed=0x55B1A5 + 59, group=4, dims=dims * 6)
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        for launch in ("execute", "execute", "graph"):
            dg.reset_c()
            if launch == "graph":
                assert p.graph_launch(torch.cuda.current_stream()) == 0
            else:
                assert p.execute() == 0
            check(G, dg.results())


def test_grouped_one_shot_entry():
    import ssblasex_amd as S

    G = make_grouped(40, seed=0x55B1A5 + 57, hi=100)
    dg = DeviceGrouped(G)
    a = G.c_args()
    st = S.ssblasGemmGroupedBatchedEx(a["group_size"], a["transa"], a["transb"], a["Atype"], a["Btype"], a["Ctype"],
                                      a["alpha"], a["beta"], a["m"], a["n"], a["k"], dg.pA, a["lda"], dg.pB,
                                      a["ldb"], dg.pC, a["ldc"])
    assert st == 0
    check(G, dg.results())


def test_grouped_host_pointer_arrays_rejected():
    import ssblasex_amd as S

    G = make_grouped(4, seed=3, hi=16)
    dg = DeviceGrouped(G)
    a = G.c_args()
    host = dg.pA.cpu()
    st, plan = S.GroupedPlan.create(a["group_size"], a["transa"], a["transb"], a["Atype"], a["Btype"], a["Ctype"],
                                    a["alpha"], a["beta"], a["m"], a["n"], a["k"], host, a["lda"], dg.pB, a["ldb"],
                                    dg.pC, a["ldc"])
    assert st == 2 and plan is None


def test_grouped_c5_full_scale():
    """10000 instances, m, n, k log-uniform in [8, 512], ld padding 0..16,
    N/T per group of 256, half fp32 N(0,1), half int8: every instance checked
    through column sums (int8 exact modulo 2^32, fp32 against an fp64 bound),
    and a seeded sample of 300 instances element by element."""
    G = make_grouped(10000, seed=0x55B1A5 + 5, group=256)
    dg = DeviceGrouped(G)
    with plan_for(dg) as p:
        assert p.graph_launch() == 0
        got = dg.results()
    for i, (b, C) in enumerate(zip(G.inst, got)):
        ra, ca = b.dims_a()
        rb, cb = b.dims_b()
        A = oracle.colmajor(b.A[0], ra, ca, b.lda)
        B = oracle.colmajor(b.B[0], rb, cb, b.ldb)
        opA = A.T if b.transa else A
        opB = B.T if b.transb else B
        Cm = oracle.colmajor(C, b.m, b.n, b.ldc)
        if b.type == "i8":
            lhs = Cm.astype(np.int64).sum(axis=0) % (1 << 32)
            rhs = (opA.astype(np.int64).sum(axis=0) @ opB.astype(np.int64)) % (1 << 32)
            assert np.array_equal(lhs, rhs), f"instance {i}: column sums differ"
        else:
            lhs = Cm.astype(np.float64).sum(axis=0)
            rhs = opA.astype(np.float64).sum(axis=0) @ opB.astype(np.float64)
            # each entry is within k u sum_l |a_il| |b_lj| (fp32 unit roundoff u)
            bound = 3 * b.k * 2.0 ** -24 * (np.abs(opA).astype(np.float64).sum(axis=0) @
                                            np.abs(opB).astype(np.float64)) + 1e-30
            assert (np.abs(lhs - rhs) <= bound).all(), f"instance {i}: column sums off"
    rng = np.random.default_rng(7)
    sample = rng.choice(len(G.inst), 300, replace=False)
    sub = type(G)()
    sub.inst = [G.inst[int(i)] for i in sample]
    check(sub, [got[int(i)] for i in sample])
