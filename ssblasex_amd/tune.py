"""Autotuner for the tile configuration per size bucket.

    python -m ssblasex_amd.tune [--out FILE] [--set bench|grid|large|all|grouped] [--quick]

For every problem of the chosen set it times each tile configuration of the
library (forced through SSBLAS_TILE_CONFIG) on the current GPU and records
the fastest one for the problem's bucket:

    form  ta  tb  mb  nb  kb  wb   config   us
      form   0 int8, 1 fp32, 2 fp64, 3 bf16x3, 4 bf16x6
      mb/nb/kb  ceil(log2(m | n | k)) clamped to [3, 13]
      wb     work class of the batch: 0 if fewer than 256 tiles of 128x128,
             1 if fewer than 2048, else 2

The table is written to ssblasex_amd/tune/<arch>.txt (or --out) and read by
the library at its first call (SSBLAS_TUNE_TABLE overrides the path).  The
library falls back to its size heuristic for buckets the table lacks.

--set grouped tunes the tile shape of grouped plans (g32, g64, g128, w32) per
(form, size class of m, n, k), size classes <= 32, <= 64, <= 128, larger:
for every bucket it builds a plan over seeded instances drawn inside the
bucket (random layouts, padded leading dimensions; enough of them to fill the
GPU several times over, so that the time is the class's throughput), times its
execution with each shape forced (SSBLAS_GROUPED_CFG) and records the best
as a row "form 0 0 cm cn ck 9 gNN us".  Rows of other sets already in the
table are kept.
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

CONFIGS = ["direct16", "direct32", "tile64", "tile128", "tile128x64", "tile64x128", "tile128k2", "tile128w8",
           "tile256x128", "tile64s2", "tile128s2", "tile128k2s2", "tile128w8s2", "tile256x128s2", "dma64",
           "dma128", "dma128k2", "dma256x128", "dma128s3", "kc256x128", "kc128", "dma256", "dma256w16", "dma256k2",
           "gdma128", "gdma64", "gdma256x128"]
FORMS = {"i8": 0, "f32": 1, "f64": 2, "bf16x3": 3, "bf16x6": 4}
COMPUTE = {"i8": 0, "f32": 0, "f64": 0, "bf16x3": 4, "bf16x6": 5}
TABLE_DIR = Path(__file__).resolve().parent / "tune"


def bucket(x: int) -> int:
    return max(3, min(13, math.ceil(math.log2(max(1, x)))))


def work_class(m: int, n: int, batch: int) -> int:
    tiles = batch * math.ceil(m / 128) * math.ceil(n / 128)
    return 0 if tiles < 256 else (1 if tiles < 2048 else 2)


def candidates(form: str, m: int, n: int):
    if m <= 16 and n <= 16:
        c = ["direct16", "direct32", "tile64"]
    elif m <= 64 and n <= 64:
        c = ["direct16", "direct32", "tile64", "tile64s2", "tile64x128", "tile128x64", "dma64", "gdma64"]
    else:
        c = CONFIGS[2:]
    # the dma* tiles are int8 only (other forms run tile128 under those names)
    return c if form == "i8" else [x for x in c if not x.startswith("dma")]


def problem_sets(name: str):
    bench = [
        ("i8", 1, 1, 512, 512, 512, 64),
        ("f32", 0, 0, 64, 64, 64, 128),
        ("f64", 0, 1, 32, 32, 32, 65536),
        ("f32", 1, 0, 256, 256, 256, 4096),
        ("bf16x3", 1, 0, 256, 256, 256, 4096),
        ("bf16x6", 1, 0, 256, 256, 256, 4096),
    ]
    grid = []
    for form in ("i8", "f32"):
        for ta in (0, 1):
            for tb in (0, 1):
                for d in (16, 32, 64, 128, 256, 512):
                    batch = max(8, min(4096, (1 << 27) // (d * d * d)))
                    grid.append((form, ta, tb, d, d, d, batch))
    # batches of several thousand 128x128 tiles, where the 256x256 int8 tiles
    # (one per CU) compete with the streamed 128x128 ones
    large = []
    for form in ("i8", "f32"):
        for ta in (0, 1):
            for tb in (0, 1):
                for d, batch in ((256, 256), (512, 64), (512, 256)):
                    large.append((form, ta, tb, d, d, d, batch))
    for ta in (0, 1):
        for tb in (0, 1):
            large.append(("f64", ta, tb, 256, 256, 256, 512))
    return {"bench": bench, "grid": grid, "large": large, "all": bench + grid + large}[name]


def time_config(S, torch, form, ta, tb, m, n, k, batch, cfg, steps, warmup):
    """Median time (us) of one call with the given config."""
    from . import DataType

    t = "i8" if form == "i8" else ("f64" if form == "f64" else "f32")
    dt = {"i8": (torch.int8, torch.int32), "f32": (torch.float32, torch.float32),
          "f64": (torch.float64, torch.float64)}[t]
    ab, cc = {"i8": (DataType.R_8I, DataType.R_32I), "f32": (DataType.R_32F, DataType.R_32F),
              "f64": (DataType.R_64F, DataType.R_64F)}[t]
    ra, ca = (m, k) if ta == 0 else (k, m)
    rb, cb = (k, n) if tb == 0 else (n, k)
    key = (t, ta, tb, m, n, k, batch)
    if time_config.cache.get("key") != key:
        time_config.cache.clear()
        g = torch.Generator(device="cuda").manual_seed(1)
        if t == "i8":
            A = torch.randint(-128, 128, (batch, ra * ca), dtype=torch.int8, device="cuda", generator=g)
            B = torch.randint(-128, 128, (batch, rb * cb), dtype=torch.int8, device="cuda", generator=g)
        else:
            A = torch.randn((batch, ra * ca), dtype=dt[0], device="cuda", generator=g)
            B = torch.randn((batch, rb * cb), dtype=dt[0], device="cuda", generator=g)
        C = torch.zeros((batch, m * n), dtype=dt[1], device="cuda")
        ptr = lambda X: torch.arange(batch, device="cuda", dtype=torch.int64) * X.stride(0) * X.element_size() + X.data_ptr()
        time_config.cache.update(key=key, bufs=(A, B, C), ptrs=(ptr(A), ptr(B), ptr(C)))
    pA, pB, pC = time_config.cache["ptrs"]
    alpha, beta = (1, 0) if t == "i8" else (1.0, 0.0)
    os.environ["SSBLAS_TILE_CONFIG"] = cfg
    stream = torch.cuda.current_stream()

    # the ctypes arguments are built once: Python's marshalling (~4 us a
    # call) would otherwise bound the small problems and hide the kernels'
    # differences
    from . import _lib

    args, keep = S._gemm_args(ta, tb, m, n, k, alpha, pA, ab, ra, pB, ab, rb, beta, pC, cc, m, batch, COMPUTE[form], 0)
    fn = _lib.load().ssblasGemmBatchedExAsync
    sh = stream.cuda_stream

    def call():
        st = fn(*args, sh)
        if st != 0:
            raise RuntimeError(f"status {st}")

    try:
        for _ in range(warmup):
            call()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(steps):
            call()
        ev[1].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) * 1000.0 / steps
    finally:
        os.environ.pop("SSBLAS_TILE_CONFIG", None)


time_config.cache = {}


GROUPED_CFGS = ["g32", "g64", "g128", "w32"]
GROUPED_WORK_CLASS = 9
CLASS_RANGE = [(8, 32), (33, 64), (65, 128), (129, 512)]


def tune_grouped(S, torch, forms, quick, log):
    """Best grouped tile shape per (form, size class m, n, k); returns table rows."""
    import math as _m

    rows = {}
    rng = np.random.default_rng(0x55B1A5)
    codes = {"i8": (0, 3, 1, 0), "f32": (4, 4, 1.0, 0.0), "f64": (5, 5, 1.0, 0.0)}
    for form in forms:
        for cm in range(4):
            for cn in range(4):
                for ck in range(4):
                    lo_hi = [CLASS_RANGE[c] for c in (cm, cn, ck)]
                    mean = [_m.sqrt(a * b) for a, b in lo_hi]
                    # enough work (~4e10 flops, at most 8000 instances) to fill the GPU several
                    # times over, so that the time measures throughput, which is what the
                    # class's share of a large mixed plan costs, not one plan's tail
                    count = int(min(8000, max(256, 4e10 / (2 * mean[0] * mean[1] * mean[2]))))
                    if quick:
                        count = min(count, 256)
                    dims = [[int(_m.floor(_m.exp(rng.uniform(_m.log(a), _m.log(b + 1))))) for a, b in lo_hi]
                            for _ in range(count)]
                    group = 16
                    ngroups = (count + group - 1) // group
                    gsize = [min(group, count - g * group) for g in range(ngroups)]
                    ta = [int(x) for x in rng.integers(0, 2, ngroups)]
                    tb = [int(x) for x in rng.integers(0, 2, ngroups)]
                    ab, cc, al, be = codes[form]
                    esz = 1 if form == "i8" else (8 if form == "f64" else 4)
                    osz = 8 if form == "f64" else 4
                    m_, n_, k_, lda, ldb, ldc, bufs = [], [], [], [], [], [], []
                    for g in range(ngroups):
                        for i in range(g * group, g * group + gsize[g]):
                            m, n, k = dims[i]
                            ra, ca = (k, m) if ta[g] else (m, k)
                            rb, cb = (n, k) if tb[g] else (k, n)
                            pa, pb, pc = (int(x) for x in rng.integers(0, 17, 3))
                            m_.append(m); n_.append(n); k_.append(k)
                            lda.append(ra + pa); ldb.append(rb + pb); ldc.append(m + pc)
                            bufs.append((torch.zeros((ra + pa) * ca * esz, dtype=torch.uint8, device="cuda"),
                                         torch.zeros((rb + pb) * cb * esz, dtype=torch.uint8, device="cuda"),
                                         torch.zeros((m + pc) * n * osz, dtype=torch.uint8, device="cuda")))
                    pA = S.pointer_array([b[0] for b in bufs])
                    pB = S.pointer_array([b[1] for b in bufs])
                    pC = S.pointer_array([b[2] for b in bufs])
                    res = {}
                    for cfg in GROUPED_CFGS:
                        os.environ["SSBLAS_GROUPED_CFG"] = cfg
                        try:
                            plan = S.GroupedPlan(gsize, ta, tb, [ab] * ngroups, [ab] * ngroups, [cc] * ngroups,
                                                 [al] * ngroups, [be] * ngroups, m_, n_, k_, pA, lda, pB, ldb, pC,
                                                 ldc)
                        finally:
                            os.environ.pop("SSBLAS_GROUPED_CFG", None)
                        stream = torch.cuda.current_stream()
                        for _ in range(3):
                            plan.execute(stream)
                        torch.cuda.synchronize()
                        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                        steps = 5 if quick else 10
                        ev[0].record(stream)
                        for _ in range(steps):
                            plan.execute(stream)
                        ev[1].record(stream)
                        torch.cuda.synchronize()
                        res[cfg] = ev[0].elapsed_time(ev[1]) * 1000.0 / steps
                        plan.close()
                    best = min(res, key=res.get)
                    rows[(FORMS[form], 0, 0, cm, cn, ck, GROUPED_WORK_CLASS)] = (best, res[best])
                    log(f"grouped {form} classes m{cm} n{cn} k{ck} ({count} instances): best {best} "
                        + " ".join(f"{c}={v:.1f}" for c, v in res.items()))
                    del bufs, pA, pB, pC
                    torch.cuda.empty_cache()
    return rows


def read_table(path: Path):
    rows = {}
    if path.exists():
        for line in path.read_text().splitlines():
            if line.startswith("#") or not line.strip():
                continue
            f = line.split()
            if len(f) >= 9:
                rows[tuple(int(x) for x in f[:7])] = (f[7], float(f[8]))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--set", default="all", choices=["bench", "grid", "large", "all", "grouped"])
    ap.add_argument("--out", default=None)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--forms", default="i8,f32,f64", help="forms of --set grouped (comma-separated)")
    a = ap.parse_args(argv)
    import torch

    import ssblasex_amd as S

    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    out = Path(a.out) if a.out else TABLE_DIR / f"{arch}.txt"
    rows = read_table(out if out.exists() else TABLE_DIR / f"{arch}.txt")
    steps = 5 if a.quick else a.steps
    t0 = time.time()
    # about a second of work before the first timing, so that the clocks
    # have left their idle state
    while time.time() - t0 < 1.0:
        time_config(S, torch, "i8", 1, 1, 512, 512, 512, 64, "tile128", 50, 0)
    t0 = time.time()
    if a.set == "grouped":
        rows.update(tune_grouped(S, torch, a.forms.split(","), a.quick, lambda msg: print(msg, flush=True)))
    for (form, ta, tb, m, n, k, batch) in (problem_sets(a.set) if a.set != "grouped" else []):
        res = {}
        for cfg in candidates(form, m, n):
            try:
                res[cfg] = time_config(S, torch, form, ta, tb, m, n, k, batch, cfg, steps, 3)
            except RuntimeError as e:
                res[cfg] = float("inf")
        best = min(res, key=res.get)
        key = (FORMS[form], ta, tb, bucket(m), bucket(n), bucket(k), work_class(m, n, batch))
        rows[key] = (best, res[best])
        print(f"{form:6s} {'NT'[ta]}{'NT'[tb]} m={m} n={n} k={k} batch={batch}: best {best} {res[best]:.1f} us | "
              + " ".join(f"{c}={v:.1f}" for c, v in sorted(res.items(), key=lambda x: x[1])), flush=True)
    out.parent.mkdir(parents=True, exist_ok=True)
    lines = ["# form ta tb mb nb kb wb config us   (written by python -m ssblasex_amd.tune on "
             f"{arch}, {time.strftime('%Y-%m-%d')})"]
    for key in sorted(rows):
        best, us = rows[key]
        lines.append(" ".join(str(x) for x in key) + f" {best} {us:.2f}")
    out.write_text("\n".join(lines) + "\n")
    print(f"wrote {len(rows)} entries to {out} in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
