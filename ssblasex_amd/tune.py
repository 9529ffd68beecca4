"""Autotuner for the tile configuration per size bucket.

    python -m ssblasex_amd.tune [--out FILE] [--set bench|grid|all] [--quick]

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
           "dma128", "dma128k2", "dma256x128", "dma128s3"]
FORMS = {"i8": 0, "f32": 1, "f64": 2, "bf16x3": 3, "bf16x6": 4}
COMPUTE = {"i8": 0, "f32": 0, "f64": 0, "bf16x3": 4, "bf16x6": 5}
TABLE_DIR = Path(__file__).resolve().parent / "tune"


def bucket(x: int) -> int:
    return max(3, min(13, math.ceil(math.log2(max(1, x)))))


def work_class(m: int, n: int, batch: int) -> int:
    tiles = batch * math.ceil(m / 128) * math.ceil(n / 128)
    return 0 if tiles < 256 else (1 if tiles < 2048 else 2)


def candidates(m: int, n: int):
    if m <= 16 and n <= 16:
        return ["direct16", "direct32", "tile64"]
    if m <= 64 and n <= 64:
        return ["direct16", "direct32", "tile64", "tile64s2", "tile64x128", "tile128x64", "dma64"]
    return CONFIGS[2:]


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
    return {"bench": bench, "grid": grid, "all": bench + grid}[name]


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

    def call():
        st = S.ssblasGemmBatchedExAsync(ta, tb, m, n, k, alpha, pA, ab, ra, pB, ab, rb, beta, pC, cc, m, batch,
                                        COMPUTE[form], 0, stream)
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--set", default="all", choices=["bench", "grid", "all"])
    ap.add_argument("--out", default=None)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--steps", type=int, default=20)
    a = ap.parse_args(argv)
    import torch

    import ssblasex_amd as S

    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    out = Path(a.out) if a.out else TABLE_DIR / f"{arch}.txt"
    rows = {}
    steps = 5 if a.quick else a.steps
    t0 = time.time()
    for (form, ta, tb, m, n, k, batch) in problem_sets(a.set):
        res = {}
        for cfg in candidates(m, n):
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
