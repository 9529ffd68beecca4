"""Benchmark of ssblasGemmBatchedEx on one MI355X.

Default workload (the headline configuration): int8 -> int32, transa = transb
= T, m = n = k = 512, batchCount = 64, lda = ldb = ldc = 512, A/B uniform
integers in [-128, 127], C uniform in [-1000, 1000], alpha = 1, beta = 0;
every A[i] / B[i] / C[i] its own allocation, pointer arrays device resident.

A step is one ssblasGemmBatchedExAsync call over the whole batch with the
inputs already in HBM.  W untimed warmup steps, then K steps bracketed by
HIP events on the launch stream and device synchronisation.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--beta 0|1]
                    [--no-cpu-baseline] [--secondary] [--grouped]

Prints one JSON line.  --secondary adds the other benchmark configurations
(fp32 64^3 x128 NN, fp64 32^3 x65536 NT native vs Ozaki, fp64 256^3 x512,
fp32 256^3 x4096 TN native vs BF16x3/x6 with ld 256 and 264, the headline
batch with beta = 1) under "secondary"; --grouped (implied by
--secondary) adds the grouped 10000-instance uneven batch under "grouped_c5".
--case ID:FORM times only one of those (c1, c3, c3l, c4, c4p, c2b1 with a
form name such as native, bf16x6 or ozaki_s9; c5 with all, small or large)
and prints its own JSON line, for profiler runs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "batched GEMM TFLOP/s|TOPS & % MFMA peak; split-precision vs native at fixed error"
PEAK = {"int8": 5033.2, "f32": 157.3, "f64": 78.6, "bf16": 2516.6}  # dense, TFLOP/s | TOPS
HBM_PEAK_GBS = 8000.0
SEED = 0x55B1A5


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--beta", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--secondary", action="store_true")
    ap.add_argument("--grouped", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only time a device-to-device copy of MIB MiB (HBM calibration for the counters)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--case", default=None, metavar="ID:FORM",
                    help="time only one other configuration, e.g. c4:bf16x6, c3:ozaki_s9, c5:all (c5 also takes small, "
                         "large, fp32, int8; profiling runs)")
    return ap.parse_args()


def gen_i8(rng, n):
    return rng.integers(-128, 128, size=n, dtype=np.int64).astype(np.int8)


def time_steps(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms per step


class Problem:
    """One batched GEMM with separately allocated instances on the device."""

    def __init__(self, t, ta, tb, m, n, k, batch, alpha, beta, seed, dist="uniform", pad=0):
        import torch
        import ssblasex_amd as S

        self.t, self.ta, self.tb, self.m, self.n, self.k, self.batch = t, ta, tb, m, n, k, batch
        self.alpha, self.beta = alpha, beta
        ra, ca = (m, k) if ta == 0 else (k, m)
        rb, cb = (k, n) if tb == 0 else (n, k)
        self.lda, self.ldb, self.ldc = ra + pad, rb + pad, m + pad
        rng = np.random.default_rng(seed)
        code = {"i8": (S.DataType.R_8I, S.DataType.R_32I), "f32": (S.DataType.R_32F, S.DataType.R_32F),
                "f64": (S.DataType.R_64F, S.DataType.R_64F)}[t]
        self.ab, self.cc = code
        tdt = {"i8": (torch.int8, torch.int32), "f32": (torch.float32, torch.float32),
               "f64": (torch.float64, torch.float64)}[t]
        dev = "cuda"
        self.A, self.B, self.C = [], [], []
        self.hostA, self.hostB, self.hostC = [], [], []
        gen_chunk = max(1, min(batch, (1 << 24) // max(1, self.lda * ca)))
        for i0 in range(0, batch, gen_chunk):
            nb = min(gen_chunk, batch - i0)
            if t == "i8":
                a = gen_i8(rng, nb * self.lda * ca)
                b = gen_i8(rng, nb * self.ldb * cb)
                c = rng.integers(-1000, 1001, size=nb * self.ldc * n, dtype=np.int64).astype(np.int32)
            else:
                npdt = np.float32 if t == "f32" else np.float64
                if dist == "normal":
                    a = rng.standard_normal(nb * self.lda * ca).astype(npdt)
                    b = rng.standard_normal(nb * self.ldb * cb).astype(npdt)
                    c = rng.standard_normal(nb * self.ldc * n).astype(npdt)
                elif dist == "wide":
                    def w(sz):
                        return (rng.uniform(-1, 1, sz) * np.exp2(rng.integers(-8, 9, sz))).astype(npdt)
                    a, b, c = w(nb * self.lda * ca), w(nb * self.ldb * cb), w(nb * self.ldc * n)
                else:
                    a = rng.uniform(-1, 1, nb * self.lda * ca).astype(npdt)
                    b = rng.uniform(-1, 1, nb * self.ldb * cb).astype(npdt)
                    c = rng.uniform(-1, 1, nb * self.ldc * n).astype(npdt)
            ta_ = torch.from_numpy(a).to(dev)
            tb_ = torch.from_numpy(b).to(dev)
            tc_ = torch.from_numpy(c).to(dev)
            for j in range(nb):
                # separate allocations, as the reference's callers make them
                self.A.append(ta_[j * self.lda * ca:(j + 1) * self.lda * ca].clone())
                self.B.append(tb_[j * self.ldb * cb:(j + 1) * self.ldb * cb].clone())
                self.C.append(tc_[j * self.ldc * n:(j + 1) * self.ldc * n].clone())
            if len(self.hostA) < 64:
                for j in range(min(nb, 64 - len(self.hostA))):
                    self.hostA.append(a[j * self.lda * ca:(j + 1) * self.lda * ca])
                    self.hostB.append(b[j * self.ldb * cb:(j + 1) * self.ldb * cb])
                    self.hostC.append(c[j * self.ldc * n:(j + 1) * self.ldc * n])
            del ta_, tb_, tc_
        self.pA = S.pointer_array(self.A)
        self.pB = S.pointer_array(self.B)
        self.pC = S.pointer_array(self.C)
        torch.cuda.synchronize()

    def call(self, compute_type=0, stream=None):
        import ssblasex_amd as S

        st = S.ssblasGemmBatchedExAsync(self.ta, self.tb, self.m, self.n, self.k, self.alpha, self.pA, self.ab,
                                        self.lda, self.pB, self.ab, self.ldb, self.beta, self.pC, self.cc,
                                        self.ldc, self.batch, compute_type, 0, stream)
        if st != 0:
            raise RuntimeError(f"ssblasGemmBatchedExAsync returned {S.Status(st).name}")

    def work(self):
        return 2.0 * self.m * self.n * self.k * self.batch

    def bytes(self):
        es = {"i8": 1, "f32": 4, "f64": 8}[self.t]
        ec = {"i8": 4, "f32": 4, "f64": 8}[self.t]
        per = (self.m * self.k + self.k * self.n) * es + self.m * self.n * ec * (2 if self.beta != 0 else 1)
        return per * self.batch + 24 * self.batch


def cpu_baseline_i8(prob: Problem, threads: int):
    """The example's OpenMP int8 loop (restated in the oracle) on the host
    cores: whole batch, repeated until >= 10 s or 5 runs."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle

    n_inst = min(prob.batch, len(prob.hostA))
    A, B = prob.hostA[:n_inst], prob.hostB[:n_inst]
    C = [c.copy() for c in prob.hostC[:n_inst]]
    oracle.naive_baseline_i8(prob.ta, prob.tb, prob.m, prob.n, prob.k, prob.alpha, A, prob.lda, B, prob.ldb,
                             prob.beta, C, prob.ldc, threads)  # warm-up
    times = []
    t_total = 0.0
    while len(times) < 5 and t_total < 10.0:
        t0 = time.perf_counter()
        oracle.naive_baseline_i8(prob.ta, prob.tb, prob.m, prob.n, prob.k, prob.alpha, A, prob.lda, B, prob.ldb,
                                 prob.beta, C, prob.ldc, threads)
        dt = time.perf_counter() - t0
        times.append(dt)
        t_total += dt
    med = float(np.median(times))
    ops = 2.0 * prob.m * prob.n * prob.k * n_inst
    return {"value": ops / med / 1e12, "unit": "TOPS", "cores": threads, "kind": "port",
            "sample": f"{n_inst} of {prob.batch} instances, the example's OpenMP int8 loop "
                      f"(example/include/naive_gemm.hpp:3-155) restated in oracle/ssblas_oracle.c; "
                      f"median of {len(times)} runs, {med:.3f} s each"}


# The other benchmark configurations: (case id, name, problem, forms).
SECONDARY_ROWS = [
    ("c1", "fp32 NN m=n=k=64 batch 128 beta=0.5", dict(t="f32", ta=0, tb=0, m=64, n=64, k=64, batch=128, alpha=1.0,
                                                       beta=0.5, seed=SEED + 1), [("native", 0, "f32")]),
    ("c3", "fp64 NT m=n=k=32 batch 65536", dict(t="f64", ta=0, tb=1, m=32, n=32, k=32, batch=65536, alpha=1.0,
                                                beta=0.0, seed=SEED + 3, dist="wide"),
     [("native", 0, "f64"), ("ozaki_s6", 0x106, "f64"), ("ozaki_s7", 0x107, "f64"), ("ozaki_s8", 0x108, "f64"),
      ("ozaki_s9", 0x109, "f64")]),
    ("c3l", "fp64 NN m=n=k=256 batch 512", dict(t="f64", ta=0, tb=0, m=256, n=256, k=256, batch=512, alpha=1.0,
                                                beta=0.0, seed=SEED + 5, dist="wide"),
     [("native", 0, "f64"), ("ozaki_s7", 0x107, "f64"), ("ozaki_s8", 0x108, "f64"), ("ozaki_s9", 0x109, "f64")]),
    ("c4", "fp32 TN m=n=k=256 batch 4096 alpha=1.5 beta=-0.5",
     dict(t="f32", ta=1, tb=0, m=256, n=256, k=256, batch=4096, alpha=1.5, beta=-0.5, seed=SEED + 4,
          dist="normal"), [("native", 0, "f32"), ("bf16x3", 4, "bf16"), ("bf16x6", 5, "bf16")]),
    ("c4p", "fp32 TN m=n=k=256 batch 4096 alpha=1.5 beta=-0.5, lda=ldb=ldc=264",
     dict(t="f32", ta=1, tb=0, m=256, n=256, k=256, batch=4096, alpha=1.5, beta=-0.5, seed=SEED + 4,
          dist="normal", pad=8), [("native", 0, "f32"), ("bf16x3", 4, "bf16"), ("bf16x6", 5, "bf16")]),
    ("c2b1", "int8 TT m=n=k=512 batch 64 beta=1",
     dict(t="i8", ta=1, tb=1, m=512, n=512, k=512, batch=64, alpha=1, beta=1, seed=SEED + 2), [("native", 0, "i8")]),
]


def secondary(args, stream, only=None):
    """Other benchmark configurations: native vs split forms.  `only` =
    (case id, form name) runs just that one (profiling runs)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import torch

    out = {}
    for cid, name, kw, forms in SECONDARY_ROWS:
        if only is not None:
            if cid != only[0]:
                continue
            forms = [f for f in forms if f[0] == only[1]]
            if not forms:
                raise SystemExit(f"case {cid} has no form {only[1]}")
        p = Problem(**kw)
        res = {}
        for fname, ct, peak_key in forms:
            try:
                ms = time_steps(lambda: p.call(ct, stream), max(5, args.steps // 10), 3, stream)
            except RuntimeError as e:
                res[fname] = {"error": str(e)}
                continue
            v = p.work() / (ms * 1e-3) / 1e12
            res[fname] = {"TFLOPs" if kw["t"] != "i8" else "TOPS": round(v, 3), "ms_per_step": round(ms, 4),
                          "frac_of_native_peak": round(v / PEAK[kw["t"] if kw["t"] != "i8" else "int8"], 4),
                          "hbm_GBs": round(p.bytes() / (ms * 1e-3) / 1e9, 1)}
        out[name] = res
        del p
        torch.cuda.empty_cache()
    return out


def grouped_c5(args, stream, only=None):
    """Grouped uneven batch: 10000 instances, m, n, k = floor(exp(U[ln 8, ln 513))),
    ld = dim + U{0..16}, transa/transb drawn per group of 256, groups alternately
    fp32 N(0,1) and int8 U{-128..127}, alpha = 1, beta = 0, buffers allocated in a
    shuffled order.  One step = one replay of the plan's captured graph.
    Also times the instances of each size class on their own (plans over the
    subsets), which shows the memory-bound small end and the MFMA-bound large end."""
    import math

    import torch

    import ssblasex_amd as S

    rng = np.random.default_rng(SEED + 5)
    count, group = 10000, 256
    ngroups = (count + group - 1) // group
    gsize = [min(group, count - g * group) for g in range(ngroups)]
    gtype = ["f32" if g % 2 == 0 else "i8" for g in range(ngroups)]
    gta = [int(x) for x in rng.integers(0, 2, ngroups)]
    gtb = [int(x) for x in rng.integers(0, 2, ngroups)]
    dims = np.floor(np.exp(rng.uniform(math.log(8), math.log(513), (count, 3)))).astype(np.int64)
    pads = rng.integers(0, 17, (count, 3))
    inst = []  # (type, ta, tb, m, n, k, lda, ldb, ldc)
    for g in range(ngroups):
        for i in range(g * group, g * group + gsize[g]):
            m, n, k = (int(x) for x in dims[i])
            ra, rb = (k if gta[g] else m), (n if gtb[g] else k)
            inst.append((gtype[g], gta[g], gtb[g], m, n, k, ra + int(pads[i, 0]), rb + int(pads[i, 1]),
                         m + int(pads[i, 2])))
    gen = torch.Generator(device="cuda").manual_seed(SEED + 5)
    bufs = [None] * (3 * count)
    for j in rng.permutation(3 * count):
        i, w = divmod(int(j), 3)
        t, ta, tb, m, n, k, lda, ldb, ldc = inst[i]
        cols = ((m if ta else k), (k if tb else n), n)[w]
        size = (lda, ldb, ldc)[w] * cols
        if t == "f32":
            bufs[j] = torch.randn(size, device="cuda", generator=gen)
        elif w < 2:
            bufs[j] = torch.randint(-128, 128, (size,), device="cuda", dtype=torch.int8, generator=gen)
        else:
            bufs[j] = torch.zeros(size, device="cuda", dtype=torch.int32)
    codes = {"f32": (S.DataType.R_32F, S.DataType.R_32F, 1.0, 0.0), "i8": (S.DataType.R_8I, S.DataType.R_32I, 1, 0)}

    def run(sel):
        """Build a plan over the instances in `sel` (kept in group order) and time its graph."""
        groups = []  # (type, ta, tb, [instances])
        for i in sel:
            t, ta, tb = inst[i][:3]
            g = i // group
            if not groups or groups[-1][0] != g:
                groups.append((g, t, ta, tb, []))
            groups[-1][4].append(i)
        ids = [i for grp in groups for i in grp[4]]
        pa = S.pointer_array([bufs[3 * i] for i in ids])
        pb = S.pointer_array([bufs[3 * i + 1] for i in ids])
        pc = S.pointer_array([bufs[3 * i + 2] for i in ids])
        plan = S.GroupedPlan([len(g[4]) for g in groups], [g[2] for g in groups], [g[3] for g in groups],
                             [codes[g[1]][0] for g in groups], [codes[g[1]][0] for g in groups],
                             [codes[g[1]][1] for g in groups], [codes[g[1]][2] for g in groups],
                             [codes[g[1]][3] for g in groups], [inst[i][3] for i in ids], [inst[i][4] for i in ids],
                             [inst[i][5] for i in ids], pa, [inst[i][6] for i in ids], pb, [inst[i][7] for i in ids],
                             pc, [inst[i][8] for i in ids])
        ms = time_steps(lambda: plan.graph_launch(stream), max(5, args.steps // 10), 3, stream)
        plan.close()
        flops = {"f32": 0, "i8": 0}
        nbytes = 0
        for i in ids:
            t, ta, tb, m, n, k = inst[i][:6]
            flops[t] += 2 * m * n * k
            e = 4 if t == "f32" else 1
            nbytes += (m * k + k * n) * e + m * n * 4 + 24
        tot = flops["f32"] + flops["i8"]
        return {"instances": len(ids), "ms_per_step": round(ms, 4), "GFLOPs_total": round(tot / (ms * 1e-3) / 1e9, 1),
                "fp32_GFLOP": round(flops["f32"] / 1e9, 3), "int8_GOP": round(flops["i8"] / 1e9, 3),
                "hbm_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "fp32_mfma_frac_if_fp32_alone": round(flops["f32"] / (ms * 1e-3) / 1e12 / PEAK["f32"], 4)}

    big = [i for i in range(count) if min(inst[i][3:6]) >= 256]
    small = [i for i in range(count) if max(inst[i][3:6]) <= 32]
    subsets = {"all": ("all", range(count)), "small": ("small (max dim <= 32)", small),
               "large": ("large (min dim >= 256)", big)}
    # only on request (profiling runs): the fp32 and the int8 instances on their own
    extra = {"fp32": ("fp32 instances", [i for i in range(count) if inst[i][0] == "f32"]),
             "int8": ("int8 instances", [i for i in range(count) if inst[i][0] == "i8"])}
    out = {}
    for key, (label, sel) in list(subsets.items()) + list(extra.items()):
        if (only is None and key in subsets) or only == key:
            out[label] = run(sel)
    del bufs
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 (batches shard by "
                                   "instance with no exchange; see DESIGN.md)"}))
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    stream = torch.cuda.current_stream()
    if args.copy_probe:
        n = args.copy_probe << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        ms = time_steps(lambda: dst.copy_(src), args.steps, args.warmup, stream)
        print(json.dumps({"copy_probe_MiB": args.copy_probe, "ms_per_copy": round(ms, 5),
                          "read_plus_write_GBs": round(2 * n / (ms * 1e-3) / 1e9, 1)}), flush=True)
        return
    if args.case:
        cid, _, form = args.case.partition(":")
        if cid == "c5":
            print(json.dumps({"case": args.case, "result": grouped_c5(args, stream, only=form or "all")}), flush=True)
        else:
            print(json.dumps({"case": args.case, "result": secondary(args, stream, only=(cid, form))}), flush=True)
        return
    prob = Problem("i8", 1, 1, 512, 512, 512, 64, alpha=1, beta=args.beta, seed=SEED + 2)
    ms = time_steps(lambda: prob.call(0, stream), args.steps, args.warmup, stream)
    sec = ms * 1e-3
    tops = prob.work() / sec / 1e12
    q = prob.bytes()
    achieved_gbs = q / sec / 1e9
    traffic = None
    prof = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "dma_stream_gemm_i8_TT_512x64_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("beta%d" % args.beta)
    line = {
        "metric": METRIC,
        "value": round(tops, 3),
        "unit": "TOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "ssblasGemmBatchedEx int8->int32 TT m=n=k=512 batchCount=64 lda=ldb=ldc=512 "
                               f"alpha=1 beta={args.beta}, separate allocations, device pointer arrays",
                   "batchCount": 64, "m": 512, "n": 512, "k": 512, "beta": args.beta},
        "mfma_frac": round(tops / PEAK["int8"], 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": q},
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline_i8(prob, threads)
    if args.secondary:
        line["secondary"] = secondary(args, stream)
    if args.grouped or args.secondary:
        line["grouped_c5"] = grouped_c5(args, stream)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
