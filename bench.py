"""Benchmark of ssblasGemmBatchedEx on one MI355X.

Headline workload: int8 -> int32, transa = transb = T, m = n = k = 512,
batchCount = 64, lda = ldb = ldc = 512, A/B uniform integers in [-128, 127],
C uniform in [-1000, 1000], alpha = 1, beta = 0; every A[i] / B[i] / C[i] its
own allocation, pointer arrays device resident.  `value` is its int8 TOPS.

A step is one ssblasGemmBatchedExAsync call over the whole batch with the
inputs already in HBM.  W untimed warmup steps, then K steps bracketed by
HIP events on the launch stream and device synchronisation.

The same line carries, under "split_vs_native", the metric's second half on
the other benchmark configurations, each timed the same way (its own warm-up,
then max(20, K/10) steps) with the normwise error of one call on a seeded
sample of instances against the double-double FP64 oracle:
  * fp32 TN m=n=k=256, batchCount 4096, N(0,1), alpha 1.5, beta -0.5:
    native fp32 vs BF16x3 vs BF16x6;
  * fp64 NT m=n=k=32, batchCount 65536, U[-1,1]*2^U{-8..8}, beta 0:
    native fp64 vs Ozaki (9 slices, and the automatic slice count);
  * fp32 NN m=n=k=64, batchCount 128, alpha 1, beta 0.5 (separate allocations);
  * the headline batch with beta = 1.
The CPU leg (rank 0 only) times the reference example's baselines on the
host cores -- its OpenMP int8 loop and its per-instance cblas_?gemm loop
(scipy's OpenBLAS) -- and evaluates the errors above with the oracle as the
checker.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--beta 0|1]
                    [--no-cpu-baseline] [--no-split] [--no-grouped] [--secondary]

The line also carries, under "grouped_c5", the grouped 10000-instance uneven
batch (one captured-graph replay per step; --no-grouped leaves it out).
--secondary adds more configurations (fp64 NN 256^3 x512 native vs Ozaki,
the fp32 TN batch at lda=ldb=ldc=264).  --case ID:FORM times
only one configuration (c1, c3, c3l, c4, c4p, c2b1 with a form name such as
native, bf16x6, ozaki_s9 or ozaki_auto; c5 with all, small, large, fp32 or
int8) and prints its own JSON line, for profiler runs.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

METRIC = "batched GEMM TFLOP/s|TOPS & % MFMA peak; split-precision vs native at fixed error"
PEAK = {"int8": 5033.2, "f32": 157.3, "f64": 78.6, "bf16": 2516.6}  # dense, TFLOP/s | TOPS
HBM_PEAK_GBS = 8000.0
SEED = 0x55B1A5
ROOT = os.path.dirname(os.path.abspath(__file__))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--beta", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="leave out the split_vs_native block")
    ap.add_argument("--secondary", action="store_true")
    ap.add_argument("--no-grouped", action="store_true", help="leave out the grouped_c5 block")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only time a device-to-device copy of MIB MiB (HBM calibration for the counters)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--case", default=None, metavar="ID:FORM",
                    help="time only one other configuration, e.g. c4:bf16x6, c3:ozaki_s9, c5:all (c5 also takes small, "
                         "large, fp32, int8; profiling runs)")
    return ap.parse_args()


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


def warm_for(fn, seconds, stream):
    """Untimed calls for about `seconds` of device time (the clocks leave
    idle before a configuration's timed steps)."""
    import torch

    t0 = time.perf_counter()
    n = 0
    while True:
        fn()
        n += 1
        if n % 8 == 0:
            torch.cuda.synchronize()
            if time.perf_counter() - t0 >= seconds:
                return


TYPECODES = {"i8": (0, 3), "f32": (4, 4), "f64": (5, 5)}  # (A/B type, C type) codes of the C ABI


class Problem:
    """One batched GEMM on the device.

    separate=True: every A[i] / B[i] / C[i] its own allocation (the
    reference's callers' layout), drawn with numpy from `seed`.
    separate=False: one allocation per operand, instance i at a fixed
    stride, drawn on the device from `seed` (a handful of dispatches however
    large the batch).  Either way the pointer arrays are device resident."""

    def __init__(self, t, ta, tb, m, n, k, batch, alpha, beta, seed, dist="uniform", pad=0, separate=True):
        import torch

        import ssblasex_amd as S

        self.t, self.ta, self.tb, self.m, self.n, self.k, self.batch = t, ta, tb, m, n, k, batch
        self.alpha, self.beta = alpha, beta
        ra, ca = (m, k) if ta == 0 else (k, m)
        rb, cb = (k, n) if tb == 0 else (n, k)
        self.ra, self.ca, self.rb, self.cb = ra, ca, rb, cb
        self.lda, self.ldb, self.ldc = ra + pad, rb + pad, m + pad
        self.ab, self.cc = TYPECODES[t]
        self.sa, self.sb, self.sc = self.lda * ca, self.ldb * cb, self.ldc * n  # elements per instance
        if separate:
            rng = np.random.default_rng(seed)
            npdt = {"i8": np.int8, "f32": np.float32, "f64": np.float64}[t]

            def draw(sz, c_operand):
                if t == "i8":
                    if c_operand:
                        return rng.integers(-1000, 1001, size=sz, dtype=np.int32)
                    return rng.integers(-128, 128, size=sz, dtype=np.int8)
                if dist == "normal":
                    return rng.standard_normal(sz).astype(npdt)
                if dist == "wide":
                    return (rng.uniform(-1, 1, sz) * np.exp2(rng.integers(-8, 9, sz))).astype(npdt)
                return rng.uniform(-1, 1, sz).astype(npdt)

            a = torch.from_numpy(draw(batch * self.sa, False)).to("cuda")
            b = torch.from_numpy(draw(batch * self.sb, False)).to("cuda")
            c = torch.from_numpy(draw(batch * self.sc, True)).to("cuda")
            # separate allocations, as the reference's callers make them
            self.A = [a[i * self.sa:(i + 1) * self.sa].clone() for i in range(batch)]
            self.B = [b[i * self.sb:(i + 1) * self.sb].clone() for i in range(batch)]
            self.C = [c[i * self.sc:(i + 1) * self.sc].clone() for i in range(batch)]
            del a, b, c
            self.pA, self.pB, self.pC = S.pointer_array(self.A), S.pointer_array(self.B), S.pointer_array(self.C)
        else:
            gen = torch.Generator(device="cuda").manual_seed(seed)
            tdt = {"i8": torch.int8, "f32": torch.float32, "f64": torch.float64}[t]

            def draw(shape, c_operand):
                if t == "i8":
                    if c_operand:
                        return torch.randint(-1000, 1001, shape, device="cuda", dtype=torch.int32, generator=gen)
                    return torch.randint(-128, 128, shape, device="cuda", dtype=torch.int8, generator=gen)
                if dist == "normal":
                    return torch.randn(shape, device="cuda", dtype=tdt, generator=gen)
                u = torch.rand(shape, device="cuda", dtype=tdt, generator=gen) * 2 - 1
                if dist == "wide":
                    e = torch.randint(-8, 9, shape, device="cuda", generator=gen).to(tdt)
                    return u * torch.exp2(e)
                return u

            self.Aflat = draw((batch, self.sa), False)
            self.Bflat = draw((batch, self.sb), False)
            self.Cflat = draw((batch, self.sc), True)
            self.A = self.B = self.C = None
            self.pA = S.pointer_array_strided(self.Aflat, batch, self.sa)
            self.pB = S.pointer_array_strided(self.Bflat, batch, self.sb)
            self.pC = S.pointer_array_strided(self.Cflat, batch, self.sc)
        torch.cuda.synchronize()

    def instance(self, which, i):
        """Device tensor of instance i of operand 'A', 'B' or 'C'."""
        if self.A is not None:
            return {"A": self.A, "B": self.B, "C": self.C}[which][i]
        return {"A": self.Aflat, "B": self.Bflat, "C": self.Cflat}[which][i]

    def host(self, which, idx):
        """Host copies (numpy) of the instances idx of one operand."""
        import torch

        if self.A is not None:
            return [self.instance(which, int(i)).cpu().numpy() for i in idx]
        flat = {"A": self.Aflat, "B": self.Bflat, "C": self.Cflat}[which]
        rows = flat[torch.as_tensor(np.asarray(list(idx), dtype=np.int64), device=flat.device)].cpu().numpy()
        return list(rows)

    def save_c(self):
        import torch

        if self.A is not None:
            return [c.clone() for c in self.C]
        return self.Cflat.clone()

    def restore_c(self, saved):
        if self.A is not None:
            for c, s in zip(self.C, saved):
                c.copy_(s)
        else:
            self.Cflat.copy_(saved)

    def call(self, compute_type=0, stream=None):
        """One ssblasGemmBatchedExAsync call through the C ABI.  The ctypes
        argument list is built once per (compute type, stream) and reused, so
        a step times the library's call, not Python's argument marshalling
        (about 4 us per call, which would bound the smallest batches)."""
        import ssblasex_amd as S
        from ssblasex_amd import _lib

        key = (compute_type, stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream or 0))
        cache = self.__dict__.setdefault("_args", {})
        if key not in cache:
            args, keep = S._gemm_args(self.ta, self.tb, self.m, self.n, self.k, self.alpha, self.pA, self.ab,
                                      self.lda, self.pB, self.ab, self.ldb, self.beta, self.pC, self.cc, self.ldc,
                                      self.batch, compute_type, 0)
            cache[key] = (args + (key[1],), keep)
        st = _lib.load().ssblasGemmBatchedExAsync(*cache[key][0])
        if st != 0:
            raise RuntimeError(f"ssblasGemmBatchedExAsync returned {S.Status(st).name}")

    def kernel(self, compute_type=0):
        import ssblasex_amd as S

        return S.kernel_config(self.ta, self.tb, self.m, self.n, self.k, self.ab, self.ab, self.cc, self.lda, self.ldb,
                               self.ldc, self.batch, compute_type)

    def work(self):
        return 2.0 * self.m * self.n * self.k * self.batch

    def bytes(self):
        es = {"i8": 1, "f32": 4, "f64": 8}[self.t]
        ec = {"i8": 4, "f32": 4, "f64": 8}[self.t]
        per = (self.m * self.k + self.k * self.n) * es + self.m * self.n * ec * (2 if self.beta != 0 else 1)
        return per * self.batch + 24 * self.batch


# ---------------------------------------------------------------------------
# CPU leg: the reference example's baselines on the host cores, and the
# oracle as the checker of the GPU results' errors
# ---------------------------------------------------------------------------
def host_info(threads):
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity_cpus": affinity, "cpu_model": model, "threads": threads,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads(args):
    """Threads for the CPU baselines: every CPU this process may run on,
    capped at the host share of one GPU (OMP_NUM_THREADS where it is set)."""
    if args.cpu_threads:
        return args.cpu_threads
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, cap) if cap > 0 else n


def _oracle():
    sys.path.insert(0, ROOT)
    import oracle

    return oracle


def timed_runs(fn, budget_s=10.0, runs=5):
    """One warm-up call, then calls until `runs` or `budget_s`; median seconds."""
    fn()
    times, total = [], 0.0
    while len(times) < runs and total < budget_s:
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        times.append(dt)
        total += dt
    return float(np.median(times)), len(times)


def cpu_baseline_i8(prob, threads):
    """The example's OpenMP int8 loop (example/include/naive_gemm.hpp:3-155,
    restated in the oracle) over the whole batch."""
    oracle = _oracle()
    idx = range(prob.batch)
    A, B, C = prob.host("A", idx), prob.host("B", idx), prob.host("C", idx)

    def run():
        oracle.naive_baseline_i8(prob.ta, prob.tb, prob.m, prob.n, prob.k, prob.alpha, A, prob.lda, B, prob.ldb,
                                 prob.beta, C, prob.ldc, threads)

    med, n = timed_runs(run)
    ops = 2.0 * prob.m * prob.n * prob.k * prob.batch
    return {"value": round(ops / med / 1e12, 5), "unit": "TOPS", "cores": threads, "kind": "port",
            "sample": f"all {prob.batch} instances, the example's OpenMP int8 loop (example/include/naive_gemm.hpp:3-155) "
                      f"restated in oracle/ssblas_oracle.c; median of {n} runs, {med:.3f} s each"}


def cpu_baseline_blas(prob, threads, max_instances):
    """The example's fp baseline (example/include/naive_gemm.hpp:195-226): a
    loop of cblas_sgemm / cblas_dgemm calls, one per instance, OpenBLAS
    threads = `threads`, on the first `max_instances` instances."""
    oracle = _oracle()
    lib = oracle.find_openblas()
    if lib is None:
        return {"error": "scipy's OpenBLAS not found"}
    nb = min(prob.batch, max_instances)
    idx = range(nb)
    A, B, C = prob.host("A", idx), prob.host("B", idx), prob.host("C", idx)

    def run():
        if oracle.blas_baseline(lib, prob.ab, prob.ta, prob.tb, prob.m, prob.n, prob.k, prob.alpha, A, prob.lda, B,
                                prob.ldb, prob.beta, C, prob.ldc, threads) != 0:
            raise RuntimeError("cblas symbols missing")

    med, n = timed_runs(run)
    flops = 2.0 * prob.m * prob.n * prob.k * nb
    return {"value": round(flops / med / 1e12, 5), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"{nb} of {prob.batch} instances, the example's per-instance cblas_{'s' if prob.t == 'f32' else 'd'}"
                      f"gemm loop (example/include/naive_gemm.hpp:195-226) into scipy's OpenBLAS "
                      f"({os.path.basename(lib)}); median of {n} runs, {med:.3f} s each"}


class ErrorCheck:
    """Normwise error ||C - C_exact||_F / (|alpha| ||A|| ||B|| + |beta| ||C0||),
    worst over a seeded sample of instances, with C_exact from the
    double-double oracle; the reference-order restatement (the oracle) on the
    same sample gives the reference's own error for the ratio."""

    def __init__(self, prob, nsample, threads):
        oracle = _oracle()
        self.prob, self.threads = prob, threads
        rng = np.random.default_rng(SEED)
        self.idx = np.sort(rng.choice(prob.batch, min(nsample, prob.batch), replace=False))
        self.A, self.B, self.C0 = prob.host("A", self.idx), prob.host("B", self.idx), prob.host("C", self.idx)
        p = prob
        self.hi, self.lo = oracle.gemm_dd_batched(p.ta, p.tb, p.m, p.n, p.k, p.alpha, self.A, p.ab, p.lda, self.B,
                                                  p.ldb, p.beta, self.C0, p.ldc, nthreads=threads)

        def frob(bufs, rows, cols, ld):
            return np.array([np.linalg.norm(oracle.colmajor(x, rows, cols, ld).astype(np.float64)) for x in bufs])

        self.scale = abs(p.alpha) * frob(self.A, p.ra, p.ca, p.lda) * frob(self.B, p.rb, p.cb, p.ldb)
        if p.beta != 0:
            self.scale = self.scale + abs(p.beta) * frob(self.C0, p.m, p.n, p.ldc)
        self.c0_dev = prob.save_c()  # every checked call starts from these C values
        ref = [c.copy() for c in self.C0]
        oracle.gemm_batched_ex(p.ta, p.tb, p.m, p.n, p.k, p.alpha, self.A, p.ab, p.lda, self.B, p.ab, p.ldb, p.beta,
                               ref, p.cc, p.ldc, nthreads=threads)
        self.ref_err = self.error(ref)

    def error(self, C):
        oracle = _oracle()
        p = self.prob
        worst = 0.0
        for j, c in enumerate(C):
            d = (oracle.colmajor(c, p.m, p.n, p.ldc).astype(np.float64) - self.hi[j]) - self.lo[j]
            worst = max(worst, float(np.sqrt(np.sum(d * d)) / self.scale[j]))
        return worst

    def gpu_error(self, ct, stream):
        """Error of one call from the initial C (timed steps with beta != 0
        change C, so it is put back first)."""
        import torch

        p = self.prob
        p.restore_c(self.c0_dev)
        p.call(ct, stream)
        torch.cuda.synchronize()
        return self.error(p.host("C", self.idx))


# ---------------------------------------------------------------------------
# the configurations
# ---------------------------------------------------------------------------
C1 = dict(t="f32", ta=0, tb=0, m=64, n=64, k=64, batch=128, alpha=1.0, beta=0.5, seed=SEED + 1)
C2B1 = dict(t="i8", ta=1, tb=1, m=512, n=512, k=512, batch=64, alpha=1, beta=1, seed=SEED + 2)
C3 = dict(t="f64", ta=0, tb=1, m=32, n=32, k=32, batch=65536, alpha=1.0, beta=0.0, seed=SEED + 3, dist="wide",
          separate=False)
C3L = dict(t="f64", ta=0, tb=0, m=256, n=256, k=256, batch=512, alpha=1.0, beta=0.0, seed=SEED + 5, dist="wide",
           separate=False)
C4 = dict(t="f32", ta=1, tb=0, m=256, n=256, k=256, batch=4096, alpha=1.5, beta=-0.5, seed=SEED + 4, dist="normal",
          separate=False)
C4P = dict(C4, pad=8)
FORMS_C4 = [("native", 0), ("bf16x3", 4), ("bf16x6", 5)]
FORMS_C3 = [("native", 0), ("ozaki_s9", 0x109), ("ozaki_auto", 6)]
# (case id, label, problem, forms, error sample size, cpu baseline instances, in the default line)
CONFIGS = [
    ("c4", "fp32 TN m=n=k=256 batch 4096 alpha=1.5 beta=-0.5", C4, FORMS_C4, 8, 512, True),
    ("c3", "fp64 NT m=n=k=32 batch 65536 U[-1,1]*2^U{-8..8}", C3,
     FORMS_C3 + [(f"ozaki_s{s}", 0x100 + s) for s in (6, 7, 8)], 1024, 65536, True),
    ("c1", "fp32 NN m=n=k=64 batch 128 beta=0.5, separate allocations", C1, [("native", 0)], 128, 128, True),
    ("c2b1", "int8 TT m=n=k=512 batch 64 beta=1", C2B1, [("native", 0)], 0, 0, True),
    ("c4p", "fp32 TN m=n=k=256 batch 4096 alpha=1.5 beta=-0.5, lda=ldb=ldc=264", C4P, FORMS_C4, 8, 0, False),
    ("c3l", "fp64 NN m=n=k=256 batch 512 U[-1,1]*2^U{-8..8}", C3L,
     [("native", 0)] + [(f"ozaki_s{s}", 0x100 + s) for s in range(3, 10)] + [("ozaki_auto", 6)], 8, 0, False),
]
DEFAULT_FORMS = {"c3": {"native", "ozaki_s9", "ozaki_auto"}}


def run_config(cid, args, stream, forms=None, cpu=True, threads=1):
    """Time each form of one configuration; with `cpu`, also its error on
    the seeded sample and the CPU baseline."""
    import torch

    _, label, kw, all_forms, nsample, cpu_inst, _ = next(c for c in CONFIGS if c[0] == cid)
    prob = Problem(**kw)
    check = ErrorCheck(prob, nsample, threads) if (cpu and nsample and kw["t"] != "i8") else None
    steps = max(20, args.steps // 10)
    res = {"label": label, "forms": {}}
    if check is not None:
        res["reference_order_error"] = check.ref_err
        res["error_sample"] = f"{len(check.idx)} of {prob.batch} instances (seeded)"
    for fname, ct in all_forms:
        if forms is not None and fname not in forms:
            continue
        err = check.gpu_error(ct, stream) if check is not None else None
        fn = (lambda: prob.call(ct, stream))
        try:
            warm_for(fn, 0.2, stream)
            ms = time_steps(fn, steps, 3, stream)
        except RuntimeError as e:
            res["forms"][fname] = {"error": str(e)}
            continue
        sec = ms * 1e-3
        rate = prob.work() / sec / 1e12
        gbs = prob.bytes() / sec / 1e9
        peak = PEAK["int8" if kw["t"] == "i8" else kw["t"]]
        ent = {"kernel": prob.kernel(ct), "ms_per_step": round(ms, 5),
               ("TOPS" if kw["t"] == "i8" else "TFLOPs"): round(rate, 3),
               "frac_of_native_mfma_peak": round(rate / peak, 4), "hbm_GBs": round(gbs, 1),
               "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        if err is not None:
            ent["normwise_error"] = err
            ent["ratio_to_reference_order"] = round(err / check.ref_err, 4) if check.ref_err > 0 else None
        if fname.startswith("ozaki"):
            # the library's own a-priori bound for the last (timed) call, every instance
            import ssblasex_amd as S

            st, rep, _, _ = S.ozaki_report(stream, per_instance=False)
            if st == S.Status.SUCCESS:
                ent["reported_bound_max"] = rep.bound_max
                ent["reported_slices"] = [rep.slices_min, rep.slices_max]
                ent["instances_over_target"] = rep.over_target
        res["forms"][fname] = ent
    if cpu and cpu_inst:
        res["cpu_baseline"] = cpu_baseline_blas(prob, threads, cpu_inst)
    del prob, check
    torch.cuda.empty_cache()
    return res


def split_vs_native(args, stream, threads, ids, cpu):
    """The metric's second half: split-precision forms against native on the
    same batch, with their errors."""
    out = {}
    for cid in ids:
        forms = DEFAULT_FORMS.get(cid) if not args.secondary else None
        out[cid] = run_config(cid, args, stream, forms=forms, cpu=cpu, threads=threads)
    # the comparisons the metric names, read off the entries above
    summ = {}
    if "c4" in out and "native" in out["c4"]["forms"]:
        f = out["c4"]["forms"]
        for s in ("bf16x3", "bf16x6"):
            if s in f and "ms_per_step" in f[s]:
                summ[f"c4 {s} speedup over native fp32"] = round(f["native"]["ms_per_step"] / f[s]["ms_per_step"], 3)
                if "ratio_to_reference_order" in f[s]:
                    summ[f"c4 {s} error / native error"] = round(f[s]["normwise_error"] / f["native"]["normwise_error"], 3)
    if "c3" in out and "native" in out["c3"]["forms"]:
        f = out["c3"]["forms"]
        for s in ("ozaki_s9", "ozaki_auto"):
            if s in f and "ms_per_step" in f[s]:
                summ[f"c3 {s} speedup over native fp64"] = round(f["native"]["ms_per_step"] / f[s]["ms_per_step"], 3)
                if "normwise_error" in f[s]:
                    summ[f"c3 {s} error / native error"] = round(f[s]["normwise_error"] / f["native"]["normwise_error"], 3)
    out["summary"] = summ
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
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 (batches shard by instance with "
                                   "no exchange step; a multi-GPU run is not part of this build)"}))
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    stream = torch.cuda.current_stream()
    threads = cpu_threads(args)
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
            forms = {form} if form else None
            print(json.dumps({"case": args.case, "result": run_config(cid, args, stream, forms=forms, cpu=False)}),
                  flush=True)
        return
    # The grouped batch goes first: freeing its 30000 buffers leaves the GPU
    # idle for a while.  The headline batch is drawn next, so that its
    # warm-up follows the other configurations' device work without an idle
    # gap in which the clocks drop.
    line_extra = {}
    if not args.no_grouped:
        line_extra["grouped_c5"] = grouped_c5(args, stream)
    prob = Problem("i8", 1, 1, 512, 512, 512, 64, alpha=1, beta=args.beta, seed=SEED + 2)
    if not args.no_split:
        ids = [c[0] for c in CONFIGS if c[6] or args.secondary]
        line_extra["split_vs_native"] = split_vs_native(args, stream, threads, ids, cpu=not args.no_cpu_baseline)

    ms = time_steps(lambda: prob.call(0, stream), args.steps, args.warmup, stream)
    sec = ms * 1e-3
    tops = prob.work() / sec / 1e12
    q = prob.bytes()
    achieved_gbs = q / sec / 1e9
    traffic, traffic_src = None, None
    prof = os.path.join(ROOT, "profiles", "dma_stream_gemm_i8_TT_512x64_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tj = json.load(f)
        traffic = tj.get("beta%d" % args.beta)
        traffic_src = "profiles/dma_stream_gemm_i8_TT_512x64_traffic.json: " + tj.get("_how", "")
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
        "kernel": prob.kernel(0),
        "mfma_frac": round(tops / PEAK["int8"], 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src, "algorithmic_bytes_per_launch": q},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_i8(prob, threads)
        line["cpu_baseline"]["host"] = host_info(threads)
    line.update(line_extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
