"""The example-compatible driver (ssblasex_amd/lib/ssblas_example): the
reference example's argument grammar (example/main.cpp:189-267) and output
lines ("MAX ERROR ...", "M,N,K,time,[s],Gflops,[Gflops]")."""
from __future__ import annotations

import re
import subprocess

import pytest

from ssblasex_amd.build import example_path

EXE = str(example_path())


def run(*args, timeout=120):
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_usage_on_wrong_argument_count():
    r = run("int", "float")
    assert r.returncode == 255
    assert "INT_TYPE SCALE_TYPE AB_TYPE C_TYPE TRANSA TRANSB BATCH alpha M N K beta" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (("short", "float", "float", "float", "N", "N", 1, 1, 8, 8, 8, 0), "UNKNOWN INT TYPE"),
    (("int", "half", "float", "float", "N", "N", 1, 1, 8, 8, 8, 0), "UNKNOWN SCALE TYPE"),
    (("int", "float", "int", "float", "N", "N", 1, 1, 8, 8, 8, 0), "UNKNOWN AB TYPE"),
    (("int", "float", "float", "float", "X", "N", 1, 1, 8, 8, 8, 0), "UNKNOWN TRANS"),
])
def test_unknown_arguments_rejected_before_device_work(args, msg):
    r = run(*args)
    assert r.returncode == 255 and msg in r.stdout


PERF = re.compile(r"^(\d+),(\d+),(\d+),([0-9.]+),\[s\],([0-9.]+),\[Gflops\]$")


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ("int", "float", "float", "float", "N", "N", 2, 1.0, 257, 129, 300, 0),
    ("int", "double", "double", "double", "T", "N", 2, 1.0, 65, 65, 1025, 0),
    ("longint", "int", "char", "int", "T", "T", 2, 1, 599, 688, 1867, 1),
    ("int", "float", "float", "float", "N", "T", 3, 1.0, 144, 256, 256, 0, 300, 300, 160),
], ids=["f32-NN", "f64-TN", "i8-TT-beta1", "f32-NT-ld"])
def test_example_shapes(args):
    r = run(*args, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[:3] == ["naive_gemm start", "naive_gemm end", "ssblasGemmEx start"]
    assert "SSBLAS_STATUS_SUCCESS" in lines
    err = [l for l in lines if l.startswith("MAX ERROR")]
    assert len(err) == 1
    value = float(err[0].split()[-1])
    if args[3] == "int":
        assert value == 0
    else:
        assert value < (1e-4 if args[3] == "float" else 1e-12)
    perf = [PERF.match(l) for l in lines if PERF.match(l)]
    assert len(perf) == 2
    for p in perf:
        assert tuple(int(x) for x in p.groups()[:3]) == tuple(int(x) for x in args[8:11])
