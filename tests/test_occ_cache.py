"""The persistent grids are sized from a per-kernel occupancy cache.  Every
batched tile kernel has the same C++ function type, so the cache must key on
the kernel's address: checked here with the host compiler on the cache
header itself (no GPU)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "ssblasex_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_distinct_kernels_of_one_type_get_distinct_entries(tmp_path):
    exe = tmp_path / "occ_cache_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{CSRC}", str(ROOT / "tests" / "native" / "occ_cache_check.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"


def test_no_type_keyed_occupancy_cache_left():
    """Grid sizing goes through kernel_occupancy() everywhere (no static
    per-type or per-function caches that could alias kernels)."""
    for src in list(CSRC.glob("*.hpp")) + list(CSRC.glob("*.hip")):
        text = src.read_text()
        if src.name == "occupancy.hpp":
            continue
        assert "hipOccupancyMaxActiveBlocksPerMultiprocessor" not in text, src.name
