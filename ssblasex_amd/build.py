"""Build libssblasGemmBatchedEx.so in-tree for gfx950.

    python -m ssblasex_amd.build            # incremental
    python -m ssblasex_amd.build --force    # rebuild everything

HIP translation units (*.hip) are compiled with hipcc for gfx950, host-only
units (*.cpp) with g++ against the HIP runtime headers, and everything is
linked into ssblasex_amd/lib/libssblasGemmBatchedEx.so.  Objects go to
ssblasex_amd/_build/.  The example-compatible driver
(ssblasex_amd/example/ssblas_example.cpp) is linked against it into
ssblasex_amd/lib/ssblas_example.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIBDIR = PKG / "lib"
LIBNAME = "libssblasGemmBatchedEx.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("SSBLAS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # packed fp32 instructions stay off: the float code is written scalar
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}",
]


def library_path() -> Path:
    return LIBDIR / LIBNAME


def _headers() -> list[Path]:
    return sorted(CSRC.glob("*.hpp")) + sorted((ROOT / "include").glob("*.h"))


_INCLUDE = re.compile(r'^\s*#\s*include\s+"([^"]+)"', re.M)


def _deps(src: Path) -> list[Path]:
    """src and every file it includes with #include "..." (transitively)."""
    seen, todo = [], [src]
    while todo:
        f = todo.pop()
        if f in seen or not f.exists():
            continue
        seen.append(f)
        for name in _INCLUDE.findall(f.read_text()):
            todo.append((f.parent / name).resolve())
    return seen


def _needs(obj: Path, deps: list[Path]) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def _compile(src: Path, force: bool) -> tuple[Path, str | None]:
    obj = OBJ / (src.name + ".o")
    if not force and not _needs(obj, _deps(src)):
        return obj, None
    if src.suffix == ".hip":
        cmd = [str(ROCM / "bin" / "hipcc"), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [os.environ.get("CXX", "g++"), *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{log}")
    return obj, log or None


def build(force: bool = False, jobs: int | None = None, verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    # biggest translation units first
    sources.sort(key=lambda p: (p.suffix != ".hip", p.name))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, log in ex.map(lambda s: _compile(s, force), sources):
            objs.append(obj)
            if log and verbose:
                print(log, file=sys.stderr)
    lib = library_path()
    if force or _needs(lib, objs):
        tmp = lib.with_suffix(".so.tmp")
        cmd = [str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC",
               *map(str, objs), "-o", str(tmp)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}{r.stderr}")
        os.replace(tmp, lib)
    _build_example(lib, force)
    return lib


def example_path() -> Path:
    return LIBDIR / "ssblas_example"


def _build_example(lib: Path, force: bool) -> None:
    """The example-compatible command-line driver, linked against the library."""
    src = PKG / "example" / "ssblas_example.cpp"
    exe = example_path()
    if not force and not _needs(exe, [src, lib] + _headers()):
        return
    cmd = [os.environ.get("CXX", "g++"), *CXX_FLAGS, "-fopenmp", str(src), "-o", str(exe),
           f"-L{LIBDIR}", "-lssblasGemmBatchedEx", f"-L{ROCM / 'lib'}", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM / 'lib'}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"example build failed\n{' '.join(cmd)}\n{r.stdout}{r.stderr}")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.jobs, verbose=a.verbose))


if __name__ == "__main__":
    main()
