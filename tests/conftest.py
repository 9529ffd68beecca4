"""pytest configuration: the `gpu` marker, import paths, built artefacts."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session", autouse=True)
def _built_artefacts():
    """Build the oracle and the library only when nothing is built yet."""
    import oracle
    from ssblasex_amd import _lib

    if not oracle._SO.exists():
        oracle.build()
    if not _lib.library_path().exists():
        from ssblasex_amd.build import build

        build()
    yield
