"""Python mirror of the grouped entry points (include/ssblasBatch.h)."""
from __future__ import annotations

from . import _lib


class GroupedPlan:
    """Placeholder: the grouped kernel is not built yet."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("grouped plans are not implemented yet")


def ssblasGemmGroupedBatchedEx(*args, **kwargs):
    raise NotImplementedError("grouped calls are not implemented yet")
