"""Status codes of the grouped entry points that are decided before any
device work (no GPU needed): the same rules as the batched call applied per
group and per instance."""
from __future__ import annotations

import pytest

import ssblasex_amd as S

R_8I, R_32I, R_32F, R_64F, R_16F = 0, 3, 4, 5, 2


def args(**over):
    a = dict(group_size=[2], transa=[0], transb=[1], Atype=[R_32F], Btype=[R_32F], Ctype=[R_32F], alpha=[1.0],
             beta=[0.0], m=[4, 4], n=[4, 4], k=[4, 4], A=0, lda=[4, 4], B=0, ldb=[4, 4], C=0, ldc=[4, 4],
             computeType=0)
    a.update(over)
    return a


def create(**over):
    a = args(**over)
    st, plan = S.GroupedPlan.create(a["group_size"], a["transa"], a["transb"], a["Atype"], a["Btype"], a["Ctype"],
                                    a["alpha"], a["beta"], a["m"], a["n"], a["k"], a["A"], a["lda"], a["B"],
                                    a["ldb"], a["C"], a["ldc"], computeType=a["computeType"])
    assert plan is None
    return st


CASES = [
    ("no groups", dict(group_size=[], transa=[], transb=[], Atype=[], Btype=[], Ctype=[], alpha=[], beta=[], m=[],
                       n=[], k=[], lda=[], ldb=[], ldc=[]), S.Status.INTERNAL_ERROR),
    ("empty groups", dict(group_size=[0], m=[], n=[], k=[], lda=[], ldb=[], ldc=[]), S.Status.INTERNAL_ERROR),
    ("negative group size", dict(group_size=[-1], m=[], n=[], k=[], lda=[], ldb=[], ldc=[]),
     S.Status.INVALID_VALUE),
    ("OP_C", dict(transa=[2]), S.Status.INVALID_VALUE),
    ("OP_C int8", dict(transb=[2], Atype=[R_8I], Btype=[R_8I], Ctype=[R_32I], alpha=[1], beta=[0]),
     S.Status.INVALID_VALUE),
    ("type triple", dict(Ctype=[R_64F]), S.Status.INVALID_VALUE),
    ("fp16", dict(Atype=[R_16F], Btype=[R_16F], Ctype=[R_16F]), S.Status.INVALID_VALUE),
    ("negative m", dict(m=[4, -1]), S.Status.INVALID_VALUE),
    ("negative k", dict(k=[-4, 4]), S.Status.INVALID_VALUE),
    ("lda below rows", dict(lda=[4, 3]), S.Status.INVALID_VALUE),
    ("ldb below rows (T)", dict(ldb=[3, 4]), S.Status.INVALID_VALUE),
    ("ldc below m", dict(ldc=[4, 2]), S.Status.INVALID_VALUE),
    ("unknown compute type", dict(computeType=77), S.Status.INVALID_VALUE),
    ("ozaki compute type", dict(computeType=6), S.Status.NOTIMPLEMENTED_ERROR),
    ("ozaki s5", dict(computeType=0x105), S.Status.NOTIMPLEMENTED_ERROR),
    ("null pointer arrays", dict(), S.Status.INVALID_VALUE),
]


@pytest.mark.parametrize("name,over,expect", CASES, ids=[c[0] for c in CASES])
def test_grouped_status(name, over, expect):
    assert create(**over) == expect


def test_one_shot_entry_same_status():
    a = args(transa=[2])
    st = S.ssblasGemmGroupedBatchedEx(a["group_size"], a["transa"], a["transb"], a["Atype"], a["Btype"],
                                      a["Ctype"], a["alpha"], a["beta"], a["m"], a["n"], a["k"], a["A"], a["lda"],
                                      a["B"], a["ldb"], a["C"], a["ldc"])
    assert st == S.Status.INVALID_VALUE


def test_null_plan_handles():
    lib = __import__("ssblasex_amd._lib", fromlist=["load"]).load()
    assert lib.ssblasGroupedPlanExecute(None, None) == S.Status.INVALID_VALUE
    assert lib.ssblasGroupedPlanGraphLaunch(None, None) == S.Status.INVALID_VALUE
    assert lib.ssblasGroupedPlanDestroy(None) == S.Status.SUCCESS


def test_argument_lengths_checked_in_python():
    with pytest.raises(ValueError):
        create(m=[4])
