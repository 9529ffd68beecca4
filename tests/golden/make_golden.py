"""Generate tests/golden/closed_forms.json.

The reference ships no golden vectors, known-answer tests or fixtures for
ssblasGemmBatchedEx (its example only prints a max error), so every vector
here is derived from what the mathematics fixes, independently of both the
oracle and the library:

* exact integer arithmetic (Python ints) for integer-valued inputs, with the
  int32 wrap written as a mod 2**32 reduction;
* hand-derived IEEE-754 round-to-nearest-even results for the cases that
  pin the reference's operation order (512-wide K blocks, the beta
  pre-multiply, the fused multiply-add chain), re-checked below with exact
  rational arithmetic and an explicit RNE rounding function.

Each case says which implementation it applies to: "oracle" (the CPU
restatement must reproduce the reference's exact rounding) and/or "gpu"
(exact cases the MI355X library must also hit bit for bit).

Run: python tests/golden/make_golden.py   (rewrites closed_forms.json)
"""
from __future__ import annotations

import json
from fractions import Fraction
from pathlib import Path

OUT = Path(__file__).resolve().parent / "closed_forms.json"


# --------------------------------------------------------------------------
# exact IEEE round-to-nearest-even of a rational to p significand bits
# --------------------------------------------------------------------------
def rne(x: Fraction, p: int) -> Fraction:
    if x == 0:
        return Fraction(0)
    s = -1 if x < 0 else 1
    a = abs(x)
    e = 0
    while a >= 2:
        a /= 2
        e += 1
    while a < 1:
        a *= 2
        e -= 1
    scaled = a * (1 << (p - 1))  # in [2^(p-1), 2^p)
    q, r = divmod(scaled.numerator, scaled.denominator)
    half = Fraction(r, scaled.denominator)
    if half > Fraction(1, 2) or (half == Fraction(1, 2) and q % 2 == 1):
        q += 1
    return s * Fraction(q) * Fraction(2) ** (e - (p - 1))


P = {"f32": 24, "f64": 53}


def fma(a, b, c, p):
    return rne(Fraction(a) * Fraction(b) + Fraction(c), p)


def hexf(x: Fraction) -> str:
    return float(x).hex()


def reference_order(a_row, b_col, alpha, beta, c0, p, block=512):
    """Exact-arithmetic statement of the documented order for one element:
    per 512-block fused chain, beta pre-multiply at the first block,
    C = fma(acc, alpha, C) per block."""
    k = len(a_row)
    c = Fraction(c0)
    for k0 in range(0, k, block):
        if k0 == 0:
            c = Fraction(0) if beta == 0 else rne(c * Fraction(beta), p)
        acc = Fraction(0)
        for kk in range(k0, min(k, k0 + block)):
            acc = fma(a_row[kk], b_col[kk], acc, p)
        c = fma(acc, alpha, c, p)
    return c


cases = []

# 1. 512-wide K blocks: 2^(p) absorbs 1.0 inside a block (ties to even), the
#    second block's 512 ones survive.  Unblocked sequential order would give
#    2^p, the reference's order 2^p + 512.
for t in ("f32", "f64"):
    p = P[t]
    big = Fraction(2) ** p
    b = [big] + [Fraction(1)] * 1023
    a = [Fraction(1)] * 1024
    expect = big + 512
    assert reference_order(a, b, 1, 0, 0, p) == expect
    cases.append(dict(name=f"block512_{t}", type=t, transa=0, transb=0, m=1, n=1, k=1024, lda=1, ldb=1024,
                      ldc=1, alpha=hexf(Fraction(1)), beta=hexf(Fraction(0)),
                      A=[[hexf(x) for x in a]], B=[[hexf(x) for x in b]], C0=[[hexf(Fraction(0))]],
                      expect=[[hexf(expect)]], applies=["oracle"],
                      why="2^p + 1 rounds to 2^p (tie, even) in block 0; block 1 sums to 512; "
                          "C = fma(512, 1, 2^p) = 2^p + 512 exactly"))

# 2. beta pre-multiply rounds on its own: fl(beta*C0) = 1 + 2^(2-p) drops
#    2^(2-2p); acc = -(1 + 2^(2-p)) cancels it exactly -> 0.  A fused
#    beta*C0 + acc would leave 2^(2-2p).
for t in ("f32", "f64"):
    p = P[t]
    u = Fraction(2) ** (1 - p)  # ulp(1)
    beta = 1 + u
    c0 = 1 + u
    bval = -(1 + 2 * u)
    assert rne(beta * c0, p) == 1 + 2 * u
    expect = Fraction(0)
    assert reference_order([Fraction(1)], [bval], 1, beta, c0, p) == expect
    cases.append(dict(name=f"beta_prepass_{t}", type=t, transa=0, transb=0, m=1, n=1, k=1, lda=1, ldb=1,
                      ldc=1, alpha=hexf(Fraction(1)), beta=hexf(beta), A=[[hexf(Fraction(1))]],
                      B=[[hexf(bval)]], C0=[[hexf(c0)]], expect=[[hexf(expect)]], applies=["oracle"],
                      why="fl(beta*C0) = 1+2u exactly cancels acc = -(1+2u); fused would give u^2"))

# 3. the K chain is fused: a1*b1 = 1 + 2h + h^2 (h = 2^-(p/2 rounded)) is
#    exact inside fma; acc0 = -(1 + 2h) cancels all but h^2.  An unfused
#    chain rounds a1*b1 first and returns 0.
for t, h in (("f32", Fraction(1, 2 ** 12)), ("f64", Fraction(1, 2 ** 27))):
    p = P[t]
    a = [Fraction(1), 1 + h]
    b = [-(1 + 2 * h), 1 + h]
    expect = h * h
    assert rne((1 + h) * (1 + h), p) == 1 + 2 * h  # the unfused product loses h^2
    assert reference_order(a, b, 1, 0, 0, p) == expect
    cases.append(dict(name=f"fused_chain_{t}", type=t, transa=0, transb=0, m=1, n=1, k=2, lda=1, ldb=2, ldc=1,
                      alpha=hexf(Fraction(1)), beta=hexf(Fraction(0)), A=[[hexf(x) for x in a]],
                      B=[[hexf(x) for x in b]], C0=[[hexf(Fraction(0))]], expect=[[hexf(expect)]],
                      applies=["oracle"], why="fma(1+h, 1+h, -(1+2h)) = h^2; unfused gives 0"))

# 4. k == 0 leaves C unchanged (no K block runs), beta not applied.
for t in ("f32", "f64", "i8"):
    if t == "i8":
        c0, alpha, beta = [[7, -9]], 3, 5
        cases.append(dict(name="k0_unchanged_i8", type=t, transa=0, transb=0, m=2, n=1, k=0, lda=2, ldb=1, ldc=2,
                          alpha=alpha, beta=beta, A=[[0, 0]], B=[[0]], C0=c0, expect=c0, applies=["oracle", "gpu"],
                          why="no K block: C untouched"))
    else:
        c0 = [[hexf(Fraction(3, 2)), hexf(Fraction(-2))]]
        cases.append(dict(name=f"k0_unchanged_{t}", type=t, transa=0, transb=0, m=2, n=1, k=0, lda=2, ldb=1,
                          ldc=2, alpha=hexf(Fraction(2)), beta=hexf(Fraction(3)), A=[[hexf(Fraction(0))] * 2],
                          B=[[hexf(Fraction(0))]], C0=c0, expect=c0, applies=["oracle", "gpu"],
                          why="no K block: C untouched"))

# 5. int32 wrap-around of alpha*S + beta*C (exact modulo 2^32).
def wrap32(v: int) -> int:
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= (1 << 31) else v


for name, alpha, beta, c0 in (("i8_wrap_alpha", 1 << 17, 3, 1000), ("i8_wrap_beta", -1, -(1 << 31), 1)):
    a = [127, 127]
    b = [127, 127]
    s = sum(x * y for x, y in zip(a, b))
    expect = wrap32(alpha * s + (beta * c0 if beta != 0 else 0))
    cases.append(dict(name=name, type="i8", transa=0, transb=0, m=1, n=1, k=2, lda=1, ldb=2, ldc=1,
                      alpha=alpha, beta=beta, A=[a], B=[b], C0=[[c0]], expect=[[expect]],
                      applies=["oracle", "gpu"], why="(alpha*S + beta*C) mod 2^32"))

# 6. beta == 0 (also -0.0 for fp) never reads C: NaN in C0 is overwritten.
for t in ("f32", "f64"):
    nan = "nan"
    for bz in ("0x0.0p+0", "-0x0.0p+0"):
        cases.append(dict(name=f"beta_zero_nan_{t}_{'neg' if bz.startswith('-') else 'pos'}", type=t, transa=0,
                          transb=0, m=1, n=1, k=1, lda=1, ldb=1, ldc=1, alpha=hexf(Fraction(2)), beta=bz,
                          A=[[hexf(Fraction(3))]], B=[[hexf(Fraction(5))]], C0=[[nan]], expect=[[hexf(Fraction(30))]],
                          applies=["oracle", "gpu"], why="beta == 0: C = alpha*acc, C0 not read"))

# 7. integer-valued inputs, every op pair, padded leading dimensions: all
#    arithmetic exact, so the result is fixed by integer arithmetic alone.
#    Distinct entries make a transposed operand or a wrong index visible.
m, n, k = 3, 2, 4
for t in ("f32", "f64", "i8"):
    for ta in (0, 1):
        for tb in (0, 1):
            ra, ca = (m, k) if ta == 0 else (k, m)
            rb, cb = (k, n) if tb == 0 else (n, k)
            lda, ldb, ldc = ra + 2, rb + 1, m + 3
            # stored values: distinct small integers, padding = sentinel 99
            Abuf = [99] * (lda * ca)
            Bbuf = [99] * (ldb * cb)
            Cbuf = [99] * (ldc * n)
            for c in range(ca):
                for r in range(ra):
                    Abuf[r + c * lda] = ((r * 7 + c * 3) % 11) - 5
            for c in range(cb):
                for r in range(rb):
                    Bbuf[r + c * ldb] = ((r * 5 + c * 2) % 9) - 4
            for c in range(n):
                for r in range(m):
                    Cbuf[r + c * ldc] = (r + 1) * (c + 2) - 3

            def opA(i, kk):
                return Abuf[i + kk * lda] if ta == 0 else Abuf[kk + i * lda]

            def opB(kk, j):
                return Bbuf[kk + j * ldb] if tb == 0 else Bbuf[j + kk * ldb]

            alpha, beta = 3, -2
            exp = list(Cbuf)
            for j in range(n):
                for i in range(m):
                    s = sum(opA(i, kk) * opB(kk, j) for kk in range(k))
                    v = alpha * s + beta * Cbuf[i + j * ldc]
                    exp[i + j * ldc] = wrap32(v) if t == "i8" else v
            enc = (lambda v: v) if t == "i8" else (lambda v: hexf(Fraction(v)))
            cases.append(dict(name=f"exact_int_{t}_{'NT'[ta]}{'NT'[tb]}", type=t, transa=ta, transb=tb, m=m, n=n,
                              k=k, lda=lda, ldb=ldb, ldc=ldc, alpha=enc(alpha), beta=enc(beta),
                              A=[[enc(v) for v in Abuf]], B=[[enc(v) for v in Bbuf]], C0=[[enc(v) for v in Cbuf]],
                              expect=[[enc(v) for v in exp]], applies=["oracle", "gpu"],
                              why="integer-valued inputs: every product and sum exact; padding untouched"))


def main():
    OUT.write_text(json.dumps({"generator": "tests/golden/make_golden.py", "cases": cases}, indent=1))
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
