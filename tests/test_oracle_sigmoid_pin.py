"""Pins the oracle's taylorSeries coefficient tables (calc_fn, functors.h:22-31)
without running the reference's device code.

calc_fn evaluates fn(val) in float for val = 2^(x - FP), then (T)(f * (1 << x)).
The reference builds without fast-math (Makefile:10), so on its GPU the float
sqrt, addition and division are IEEE-rounded, and its expf is documented to be
within 2 ulp of exp over the full range.  For sqrt and 1/x the table is then
fixed bit for bit; for the sigmoid 1 / (1 + exp(-val)) the test enumerates
every float a 2-ulp expf may return for the exact exp(-val) (decimal arithmetic
to 60 digits) and runs the rest of the functor in IEEE float32.  Where all of
them give one table entry (every entry at FP 13, all but one at FP 20, all
but five at FP 26) it must be the oracle's: those entries are pinned for any
expf within its documented bound.  For the six others the bound leaves two
adjacent float products open (OPEN below); the oracle's entry must be one of
the two, and which one the reference's expf gives stays unpinned.  A planted
mistake in the oracle's functor (exp(val) for exp(-val), the product taken in
double, a 64-bit `1 << x`) changes pinned entries and fails the test.
"""
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle as O

getcontext().prec = 60
F32 = np.float32


def int_pow2(x):
    """(float)(1 << x) with the reference's 32-bit int shift: 1 << 31 is INT_MIN, 0 from 32 on."""
    if x >= 32:
        return F32(0.0)
    v = 1 << x
    return F32(v - (1 << 32) if v >= 1 << 31 else v)


def to_u64(v):
    """float -> uint64 as the reference's GPU converts: negatives and NaN to 0, saturating above."""
    v = float(v)
    if not v > 0.0:
        return 0
    if v >= 2.0 ** 64:
        return (1 << 64) - 1
    return int(v)


def val_of(x, fp):
    return F32(np.inf) if x > fp + 127 else F32(2.0 ** (x - fp))


def exp_candidates(v):
    """Every float32 within 2 ulp of exp(-v) (v a float32 >= 0)."""
    if np.isinf(v):
        return [F32(0.0)]  # expf(-inf) is exactly 0
    exact = (-Decimal(float(v))).exp()
    c = F32(float(exact))
    ulp = Decimal(float(np.spacing(c if c > 0 else F32(np.finfo(np.float32).smallest_subnormal))))
    out = []
    lo = c
    for _ in range(4):
        lo = np.nextafter(lo, F32(-1.0), dtype=np.float32) if lo > 0 else lo
    cand = lo
    for _ in range(9):
        if cand >= 0 and abs(Decimal(float(cand)) - exact) <= 2 * ulp:
            out.append(F32(cand))
        cand = np.nextafter(cand, F32(np.inf), dtype=np.float32)
    assert out, "no candidate within 2 ulp"
    return out


def sigmoid_entries(x, fp):
    v = val_of(x, fp)
    res = set()
    with np.errstate(over="ignore"):
        for e in exp_candidates(v):
            f = F32(1.0) / (F32(1.0) + e)
            res.add(to_u64(F32(f * int_pow2(x))))
    return res


# table arguments x whose entry a 2-ulp expf leaves open between two values
OPEN = {13: [], 20: [13], 26: [15, 19, 24, 26, 27]}


@pytest.mark.parametrize("fp", [13, 20, 26])
def test_sigmoid_table_pinned_by_expf_error_bound(fp):
    for x in range(0, 65):
        entries = sigmoid_entries(x, fp)
        if x in OPEN[fp]:
            assert len(entries) == 2 and O.calc_fn(2, x, fp) in entries, f"x={x}: {sorted(entries)}"
        else:
            assert entries == {O.calc_fn(2, x, fp)}, f"x={x}: {sorted(entries)}"


@pytest.mark.parametrize("fp", [13, 20, 26])
def test_sqrt_and_inverse_tables_exact(fp):
    with np.errstate(over="ignore", divide="ignore"):
        for x in range(0, 65):
            v = val_of(x, fp)
            assert O.calc_fn(0, x, fp) == to_u64(F32(np.sqrt(v) * int_pow2(x))), f"sqrt x={x}"
            assert O.calc_fn(1, x, fp) == to_u64(F32((F32(1.0) / v) * int_pow2(x))), f"inverse x={x}"


def test_planted_sign_error_is_caught():
    # exp(+val) instead of exp(-val) gives a different table at FP 26
    fp = 26
    wrong = []
    for x in range(0, 40):
        v = val_of(x, fp)
        with np.errstate(over="ignore"):
            f = F32(1.0) / (F32(1.0) + F32(np.exp(v, dtype=np.float32)))
            wrong.append(to_u64(F32(f * int_pow2(x))))
    assert wrong != [O.calc_fn(2, x, fp) for x in range(0, 40)]
