"""The exactness test behind the sequential-sums mode (host-side check of
its mathematics). A reach row's nonzero f32 entries are all multiples of
u, the smallest of their lowest set bits; when the row total S satisfies
S < 2^53 u, every f64 sum of any subset in any order is exact, so the GPU's
chunked sums equal the CPU engine's sequential ones and the mode may skip
its sequential path. Checked here on seeded reach-like rows: every row
that passes the test sums to the same f64 bits in sequential, reversed,
shuffled and chunked orders (and its prefixes and differences are exact);
rows that fail it can differ, which is why the mode then takes the CPU's
order."""
import math

import numpy as np


def _lsb_exp(x):
    """Exponent of the lowest set bit of a positive f32 (normal or
    subnormal), from its bit pattern."""
    b = int(np.float32(x).view(np.uint32))
    e, m = (b >> 23) & 0xFF, b & 0x7FFFFF
    if e:
        m |= 0x800000
    ctz = (m & -m).bit_length() - 1
    return (e if e else 1) - 150 + ctz


def _exp_floor(x):
    """The fold's cheaper bound on the lowest set bit: 2^(exponent field - 150)."""
    b = int(np.float32(x).view(np.uint32))
    return max((b >> 23) & 0xFF, 1) - 150


def _passes(row, lsb):
    nz = [float(v) for v in row if v != 0]
    if not nz:
        return True
    total = math.fsum(nz)
    return total * (1.0 + 2.0 ** -40) < math.ldexp(1.0, 53 + min(lsb(v) for v in nz))


def _seq(vals):
    t = 0.0
    for v in vals:
        t += float(v)
    return t


def _chunked(vals, chunk):
    parts = [_seq(vals[i:i + chunk]) for i in range(0, len(vals), chunk)]
    t = 0.0
    for p in reversed(parts):
        t += p
    return t


def _reach_rows(rng, n_rows):
    """Products of 1-12 strategy probabilities times range weights (f32, as
    the down pass forms them), some entries zero, some rows with tiny
    probabilities."""
    rows = []
    for _ in range(n_rows):
        n = int(rng.integers(1, 1327))
        r = rng.uniform(0.2, 1.0, n).astype(np.float32)
        for _ in range(int(rng.integers(1, 13))):
            p = rng.uniform(0.0, 1.0, n).astype(np.float32)
            if rng.random() < 0.3:
                p[rng.random(n) < 0.1] = np.float32(rng.uniform(1e-9, 1e-4))
            r = (r * p).astype(np.float32)
        r[rng.random(n) < 0.2] = 0
        rows.append(r)
    return rows


def test_rows_that_pass_sum_exactly_in_any_order():
    rng = np.random.default_rng(20261204)
    passed = failed = 0
    for row in _reach_rows(rng, 300):
        if not _passes(row, _lsb_exp):
            failed += 1
            continue
        passed += 1
        exact = math.fsum(float(v) for v in row)
        assert _seq(row) == exact
        assert _seq(row[::-1]) == exact
        assert _seq(rng.permutation(row)) == exact
        for chunk in (3, 64, 256):
            assert _chunked(row, chunk) == exact
        # prefixes and the "stronger" totals formed as differences
        k = int(rng.integers(0, len(row) + 1))
        lo, hi = _seq(row[:k]), _seq(row[k:][::-1])
        assert lo == math.fsum(float(v) for v in row[:k])
        assert exact - lo == hi
    assert passed > 50 and failed > 10, (passed, failed)


def test_exponent_floor_is_a_weaker_test():
    rng = np.random.default_rng(7)
    for row in _reach_rows(rng, 200):
        if _passes(row, _exp_floor):
            assert _passes(row, _lsb_exp)
        for v in row[row != 0][:50]:
            assert _exp_floor(v) <= _lsb_exp(v)


def test_rows_that_fail_can_depend_on_the_order():
    """A large entry followed by many tiny ones: adding from the top loses
    every tiny entry, from the bottom keeps their sum."""
    row = np.array([1.0] + [2.0 ** -60] * 1000, dtype=np.float32)
    assert not _passes(row, _lsb_exp)
    assert _seq(row) != _seq(row[::-1])


def test_lowest_set_bit():
    assert _lsb_exp(1.0) == 0
    assert _lsb_exp(0.75) == -2
    assert _lsb_exp(3.0) == 0
    assert _lsb_exp(np.float32(2.0 ** -149)) == -149  # smallest subnormal
    assert _lsb_exp(np.float32(2.0 ** -126)) == -126  # smallest normal
    x = np.float32(1.0) + np.float32(2.0 ** -23)
    assert _lsb_exp(x) == -23
