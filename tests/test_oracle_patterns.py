"""Pins of the oracle's selected-area and centre-shifted patterns (oracle/cubo_oracle.c oc_pattern).

Each check is independent of the oracle's own loop:
- the reference's scalar Neumaier transcription (CompensatedSumTests.swift:10-21), which its TwoSum mean must match
  bit for bit on random values with NaN, +-Inf and -0 columns (CompensatedSumTests.swift:150-211);
- exact rational sums (fractions.Fraction) that the compensated mean must meet to a few ulps;
- NumPy's maximum over the selected (weighted) patterns, and hand-built +-0 ties for the two tie rules;
- a closed form for the rounded shift: beams planted at the origins land on the map's rounded mean;
- NumPy's np.roll of each pattern for random integral shifts;
- the refusals of SelectedDiffraction.swift:60-102.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O


def _lcg_floats(seed: int, n: int) -> np.ndarray:
    """CompensatedSumTests.swift:28-35: Float(Double(state >> 11) / 2^53 - 0.35) * 3e3 with the 64-bit LCG, then
    every 7th value times 1e-6 and every 11th from 3 times 1e6 (Float32 arithmetic)."""
    a, b, m = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    state = seed
    out = np.empty(n, np.float32)
    for k in range(n):
        state = (state * a + b) & m
        out[k] = np.float32((state >> 11) / float(1 << 53) - 0.35) * np.float32(3e3)
    out[0::7] *= np.float32(1e-6)
    out[3::11] *= np.float32(1e6)
    return out


def _specials(v: np.ndarray, dc: int, scan_count: int) -> None:
    """CompensatedSumTests.swift:38-47."""
    v[2 * dc + 0] = np.nan
    v[1 * dc + 1] = np.inf
    v[0 * dc + 2] = -np.inf
    v[3 * dc + 3] = np.inf
    v[5 * dc + 3] = -np.inf
    v[4 * dc + 4] = np.inf
    v[6 * dc + 4] = 5
    v[np.arange(scan_count) * dc + 5] = -0.0
    v[np.arange(scan_count) * dc + 6] = 0.0
    v[3 * dc + 7] = np.finfo(np.float32).max
    v[4 * dc + 7] = np.finfo(np.float32).max


class _OldNeumaier:
    """The scalar loop of CompensatedSumTests.swift:10-21."""

    def __init__(self, start: float):
        self.sum = start
        self.corr = start

    def add(self, x: float) -> None:
        a = self.sum
        t = a + x
        if math.isfinite(t):
            e = (a - t) + x if abs(a) >= abs(x) else (x - t) + a
            if e != 0:
                self.corr += e
        self.sum = t

    @property
    def value(self) -> float:
        return self.sum + self.corr


def _same_bits(a, b) -> bool:
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return a.shape == b.shape and bool(np.all((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))))


def test_mask_and_weighted_means_equal_the_scalar_neumaier():
    """CompensatedSumTests.swift:184-211 (maskAndWeightedMeansEqualTheScalarNeumaier), on the oracle."""
    npos, n = 20, 143
    v = _lcg_floats(2026, npos * n)
    _specials(v, n, npos)
    cube = v.reshape(4, 5, 13, 11)
    mask = np.arange(npos) % 3 != 1
    weights = (np.arange(npos) % 7) * 0.37 + 0.1
    plain = [_OldNeumaier(0.0) for _ in range(n)]
    weighted = [_OldNeumaier(0.0) for _ in range(n)]
    wsum = _OldNeumaier(0.0)
    count = 0.0
    for k in range(npos):
        for i in range(n):
            weighted[i].add(float(v[k * n + i]) * float(weights[k]))
        wsum.add(float(weights[k]))
        if not mask[k]:
            continue
        count += 1
        for i in range(n):
            plain[i].add(float(v[k * n + i]))
    assert _same_bits(O.pattern(cube, "mean", take=mask), [p.value / count for p in plain])
    assert _same_bits(O.pattern(cube, "mean", weights=weights), [w.value / wsum.value for w in weighted])


def test_mask_and_weighted_means_with_zero_weights_and_tails():
    """CompensatedSumTests.swift:148-182 (437 pixels, a scan row with no picked position, weights with exact zeros so
    0 * Inf = NaN), against the scalar Neumaier."""
    R, Cs, qy, qx = 5, 7, 19, 23
    npos, n = R * Cs, qy * qx
    v = _lcg_floats(2026, npos * n)
    _specials(v, n, npos)
    v[9 * n + n - 1] = np.nan
    v[10 * n + n - 2] = np.inf
    v[12 * n + n - 2] = -np.inf
    cube = v.reshape(R, Cs, qy, qx)
    ks = np.arange(npos)
    mask = ((ks * 7) % 5 != 2) & (ks // Cs != 2)
    pat = [0.37, 0, -1.25, 2.5, 3.0, 0.125]
    weights = np.array([0.0 if k % 6 == 1 else pat[k % 6] + k * 1e-3 for k in range(npos)])
    plain = [_OldNeumaier(0.0) for _ in range(n)]
    weighted = [_OldNeumaier(0.0) for _ in range(n)]
    for k in range(npos):
        for i in range(n):
            weighted[i].add(float(v[k * n + i]) * float(weights[k]))
            if mask[k]:
                plain[i].add(float(v[k * n + i]))
    total = _OldNeumaier(0.0)
    for w in weights:
        total.add(float(w))
    assert _same_bits(O.pattern(cube, "mean", take=mask), [p.value / mask.sum() for p in plain])
    assert _same_bits(O.pattern(cube, "mean", weights=weights), [w.value / total.value for w in weighted])


def test_compensated_means_are_the_exact_sums_to_a_few_ulps():
    rng = np.random.default_rng(3)
    R, Cs, qy, qx = 6, 5, 4, 6
    npos = R * Cs
    x = (rng.normal(size=(npos, qy * qx)) * 10.0 ** rng.integers(-6, 7, size=(npos, qy * qx))).astype(np.float32)
    cube = x.reshape(R, Cs, qy, qx)
    take = rng.random(npos) < 0.6
    w = rng.normal(size=npos)
    m = O.pattern(cube, "mean", take=take)
    mw = O.pattern(cube, "mean", weights=w)
    tw = sum(Fraction(float(a)) for a in w)
    for i in range(qy * qx):
        exact = sum(Fraction(float(x[k, i])) for k in range(npos) if take[k]) / int(take.sum())
        assert abs(Fraction(m[i]) - exact) <= abs(exact) * Fraction(4, 2 ** 52) + Fraction(1, 2 ** 1000)
        # the products are rounded to Double first (Double(x) * w), then summed exactly; the divisor is the
        # Neumaier sum of the weights, itself within an ulp of the exact sum
        exact_w = sum(Fraction(float(x[k, i]) * float(w[k])) for k in range(npos))
        assert abs(Fraction(mw[i]) * tw - exact_w) <= (abs(exact_w) + sum(abs(Fraction(float(x[k, i]) * float(w[k])))
                                                                        for k in range(npos)) * Fraction(1, 2 ** 40)) * Fraction(4, 2 ** 52) * 4


def test_max_is_numpys_over_the_selected_weighted_patterns():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(24, 30)).astype(np.float32)
    x[7, 3] = np.nan
    cube = x.reshape(4, 6, 5, 6)
    take = rng.random(24) < 0.5
    take[7] = True
    w = rng.normal(size=24)
    np.testing.assert_array_equal(O.pattern(cube, "max", take=take), np.max(x[take].astype(np.float64), axis=0))
    np.testing.assert_array_equal(O.pattern(cube, "max", weights=w), np.max(x.astype(np.float64) * w[:, None], axis=0))
    neg = -np.abs(x) - 1
    neg[7, 3] = -5
    assert np.all(O.pattern(neg.reshape(cube.shape), "max", take=take) < 0)     # seeded at -inf


def test_signed_zero_ties_first_for_patterns_later_for_mean_and_max_shifted():
    # pixel 0 sees -0 then +0, pixel 1 +0 then -0
    x = np.zeros((2, 2), np.float32)
    x[0, 0], x[1, 0] = -0.0, 0.0
    x[0, 1], x[1, 1] = 0.0, -0.0
    cube = x.reshape(1, 2, 1, 2)
    first = O.pattern(cube, "max", take=[True, True])
    assert np.signbit(first[0]) and not np.signbit(first[1])
    zero = (np.zeros(2), np.zeros(2))
    later = O.pattern(cube, "max", origin=zero, ties="later")
    assert not np.signbit(later[0]) and np.signbit(later[1])
    assert np.array_equal(np.signbit(O.pattern(cube, "max", origin=zero)), np.signbit(first))


def test_rounded_shift_centres_planted_beams():
    """Each pattern holds one bright pixel at its origin; moved by the map's rounded mean minus the origin, every
    beam lands on round(mean): the mean pattern is one pixel of 1 and the max one pixel of 7."""
    rng = np.random.default_rng(9)
    R, Cs, qy, qx = 5, 6, 16, 24
    npos = R * Cs
    orow = rng.integers(3, 12, size=npos).astype(np.float64)
    ocol = rng.integers(4, 20, size=npos).astype(np.float64)
    # a mean at a half integer would round by the parity of each offset; keep it off the tie
    for o in (orow, ocol):
        while (2 * o.sum()) % (2 * npos) == npos:
            o[0] += 1
    x = np.zeros((npos, qy, qx), np.float32)
    x[np.arange(npos), orow.astype(int), ocol.astype(int)] = 7
    cube = x.reshape(R, Cs, qy, qx)
    cr, cc = round(np.mean(orow)), round(np.mean(ocol))
    m = O.pattern(cube, "mean", origin=(orow, ocol)).reshape(qy, qx)
    mx = O.pattern(cube, "max", origin=(orow, ocol)).reshape(qy, qx)
    want = np.zeros((qy, qx))
    want[cr, cc] = 7
    np.testing.assert_array_equal(mx, want)
    np.testing.assert_allclose(m, want, rtol=1e-15, atol=0)


def test_rounded_shift_is_nprolls_of_each_pattern():
    rng = np.random.default_rng(11)
    R, Cs, qy, qx = 4, 5, 6, 8
    npos = R * Cs
    x = rng.integers(0, 1000, size=(npos, qy, qx)).astype(np.float32)
    orow = rng.normal(scale=9, size=npos)
    ocol = rng.normal(scale=9, size=npos)
    take = rng.random(npos) < 0.7
    sr = np.rint(np.mean(orow) - orow).astype(int)           # half to even, as Double.rounded(.toNearestOrEven)
    sc = np.rint(np.mean(ocol) - ocol).astype(int)
    moved = np.stack([np.roll(x[k].astype(np.float64), (sr[k], sc[k]), axis=(0, 1)) for k in range(npos)])
    cube = x.reshape(R, Cs, qy, qx)
    np.testing.assert_array_equal(O.pattern(cube, "max", take=take, origin=(orow, ocol)).reshape(qy, qx),
                                  moved[take].max(axis=0))
    np.testing.assert_array_equal(O.pattern(cube, "mean", take=take, origin=(orow, ocol)).reshape(qy, qx),
                                  moved[take].sum(axis=0) / take.sum())     # integers: every order is exact
    np.testing.assert_array_equal(O.pattern(cube, "mean", origin=(orow, ocol), ties="later").reshape(qy, qx),
                                  moved.sum(axis=0) / npos)


def test_zero_shift_equals_the_unshifted_selection():
    npos, n = 20, 143
    v = _lcg_floats(7, npos * n)
    _specials(v, n, npos)
    cube = v.reshape(4, 5, 13, 11)
    take = np.arange(npos) % 4 != 0
    w = np.linspace(-1.5, 2.0, npos)
    zero = (np.full(npos, 3.0), np.full(npos, -2.0))
    for stat in ("mean", "max"):
        assert _same_bits(O.pattern(cube, stat, take=take, origin=zero), O.pattern(cube, stat, take=take))
        assert _same_bits(O.pattern(cube, stat, weights=w, origin=zero), O.pattern(cube, stat, weights=w))


def test_refusals():
    cube = np.ones((2, 3, 4, 4), np.float32)
    with pytest.raises(O.OracleError):
        O.pattern(cube, "mean", take=np.zeros(6, bool))
    with pytest.raises(O.OracleError):
        O.pattern(cube, "mean", weights=np.zeros(6))
    O.pattern(cube, "max", weights=np.zeros(6))                 # a max over zero weights is defined
    with pytest.raises(O.OracleError):
        O.pattern(cube, "mean", weights=np.array([1, 2, np.nan, 0, 0, 0.0]))
    with pytest.raises(O.OracleError):
        O.pattern(cube, "mean", take=np.ones(6, bool), weights=np.ones(6))
    with pytest.raises(O.OracleError):
        O.pattern(cube, "mean", origin=(np.array([0, 0, 0, 0, 0, np.inf]), np.zeros(6)))
