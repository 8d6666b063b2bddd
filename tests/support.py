"""Shared test helpers: the vendored CuboKit pins and the reference's two comparisons.

assert_relative follows Tests/CuboKitTestSupport/TestSupport.swift:135-151 (relative with the floor
1e-6 * max|want|); assert_identical follows :155-163 (bit patterns); near/of_max follow
Tests/CuboKitTestSupport/PhaseSupport.swift:21-45.
"""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cubokit")


def golden(name: str) -> str:
    return os.path.join(GOLDEN, name)


_cache: dict = {}


def virtual_pins() -> dict:
    if "v" not in _cache:
        with open(golden("virtual-pins.json")) as f:
            _cache["v"] = json.load(f)
    return _cache["v"]


def com_pins() -> dict:
    if "c" not in _cache:
        with open(golden("com-dpc-pins.json")) as f:
            _cache["c"] = json.load(f)
    return _cache["c"]


def origin_fit_pins() -> dict:
    if "o" not in _cache:
        with open(golden("origin-fit-pins.json")) as f:
            _cache["o"] = json.load(f)
    return _cache["o"]


def doubles(v) -> np.ndarray:
    return np.array([float(s) for s in v], dtype=np.float64)


def load_cube(name: str) -> np.ndarray:
    return np.load(golden(name if name.endswith(".npy") else name + ".npy"), allow_pickle=False)


def is_integer_dtype(dtype: str) -> bool:
    return dtype.startswith("uint") or dtype.startswith("int")


def geometry_kwargs(g: dict) -> tuple[str, dict]:
    kind = g["kind"]
    kw = {k: v for k, v in g.items() if k != "kind"}
    if kind == "mask":
        kw["weights"] = np.array([np.float32(float(s)) for s in g["weights"]], np.float32)
    return kind, kw


def assert_relative(got, want, t, msg=""):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    assert got.shape == want.shape, f"{msg}: shape {got.shape} vs {want.shape}"
    if want.size == 0:
        return 0.0
    floor = 1e-6 * np.max(np.abs(want))
    den = np.maximum(np.abs(want), floor)
    with np.errstate(invalid="ignore", divide="ignore"):
        d = np.where(den > 0, np.abs(got - want) / den, np.abs(got - want))
    d = np.where(np.isnan(d), np.inf, d)
    worst = float(np.max(d))
    k = int(np.argmax(d))
    assert worst <= t, f"{msg}: worst relative distance {worst} at {k} (got {got[k]!r}, want {want[k]!r})"
    return worst


def assert_identical(got, want, msg=""):
    got = np.ascontiguousarray(got, np.float64)
    want = np.ascontiguousarray(want, np.float64)
    assert got.shape == want.shape, f"{msg}: shape {got.shape} vs {want.shape}"
    gb, wb = got.view(np.uint64), want.view(np.uint64)
    bad = np.nonzero(gb != wb)[0]
    assert bad.size == 0, f"{msg}: differs at {bad[0]}: got {got[bad[0]]!r}, want {want[bad[0]]!r}"


def near(got, want, t, msg=""):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    assert got.shape == want.shape, f"{msg}: count {got.shape} vs {want.shape}"
    if want.size == 0:
        return 0.0
    floor = 1e-6 * np.max(np.abs(want))
    den = np.maximum(np.maximum(np.abs(want), floor), np.finfo(np.float64).tiny)
    d = np.abs(got - want) / den
    worst = float(np.max(np.where(np.isnan(d), np.inf, d)))
    assert worst <= t, f"{msg}: {worst} > {t}"
    return worst


def of_max(got, want, t, msg=""):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    d = float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), np.finfo(np.float64).tiny))
    assert d <= t, f"{msg}: {d} > {t}"
    return d


def lcg_values(seed: int, n: int, modulus: int) -> np.ndarray:
    """Float32((state >> 33) % modulus) of the 64-bit LCG state = state * 6364136223846793005 + 1442695040888963407,
    the generator of CenterOfMassProbeSizeTests.swift:168-172, vectorised: state_k = a^k s + b (1 + a + ... + a^(k-1))
    mod 2^64."""
    a = np.uint64(6364136223846793005)
    b = np.uint64(1442695040888963407)
    with np.errstate(over="ignore"):
        powers = np.cumprod(np.full(n, a, dtype=np.uint64), dtype=np.uint64)          # a^1 .. a^n
        geo = np.cumsum(np.concatenate([[np.uint64(1)], powers[:-1]]), dtype=np.uint64)  # 1 + ... + a^(k-1)
        states = powers * np.uint64(seed) + b * geo
    return ((states >> np.uint64(33)) % np.uint64(modulus)).astype(np.float32)
