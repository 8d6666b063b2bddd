"""GPU parity of the selected-area and centre-shifted patterns (cubo_hip_pattern_*) against the oracle.

Tolerances (stated per assertion):
- max: bit-identical, NaN where the oracle has NaN (SelectedDiffraction.swift:161; the reference's Metal path is
  held to 1e-5, SelectedAreaMetalTests.swift:38);
- mean on the exact integer kernel (uint16 cube, a mask or no selection): bit-identical;
- mean on the FP64 kernel (Float32 cubes, weights): relative 1e-12 with the reference's floor 1e-6 * max|want|
  (TestSupport.swift:135-151) — compensated sums on both sides, in different orders; non-finite results equal;
- rounded shifts: the same bars, and a closed form (beams planted at their origins land on the rounded mean).
"""
import ctypes as ct

import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import _lib as L
from cubokit_amd import synthetic
from oracle import oracle as O
from tests.support import assert_identical, assert_relative

pytestmark = pytest.mark.gpu

TILINGS = [C.ScanTiling.rows(1), C.ScanTiling.rows(4), C.ScanTiling.automatic]
MEAN, MAX = C.DiffractionStatistic.mean, C.DiffractionStatistic.max


def _lcg_uniform(seed: int, n: int) -> np.ndarray:
    a, b, m = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    out = np.empty(n)
    s = seed
    for k in range(n):
        s = (s * a + b) & m
        out[k] = (s >> 11) / float(1 << 53)
    return out


def _metal_test_cube(rows, columns, qr, qc, seed):
    """SelectedAreaMetalTests.swift:8-18: a Gaussian blob plus noise minus an offset, so sums cancel."""
    u = _lcg_uniform(seed, rows * columns * qr * qc).reshape(rows * columns, qr, qc)
    r = np.arange(qr)[:, None] - qr // 2
    c = np.arange(qc)[None, :] - qc // 2
    d = (r * r + c * c).astype(np.float64)
    p = np.arange(rows * columns, dtype=np.float64)[:, None, None]
    v = 400 * np.exp(-d / 30) * (1 + 0.3 * np.sin(p)) + 20 * (u - 0.5) - 30
    return v.astype(np.float32).reshape(rows, columns, qr, qc)


def _selection(n):
    """SelectedAreaMetalTests.swift:23-27."""
    k = np.arange(n)
    mask = (k * 7 + 3) % 5 != 0
    weights = np.where(mask, 0.25 + ((k * 13) % 17) / 8 - np.where(k % 6 == 0, 1.5, 0.0), 0.0)
    return mask, weights


def _scan(kind, mask, weights):
    return C.ScanSelection.mask(mask) if kind == "mask" else C.ScanSelection.weights(weights)


def _same_max(got, want, msg):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan), f"{msg}: NaN positions"
    assert_identical(np.where(nan, 0, got), np.where(nan, 0, want), msg)


def _same_mean(got, want, tol, msg):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    fin = np.isfinite(want)
    assert np.array_equal(np.isnan(got), np.isnan(want)), f"{msg}: NaN positions"
    assert np.array_equal(got[~fin & ~np.isnan(want)], want[~fin & ~np.isnan(want)]), f"{msg}: infinities"
    if tol == 0:
        assert_identical(got[fin], want[fin], msg)
    else:
        assert_relative(got[fin], want[fin], tol, msg)


def test_weighted_and_masked_patterns_agree_with_the_oracle():
    """SelectedAreaMetalTests.swift:29-45 on the device, held to 1e-12 (mean) and bit identity (max)."""
    cube = _metal_test_cube(11, 9, 20, 24, 11)
    mask, weights = _selection(99)
    res = C.HIPResidentCube(cube)
    assert C.pattern_path(res, take=mask) == 2 and C.pattern_path(res, weights=weights) == 2
    for kind in ("mask", "weights"):
        want_mean = O.pattern(cube, "mean", **({"take": mask} if kind == "mask" else {"weights": weights}))
        want_max = O.pattern(cube, "max", **({"take": mask} if kind == "mask" else {"weights": weights}))
        for t in TILINGS:
            print(f"case: reference cube {kind} tiling {t}", flush=True)
            m = C.VirtualDiffraction.pattern(res, MEAN, _scan(kind, mask, weights), tiling=t)
            x = C.VirtualDiffraction.pattern(res, MAX, _scan(kind, mask, weights), tiling=t)
            assert m.path == C.ComputePath.hip
            assert_relative(m.values, want_mean, 1e-12, f"mean {kind} {t}")
            _same_max(x.values, want_max, f"max {kind} {t}")
        # a host array is staged through a temporary resident cube: the same kernels, the same bits
        assert_identical(C.VirtualDiffraction.pattern(cube, MEAN, _scan(kind, mask, weights)).values,
                         C.VirtualDiffraction.pattern(res, MEAN, _scan(kind, mask, weights)).values, "staged")


def test_the_comparison_catches_a_wrong_selection():
    """SelectedAreaMetalTests.swift:49-65: one dropped position, a mask where weights were meant and 1 % in every
    weight all move the device result by far more than the tolerance."""
    cube = _metal_test_cube(11, 9, 20, 24, 11)
    mask, weights = _selection(99)
    res = C.HIPResidentCube(cube)

    def dist(a, b):
        return np.max(np.abs(a - b)) / max(1e-12, np.max(np.abs(b)))
    cpu = O.pattern(cube, "mean", take=mask)
    off = mask.copy()
    off[np.argmax(mask)] = False
    assert dist(C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.mask(off)).values, cpu) > 1e-5
    cpu_w = O.pattern(cube, "mean", weights=weights)
    assert dist(C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.mask(weights != 0)).values, cpu_w) > 1e-3
    cpu_max = O.pattern(cube, "max", weights=weights)
    assert dist(C.VirtualDiffraction.pattern(res, MAX, C.ScanSelection.weights(weights * 1.01)).values, cpu_max) > 1e-3


def test_refusals_and_edges():
    """SelectedAreaMetalTests.swift:67-84."""
    cube = _metal_test_cube(4, 4, 6, 6, 3)
    res = C.HIPResidentCube(cube)
    all_ = C.ScanSelection.mask(np.ones(16, bool))
    with pytest.raises(C.CuboVirtualError):
        C.VirtualDiffraction.pattern(res, C.DiffractionStatistic.median, all_)
    with pytest.raises(C.CuboVirtualError):
        C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.mask([True]))
    with pytest.raises(C.CuboVirtualError):
        C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.mask(np.zeros(16, bool)))
    with pytest.raises(C.CuboVirtualError):
        C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.weights(np.zeros(16)))
    with pytest.raises(C.CuboVirtualError):
        C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.weights(np.r_[np.ones(15), np.inf]))
    # the C ABI refuses what the mirror would never send
    req = L.PatternRequest(statistic=7, take=None, weights=None, origin_row=None, origin_col=None, max_ties=0)
    out = np.zeros(36)
    assert L.lib().cubo_hip_pattern(res.handle, ct.byref(req), out.ctypes.data) == L.ERR_INVALID_ARGUMENT
    t = np.ones(16, np.uint8)
    w = np.ones(16)
    req = L.PatternRequest(statistic=0, take=t.ctypes.data, weights=w.ctypes.data, origin_row=None, origin_col=None,
                           max_ties=0)
    assert L.lib().cubo_hip_pattern(res.handle, ct.byref(req), out.ctypes.data) == L.ERR_INVALID_ARGUMENT
    # .all without a shift is the existing mean and max
    m = C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.all())
    assert_identical(m.values, C.VirtualDiffraction.mean(res).values, ".all mean")
    # a negative-only cube: the maximum is negative (seeded at -infinity)
    neg = C.HIPResidentCube(-np.abs(cube) - 1)
    assert np.all(C.VirtualDiffraction.pattern(neg, MAX, all_).values < 0)


def _lcg_specials(shape):
    """The CompensatedSumTests.swift:28-47 values and special columns (NaN, +-Inf, +Inf with -Inf, -0, 0, FLT_MAX)."""
    R, Cs, qy, qx = shape
    npos, n = R * Cs, qy * qx
    u = _lcg_uniform(2026, npos * n)
    v = (np.float32(1) * (u - 0.35).astype(np.float32)) * np.float32(3e3)
    v[0::7] *= np.float32(1e-6)
    v[3::11] *= np.float32(1e6)
    v[2 * n + 0] = np.nan
    v[1 * n + 1] = np.inf
    v[0 * n + 2] = -np.inf
    v[3 * n + 3], v[5 * n + 3] = np.inf, -np.inf
    v[4 * n + 4], v[6 * n + 4] = np.inf, 5
    v[np.arange(npos) * n + 5] = -0.0
    v[np.arange(npos) * n + 6] = 0.0
    v[3 * n + 7] = v[4 * n + 7] = np.finfo(np.float32).max
    v[9 * n + n - 1] = np.nan
    v[10 * n + n - 2], v[12 * n + n - 2] = np.inf, -np.inf
    return v.reshape(shape)


@pytest.mark.parametrize("shape", [(5, 7, 19, 23), (5, 7, 16, 24)])
def test_non_finite_values_and_signed_zeros(shape):
    """CompensatedSumTests.swift:148-182 shapes and specials: 437 pixels (one pixel per lane) and 384 (four)."""
    cube = _lcg_specials(shape)
    npos = shape[0] * shape[1]
    k = np.arange(npos)
    mask = ((k * 7) % 5 != 2) & (k // shape[1] != 2)
    pat = np.array([0.37, 0, -1.25, 2.5, 3.0, 0.125])
    weights = np.where(k % 6 == 1, 0.0, pat[k % 6] + k * 1e-3)          # exact zeros: 0 * Inf = NaN
    res = C.HIPResidentCube(cube)
    for t in (C.ScanTiling.rows(1), C.ScanTiling.automatic):
        for kind, kw in (("mask", {"take": mask}), ("weights", {"weights": weights})):
            print(f"case: specials {shape} {kind} {t}", flush=True)
            scan = _scan(kind, mask, weights)
            _same_mean(C.VirtualDiffraction.pattern(res, MEAN, scan, tiling=t).values, O.pattern(cube, "mean", **kw),
                       1e-12, f"mean {kind} {t}")
            _same_max(C.VirtualDiffraction.pattern(res, MAX, scan, tiling=t).values, O.pattern(cube, "max", **kw),
                      f"max {kind} {t}")


@pytest.mark.parametrize("det", [(24, 32), (13, 11)])
def test_uint16_mask_is_exact_and_weights_are_fp64(det):
    rng = np.random.default_rng(41)
    qy, qx = det
    cube = rng.integers(0, 65536, size=(9, 11, qy, qx)).astype(np.uint16)
    npos = 99
    mask = rng.random(npos) < 0.4
    weights = rng.normal(size=npos)
    weights[5] = -0.0
    res = C.HIPResidentCube(cube)
    assert C.pattern_path(res, take=mask) == 1 and C.pattern_path(res, weights=weights) == 2
    for t in (C.ScanTiling.rows(1), C.ScanTiling.rows(3), C.ScanTiling.automatic):
        print(f"case: uint16 {det} tiling {t}", flush=True)
        m = C.ScanSelection.mask(mask)
        assert_identical(C.VirtualDiffraction.pattern(res, MEAN, m, tiling=t).values, O.pattern(cube, "mean", take=mask),
                         "uint16 mask mean")
        assert_identical(C.VirtualDiffraction.pattern(res, MAX, m, tiling=t).values, O.pattern(cube, "max", take=mask),
                         "uint16 mask max")
        w = C.ScanSelection.weights(weights)
        assert_relative(C.VirtualDiffraction.pattern(res, MEAN, w, tiling=t).values,
                        O.pattern(cube, "mean", weights=weights), 1e-12, "uint16 weights mean")
        _same_max(C.VirtualDiffraction.pattern(res, MAX, w, tiling=t).values, O.pattern(cube, "max", weights=weights),
                  "uint16 weights max")


@pytest.mark.parametrize("dtype,det", [(np.uint16, (16, 24)), (np.uint16, (15, 23)), (np.float32, (16, 24)),
                                       (np.float32, (12, 10))])
def test_rounded_shift_centres_planted_beams(dtype, det):
    """Each pattern is one bright pixel at its origin: moved by the rounded mean minus the origin, every beam lands
    on round(mean) (the closed form of the centring)."""
    rng = np.random.default_rng(13)
    qy, qx = det
    R, Cs = 6, 7
    npos = R * Cs
    orow = rng.integers(0, qy, size=npos).astype(np.float64)
    ocol = rng.integers(0, qx, size=npos).astype(np.float64)
    # a mean at a half integer would round by the parity of each offset; keep it off the tie
    for o, n in ((orow, qy), (ocol, qx)):
        while (2 * o.sum()) % (2 * npos) == npos:
            o[0] = (o[0] + 1) % n
    x = np.zeros((npos, qy, qx), dtype)
    x[np.arange(npos), orow.astype(int), ocol.astype(int)] = 7
    res = C.HIPResidentCube(x.reshape(R, Cs, qy, qx))
    shift = C.PatternShift.rounded(C.OriginMap(orow, ocol))
    want = np.zeros((qy, qx))
    want[int(np.rint(orow.mean())), int(np.rint(ocol.mean()))] = 7
    assert_identical(C.VirtualDiffraction.pattern(res, MAX, C.ScanSelection.all(), shift=shift).values, want.ravel(),
                     "max")
    np.testing.assert_allclose(C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.all(), shift=shift).values,
                               want.ravel(), rtol=1e-15, atol=0)


@pytest.mark.parametrize("dtype,det", [(np.uint16, (16, 24)), (np.uint16, (15, 23)), (np.float32, (16, 24)),
                                       (np.float32, (12, 10))])
def test_rounded_shift_matches_the_oracle(dtype, det):
    rng = np.random.default_rng(17)
    qy, qx = det
    R, Cs = 7, 6
    npos = R * Cs
    if dtype == np.uint16:
        cube = rng.integers(0, 65536, size=(R, Cs, qy, qx)).astype(np.uint16)
    else:
        cube = _lcg_specials((R, Cs, qy, qx))
    orow = rng.normal(scale=3 * qy, size=npos)                      # shifts wrap several times both ways
    ocol = rng.normal(scale=3 * qx, size=npos)
    mask = rng.random(npos) < 0.6
    weights = rng.normal(size=npos)
    res = C.HIPResidentCube(cube)
    shift = C.PatternShift.rounded(C.OriginMap(orow, ocol))
    exact = 0 if dtype == np.uint16 else 1e-12
    for t in (C.ScanTiling.rows(2), C.ScanTiling.automatic):
        print(f"case: shift {dtype.__name__} {det} {t}", flush=True)
        for kind, scan, kw in (("mask", C.ScanSelection.mask(mask), {"take": mask}),
                               ("weights", C.ScanSelection.weights(weights), {"weights": weights}),
                               ("all", C.ScanSelection.all(), {})):
            tol = exact if kind != "weights" else 1e-12
            _same_mean(C.VirtualDiffraction.pattern(res, MEAN, scan, shift=shift, tiling=t).values,
                       O.pattern(cube, "mean", origin=(orow, ocol), **kw), tol, f"mean {kind}")
            _same_max(C.VirtualDiffraction.pattern(res, MAX, scan, shift=shift, tiling=t).values,
                      O.pattern(cube, "max", origin=(orow, ocol), **kw), f"max {kind}")
        mean, mx = C.VirtualDiffraction.mean_and_max_shifted(res, shift, tiling=t)
        _same_mean(mean.values, O.pattern(cube, "mean", origin=(orow, ocol), ties="later"), exact, "meanAndMax mean")
        _same_max(mx.values, O.pattern(cube, "max", origin=(orow, ocol), ties="later"), "meanAndMax max")


def test_signed_zero_ties_follow_each_functions_rule():
    x = np.zeros((1, 2, 4, 4), np.float32)
    x[0, 0, 0, 0], x[0, 1, 0, 0] = -0.0, 0.0
    x[0, 0, 0, 1], x[0, 1, 0, 1] = 0.0, -0.0
    res = C.HIPResidentCube(x)
    shift = C.PatternShift.rounded(C.OriginMap(np.zeros(2), np.zeros(2)))
    first = C.VirtualDiffraction.pattern(res, MAX, C.ScanSelection.mask([True, True])).values
    assert np.signbit(first[0]) and not np.signbit(first[1])
    _, later = C.VirtualDiffraction.mean_and_max_shifted(res, shift)
    assert not np.signbit(later.values[0]) and np.signbit(later.values[1])


def test_session_order_reuse_and_progress():
    cube = _metal_test_cube(6, 5, 8, 8, 2)
    res = C.HIPResidentCube(cube)
    mask, _ = _selection(30)
    t = np.ascontiguousarray(mask.astype(np.uint8))
    req = L.PatternRequest(statistic=0, take=t.ctypes.data, weights=None, origin_row=None, origin_col=None,
                           max_ties=0)
    h = ct.c_void_p()
    lib = L.lib()
    assert lib.cubo_hip_pattern_begin(res.handle, ct.byref(req), ct.byref(h)) == 0
    out1, out2 = np.zeros(64), np.zeros(64)
    assert lib.cubo_hip_pattern_rows(h, 2, 1) == L.ERR_INVALID_ARGUMENT        # tiles come in scan order
    assert lib.cubo_hip_pattern_rows(h, 0, 3) == 0
    assert lib.cubo_hip_pattern_results(h, out1.ctypes.data) == L.ERR_INCOMPLETE
    assert lib.cubo_hip_pattern_rows(h, 3, 3) == 0
    assert lib.cubo_hip_pattern_results(h, out1.ctypes.data) == 0
    assert lib.cubo_hip_pattern_reset(h) == 0
    assert lib.cubo_hip_pattern_rows(h, 0, 6) == 0
    assert lib.cubo_hip_pattern_end(h, out2.ctypes.data) == 0
    assert_relative(out1, O.pattern(cube, "mean", take=mask), 1e-12, "two tiles")
    assert_relative(out2, out1, 1e-15, "reused session")
    seen = []
    C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.mask(mask), tiling=C.ScanTiling.rows(2),
                                 progress=seen.append)
    assert seen == [2 / 6, 4 / 6, 1.0]
    with pytest.raises(C.CancellationError):
        C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.mask(mask), tiling=C.ScanTiling.rows(2),
                                     is_cancelled=lambda: len(seen) > 3, progress=seen.append)


def test_bandwidth_cube_selected_and_shifted_patterns():
    """256x256x128x128 uint16 (2 GiB): every position selected gives the fused pass's mean and max bit for bit;
    unit weights give the same mean on the FP64 kernel; a disk of scan positions, moved by the rounded centre of
    mass, matches the oracle bit for bit (integer kernel)."""
    shape, dtype, params = synthetic.bandwidth_u16()
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    npos = shape[0] * shape[1]
    mean, mx = C.VirtualDiffraction.mean_and_max(res)
    every = C.ScanSelection.mask(np.ones(npos, bool))
    print("case: full-size all-selected", flush=True)
    assert_identical(C.VirtualDiffraction.pattern(res, MEAN, every).values, mean.values, "all-selected mean")
    assert_identical(C.VirtualDiffraction.pattern(res, MAX, every).values, mx.values, "all-selected max")
    print("case: full-size unit weights", flush=True)
    ones = C.ScanSelection.weights(np.ones(npos))
    assert_identical(C.VirtualDiffraction.pattern(res, MEAN, ones).values, mean.values, "unit-weight mean")
    assert_identical(C.VirtualDiffraction.pattern(res, MAX, ones).values, mx.values, "unit-weight max")
    print("case: full-size shifted disk", flush=True)
    com = C.CenterOfMass.measure(res)
    r, c = np.divmod(np.arange(npos), shape[1])
    disk = (r - 128.0) ** 2 + (c - 100.0) ** 2 < 40.0 ** 2
    shift = C.PatternShift.rounded(C.OriginMap(com.row.values, com.column.values))
    host = res.read_scan_tile(range(0, shape[0]))
    orig = (com.row.values, com.column.values)
    got_mean = C.VirtualDiffraction.pattern(res, MEAN, C.ScanSelection.mask(disk), shift=shift).values
    got_max = C.VirtualDiffraction.pattern(res, MAX, C.ScanSelection.mask(disk), shift=shift).values
    assert_identical(got_mean, O.pattern(host, "mean", take=disk, origin=orig), "shifted disk mean")
    assert_identical(got_max, O.pattern(host, "max", take=disk, origin=orig), "shifted disk max")


@pytest.mark.parametrize("dtype,det", [(np.float32, (16, 24)), (np.uint16, (15, 12)),          # batched FFTs
                                       (np.uint16, (32, 64)), (np.float32, (64, 32)), (np.uint16, (128, 128)),
                                       (np.uint16, (16, 256)), (np.float32, (256, 32))])
def test_fourier_shift_matches_the_oracle(dtype, det):
    """PatternShift.fourier over a mask, weights and every position (first-kept maxima) and meanAndMax(of:shift:)
    (later-kept): FP64 FFTs on both sides, normwise within 1e-12 of max|oracle|, through several tiles. Power-of-two
    detectors take the one-kernel path (half spectrum in LDS), the others batches of 2-D FFTs."""
    rng = np.random.default_rng(23)
    qy, qx = det
    R, Cs = 7, 6
    npos = R * Cs
    if dtype == np.uint16:
        cube = rng.integers(0, 65536, size=(R, Cs, qy, qx)).astype(np.uint16)
    else:
        cube = rng.normal(30, 10, size=(R, Cs, qy, qx)).astype(np.float32)
    orow, ocol = rng.normal(7, 2.5, npos), rng.normal(9, 2.5, npos)
    mask = rng.random(npos) < 0.6
    weights = rng.normal(size=npos)
    res = C.HIPResidentCube(cube)
    shift = C.PatternShift.fourier_shift(C.OriginMap(orow, ocol))

    def close(got, want, msg):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), msg
    for t in (C.ScanTiling.rows(2), C.ScanTiling.automatic):
        for kind, scan, kw in (("mask", C.ScanSelection.mask(mask), {"take": mask}),
                               ("weights", C.ScanSelection.weights(weights), {"weights": weights}),
                               ("all", C.ScanSelection.all(), {})):
            print(f"case: fourier {dtype.__name__} {det} {kind} {t}", flush=True)
            for stat in (MEAN, MAX):
                got = C.VirtualDiffraction.pattern(res, stat, scan, shift=shift, tiling=t).values
                close(got, O.pattern(cube, stat.value, origin=(orow, ocol), fourier=True, **kw), f"{kind} {stat}")
        mean, mx = C.VirtualDiffraction.mean_and_max_shifted(res, shift, tiling=t)
        close(mean.values, O.pattern(cube, "mean", origin=(orow, ocol), ties="later", fourier=True), "meanAndMax mean")
        close(mx.values, O.pattern(cube, "max", origin=(orow, ocol), ties="later", fourier=True), "meanAndMax max")


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
def test_fourier_shift_groups_of_many_entries(dtype):
    """The one-kernel path with more entries than blocks: each block adds a run of patterns in order before the
    groups are combined; a signed weight map and a mask; within 1e-12 of max|oracle|."""
    rng = np.random.default_rng(29)
    R, Cs, qy, qx = 24, 23, 32, 32
    npos = R * Cs
    if dtype == np.uint16:
        cube = rng.integers(0, 65536, size=(R, Cs, qy, qx)).astype(np.uint16)
    else:
        cube = rng.normal(30, 10, size=(R, Cs, qy, qx)).astype(np.float32)
    orow, ocol = rng.normal(15, 4, npos), rng.normal(16, 4, npos)
    mask = rng.random(npos) < 0.8
    weights = rng.normal(size=npos)
    res = C.HIPResidentCube(cube)
    shift = C.PatternShift.fourier_shift(C.OriginMap(orow, ocol))
    for t in (C.ScanTiling.automatic, C.ScanTiling.rows(11)):
        for kind, scan, kw in (("mask", C.ScanSelection.mask(mask), {"take": mask}),
                               ("weights", C.ScanSelection.weights(weights), {"weights": weights})):
            print(f"case: fourier groups {dtype.__name__} {kind} {t}", flush=True)
            for stat in (MEAN, MAX):
                got = C.VirtualDiffraction.pattern(res, stat, scan, shift=shift, tiling=t).values
                want = O.pattern(cube, stat.value, origin=(orow, ocol), fourier=True, **kw)
                assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), f"{kind} {stat}"


@pytest.mark.parametrize("det", [(32, 64), (16, 24)])          # the one-kernel path, batched FFTs
def test_fourier_shift_with_a_nan_pixel(det):
    """A NaN pixel spreads through its pattern's transform: selected, the Fourier-shifted mean and max are NaN
    everywhere, as the oracle's; masked out, the others give the oracle's values within 1e-12 of max|oracle|."""
    rng = np.random.default_rng(31)
    qy, qx = det
    R, Cs = 4, 5
    npos = R * Cs
    cube = rng.normal(30, 10, size=(R, Cs, qy, qx)).astype(np.float32)
    cube[2, 3, qy // 2, 1] = np.nan
    orow, ocol = rng.normal(qy / 2, 2, npos), rng.normal(qx / 2, 2, npos)
    shift = C.PatternShift.fourier_shift(C.OriginMap(orow, ocol))
    res = C.HIPResidentCube(cube)
    bad = 2 * Cs + 3
    for stat in (MEAN, MAX):
        print(f"case: fourier nan {det} {stat}", flush=True)
        got = C.VirtualDiffraction.pattern(res, stat, C.ScanSelection.all(), shift=shift).values
        want = O.pattern(cube, stat.value, origin=(orow, ocol), fourier=True)
        assert np.all(np.isnan(want)) and np.all(np.isnan(got)), stat
        keep = np.arange(npos) != bad
        got = C.VirtualDiffraction.pattern(res, stat, C.ScanSelection.mask(keep), shift=shift).values
        want = O.pattern(cube, stat.value, take=keep, origin=(orow, ocol), fourier=True)
        assert np.all(np.isfinite(want)) and np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), stat
