"""Origins on the device against CuboKit's origin-fit pins and the oracle.

Bars: the default radius (the probe size of the device's max pattern) bit for bit; get_origin's origins within
1e-9 px of the pinned float64 values (OriginFitSupport.swift:44-46) and within 1e-12 px of the oracle; the filtered
pattern within 1e-12 (:48-50); the fast centre (the filtered argmax) exact; the iterated window within relative 1e-5
of the oracle's Float32 loop, the reference's own bar between its CPU and Metal implementations
(IteratedWindowOriginTests.swift:36-50).
"""
import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import origins
from oracle import oracle as O
from tests.support import assert_relative, doubles, load_cube, origin_fit_pins
from tests.test_oracle_origin_pins import iterated_window_cube, iterated_window_hand_cases

pytestmark = pytest.mark.gpu


def test_pinned_radii_origins_and_filters():
    for e in origin_fit_pins()["cubes"]:
        print(f"case: origin pins {e['cube']}", flush=True)
        arr = load_cube(e["cube"])
        res = C.HIPResidentCube(arr)
        omap, radius = C.OriginMeasurement.measure(res, tiling=C.ScanTiling.rows(2))
        assert radius == float(e["radius"]), e["cube"]
        assert np.max(np.abs(omap.row - doubles(e["origin_row"]))) <= 1e-9
        assert np.max(np.abs(omap.column - doubles(e["origin_column"]))) <= 1e-9
        r, c = O.origin_py4dstem(arr, radius, 1.2)
        assert np.max(np.abs(omap.row - r)) <= 1e-12 and np.max(np.abs(omap.column - c)) <= 1e-12
        f = origins.gaussian_filter(res, 0, 0, radius).ravel()
        assert_relative(f, doubles(e["filtered_00"]), 1e-12, f"{e['cube']} filtered")
        fast, _ = C.OriginMeasurement.measure(res, radius, fast_center=True)
        fr, fc = O.origin_py4dstem(arr, radius, 1.2, fast_center=True)
        np.testing.assert_array_equal(fast.row, fr)
        np.testing.assert_array_equal(fast.column, fc)
        n = arr.shape[0] * arr.shape[1]
        mask = np.arange(n) % 3 != 1
        part, _ = C.OriginMeasurement.measure(res, radius, scan_mask=mask)
        assert np.all(np.isnan(part.row[~mask])) and np.max(np.abs(part.row[mask] - r[mask])) <= 1e-12
        res.release()


@pytest.mark.parametrize("dtype,det,radius", [(np.uint16, (13, 21), 2.6),     # 546 B patterns: staged per element
                                              (np.float32, (70, 33), 6.4),    # several axis-0 strips, odd width
                                              (np.float32, (64, 64), 9.9),    # power-of-two rows, Float32
                                              (np.uint16, (40, 48), 17.3)])   # taps longer than the detector
def test_filter_shapes_against_the_oracle(dtype, det, radius):
    """get_origin on detector shapes the pins do not have: patterns staged element by element, several strips of
    the axis-0 pass, odd and power-of-two rows, a filter longer than the pattern (every tap clamped); origins
    within 1e-12 px of the oracle, the fast centre and the filtered first pattern as the oracle's."""
    rng = np.random.default_rng(41)
    qy, qx = det
    R, Cs = 3, 4
    ii, jj = np.meshgrid(np.arange(qy), np.arange(qx), indexing="ij")
    cube = np.empty((R, Cs, qy, qx), dtype)
    for p in range(R * Cs):
        cy, cx = rng.uniform(0.3, 0.7) * qy, rng.uniform(0.3, 0.7) * qx
        disk = np.where((ii - cy) ** 2 + (jj - cx) ** 2 < (0.25 * min(qy, qx)) ** 2, 900.0, 0.0)
        cube.reshape(-1, qy, qx)[p] = (disk + rng.integers(0, 40, size=(qy, qx))).astype(dtype)
    print(f"case: origin shape {dtype.__name__} {det} r={radius}", flush=True)
    res = C.HIPResidentCube(cube)
    omap, _ = C.OriginMeasurement.measure(res, radius, 1.2, tiling=C.ScanTiling.rows(2))
    r, c = O.origin_py4dstem(cube, radius, 1.2)
    assert np.max(np.abs(omap.row - r)) <= 1e-12 and np.max(np.abs(omap.column - c)) <= 1e-12
    fast, _ = C.OriginMeasurement.measure(res, radius, fast_center=True)
    fr, fc = O.origin_py4dstem(cube, radius, 1.2, fast_center=True)
    np.testing.assert_array_equal(fast.row, fr)
    np.testing.assert_array_equal(fast.column, fc)
    got = origins.gaussian_filter(res, 0, 0, radius)
    np.testing.assert_array_equal(got, O.gaussian_filter(cube[0, 0].astype(np.float64), radius))
    res.release()


def _mfma_origin_cube():
    """128 x 128 uint16 patterns for the matrix-core get_origin: disks of several radii and doses (some with counts
    of 2048 and more, the data's FP16 high parts), a single bright pixel, two equal disks placed symmetrically, and
    the patterns it must leave to the exact kernel: all zero, flat, a dose of a few counts."""
    rng = np.random.default_rng(61)
    n = 128
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pats = []
    for k in range(40):
        cy, cx = rng.uniform(40, 88, 2)
        rad = rng.uniform(8, 22)
        dose = [500.0, 30000.0, 60.0][k % 3]
        disk = dose / (1 + np.exp(((ii - cy) ** 2 + (jj - cx) ** 2) ** 0.5 - rad))
        pats.append(rng.poisson(disk + 2.0))
    one = np.zeros((n, n)); one[70, 45] = 1000; pats.append(one)
    two = np.zeros((n, n))
    for cy in (40.0, 87.0):
        two += 800.0 / (1 + np.exp(((ii - cy) ** 2 + (jj - 63.5) ** 2) ** 0.5 - 10))
    pats.append(np.round(two))
    pats.append(np.zeros((n, n)))
    pats.append(np.full((n, n), 7.0))
    pats.append(rng.poisson(0.05, (n, n)).astype(float))
    pats.append(rng.poisson(np.full((n, n), 3.0)))
    pats += [rng.poisson(np.full((n, n), 2.0)) * 0 + pats[k] for k in range(2)]
    arr = np.clip(np.array(pats), 0, 65535).astype(np.uint16)
    return arr.reshape(6, 8, n, n)


def test_matrix_core_get_origin_against_the_exact_kernel_and_the_oracle(monkeypatch):
    """uint16 128 x 128 patterns take the matrix-core get_origin (the filter on the matrix cores, the argmax's
    candidates recomputed in the reference's Double order): the origins within 1e-12 px of the exact kernel and of
    the oracle (NaN where the pattern is empty), the fast centre (the filtered argmax) exactly the oracle's, with a
    scan mask, and through scan-row tiles."""
    cube = _mfma_origin_cube()
    print("case: matrix-core get_origin, 48 patterns", flush=True)
    res = C.HIPResidentCube(cube)
    for radius in (14.1, 5.3, 31.0):
        r, c = O.origin_py4dstem(cube, radius, 1.2)
        omap, _ = C.OriginMeasurement.measure(res, radius, 1.2, tiling=C.ScanTiling.rows(4))
        np.testing.assert_allclose(omap.row, r, rtol=0, atol=1e-12, equal_nan=True)
        np.testing.assert_allclose(omap.column, c, rtol=0, atol=1e-12, equal_nan=True)
        monkeypatch.setenv("CUBO_HIP_ORIGIN_KERNEL", "exact")
        ex, _ = C.OriginMeasurement.measure(res, radius, 1.2)
        monkeypatch.delenv("CUBO_HIP_ORIGIN_KERNEL")
        np.testing.assert_allclose(omap.row, ex.row, rtol=0, atol=1e-12, equal_nan=True)
        np.testing.assert_allclose(omap.column, ex.column, rtol=0, atol=1e-12, equal_nan=True)
        fast, _ = C.OriginMeasurement.measure(res, radius, fast_center=True)
        fr, fc = O.origin_py4dstem(cube, radius, 1.2, fast_center=True)
        np.testing.assert_array_equal(fast.row, fr)
        np.testing.assert_array_equal(fast.column, fc)
    mask = np.arange(48) % 4 != 2
    part, _ = C.OriginMeasurement.measure(res, 14.1, scan_mask=mask)
    r, c = O.origin_py4dstem(cube, 14.1, 1.2)
    assert np.all(np.isnan(part.row[~mask]))
    np.testing.assert_allclose(part.row[mask], r[mask], rtol=0, atol=1e-12, equal_nan=True)
    res.release()


def test_argmax_ties_boundaries_and_nan():
    """OriginFitTests.swift:50-73 and 88-105 on the device."""
    q = 9
    flat = C.HIPResidentCube(np.full((1, 1, q, q), 2.0, np.float32))
    o, _ = C.OriginMeasurement.measure(flat, 3.0, 1.0)
    ii, jj = np.meshgrid(np.arange(q), np.arange(q), indexing="ij")
    inside = np.sqrt(ii ** 2 + jj ** 2) < 3
    assert abs(o.row[0] - ii[inside].mean()) <= 1e-12 and abs(o.column[0] - jj[inside].mean()) <= 1e-12
    i = np.arange(24)[:, None] - 11.5
    j = np.arange(25)[None, :] - 12.5
    v = np.where(i * i + j * j < 6.1 * 6.1, 1000.0, 0.0).astype(np.float32).reshape(1, 1, 24, 25)
    o, _ = C.OriginMeasurement.measure(C.HIPResidentCube(v), 3.3, 1.2)
    assert o.row[0] == 11 and o.column[0] == 12
    v[0, 0, 0, 0] = np.nan
    o, _ = C.OriginMeasurement.measure(C.HIPResidentCube(v), 3.3)
    assert np.isnan(o.row[0])
    with pytest.raises(C.CuboVirtualError):
        C.OriginMeasurement.measure(flat, -1.0)


@pytest.mark.parametrize("ring,nan", [(False, True), (True, False)])
def test_iterated_window_matches_the_oracle(ring, nan):
    cube, truth = iterated_window_cube(ring=ring, nan=nan)
    res = C.HIPResidentCube(cube)
    for est, kw in ((C.OriginEstimator.iterated_window(), {}),
                    (C.OriginEstimator.iterated_window(passes=2, minimum_margin=0.5, binned_coarse=False),
                     {"passes": 2, "margin": 0.5, "binned": False})):
        print(f"case: iterated window ring={ring} nan={nan} {est}", flush=True)
        m = C.OriginMeasurement.measure_with(res, 6.2, estimator=est, tiling=C.ScanTiling.rows(3))
        r, c = O.origin_iterated(cube, 6.2, **kw)
        for a, b in ((m.row, r), (m.column, c)):
            assert np.all(np.abs(a - b) <= 1e-5 * np.maximum(np.abs(b), 1)), (a, b)
    if not ring:
        m = C.OriginMeasurement.measure_with(res, 6.2)
        for k, (tr, tc) in enumerate(truth):
            assert abs(m.row[k] - tr) < 0.1 and abs(m.column[k] - tc) < 0.1
    with pytest.raises(C.CuboVirtualError):
        C.OriginMeasurement.measure_with(res, 6.2, estimator=C.OriginEstimator.iterated_window(passes=1001))


def test_iterated_window_by_hand():
    """The by-hand cases of the oracle's pin (pass count, margin, binned and unclamped seed, clamped pass, the median
    fill of non-finite pixels, an all-NaN pattern) on the device, each pattern one position of a 1x1 scan."""
    for name, pat, kw, (wr, wc) in iterated_window_hand_cases():
        print(f"case: iterated window by hand, {name}", flush=True)
        res = C.HIPResidentCube(np.ascontiguousarray(pat.reshape(1, 1, 16, 16)))
        est = C.OriginEstimator.iterated_window(passes=kw["passes"], minimum_margin=kw["margin"],
                                                binned_coarse=kw["binned"])
        m = C.OriginMeasurement.measure_with(res, 2.0, radius_scale=1.2, estimator=est)
        assert abs(m.row[0] - wr) <= 1e-5 * max(1, abs(wr)) and abs(m.column[0] - wc) <= 1e-5 * max(1, abs(wc)), (
            name, m.row[0], m.column[0])


def test_bandwidth_cube_origins_against_the_oracle_on_two_scan_rows():
    """The 256x256x128x128 uint16 cube: every position's origin on the device; two scan rows (512 positions)
    against the oracle, and the recovered origins follow the planted deflection."""
    from cubokit_amd import synthetic
    shape, dtype, params = synthetic.bandwidth_u16()
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    print("case: full-size get_origin", flush=True)
    omap, radius = C.OriginMeasurement.measure(res, tiling=C.ScanTiling.rows(shape[0]))
    rows = range(100, 102)
    sub = res.read_scan_tile(rows)
    r, c = O.origin_py4dstem(sub, radius, 1.2)
    sl = slice(rows.start * shape[1], rows.stop * shape[1])
    assert np.max(np.abs(omap.row[sl] - r)) <= 1e-12 and np.max(np.abs(omap.column[sl] - c)) <= 1e-12
    print("case: full-size iterated window", flush=True)
    iw = C.OriginMeasurement.measure_with(res, radius, tiling=C.ScanTiling.rows(shape[0]))
    ir, ic = O.origin_iterated(sub, radius)
    assert np.all(np.abs(iw.row[sl] - ir) <= 1e-5 * np.maximum(np.abs(ir), 1))
    assert np.all(np.abs(iw.column[sl] - ic) <= 1e-5 * np.maximum(np.abs(ic), 1))
    # both estimators see the same beam: their maps agree to a fraction of a pixel
    assert np.median(np.abs(iw.row - omap.row)) < 0.5


def _oracle_window_sensitivity(cube, radius, idx, shipped=1.2, wide=2.5):
    """The same measurement on the oracle: the sampled patterns as a 1 x n scan, the Float32 iterated window at both
    scales, the Float32 distances and their median."""
    qy, qx = cube.shape[2:]
    sample = np.ascontiguousarray(cube.reshape(-1, qy, qx)[idx]).reshape(1, len(idx), qy, qx)
    r0, c0 = O.origin_iterated(sample, radius, shipped)
    r1, c1 = O.origin_iterated(sample, radius, wide)
    f = np.float32
    dr, dc = r1.astype(f) - r0.astype(f), c1.astype(f) - c0.astype(f)
    d = np.sort(np.sqrt(dr * dr + dc * dc, dtype=f))
    m = d.size // 2
    return float(d[m] if d.size % 2 else (d[m - 1] + d[m]) / f(2))


def test_window_sensitivity():
    """IteratedWindowOriginTests.swift:66-84: a compact probe reads near 0, a ringed one walks with the window (and
    more than 20 times the compact one), equal scales give exactly 0, a radius of -1 gives no answer; every value
    agrees with the oracle's Float32 estimate on the same sample within 1e-4 (the reference's bar between its CPU and
    Metal paths), also with a sample limit of 7 (every third of 20 positions)."""
    from cubokit_amd.origins import sensitivity_sample_indices
    compact, _ = iterated_window_cube()
    ringed, _ = iterated_window_cube(ring=True)
    res_c, res_r = C.HIPResidentCube(compact), C.HIPResidentCube(ringed)
    print("case: window sensitivity", flush=True)
    small = C.OriginMeasurement.window_sensitivity(res_c, 6.2)
    large = C.OriginMeasurement.window_sensitivity(res_r, 6.2)
    assert small is not None and small < 0.1, small
    assert large is not None and large > 0.2 and large > 20 * small, large
    assert C.OriginMeasurement.window_sensitivity(res_r, 6.2, shipped_scale=1.2, wide_scale=1.2) == 0
    assert C.OriginMeasurement.window_sensitivity(res_r, -1.0) is None
    all20 = list(range(20))
    for cube, res, val in ((compact, res_c, small), (ringed, res_r, large)):
        want = _oracle_window_sensitivity(cube, 6.2, all20)
        assert abs(val - want) <= 1e-4 * max(want, 1), (val, want)
    seven = C.OriginMeasurement.window_sensitivity(res_r, 6.2, sample_limit=7)
    idx = sensitivity_sample_indices(20, 7, 40 * 40 * 4, 256 << 20)
    assert idx == [0, 3, 6, 9, 12, 15, 18]
    want = _oracle_window_sensitivity(ringed, 6.2, idx)
    assert abs(seven - want) <= 1e-4 * max(want, 1), (seven, want)
    # a host array takes the same path (its sampled patterns gathered on the host)
    assert abs(C.OriginMeasurement.window_sensitivity(ringed, 6.2) - large) <= 1e-6 * max(large, 1)
    pats = res_r.read_patterns([0, 7, 19])
    assert np.array_equal(pats, ringed.reshape(20, 40, 40)[[0, 7, 19]])
    with pytest.raises(C.CuboVirtualError):
        res_r.read_patterns([20])
