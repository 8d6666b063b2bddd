"""The oracle against CuboKit's committed golden vectors (CPU only).

Mirrors the reference's own pins and truth tests:
- Tests/CuboVirtualTests/PinsTests.swift:38-85 (masks, virtual images, mean and max),
- Tests/CuboVirtualTests/TruthTests.swift:15-150 (closed forms and refusals),
- Tests/CuboVirtualTests/CenterOfMassProbeSizeTests.swift:39-70, 119-148 (centre of mass),
- Tests/CuboKitTestSupport/PhaseSupport.swift:49-73 and Tests/CuboPhaseTests/CuboPhaseTests.swift:16-170 (phase chain).
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests.support import (lcg_values, assert_identical, assert_relative, com_pins, doubles, geometry_kwargs, golden,
                           is_integer_dtype, load_cube, near, of_max, virtual_pins)


def _cubes():
    return virtual_pins()["cubes"]


def test_pins_carry_their_versions_and_every_cube():
    p = virtual_pins()
    assert p["versions"]["py4dstem"] == "0.14.19"
    assert p["versions"]["py4dstem_commit"] == "f050d207851f8accd3aa550365baaf1f38230428"
    assert p["seed"] == 20261014
    assert [c["name"] for c in p["cubes"]] == ["digits", "gaussian", "counts16", "counts32", "thirds", "negative",
                                               "boundary"]
    assert sum(len(c["geometries"]) for c in p["cubes"]) == 71
    assert p["boundary_pixels_float32_disagrees"] >= 1


def test_fixtures_are_the_generators_bytes():
    for c in _cubes():
        with open(golden(c["file"]), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == c["sha256"], c["name"]
        a = load_cube(c["file"])
        assert list(a.shape) == c["shape"]
        assert str(a.dtype) == c["dtype"]


def test_masks_are_make_detectors():
    for c in _cubes():
        qy, qx = c["shape"][2], c["shape"][3]
        for g in c["geometries"]:
            kind, kw = geometry_kwargs(g["geometry"])
            m = O.mask(kind, qy, qx, **kw)
            sel = np.nonzero(m != 0)[0].tolist()
            assert sel == g["mask_pixels"], f"{c['name']}/{g['name']}"
            assert len(sel) == g["mask_count"]


def test_virtual_images_agree_with_float64_and_py4dstem():
    for c in _cubes():
        cube = load_cube(c["file"])
        qy, qx = c["shape"][2], c["shape"][3]
        for g in c["geometries"]:
            kind, kw = geometry_kwargs(g["geometry"])
            img = O.virtual_image(cube, O.mask(kind, qy, qx, **kw))
            assert_relative(img, doubles(g["image_f64"]), 1e-10, f"{c['name']}/{g['name']} vs float64")
            if is_integer_dtype(c["dtype"]) and c["exact_in_float32"]:
                assert_identical(img, doubles(g["image_py4dstem_native"]), f"{c['name']}/{g['name']} vs py4DSTEM")


def test_mean_and_max_agree_with_py4dstem():
    for c in _cubes():
        cube = load_cube(c["file"])
        mean, mx = O.mean_max(cube)
        assert_identical(mx, doubles(c["max"]), f"{c['name']} max")
        assert_relative(mean, doubles(c["mean_f64"]), 1e-10, f"{c['name']} mean")
        if is_integer_dtype(c["dtype"]) and c["exact_in_float32"]:
            assert_identical(mean, doubles(c["mean_py4dstem_native"]), f"{c['name']} mean vs py4DSTEM")
            assert_identical(mean, doubles(c["mean_f64"]), f"{c['name']} mean vs float64")


# --- truth tests (TruthTests.swift) ------------------------------------------------------------------
def test_digits_closed_forms():
    cube = load_cube("digits")
    assert cube.shape == (2, 3, 5, 7)
    n = 6.0
    mean, mx = O.mean_max(cube)
    for i in range(5):
        for j in range(7):
            assert mean[i * 7 + j] == (n - 1) / 2 * 10000 + i * 100 + j
            assert mx[i * 7 + j] == (n - 1) * 10000 + i * 100 + j
    img = O.virtual_image(cube, O.mask("circle", 5, 7, centerRow=2.3, centerColumn=4.7, radius=2.9))
    for s in range(6):
        want = 0.0
        for i in range(5):
            for j in range(7):
                di, dj = i - 2.3, j - 4.7
                if di * di + dj * dj < 2.9 * 2.9:
                    want += s * 10000 + i * 100 + j
        assert img[s] == want


def test_ones_cube():
    cube = np.ones((3, 2, 9, 11), np.float32)
    weights = np.array([[0, 0.5, 1, 2, -1][k % 5] for k in range(99)], np.float32)
    for kind, kw in [("circle", dict(centerRow=4, centerColumn=5, radius=3.5)),
                     ("annulus", dict(centerRow=4.5, centerColumn=5.5, inner=1.2, outer=4.1)),
                     ("rectangle", dict(rowMin=1, rowMax=7, columnMin=2, columnMax=30)),
                     ("point", dict(row=8, column=10)), ("mask", dict(weights=weights))]:
        m = O.mask(kind, 9, 11, **kw)
        want = float(np.sum(m.astype(np.float64)))
        assert np.all(O.virtual_image(cube, m) == want)
    mean, mx = O.mean_max(cube)
    assert np.all(mean == 1) and np.all(mx == 1)


def test_boundary_counts_by_hand():
    for qy, qx in [(16, 16), (32, 32)]:
        c, d = qy // 2, qx // 2
        assert np.count_nonzero(O.mask("circle", qy, qx, centerRow=c, centerColumn=d, radius=2)) == 9
        assert np.count_nonzero(O.mask("circle", qy, qx, centerRow=c, centerColumn=d, radius=3)) == 25
        assert np.count_nonzero(O.mask("annulus", qy, qx, centerRow=c, centerColumn=d, inner=1, outer=2.5)) == 16


def test_rectangle_rounds_half_to_even_and_clips():
    m = O.mask("rectangle", 5, 7, rowMin=0.5, rowMax=2.5, columnMin=1.5, columnMax=3.5)
    assert np.nonzero(m)[0].tolist() == [2, 3, 9, 10]
    m = O.mask("rectangle", 5, 7, rowMin=3, rowMax=99, columnMin=5, columnMax=99)
    assert np.nonzero(m)[0].tolist() == [26, 27, 33, 34]
    m = O.mask("rectangle", 5, 7, rowMin=4, rowMax=2, columnMin=0, columnMax=7)
    assert np.count_nonzero(m) == 0


def test_scaled_pattern_cube():
    p = np.array([((k * 37) % 23) - 5 for k in range(42)], np.float32)
    cube = np.stack([np.float32(s) * p for s in range(20)]).reshape(4, 5, 6, 7)
    mean, _ = O.mean_max(cube)
    assert np.all(mean == np.array([(20 - 1) / 2 * float(v) for v in p]))
    m = O.mask("annulus", 6, 7, centerRow=2.5, centerColumn=3, inner=1, outer=3)
    pm = float(np.sum(p.astype(np.float64) * m.astype(np.float64)))
    assert np.all(O.virtual_image(cube, m) == np.arange(20) * pm)


def test_all_negative_maximum():
    _, mx = O.mean_max(load_cube("negative"))
    assert np.all(mx < 0)


def test_point_reads_its_own_pixel():
    cube = load_cube("digits")
    for row, col in [(1, 5), (3, 2), (4, 0), (0, 6), (2, 4)]:
        img = O.virtual_image(cube, O.mask("point", 5, 7, row=row, column=col))
        assert np.all(img == np.array([i * 10000 + row * 100 + col for i in range(6)], np.float64))


def test_refusals():
    with pytest.raises(O.OracleError):
        O.mask("annulus", 8, 8, centerRow=4, centerColumn=4, inner=3, outer=3)
    with pytest.raises(O.OracleError):
        O.mask("circle", 8, 8, centerRow=4, centerColumn=4, radius=-1)
    with pytest.raises(O.OracleError):
        O.mask("circle", 8, 8, centerRow=float("nan"), centerColumn=4, radius=1)
    with pytest.raises(O.OracleError):
        O.mask("rectangle", 8, 8, rowMin=-1, rowMax=3, columnMin=0, columnMax=2)
    with pytest.raises(O.OracleError):
        O.mask("point", 8, 8, row=8, column=0)


def test_non_finite_and_signed_zero_behaviour():
    """RefuterRegressionTests.swift:101-112."""
    cube = np.array([np.inf, -0.0, np.nan, 1, 2, -0.0, 3, -np.inf], np.float32).reshape(2, 1, 1, 4)
    mean, mx = O.mean_max(cube)
    assert mean[0] == np.inf
    assert mean[1] == 0 and np.signbit(mean[1])
    assert np.isnan(mean[2]) and np.isnan(mx[2])
    assert mean[3] == -np.inf
    assert mx[0] == np.inf and mx[3] == 1


def test_max_propagates_nan_at_any_position():
    """NaNMaxAndPolarThresholdTests.swift:12-26."""
    shape = (3, 3, 6, 7)
    n = 42
    for position in [0, 4, 8]:
        for pixel in [0, 17, 41]:
            v = (np.arange(9 * n) % 13 + 1).astype(np.float32)
            v[position * n + pixel] = np.nan
            _, mx = O.mean_max(v.reshape(shape))
            assert np.isnan(mx[pixel])
            assert np.all(np.isfinite(np.delete(mx, pixel)))


# --- centre of mass (CenterOfMassProbeSizeTests.swift) ----------------------------------------------
def _com_weights(e):
    w = e.get("weights")
    return None if w is None else np.array([np.float32(float(s)) for s in w], np.float32)


def test_com_pins_carry_their_versions():
    p = com_pins()
    assert p["versions"]["py4dstem_commit"] == "f050d207851f8accd3aa550365baaf1f38230428"
    assert len(p["com"]) == 24 and len(p["probe"]) == 20


def test_center_of_mass_agrees_with_the_float64_definition():
    for e in com_pins()["com"]:
        name = e["cube"]
        cube = load_cube(name)
        w = _com_weights(e)
        got = O.com(cube, w)
        s, r, c = doubles(e["intensity"]), doubles(e["row"]), doubles(e["column"])
        if name in ("counts16", "counts32"):
            assert_identical(got["intensity"], s, f"{name} intensity")
            assert_identical(got["moment_row"], doubles(e["moment_row"]), f"{name} moment row")
            assert_identical(got["row"], r, f"{name} row")
            assert_identical(got["column"], c, f"{name} column")
        else:
            assert_relative(got["intensity"], s, 1e-10, f"{name} intensity")
            assert_relative(got["moment_row"], doubles(e["moment_row"]), 1e-10, f"{name} moment row")
            assert_relative(got["moment_column"], doubles(e["moment_column"]), 1e-10, f"{name} moment col")
        x = cube.reshape(-1, cube.shape[2] * cube.shape[3]).astype(np.float64)
        wd = np.ones(x.shape[1]) if w is None else w.astype(np.float64)
        abs_sum = np.abs(x * wd).sum(axis=1)
        cond = np.abs(s) >= 1e-6 * abs_sum
        for k in np.nonzero(cond)[0]:
            if np.isnan(r[k]):
                assert np.isnan(got["row"][k])
                continue
            assert abs(got["row"][k] - r[k]) <= 1e-9 and abs(got["column"][k] - c[k]) <= 1e-9


def test_planted_shift_is_recovered():
    got = O.com(load_cube("shift"))
    t = com_pins()["shift_truth"]
    assert_relative(got["row"], doubles(t["row"]), 1e-12, "shift row")
    assert_relative(got["column"], doubles(t["column"]), 1e-12, "shift column")


def test_zero_pattern_is_nan():
    cube = np.array([0] * 9 + [0, 0, 0, 0, 5, 0, 0, 0, 0], np.float32).reshape(1, 2, 3, 3)
    got = O.com(cube)
    assert np.isnan(got["row"][0]) and np.isnan(got["column"][0])
    assert got["row"][1] == 1 and got["column"][1] == 1


def test_large_integer_cube_moments_are_exact():
    """CenterOfMassProbeSizeTests.swift:165-179: the 1024^2 LCG cube; exact integer sums."""
    vals = lcg_values(7, 4 * 1024 * 1024, 4001)
    cube = vals.reshape(2, 2, 1024, 1024)
    got = O.com(cube)
    x = cube.reshape(4, -1).astype(np.int64)
    ii, jj = np.indices((1024, 1024))
    assert np.all(got["intensity"] == x.sum(axis=1).astype(np.float64))
    assert np.all(got["moment_row"] == (x * ii.reshape(-1)).sum(axis=1).astype(np.float64))
    assert np.all(got["moment_column"] == (x * jj.reshape(-1)).sum(axis=1).astype(np.float64))
    assert got["moment_row"].max() > 3e10


# --- phase chain (PhaseSupport.swift, CuboPhaseTests.swift) -----------------------------------------
def _check_chain(mx, my, R, C, reference, ch, name):
    fx, fy, nx, ny = O.normalize(mx, my, R, C, reference)
    if reference == "plane":
        near(fx, doubles(ch["fitted_x"]), 1e-10, f"{name} fitted row")
        near(fy, doubles(ch["fitted_y"]), 1e-10, f"{name} fitted column")
    dn = np.max(np.abs(nx - doubles(ch["normalized_x"])))
    assert dn <= 1e-10, f"{name}: normalized {dn}"
    sol = O.rotation_solve(nx, ny, R, C)
    assert sol["degrees"] == ch["degrees"], f"{name}: {sol['degrees']} vs {ch['degrees']}"
    assert sol["transposed"] == ch["transposed"]
    assert sol["degrees"] == round(ch["degrees_py4dstem"])
    assert sol["transposed"] == ch["transposed_py4dstem"]
    near(sol["curve"], doubles(ch["curve"]), 1e-10, f"{name} curve")
    near(sol["curve_transposed"], doubles(ch["curve_transposed"]), 1e-10, f"{name} curve transposed")
    dpc = O.dpc_reconstruct(sol["cx"], sol["cy"], R, C)
    assert dpc["accepted"] == ch["accepted"], f"{name}: backtracking sequence"
    near(dpc["errors"], doubles(ch["errors"]), 1e-9, f"{name} errors")
    of_max(dpc["phase"], doubles(ch["phase"]), 1e-8, f"{name} phase")


def test_chain_on_cubes_agrees_with_the_float64_chain():
    n = 0
    for e in com_pins()["com"]:
        if "chain" not in e:
            continue
        cube = load_cube(e["cube"])
        R, C = cube.shape[:2]
        got = O.com(cube)
        _check_chain(got["row"], got["column"], R, C, "plane", e["chain"], e["cube"])
        n += 1
    assert n == 4


def test_chain_on_planted_fields():
    for f in com_pins()["fields"]:
        R, C = f["shape"]
        mx, my = doubles(f["measured_x"]), doubles(f["measured_y"])
        _check_chain(mx, my, R, C, "none", f, f"field {f['theta']}")
        sol = O.rotation_solve(mx, my, R, C)
        assert sol["degrees"] == f["theta"] and sol["transposed"] == f["transpose"]


def test_known_phase_is_reconstructed():
    f = next(f for f in com_pins()["fields"] if f["theta"] == 0)
    R, C = f["shape"]
    dpc = O.dpc_reconstruct(doubles(f["measured_x"]), doubles(f["measured_y"]), R, C)
    phi = doubles(f["phi"])
    p, q = dpc["phase"] - dpc["phase"].mean(), phi - phi.mean()
    assert (p @ q) / np.sqrt((p @ p) * (q @ q)) >= 0.99
    e = dpc["errors"]
    assert np.all(e[1:] <= e[:-1])


def test_curl_of_a_gradient_vanishes():
    f = next(f for f in com_pins()["fields"] if f["theta"] == 23)
    R, C = f["shape"]
    sol = O.rotation_solve(doubles(f["measured_x"]), doubles(f["measured_y"]), R, C)
    k = list(O.DEFAULT_ANGLES).index(23.0)
    assert sol["curve"][k] <= 1e-12 * sol["curve"].max()


def test_references():
    p = com_pins()["plane"]
    R, C = p["shape"]
    x, y = doubles(p["x"]), doubles(p["y"])
    _, _, nx, ny = O.normalize(x, y, R, C, "plane")
    assert np.all(np.abs(np.concatenate([nx, ny])) <= 1e-12)
    _, _, cx, _ = O.normalize(x, y, R, C, "constant", sampling=2)
    assert abs(cx.sum()) <= 1e-12
    xn = x.copy()
    xn[4] = np.nan
    fx, _, mx, _ = O.normalize(xn, y, R, C, "plane")
    assert mx[4] == 0 and np.all(np.abs(mx) <= 1e-12)
    assert np.all(np.abs(fx - x) <= 1e-12)


def test_ties_keep_the_plain_orientation_and_the_first_angle():
    z = np.zeros(20)
    s = O.rotation_solve(z, z, 4, 5)
    assert not s["transposed"] and s["degrees"] == -89
    d = O.rotation_solve(z, z, 4, 5, maximize=True)
    assert not d["transposed"] and d["degrees"] == -89


def test_backtracking_halves_the_step():
    f = next(f for f in com_pins()["fields"] if f["theta"] == 0)
    R, C = f["shape"]
    mx, my = doubles(f["measured_x"]), doubles(f["measured_y"])
    r = O.dpc_reconstruct(mx, my, R, C, max_iter=40, step=8)
    rejected = sum(1 for a in r["accepted"] if not a)
    assert rejected > 0
    assert r["final_step"] == 8 / (1 << rejected)
    assert np.all(r["errors"][1:] <= r["errors"][:-1])
    off = O.dpc_reconstruct(mx, my, R, C, max_iter=40, step=8, backtrack=False)
    assert all(off["accepted"]) and off["final_step"] == 8


def test_phase_refusals():
    with pytest.raises(O.OracleError):
        O.rotation_solve(np.ones(10), np.zeros(10), 2, 5)
    with pytest.raises(O.OracleError):
        O.dpc_reconstruct(np.zeros(9), np.zeros(9), 3, 3)
    with pytest.raises(O.OracleError):
        O.plane_fit(np.array([1.0, 2, 3, 4]), 1, 4)


def _planted(theta):
    f = next(f for f in com_pins()["fields"] if f["theta"] == theta)
    R, C = f["shape"]
    return doubles(f["measured_x"]), doubles(f["measured_y"]), R, C


def test_forced_angle_and_orientation():
    """CuboPhaseTests.swift:59-75 on the theta = 23 planted field: a forced angle scores both orientations there (one
    entry each, the same values as the full search's entries at that angle); a forced angle and orientation evaluate
    no curve; a forced orientation evaluates only its own curve and picks the planted angle."""
    mx, my, R, C = _planted(23)
    full = O.rotation_solve(mx, my, R, C)
    k = list(O.DEFAULT_ANGLES).index(23.0)
    forced = O.rotation_solve(mx, my, R, C, force_degrees=23)
    assert forced["degrees"] == 23 and not forced["transposed"]
    assert forced["curve"].tolist() == [full["curve"][k]] and forced["curve_transposed"].tolist() == [
        full["curve_transposed"][k]]
    both = O.rotation_solve(mx, my, R, C, force_degrees=23, force_transpose=True)
    assert both["transposed"] and both["degrees"] == 23
    assert both["curve"].size == 0 and both["curve_transposed"].size == 0
    only = O.rotation_solve(mx, my, R, C, force_transpose=False)
    assert only["degrees"] == 23 and not only["transposed"] and only["curve_transposed"].size == 0
    assert np.array_equal(only["curve"], full["curve"])
    only_t = O.rotation_solve(mx, my, R, C, force_transpose=True)
    assert only_t["transposed"] and only_t["curve"].size == 0
    assert np.array_equal(only_t["curve_transposed"], full["curve_transposed"])
    assert only_t["degrees"] == O.DEFAULT_ANGLES[int(np.argmin(full["curve_transposed"]))]
    # the corrected field is the input rotated by deg * (pi / 180) (CuboPhaseTests.swift:127-137), components swapped
    # first when transposed
    for deg in (-89.0, -61.0, 0.0, 23.0, 37.0):
        s = O.rotation_solve(mx, my, R, C, force_degrees=deg, force_transpose=False)
        rad = deg * (np.pi / 180)
        c, si = np.cos(rad), np.sin(rad)
        assert np.array_equal(s["cx"], c * mx - si * my) and np.array_equal(s["cy"], si * mx + c * my), deg
    s = O.rotation_solve(mx, my, R, C, force_degrees=10.0, force_transpose=True)
    rad = 10.0 * (np.pi / 180)
    assert np.array_equal(s["cx"], np.cos(rad) * my - np.sin(rad) * mx)


def test_rotation_refuses_non_finite_angles():
    mx, my, R, C = _planted(0)
    for bad in (dict(angles=np.array([1.0, np.nan])), dict(angles=np.array([np.inf])), dict(force_degrees=np.nan),
                dict(angles=np.zeros(0))):
        with pytest.raises(O.OracleError):
            O.rotation_solve(mx, my, R, C, **bad)
    # a forced angle does not look at the search list
    s = O.rotation_solve(mx, my, R, C, angles=np.array([np.nan]), force_degrees=5.0)
    assert s["degrees"] == 5.0


def _linear_field(R, C, a, b):
    """x = a * row, y = b * column (flattened, row-major)."""
    r, c = np.meshgrid(np.arange(R, dtype=np.float64), np.arange(C, dtype=np.float64), indexing="ij")
    return (a * r).ravel(), (b * c).ravel()


def test_divergence_and_curl_of_linear_fields():
    """Closed forms: for x = a*row, y = b*column every interior central difference is exact, so
    div = |2a + 2b| and curl = 0; for x = a*column, y = b*row, div = 0 and curl = |2b - 2a|."""
    R, C = 6, 7
    for a, b in ((1.0, 2.0), (0.5, -3.0), (-1.25, 0.25)):
        x, y = _linear_field(R, C, a, b)
        assert O.lib().oc_divergence(x, y, R, C) == abs(2 * a + 2 * b)
        assert O.lib().oc_curl(x, y, R, C) == 0.0
        # the other orientation of the same field: x = a*column, y = b*row
        r, c = np.meshgrid(np.arange(R, dtype=np.float64), np.arange(C, dtype=np.float64), indexing="ij")
        x2, y2 = (a * c).ravel(), (b * r).ravel()
        assert O.lib().oc_divergence(x2, y2, R, C) == 0.0
        assert O.lib().oc_curl(x2, y2, R, C) == abs(2 * b - 2 * a)
    # a wrong sign or a swapped axis changes the value: x = row, y = -column has zero divergence
    x, y = _linear_field(R, C, 1.0, -1.0)
    assert O.lib().oc_divergence(x, y, R, C) == 0.0


def _rotate(x, y, deg):
    rad = deg * (np.pi / 180)
    return np.cos(rad) * x - np.sin(rad) * y, np.sin(rad) * x + np.cos(rad) * y


def divergence_field(R=12, C=15, a=1.0, b=2.0, theta0=31.0):
    """The linear field (a*row, b*column) rotated by -theta0: rotating it back by theta gives the linear field
    rotated by theta - theta0, whose divergence is 2|a + b||cos(theta - theta0)| (plain) and whose transposed
    divergence is 2|b - a||sin(theta + theta0)|; the curl is 2|a + b||sin(theta - theta0)| and 2|b - a||cos(theta +
    theta0)|."""
    x, y = _linear_field(R, C, a, b)
    return _rotate(x, y, -theta0)


def test_divergence_criterion_on_a_rotated_linear_field():
    R, C, a, b, t0 = 12, 15, 1.0, 2.0, 31.0
    mx, my = divergence_field(R, C, a, b, t0)
    ang = O.DEFAULT_ANGLES
    d = O.rotation_solve(mx, my, R, C, maximize=True)
    phi = np.deg2rad(ang - t0)
    psi = np.deg2rad(ang + t0)
    np.testing.assert_allclose(d["curve"], 2 * abs(a + b) * np.abs(np.cos(phi)), rtol=0, atol=1e-12)
    np.testing.assert_allclose(d["curve_transposed"], 2 * abs(b - a) * np.abs(np.sin(psi)), rtol=0, atol=1e-12)
    assert d["degrees"] == t0 and not d["transposed"]
    c = O.rotation_solve(mx, my, R, C)
    np.testing.assert_allclose(c["curve"], 2 * abs(a + b) * np.abs(np.sin(phi)), rtol=0, atol=1e-12)
    np.testing.assert_allclose(c["curve_transposed"], 2 * abs(b - a) * np.abs(np.cos(psi)), rtol=0, atol=1e-12)
    assert c["degrees"] == t0 and not c["transposed"]
    # the corrected field is the linear field again
    x, y = _linear_field(R, C, a, b)
    assert np.max(np.abs(d["cx"] - x)) <= 1e-12 and np.max(np.abs(d["cy"] - y)) <= 1e-12
    # transposed optimum: the components swapped before the rotation make (b*column, a*row) the field to find
    tx, ty = _rotate(*_linear_field(R, C, a, b)[::-1], -t0)
    dt = O.rotation_solve(tx, ty, R, C, maximize=True)
    assert dt["transposed"] and dt["degrees"] == -t0


def test_given_reference_and_tolerance_refusals():
    """CuboPhaseTests.swift:95-96: .given(plane) subtracts the field itself, exactly zero; the tolerance must be
    finite and not negative (CenterOfMassReference.swift:30-32)."""
    p = com_pins()["plane"]
    R, C = p["shape"]
    x, y = doubles(p["x"]), doubles(p["y"])
    fx, fy, nx, ny = O.normalize(x, y, R, C, "given", given=(x, y))
    assert np.all(nx == 0) and np.all(ny == 0)
    assert np.array_equal(fx, x) and np.array_equal(fy, y)
    g = np.arange(R * C, dtype=np.float64)
    xn = x.copy()
    xn[3] = np.nan
    xn[5] = np.inf
    _, _, nx, ny = O.normalize(xn, y, R, C, "given", sampling=2.0, given=(g, -g))
    want = (xn - g) * 2.0
    want[3] = 0.0
    want[5] = np.inf                       # the greatest finite value, times the sampling 2, overflows
    assert np.array_equal(nx, want) and np.array_equal(ny, (y + g) * 2.0)
    for tol in (-1e-12, np.nan, np.inf):
        with pytest.raises(O.OracleError):
            O.normalize(x, y, R, C, "plane", tol=tol)


# --- the CPU-baseline pass ------------------------------------------------------------------------------
def _first_processing_against_pinned(cube_u16, bf, adf, rows_per_tile, threads):
    R, C, qy, qx = cube_u16.shape
    got = O.cpu_first_processing_u16(cube_u16, qy, qx, C, bf, adf, threads=threads, rows_per_tile=rows_per_tile)
    host = cube_u16.astype(np.float32)
    assert_identical(got["vi_bf"], O.virtual_image(host, bf), "BF")
    assert_identical(got["vi_adf"], O.virtual_image(host, adf), "ADF")
    mean, mx = O.mean_max(host)
    assert_identical(got["mean"], mean, "mean")
    assert_identical(got["max"], mx, "max")
    s = O.com_sums(host)
    for k, want in zip(("intensity", "moment_row", "moment_column"), s):
        assert_identical(got[k], want, k)
    return got


def test_cpu_first_processing_is_the_pinned_functions_tile_by_tile():
    """The CPU baseline's pass (and the checker of the 2 GiB bit-exact test) gives the bits of the pinned
    oc_virtual_image / oc_mean_max / oc_com_sums on uint16 cubes, whatever the tiling and thread count: scan rows
    in tiles of 1, 2 and 4 rows of a 9-row scan (the last tile partial), a 192-pixel pattern (a partial 128-pixel
    block of the mean), counts up to 65535."""
    rng = np.random.default_rng(7)
    cube = rng.integers(0, 65536, size=(9, 5, 12, 16), dtype=np.uint16)
    cube[2, 3] = 65535
    cube[4, 0] = 0
    bf = O.mask("circle", 12, 16, centerRow=5.5, centerColumn=7.2, radius=4.1)
    adf = O.mask("annulus", 12, 16, centerRow=5.5, centerColumn=7.2, inner=2.0, outer=6.5)
    for rows_per_tile, threads in ((1, 1), (2, 3), (4, 8), (9, 2)):
        _first_processing_against_pinned(cube, bf, adf, rows_per_tile, threads)


def test_cpu_first_processing_meets_the_counts16_pins():
    """counts16 (uint16, CuboKit's golden cube): py4DSTEM's own images of two 0/1 geometries, mean and max, bit for
    bit, with 1-row tiles."""
    c = next(c for c in _cubes() if c["name"] == "counts16")
    cube = load_cube(c["file"])
    qy, qx = c["shape"][2], c["shape"][3]
    geo = {g["name"]: g for g in c["geometries"]}
    g1, g2 = geo["bf_centre"], geo["annulus_fractional"]
    m1 = O.mask(*geometry_kwargs(g1["geometry"])[:1], qy, qx, **geometry_kwargs(g1["geometry"])[1])
    m2 = O.mask(*geometry_kwargs(g2["geometry"])[:1], qy, qx, **geometry_kwargs(g2["geometry"])[1])
    got = _first_processing_against_pinned(cube, m1, m2, 1, 4)
    assert_identical(got["vi_bf"], doubles(g1["image_py4dstem_native"]), "BF vs py4DSTEM")
    assert_identical(got["vi_adf"], doubles(g2["image_py4dstem_native"]), "ADF vs py4DSTEM")
    assert_identical(got["mean"], doubles(c["mean_py4dstem_native"]), "mean vs py4DSTEM")
    assert_identical(got["max"], doubles(c["max"]), "max vs py4DSTEM")
