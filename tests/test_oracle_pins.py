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
