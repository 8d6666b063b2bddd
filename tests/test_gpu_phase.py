"""GPU parity of the CuboPhase chain on the device against CuboKit's golden chains (com-dpc-pins.json).

For the 4 cubes (plane reference) and the 3 planted fields (no reference), as CuboPhaseTests.swift:16-139 and
the oracle's own pin tests check them: fitted plane and normalised field within 1e-10, the rotation angle and
orientation exact, both curves within 1e-10, the corrected field within 1e-10 absolute, then DPC on the device from the
device's own corrected field: the backtracking sequence exact, errors within 1e-9, phase within 1e-8 of max.
"""
import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import phase as P
from cubokit_amd.virtual import CuboVirtualError
from tests.support import com_pins, doubles, load_cube, near, of_max, origin_fit_pins

pytestmark = pytest.mark.gpu


def _check_chain(mx, my, R, C_, reference, ch, name):
    fx, fy, nx, ny = P.normalize(mx, my, R, C_, reference)
    if reference == "plane":
        near(fx, doubles(ch["fitted_x"]), 1e-10, f"{name} fitted row")
        near(fy, doubles(ch["fitted_y"]), 1e-10, f"{name} fitted column")
    assert np.max(np.abs(nx - doubles(ch["normalized_x"]))) <= 1e-10, f"{name}: normalized row"
    assert np.max(np.abs(ny - doubles(ch["normalized_y"]))) <= 1e-10, f"{name}: normalized column"
    sol = P.rotation_solve(nx, ny, R, C_)
    assert sol["degrees"] == float(ch["degrees"]) and sol["transposed"] == bool(ch["transposed"]), name
    near(sol["curve"], doubles(ch["curve"]), 1e-10, f"{name} curve")
    near(sol["curve_transposed"], doubles(ch["curve_transposed"]), 1e-10, f"{name} curve transposed")
    # the corrected field, like the normalised one, to 1e-10 absolute (its entries near zero make a relative
    # bar meaningless)
    assert np.max(np.abs(sol["cx"] - doubles(ch["corrected_x"]))) <= 1e-10, f"{name}: corrected row"
    assert np.max(np.abs(sol["cy"] - doubles(ch["corrected_y"]))) <= 1e-10, f"{name}: corrected column"
    dpc = P.DPC(R, C_).reconstruct(sol["cx"], sol["cy"])
    assert dpc["accepted"] == ch["accepted"], f"{name}: backtracking sequence"
    near(dpc["errors"], doubles(ch["errors"]), 1e-9, f"{name} errors")
    of_max(dpc["phase"], doubles(ch["phase"]), 1e-8, f"{name} phase")


def test_chain_on_cubes_from_the_device_center_of_mass():
    """The whole first-processing chain on the device: the fused pass's centre of mass, the plane reference,
    the rotation and DPC."""
    n = 0
    for e in com_pins()["com"]:
        if "chain" not in e:
            continue
        cube = load_cube(e["cube"])
        R, C_ = cube.shape[:2]
        com = C.CenterOfMass.measure(C.HIPResidentCube(cube))
        _check_chain(com.row.values, com.column.values, R, C_, "plane", e["chain"], e["cube"])
        n += 1
    assert n == 4


def test_chain_on_planted_fields():
    for f in com_pins()["fields"]:
        R, C_ = f["shape"]
        _check_chain(doubles(f["measured_x"]), doubles(f["measured_y"]), R, C_, "none", f, f"field {f['theta']}")
        assert float(f["degrees"]) == f["theta"]


def test_reference_variants_and_refusals():
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    R, C_ = 7, 9
    mx, my = rng.normal(size=R * C_), rng.normal(size=R * C_)
    mx[5] = np.nan
    my[7] = np.inf
    for ref in ("constant", "none"):
        got = P.normalize(mx, my, R, C_, ref, sampling=2.0)
        want = O.normalize(mx, my, R, C_, ref, sampling=2.0)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)                  # elementwise: the same bits, NaN where NaN
    with pytest.raises(CuboVirtualError):
        P.normalize(np.full(6, np.nan), np.zeros(6), 2, 3, "plane")
    with pytest.raises(CuboVirtualError):
        P.rotation_solve(np.ones(6), np.zeros(6), 2, 3)
    z = np.zeros(20)
    s = P.rotation_solve(z, z, 4, 5)
    assert s["degrees"] == -89 and not s["transposed"]            # all ties: the first angle, plain orientation


def _planted(theta):
    f = next(f for f in com_pins()["fields"] if f["theta"] == theta)
    R, C_ = f["shape"]
    return doubles(f["measured_x"]), doubles(f["measured_y"]), R, C_


def _same_solution(got, want, name, scale=None):
    """The same angle and orientation, curves within 1e-10 (relative with CuboKit's 1e-6-of-max floor, or of
    `scale` for curves whose entries can all be rounding noise, such as the curl at a forced planted angle), and
    the corrected field within 1e-12 of its largest entry."""
    assert got["degrees"] == want["degrees"] and got["transposed"] == want["transposed"], name
    assert got["curve"].size == want["curve"].size and got["curve_transposed"].size == want["curve_transposed"].size
    for key in ("curve", "curve_transposed"):
        if not want[key].size:
            continue
        if scale is None:
            near(got[key], want[key], 1e-10, f"{name} {key}")
        else:
            assert np.max(np.abs(got[key] - want[key])) <= 1e-10 * scale, f"{name} {key}"
    assert np.max(np.abs(got["cx"] - want["cx"])) <= 1e-12 * max(1, np.max(np.abs(want["cx"]))), name
    assert np.max(np.abs(got["cy"] - want["cy"])) <= 1e-12 * max(1, np.max(np.abs(want["cy"]))), name


def test_forced_angle_and_orientation():
    """CuboPhaseTests.swift:59-75 on the device: a forced angle keeps the plain orientation on the theta = 23 field;
    a forced angle and orientation evaluate no curve; a forced orientation evaluates only its own curve and finds
    the planted angle. Every branch against the oracle, on both criteria."""
    from oracle import oracle as O
    mx, my, R, C_ = _planted(23)
    for maximize in (False, True):
        full = O.rotation_solve(mx, my, R, C_, maximize=maximize)
        scale = max(full["curve"].max(), full["curve_transposed"].max())
        for kw in (dict(force_degrees=23.0), dict(force_degrees=23.0, force_transpose=True),
                   dict(force_transpose=False), dict(force_transpose=True), dict(force_degrees=-40.5)):
            print(f"case: rotation maximize={maximize} {kw}", flush=True)
            _same_solution(P.rotation_solve(mx, my, R, C_, maximize=maximize, **kw),
                           O.rotation_solve(mx, my, R, C_, maximize=maximize, **kw), f"{maximize} {kw}", scale)
    forced = P.rotation_solve(mx, my, R, C_, force_degrees=23)
    assert forced["degrees"] == 23 and not forced["transposed"] and forced["curve"].size == 1
    both = P.rotation_solve(mx, my, R, C_, force_degrees=23, force_transpose=True)
    assert both["transposed"] and both["curve"].size == 0 and both["curve_transposed"].size == 0
    only = P.rotation_solve(mx, my, R, C_, force_transpose=False)
    assert only["degrees"] == 23 and only["curve_transposed"].size == 0
    np.testing.assert_array_equal(P.rotation_curve(mx, my, R, C_, transpose=False), only["curve"])
    full = P.rotation_solve(mx, my, R, C_)
    k = list(P.ANGLES).index(23.0)
    assert full["curve"][k] <= 1e-12 * full["curve"].max()


def test_divergence_criterion():
    """maximize (RotationCriterion.maximizeDivergence) on the planted fields against the oracle, and on a linear
    field rotated by -31 degrees against the closed forms 2|a+b||cos(theta - 31)| (plain) and
    2|b-a||sin(theta + 31)| (transposed): the optimum is the planted angle in the plain orientation."""
    from oracle import oracle as O
    from tests.test_oracle_pins import divergence_field
    for f in com_pins()["fields"]:
        R, C_ = f["shape"]
        mx, my = doubles(f["measured_x"]), doubles(f["measured_y"])
        print(f"case: divergence on the planted field {f['theta']}", flush=True)
        _same_solution(P.rotation_solve(mx, my, R, C_, maximize=True), O.rotation_solve(mx, my, R, C_, maximize=True),
                       f"divergence {f['theta']}")
    R, C_, a, b, t0 = 12, 15, 1.0, 2.0, 31.0
    mx, my = divergence_field(R, C_, a, b, t0)
    d = P.rotation_solve(mx, my, R, C_, maximize=True)
    np.testing.assert_allclose(d["curve"], 2 * abs(a + b) * np.abs(np.cos(np.deg2rad(P.ANGLES - t0))), rtol=0,
                               atol=1e-12)
    np.testing.assert_allclose(d["curve_transposed"], 2 * abs(b - a) * np.abs(np.sin(np.deg2rad(P.ANGLES + t0))),
                               rtol=0, atol=1e-12)
    assert d["degrees"] == t0 and not d["transposed"]
    z = np.zeros(20)
    s = P.rotation_solve(z, z, 4, 5, maximize=True)
    assert s["degrees"] == -89 and not s["transposed"]            # all ties: the first angle, plain orientation


def test_given_reference():
    """CuboPhaseTests.swift:95-96: .given(plane) normalises the plane to exactly zero; a given field with NaN and
    infinite measured entries gives the oracle's bits."""
    from oracle import oracle as O
    p = com_pins()["plane"]
    R, C_ = p["shape"]
    x, y = doubles(p["x"]), doubles(p["y"])
    fx, fy, nx, ny = P.normalize(x, y, R, C_, "given", given=(x, y))
    assert np.all(nx == 0) and np.all(ny == 0) and np.array_equal(fx, x) and np.array_equal(fy, y)
    rng = np.random.default_rng(3)
    mx, my, gx, gy = (rng.normal(size=R * C_) for _ in range(4))
    mx[2], my[4], mx[6] = np.nan, np.inf, -np.inf
    got = P.normalize(mx, my, R, C_, "given", sampling=0.5, given=(gx, gy))
    want = O.normalize(mx, my, R, C_, "given", sampling=0.5, given=(gx, gy))
    for a_, b_ in zip(got, want):
        np.testing.assert_array_equal(a_, b_)


def test_cancellation():
    mx, my, R, C_ = _planted(0)
    with pytest.raises(C.CancellationError):
        P.rotation_solve(mx, my, R, C_, is_cancelled=lambda: True)
    s = P.rotation_solve(mx, my, R, C_, is_cancelled=lambda: False)
    assert s["degrees"] == 0


def test_chain_from_measured_origins():
    """Origins measured on the device (py4DSTEM's get_origin) are the given reference of the device's own centre of
    mass (py4DSTEM's force_com_shifts), then the rotation and DPC: each stage against the oracle fed the device's
    output of the stage before, on CuboKit's origin-fit cubes."""
    from oracle import oracle as O
    for e in origin_fit_pins()["cubes"]:
        print(f"case: origins -> given -> rotation -> DPC on {e['cube']}", flush=True)
        arr = load_cube(e["cube"])
        R, C_ = arr.shape[:2]
        res = C.HIPResidentCube(arr)
        omap, radius = C.OriginMeasurement.measure(res)
        com = C.CenterOfMass.measure(res)
        r, c = O.origin_py4dstem(arr, radius, 1.2)
        assert np.max(np.abs(omap.row - r)) <= 1e-12 and np.max(np.abs(omap.column - c)) <= 1e-12
        got = P.normalize(com.row.values, com.column.values, R, C_, "given", given=(omap.row, omap.column))
        want = O.normalize(com.row.values, com.column.values, R, C_, "given", given=(omap.row, omap.column))
        for a_, b_ in zip(got, want):
            np.testing.assert_array_equal(a_, b_)
        nx, ny = got[2], got[3]
        sol = P.rotation_solve(nx, ny, R, C_)
        _same_solution(sol, O.rotation_solve(nx, ny, R, C_), e["cube"])
        dpc = P.DPC(R, C_).reconstruct(sol["cx"], sol["cy"])
        ref = O.dpc_reconstruct(sol["cx"], sol["cy"], R, C_)
        assert dpc["accepted"] == ref["accepted"], e["cube"]
        near(dpc["errors"], ref["errors"], 1e-9, f"{e['cube']} errors")
        of_max(dpc["phase"], ref["phase"], 1e-8, f"{e['cube']} phase")
