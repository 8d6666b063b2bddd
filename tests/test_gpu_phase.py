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
from tests.support import com_pins, doubles, load_cube, near, of_max

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
