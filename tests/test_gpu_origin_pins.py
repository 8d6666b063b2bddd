"""The device path against CuboKit's origin-fit pins (origin-fit-pins.json, OriginFitSupport.swift:31-96):
centred and calibrated images (the fused pass), rounded- and bilinear-shifted images (the shifted stack: on the
matrix cores for uint16 cubes, on the vector ALUs in Double for Float32 cubes) and the rounded-shift mean and max patterns (the pattern kernels), on the pinned
float64 origin maps.

Bars as the reference's: 1e-10 relative against the float64 definition; py4DSTEM's own values bit for bit on the
integer cubes (centred, rounded, calibrated: 0/1 masks on uint16 data are exact on every path here); the rounded
max bit for bit."""
import numpy as np
import pytest

import cubokit_amd as C
from tests.support import assert_identical, assert_relative, doubles, load_cube, origin_fit_pins

pytestmark = pytest.mark.gpu

INTEGER = {"planted_u16", "counts16"}


def _geo(g, scale=1.0):
    if g["kind"] == "circle":
        return C.DetectorGeometry.circle(g["centerRow"], g["centerColumn"], g["radius"])
    return C.DetectorGeometry.annulus(g["centerRow"], g["centerColumn"], g["inner"], g["outer"])


def test_images_and_patterns_on_the_pinned_origins():
    n = 0
    for e in origin_fit_pins()["cubes"]:
        arr = load_cube(e["cube"])
        res = C.HIPResidentCube(arr)
        orow, ocol = doubles(e["origin_row"]), doubles(e["origin_column"])
        omap = C.OriginMap(orow, ocol)
        centre = (float(np.mean(orow)), float(np.mean(ocol)))
        for im in e["images"]:
            name = f"{e['cube']} {im['geometry']['kind']}"
            print(f"case: {name}", flush=True)
            geo = _geo(im["geometry"])
            centred = C.image_centered(res, geo, centered_on=centre).values
            cal = C.image_centered(res, _geo(im["calibrated_geometry"]), pixel_size=im["pixel_size"]).values
            assert_relative(centred, doubles(im["centered"]), 1e-10, f"{name} centred")
            assert_relative(cal, doubles(im["calibrated"]), 1e-10, f"{name} calibrated")
            # uint16 cubes: the matrix-core shifted stack; Float32 cubes: the vector-ALU pass in Double
            rounded = C.image_centered(res, geo, centered_on=centre, shift=C.MaskShift.rounded(omap)).values
            bilinear = C.image_centered(res, geo, centered_on=centre, shift=C.MaskShift.bilinear(omap)).values
            assert_relative(rounded, doubles(im["rounded"]), 1e-10, f"{name} rounded")
            assert_relative(bilinear, doubles(im["bilinear"]), 1e-10, f"{name} bilinear")
            if e["cube"] in INTEGER:
                assert_identical(centred, doubles(im["centered_py4dstem"]), f"{name} centred vs py4DSTEM")
                assert_identical(cal, doubles(im["calibrated_py4dstem"]), f"{name} calibrated vs py4DSTEM")
                assert_identical(rounded, doubles(im["rounded_py4dstem"]), f"{name} rounded vs py4DSTEM")
            n += 1
        pm = e["patterns"]["rounded"]
        mean, mx = C.VirtualDiffraction.mean_and_max_shifted(res, C.PatternShift.rounded(omap))
        assert_relative(mean.values, doubles(pm["mean"]), 1e-10, f"{e['cube']} rounded mean")
        assert_relative(mean.values, doubles(pm["mean_py4dstem"]), 1e-10, f"{e['cube']} rounded mean vs py4DSTEM")
        assert_identical(mx.values, doubles(pm["max"]), f"{e['cube']} rounded max")
        # the Fourier shift (FP64 FFTs on the device): the reference's normwise 1e-12 against the pins
        pf = e["patterns"]["fourier"]
        fmean, fmax = C.VirtualDiffraction.mean_and_max_shifted(res, C.PatternShift.fourier_shift(omap))
        for got, key in ((fmean.values, "mean"), (fmax.values, "max")):
            want = doubles(pf[key])
            assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), f"{e['cube']} fourier {key}"
        res.release()
    assert n == 8


def test_planted_beam_shift_is_recovered():
    """OriginFitTests.swift:21-33 from the pinned truth: the rounded shifted mean is the first pattern rolled onto
    the map's rounded mean."""
    for name in ("planted_f32", "planted_u16"):
        e = next(c for c in origin_fit_pins()["cubes"] if c["cube"] == name)
        arr = load_cube(name)
        orow, ocol = doubles(e["truth_row"]), doubles(e["truth_column"])
        mean, _ = C.VirtualDiffraction.mean_and_max_shifted(C.HIPResidentCube(arr),
                                                            C.PatternShift.rounded(C.OriginMap(orow, ocol)))
        sr = int(np.rint(np.mean(orow) - orow[0]))
        sc = int(np.rint(np.mean(ocol) - ocol[0]))
        disk = np.roll(arr[0, 0].astype(np.float64), (sr, sc), axis=(0, 1)).ravel()
        assert_identical(mean.values, disk, f"{name} shifted mean")


def test_half_shifts_round_to_even():
    """OriginFitTests.swift:75-86 on uint16 data."""
    arr = load_cube("counts16")
    n = arr.shape[0] * arr.shape[1]
    rows = (np.arange(n) % 2).astype(np.float64)
    res = C.HIPResidentCube(arr)
    g = C.DetectorGeometry.circle(5, 6, 3.5)
    shifted = C.image_centered(res, g, shift=C.MaskShift.rounded(C.OriginMap(rows, rows)))
    assert_identical(shifted.values, C.VirtualImaging.image(res, g).values, "half shifts")
