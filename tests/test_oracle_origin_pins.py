"""The oracle against CuboKit's origin-fit pins (origin-fit-pins.json, OriginFitSupport.swift:31-96): centred,
calibrated, rounded- and bilinear-shifted virtual images and the rounded-shift mean and max patterns of four cubes,
on the pinned float64 origin maps. Bars as the reference's: 1e-10 relative against the float64 definition,
py4DSTEM's own values bit for bit on the integer cubes, the rounded max bit for bit."""
import numpy as np

from oracle import oracle as O
from tests.support import assert_identical, assert_relative, doubles, load_cube, origin_fit_pins

INTEGER = {"planted_u16", "counts16"}


def _mask(g, qy, qx, dr=0.0, dc=0.0, scale=1.0):
    if g["kind"] == "circle":
        return O.mask("circle", qy, qx, centerRow=(g["centerRow"] + dr) / scale,
                      centerColumn=(g["centerColumn"] + dc) / scale, radius=g["radius"] / scale)
    return O.mask("annulus", qy, qx, centerRow=(g["centerRow"] + dr) / scale, centerColumn=(g["centerColumn"] + dc) / scale,
                  inner=g["inner"] / scale, outer=g["outer"] / scale)


def test_images_and_patterns_on_the_pinned_origins():
    n = 0
    for e in origin_fit_pins()["cubes"]:
        cube = load_cube(e["cube"])
        qy, qx = cube.shape[2:]
        orow, ocol = doubles(e["origin_row"]), doubles(e["origin_column"])
        mr, mc = np.mean(orow), np.mean(ocol)            # NumPy's pairwise mean, as OriginMap.meanRow
        for im in e["images"]:
            name = f"{e['cube']} {im['geometry']['kind']}"
            m = _mask(im["geometry"], qy, qx, mr, mc)
            centred = O.virtual_image(cube, m)
            rounded = O.vi_rounded(cube, m.astype(np.float64), orow, ocol)
            bilinear = O.vi_bilinear(cube, m.astype(np.float64), orow, ocol)
            cal = O.virtual_image(cube, _mask(im["calibrated_geometry"], qy, qx, scale=im["pixel_size"]))
            assert_relative(centred, doubles(im["centered"]), 1e-10, f"{name} centred")
            assert_relative(rounded, doubles(im["rounded"]), 1e-10, f"{name} rounded")
            assert_relative(bilinear, doubles(im["bilinear"]), 1e-10, f"{name} bilinear")
            assert_relative(cal, doubles(im["calibrated"]), 1e-10, f"{name} calibrated")
            if e["cube"] in INTEGER:
                assert_identical(centred, doubles(im["centered_py4dstem"]), f"{name} centred vs py4DSTEM")
                assert_identical(rounded, doubles(im["rounded_py4dstem"]), f"{name} rounded vs py4DSTEM")
                assert_identical(cal, doubles(im["calibrated_py4dstem"]), f"{name} calibrated vs py4DSTEM")
            n += 1
        pm = e["patterns"]["rounded"]
        mean = O.pattern(cube, "mean", origin=(orow, ocol), ties="later")
        mx = O.pattern(cube, "max", origin=(orow, ocol), ties="later")
        assert_relative(mean, doubles(pm["mean"]), 1e-10, f"{e['cube']} rounded mean")
        assert_identical(mx, doubles(pm["max"]), f"{e['cube']} rounded max")
    assert n == 8


def test_half_shifts_round_to_even():
    """OriginFitTests.swift:75-86: every shift is +-0.5, which rounds to 0: the rounded image is the plain one."""
    cube = load_cube("gaussian")
    rows = (np.arange(30) % 2).astype(np.float64)
    m = O.mask("circle", 16, 16, centerRow=7, centerColumn=8, radius=4.5)
    assert_relative(O.vi_rounded(cube, m.astype(np.float64), rows, rows), O.virtual_image(cube, m), 1e-12, "half shifts")
