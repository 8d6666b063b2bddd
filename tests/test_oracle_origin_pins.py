"""The oracle against CuboKit's origin-fit pins (origin-fit-pins.json, OriginFitSupport.swift:31-96): centred,
calibrated, rounded- and bilinear-shifted virtual images and the rounded- and Fourier-shift mean and max patterns of
four cubes,
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
        # the Fourier shift: its error is normwise, |d| <= 1e-12 max|want| (OriginFitSupport.swift:78-82, 89-91)
        pf = e["patterns"]["fourier"]
        for stat in ("mean", "max"):
            got = O.pattern(cube, stat, origin=(orow, ocol), ties="later", fourier=True)
            want = doubles(pf[stat])
            assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), f"{e['cube']} fourier {stat}"
    assert n == 8


def test_half_shifts_round_to_even():
    """OriginFitTests.swift:75-86: every shift is +-0.5, which rounds to 0: the rounded image is the plain one."""
    cube = load_cube("gaussian")
    rows = (np.arange(30) % 2).astype(np.float64)
    m = O.mask("circle", 16, 16, centerRow=7, centerColumn=8, radius=4.5)
    assert_relative(O.vi_rounded(cube, m.astype(np.float64), rows, rows), O.virtual_image(cube, m), 1e-12, "half shifts")


def _max_pattern(cube):
    x = np.asarray(cube).astype(np.float32).reshape(-1, cube.shape[2] * cube.shape[3])
    return x.max(axis=0).astype(np.float64).reshape(cube.shape[2:])


def test_probe_radius_origins_and_filter_on_the_pins():
    """OriginFitSupport.swift:36-49: the default radius (probe size of the max pattern) exact, the origins within
    1e-9 px, the filtered first pattern within 1e-12; the planted cubes' truth within 1e-9 px
    (OriginFitTests.swift:21-27)."""
    import cubokit_amd as C
    for e in origin_fit_pins()["cubes"]:
        cube = load_cube(e["cube"])
        mx = _max_pattern(cube)
        radius = O.probe_size(mx)[0]
        assert radius == float(e["radius"]), e["cube"]
        # the library's host ProbeSize gives the same bits (no device needed)
        lib_r = C.ProbeSize.estimate(C.DetectorImage(mx.shape[0], mx.shape[1], mx.ravel(), C.ComputePath.hip)).radius
        assert lib_r == radius
        r, c = O.origin_py4dstem(cube, radius, 1.2)
        assert np.max(np.abs(r - doubles(e["origin_row"]))) <= 1e-9
        assert np.max(np.abs(c - doubles(e["origin_column"]))) <= 1e-9
        f = O.gaussian_filter(cube[0, 0].astype(np.float32).astype(np.float64), radius).ravel()
        assert_relative(f, doubles(e["filtered_00"]), 1e-12, f"{e['cube']} filtered")
        if e.get("truth_row"):
            assert np.max(np.abs(r - doubles(e["truth_row"]))) <= 1e-9
            assert np.max(np.abs(c - doubles(e["truth_column"]))) <= 1e-9


def test_argmax_ties_and_the_strict_mask_boundary():
    """OriginFitTests.swift:50-73 by hand: a constant pattern's argmax is pixel (0, 0) and the disk keeps
    |q - argmax| < r * rscale strictly; a bright pixel on a gradient."""
    q = 9
    flat = np.full((1, 1, q, q), 2.0, np.float32)
    r, c = O.origin_py4dstem(flat, 3.0, 1.0)
    ii, jj = np.meshgrid(np.arange(q), np.arange(q), indexing="ij")
    inside = np.sqrt(ii ** 2 + jj ** 2) < 3
    assert abs(r[0] - ii[inside].mean()) <= 1e-12 and abs(c[0] - jj[inside].mean()) <= 1e-12
    v = (1 + np.arange(q * q) // q).astype(np.float32)
    v[4 * q + 4] = 1000
    cube = v.reshape(1, 1, q, q)
    r, c = O.origin_py4dstem(cube, 3.0, 1.0)
    d = np.sqrt((ii - 4.0) ** 2 + (jj - 4.0) ** 2) < 3
    w = v.reshape(q, q).astype(np.float64)
    assert abs(r[0] - (ii * w)[d].sum() / w[d].sum()) <= 1e-12


def test_symmetric_ties_keep_the_first_argmax():
    """OriginFitTests.swift:88-105: a disk centred at (11.5, 12.5) filters to exactly symmetric values, so the first
    of the tied maxima wins and the origin is (11, 12); a NaN anywhere makes the origin NaN."""
    i = np.arange(24)[:, None] - 11.5
    j = np.arange(25)[None, :] - 12.5
    v = np.where(i * i + j * j < 6.1 * 6.1, 1000.0, 0.0).astype(np.float32).reshape(1, 1, 24, 25)
    r, c = O.origin_py4dstem(v, 3.3, 1.2)
    assert r[0] == 11 and c[0] == 12
    v2 = v.copy()
    v2[0, 0, 0, 0] = np.nan
    r, _ = O.origin_py4dstem(v2, 3.3)
    assert np.isnan(r[0])


def iterated_window_cube(rows=4, columns=5, q=40, ring=False, nan=False):
    """IteratedWindowOriginTests.swift:13-34: Gaussian-edged disks of radius 6 at per-position sub-pixel centres."""
    values, truth = [], []
    i = np.arange(q, dtype=np.float64)[:, None]
    j = np.arange(q, dtype=np.float64)[None, :]
    for sr in range(rows):
        for sc in range(columns):
            cr, cc = 19.3 + 0.21 * sr - 0.07 * sc, 20.6 - 0.13 * sr + 0.17 * sc
            truth.append((cr, cc))
            d = np.sqrt((i - cr) ** 2 + (j - cc) ** 2)
            ii, jj = np.meshgrid(np.arange(q), np.arange(q), indexing="ij")
            v = 100 / (1 + np.exp((d - 6) / 0.5)) + 1 + 0.01 * ((ii * 7 + jj * 3 + sr + sc) % 5)
            if ring:
                v = v + 40 / (1 + np.exp((np.abs(d - 14) - 1.5) / 0.4))
            v = v.astype(np.float32).ravel()
            if nan:
                v[19 * q + 20] = np.nan
                v[21 * q + 18] = np.inf
                v[3] = -np.inf
                v[0] = np.nan
            values.append(v)
    return np.array(values).reshape(rows, columns, q, q), truth


def test_iterated_window_recovers_the_planted_origin():
    """IteratedWindowOriginTests.swift:52-60 on the oracle: within 0.1 px with non-finite pixels."""
    cube, truth = iterated_window_cube(nan=True)
    r, c = O.origin_iterated(cube, 6.2)
    for k, (tr, tc) in enumerate(truth):
        assert abs(r[k] - tr) < 0.1 and abs(c[k] - tc) < 0.1, (k, r[k], c[k], tr, tc)


def iterated_window_hand_cases():
    """Tiny 16x16 patterns whose iterated-window origins follow by hand from IteratedWindowOrigin.swift:49-168 (radius
    2, scale 1.2: the window is max(2.4, 2 + margin)). Each case: (name, pattern, keyword arguments, (row, column)).
    - pass count: pixels 10 at (8,8), 10 at (8,11), 5 at (8,12), seed the first brightest pixel (8,8); a window of 3
      takes (8,11) (distance 3 is inside) and not (8,12): pass 1 gives column (80+110)/20 = 9.5, pass 2 from there
      also takes (8,12): (80+110+60)/25 = 10, pass 3 moves by 0 and stops;
    - margin: a window of 2.5 around (8,8) holds only (8,8), so the centre stays;
    - binned seed: 2x2 blocks, the first brightest [8,9]x[8,9], centre (8.5, 8.5);
    - clamping: a -4 at (8,9) lowers that block's sum to 6 (the coarse sums are not clamped), so the seed is block
      [8,9]x[10,11], centre (8.5, 10.5); the pass clamps it to 0: (8, (80+110+60)/25);
    - non-finite fill: a NaN at (8,8) reads as the median of its 8 neighbours 1,1,1,1,2,50,60,70 = 1.5 (not their mean
      23.25); the seed is the 70 at (9,9) and one pass takes the 3x3 block: mass 187.5, row 1677/187.5,
      column 1521/187.5; with an infinity at (7,7) too, (8,8) reads the median of 7 values (2) and (7,7) the median of
      its finite neighbours 0,0,0,0,0,1,1 (0): mass 187, row 1674/187, column 1518/187;
    - no finite neighbour: every pixel NaN reads 0, every block sums to 0 and the first block wins, (0.5, 0.5), and a
      pass with no mass keeps it."""
    cases = []
    base = np.zeros((16, 16), np.float32)
    p = base.copy()
    p[8, 8], p[8, 11], p[8, 12] = 10, 10, 5
    for passes, col in ((0, 8.0), (1, 9.5), (2, 10.0), (3, 10.0), (4, 10.0)):
        cases.append((f"passes {passes}", p, dict(passes=passes, margin=1.0, binned=False), (8.0, col)))
    cases.append(("margin 0.5", p, dict(passes=4, margin=0.5, binned=False), (8.0, 8.0)))
    cases.append(("binned seed", p, dict(passes=0, margin=1.0, binned=True), (8.5, 8.5)))
    q = p.copy()
    q[8, 9] = -4
    cases.append(("unclamped seed", q, dict(passes=0, margin=1.0, binned=True), (8.5, 10.5)))
    cases.append(("clamped pass", q, dict(passes=1, margin=1.0, binned=True), (8.0, 10.0)))
    f = base.copy()
    for (i, j), v in {(7, 7): 1, (7, 8): 1, (7, 9): 1, (8, 7): 1, (8, 9): 2, (9, 7): 50, (9, 8): 60, (9, 9): 70}.items():
        f[i, j] = v
    f[8, 8] = np.nan
    cases.append(("NaN median of 8", f, dict(passes=1, margin=1.0, binned=False), (1677 / 187.5, 1521 / 187.5)))
    g = f.copy()
    g[7, 7] = np.inf
    cases.append(("NaN median of 7", g, dict(passes=1, margin=1.0, binned=False), (1674 / 187, 1518 / 187)))
    cases.append(("all NaN", np.full((16, 16), np.nan, np.float32), dict(passes=4, margin=1.0, binned=True),
                  (0.5, 0.5)))
    return cases


def test_iterated_window_by_hand():
    for name, pat, kw, (wr, wc) in iterated_window_hand_cases():
        r, c = O.origin_iterated(pat.reshape(1, 1, 16, 16), 2.0, 1.2, **kw)
        assert abs(r[0] - wr) <= 1e-5 * max(1, abs(wr)) and abs(c[0] - wc) <= 1e-5 * max(1, abs(wc)), (name, r, c)
