"""GPU parity of the fused streaming reduction against the oracle and CuboKit's golden vectors.

Tolerances (stated per assertion):
- virtual images, mean, CoM sums: relative 1e-10 against the float64 definition with the reference's floor
  (1e-6 * max|want|, Tests/CuboKitTestSupport/TestSupport.swift:135-151) — stricter than the 1e-5 the
  reference holds its Metal path to (PinsTests.swift:88-107);
- max: bit-identical (PinsTests.swift:96);
- integer cubes (uint16, and uint32 counts read as Float32): images, mean and CoM bit-identical
  (PinsTests.swift:62-64, CenterOfMassProbeSizeTests.swift:45-50, 165-179);
- CoM quotient: 1e-9 px where |S| >= 1e-6 sum|wP| (CenterOfMassProbeSizeTests.swift:57-69).
"""
import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import synthetic
from oracle import oracle as O
from tests.support import (assert_identical, assert_relative, com_pins, doubles, geometry_kwargs, is_integer_dtype,
                           lcg_values, load_cube, virtual_pins)

pytestmark = pytest.mark.gpu

TILINGS = [C.ScanTiling.automatic, C.ScanTiling.rows(1), C.ScanTiling.rows(2)]


def _geo(g):
    kind, kw = geometry_kwargs(g)
    if kind == "point":
        return C.DetectorGeometry.point(kw["row"], kw["column"])
    if kind == "circle":
        return C.DetectorGeometry.circle(kw["centerRow"], kw["centerColumn"], kw["radius"])
    if kind == "annulus":
        return C.DetectorGeometry.annulus(kw["centerRow"], kw["centerColumn"], kw["inner"], kw["outer"])
    if kind == "rectangle":
        return C.DetectorGeometry.rectangle(kw["rowMin"], kw["rowMax"], kw["columnMin"], kw["columnMax"])
    return C.DetectorGeometry.mask(kw["weights"])


def _resident(arr):
    return C.HIPResidentCube(arr)


def test_device_is_gfx950():
    info = C.device_info(0)
    print("device", info.name, info.arch, info.compute_units, info.hbm_bytes)
    assert info.arch.decode().startswith("gfx950")
    assert info.wavefront_size == 64


def test_pins_virtual_images():
    for c in virtual_pins()["cubes"]:
        arr = load_cube(c["file"])
        cube = _resident(arr)
        for g in c["geometries"]:
            geo = _geo(g["geometry"])
            ref = None
            for t in TILINGS:
                img = C.VirtualImaging.image(cube, geo, tiling=t)
                assert img.path == C.ComputePath.hip
                if ref is None:
                    ref = img.values
                    assert_relative(img.values, doubles(g["image_f64"]), 1e-10, f"{c['name']}/{g['name']}")
                    if is_integer_dtype(c["dtype"]) and c["exact_in_float32"]:
                        assert_identical(img.values, doubles(g["image_py4dstem_native"]), f"{c['name']}/{g['name']}")
                else:
                    assert_identical(img.values, ref, f"{c['name']}/{g['name']} tiling {t}")


def test_pins_mean_and_max():
    for c in virtual_pins()["cubes"]:
        cube = _resident(load_cube(c["file"]))
        for t in TILINGS:
            mean, mx = C.VirtualDiffraction.mean_and_max(cube, tiling=t)
            assert_identical(mx.values, doubles(c["max"]), f"{c['name']} max {t}")
            assert_relative(mean.values, doubles(c["mean_f64"]), 1e-10, f"{c['name']} mean {t}")
            if is_integer_dtype(c["dtype"]) and c["exact_in_float32"]:
                assert_identical(mean.values, doubles(c["mean_py4dstem_native"]), f"{c['name']} mean {t}")


def test_pins_center_of_mass():
    for e in com_pins()["com"]:
        name = e["cube"]
        arr = load_cube(name)
        cube = _resident(arr)
        w = e.get("weights")
        geo = None if w is None else C.DetectorGeometry.mask(np.array([np.float32(float(s)) for s in w]))
        s, r, col = doubles(e["intensity"]), doubles(e["row"]), doubles(e["column"])
        for t in TILINGS[:2]:
            com = C.CenterOfMass.measure(cube, mask=geo, tiling=t)
            if name in ("counts16", "counts32"):
                assert_identical(com.intensity.values, s, f"{name} intensity")
                assert_identical(com.moment_row.values, doubles(e["moment_row"]), f"{name} moment row")
                assert_identical(com.row.values, r, f"{name} row")
                assert_identical(com.column.values, col, f"{name} column")
            else:
                assert_relative(com.intensity.values, s, 1e-10, f"{name} intensity")
                assert_relative(com.moment_row.values, doubles(e["moment_row"]), 1e-10, f"{name} Mr")
                assert_relative(com.moment_column.values, doubles(e["moment_column"]), 1e-10, f"{name} Mc")
            x = arr.reshape(-1, arr.shape[2] * arr.shape[3]).astype(np.float64)
            wd = np.ones(x.shape[1]) if w is None else np.array([float(v) for v in w])
            cond = np.abs(s) >= 1e-6 * np.abs(x * wd).sum(axis=1)
            for k in np.nonzero(cond)[0]:
                if np.isnan(r[k]):
                    assert np.isnan(com.row.values[k])
                else:
                    assert abs(com.row.values[k] - r[k]) <= 1e-9 and abs(com.column.values[k] - col[k]) <= 1e-9


def test_planted_shift_and_zero_pattern():
    t = com_pins()["shift_truth"]
    com = C.CenterOfMass.measure(_resident(load_cube("shift")))
    assert_relative(com.row.values, doubles(t["row"]), 1e-12, "shift row")
    assert_relative(com.column.values, doubles(t["column"]), 1e-12, "shift column")
    z = np.array([0] * 9 + [0, 0, 0, 0, 5, 0, 0, 0, 0], np.float32).reshape(1, 2, 3, 3)
    com = C.CenterOfMass.measure(_resident(z))
    assert np.isnan(com.row.values[0]) and np.isnan(com.column.values[0])
    assert com.row.values[1] == 1 and com.column.values[1] == 1


def test_non_finite_and_signed_zero():
    """RefuterRegressionTests.swift:101-112."""
    cube = np.array([np.inf, -0.0, np.nan, 1, 2, -0.0, 3, -np.inf], np.float32).reshape(2, 1, 1, 4)
    mean, mx = C.VirtualDiffraction.mean_and_max(_resident(cube))
    assert mean.values[0] == np.inf
    assert mean.values[1] == 0 and np.signbit(mean.values[1])
    assert np.isnan(mean.values[2]) and np.isnan(mx.values[2])
    assert mean.values[3] == -np.inf
    assert mx.values[0] == np.inf and mx.values[3] == 1


def _hot_cube():
    """MetalNonFiniteTests.swift:11-27: infinities in selected patterns."""
    rows, cols, qr, qc = 6, 5, 8, 10
    n = qr * qc
    p = np.arange(rows * cols)[:, None]
    i = np.arange(n)[None, :]
    v = (np.sin((p * 31 + i * 7).astype(np.float64)) * 50 + (i % 9) - 3).astype(np.float32).reshape(-1)
    plus, minus, both = 3 * qc + 4, 5 * qc + 1, 6 * qc + 7
    v[7 * n + plus] = np.inf
    v[22 * n + minus] = -np.inf
    v[3 * n + both] = np.inf
    v[18 * n + both] = -np.inf
    return v.reshape(rows, cols, qr, qc)


def _same_special(got, want, tol, msg):
    got, want = np.asarray(got), np.asarray(want)
    fin = np.isfinite(want)
    assert np.all(np.isnan(got[np.isnan(want)])), msg
    inf = np.isinf(want)
    assert np.array_equal(got[inf], want[inf]), msg
    assert_relative(got[fin], want[fin], tol, msg)


def test_infinities_match_the_oracle():
    cube = _hot_cube()
    res = _resident(cube)
    mean, mx = C.VirtualDiffraction.mean_and_max(res)
    om, ox = O.mean_max(cube)
    _same_special(mean.values, om, 1e-10, "mean")
    assert_identical(mx.values, ox, "max")
    disk = C.DetectorGeometry.circle(4, 4, 3)
    img = C.VirtualImaging.image(res, disk)
    _same_special(img.values, O.virtual_image(cube, C.DetectorMask.make(disk, res.shape)), 1e-10, "image")
    for mask in (None, disk):
        com = C.CenterOfMass.measure(res, mask=mask)
        w = None if mask is None else C.DetectorMask.make(mask, res.shape)
        ref = O.com(cube, w)
        for k in ("intensity", "moment_row", "moment_column", "row", "column"):
            _same_special(getattr(com, k).values, ref[k], 1e-10, f"com {k}")


def test_large_integer_cube_center_of_mass_is_exact():
    """CenterOfMassProbeSizeTests.swift:165-179 on the exact integer kernel and on the FP64 kernel."""
    vals = lcg_values(7, 4 * 1024 * 1024, 4001).reshape(2, 2, 1024, 1024)
    ref = O.com(vals)
    for arr in (vals.astype(np.uint16), vals):
        res = _resident(arr)
        com = C.CenterOfMass.measure(res)
        assert_identical(com.intensity.values, ref["intensity"], "1024^2 intensity")
        assert_identical(com.moment_row.values, ref["moment_row"], "1024^2 row moment")
        assert_identical(com.moment_column.values, ref["moment_column"], "1024^2 column moment")
        assert com.moment_row.values.max() > 3e10


def test_kernel_families_are_the_intended_ones():
    res = _resident(np.zeros((2, 2, 16, 16), np.uint16))
    bf = C.DetectorMask.make(C.DetectorGeometry.circle(8, 8, 4), res.shape)
    assert C.reduce_path(res, bf[None], want_com=True) == 1
    assert C.reduce_path(res, (bf * 0.5)[None]) == 2
    res32 = _resident(np.zeros((2, 2, 16, 16), np.float32))
    assert C.reduce_path(res32, bf[None]) == 2


def test_progress_and_cancellation():
    res = _resident(np.ones((5, 3, 8, 8), np.uint16))
    seen = []
    C.VirtualDiffraction.mean_and_max(res, tiling=C.ScanTiling.rows(2), progress=seen.append)
    assert seen == sorted(seen) and seen[-1] == 1.0 and len(seen) == 3
    calls = []

    def cancel_at_second():
        calls.append(1)
        return len(calls) >= 3          # the up-front check, tile 1, then cancel before tile 2

    with pytest.raises(C.CancellationError):
        C.VirtualDiffraction.mean_and_max(res, tiling=C.ScanTiling.rows(2), is_cancelled=cancel_at_second)
    with pytest.raises(C.CancellationError):
        C.VirtualImaging.image(res, C.DetectorGeometry.point(1, 1), is_cancelled=lambda: True)


def test_small_float32_workload_matches_the_oracle():
    """20x20x512x512 Float32 (disk r=30, 2 px sigmoid edge, Poisson): images, mean, max and CoM."""
    shape, dtype, params = synthetic.small_f32()
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    host = res.read_scan_tile(range(0, shape[0]))
    bf = C.DetectorGeometry.circle(256, 256, 32)
    adf = C.DetectorGeometry.annulus(256, 256, 40, 200)
    out = C.first_look(res, [bf, adf])
    for geo, img in zip((bf, adf), out.images):
        assert_relative(img.values, O.virtual_image(host, C.DetectorMask.make(geo, res.shape)), 1e-10, "image")
    om, ox = O.mean_max(host)
    assert_relative(out.mean.values, om, 1e-10, "mean")
    assert_identical(out.max.values, ox, "max")
    ref = O.com(host)
    for k in ("intensity", "moment_row", "moment_column"):
        assert_relative(getattr(out.com, k).values, ref[k], 1e-10, k)
    assert np.max(np.abs(out.com.row.values - ref["row"])) <= 1e-9


def test_bandwidth_workload_is_bit_exact():
    """256x256x128x128 uint16 (2 GiB): one fused pass gives BF r=14, ADF 20-60, mean, max and CoM; every
    output is an exact integer result, so it must equal the oracle's bit for bit."""
    shape, dtype, params = synthetic.bandwidth_u16()
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    bf = C.DetectorGeometry.circle(64, 64, 14)
    adf = C.DetectorGeometry.annulus(64, 64, 20, 60)
    assert C.reduce_path(res, np.stack([C.DetectorMask.make(g, res.shape) for g in (bf, adf)]), want_com=True) == 1
    out = C.first_look(res, [bf, adf])
    host = res.read_scan_tile(range(0, shape[0]))
    ref = O.cpu_first_processing_u16(host, 128, 128, shape[1], C.DetectorMask.make(bf, res.shape),
                                     C.DetectorMask.make(adf, res.shape), threads=8)
    assert_identical(out.images[0].values, ref["vi_bf"], "BF")
    assert_identical(out.images[1].values, ref["vi_adf"], "ADF")
    assert_identical(out.mean.values, ref["mean"], "mean")
    assert_identical(out.max.values, ref["max"], "max")
    for k in ("intensity", "moment_row", "moment_column"):
        assert_identical(getattr(out.com, k).values, ref[k], k)
    # the synthetic data are what the workload describes: a bright disk and real counts
    assert out.max.values.max() > 400 and out.mean.values[64 * 128 + 64] > 400


def test_tilings_give_the_same_bits_on_the_integer_kernel():
    shape, dtype, params = synthetic.bandwidth_u16(scan=(24, 20))
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    geos = [C.DetectorGeometry.circle(64, 64, 14), C.DetectorGeometry.annulus(64, 64, 20, 60)]
    a = C.first_look(res, geos)
    for t in (C.ScanTiling.rows(1), C.ScanTiling.rows(5)):
        b = C.first_look(res, geos, tiling=t)
        for x, y in zip(a.images, b.images):
            assert_identical(y.values, x.values, f"image {t}")
        assert_identical(b.mean.values, a.mean.values, f"mean {t}")
        assert_identical(b.max.values, a.max.values, f"max {t}")
        assert_identical(b.com.moment_column.values, a.com.moment_column.values, f"com {t}")


@pytest.mark.parametrize("masks", [2, 4])
def test_both_integer_block_layouts_give_the_same_bits(monkeypatch, masks):
    """The integer kernel's two block layouts (one block per position group with a wave per slice, rows
    written by the kernel; or blocks per (slice, group) with a separate slice sum) give the same bits, also
    when the scan is not a multiple of the group size and over several tiles."""
    shape, dtype, params = synthetic.bandwidth_u16(scan=(37, 23))
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    geos = [C.DetectorGeometry.circle(64, 64, 14), C.DetectorGeometry.annulus(64, 64, 20, 60),
            C.DetectorGeometry.circle(40, 70, 9), C.DetectorGeometry.rectangle(10, 30, 50, 100)][:masks]
    runs = {}
    for layout in ("full", "split"):
        monkeypatch.setenv("CUBO_HIP_INT_LAYOUT", layout)
        for t in (None, C.ScanTiling.rows(7)):
            runs[(layout, t is None)] = C.first_look(res, geos, tiling=t)
    ref = runs[("full", True)]
    host = res.read_scan_tile(range(0, 37)).astype(np.float32)
    want = [O.virtual_image(host, C.DetectorMask.make(g, res.shape)) for g in geos]
    for k, r in runs.items():
        for i, (x, y) in enumerate(zip(ref.images, r.images)):
            assert_identical(y.values, x.values, f"image {i} {k}")
            assert_identical(y.values, want[i], f"image {i} {k} vs oracle")
        assert_identical(r.mean.values, ref.mean.values, f"mean {k}")
        assert_identical(r.max.values, ref.max.values, f"max {k}")
        for f in ("row", "column", "intensity", "moment_row", "moment_column"):
            assert_identical(getattr(r.com, f).values, getattr(ref.com, f).values, f"com {f} {k}")


def test_view_and_results_agree_and_reset_reuses_the_reduction():
    """cubo_hip_reduce_view (pinned, no copy) and cubo_hip_reduce_results give the same bits; reset runs the same
    request again without reallocation and gives the same bits again."""
    import ctypes as ct
    from cubokit_amd import _lib as L
    shape, dtype, params = synthetic.bandwidth_u16(scan=(16, 12))
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    masks = np.stack([C.DetectorMask.make(C.DetectorGeometry.circle(64, 64, 14), res.shape)])
    req = L.ReduceRequest(masks=masks.ctypes.data, n_masks=1, want_mean_max=1, want_com=1, com_weights=None)
    h = ct.c_void_p()
    assert L.lib().cubo_hip_reduce_begin(res.handle, ct.byref(req), ct.byref(h)) == 0
    npos = 16 * 12
    outs = []
    for rep in range(2):
        if rep:
            assert L.lib().cubo_hip_reduce_reset(h) == 0
        for r0 in range(0, 16, 5):
            assert L.lib().cubo_hip_reduce_rows(h, r0, min(5, 16 - r0)) == 0
        v = L.ReduceView()
        assert L.lib().cubo_hip_reduce_view(h, ct.byref(v)) == 0
        got = {k: np.ctypeslib.as_array(getattr(v, k), (n,)).copy() for k, n in
               (("images", npos), ("mean", 128 * 128), ("max", 128 * 128), ("com_moment_row", npos))}
        r = {k: np.empty(n) for k, n in (("images", npos), ("mean", 128 * 128), ("max", 128 * 128),
                                          ("com_moment_row", npos))}
        rr = L.ReduceResult(**{k: a.ctypes.data for k, a in r.items()})
        assert L.lib().cubo_hip_reduce_results(h, ct.byref(rr)) == 0
        for k in r:
            assert_identical(r[k], got[k], f"view vs results {k}")
        outs.append(got)
    L.lib().cubo_hip_reduce_release(h)
    for k in outs[0]:
        assert_identical(outs[1][k], outs[0][k], f"reset {k}")
    ref = C.first_look(res, [C.DetectorGeometry.circle(64, 64, 14)])
    assert_identical(outs[0]["images"], ref.images[0].values, "images vs first_look")
    assert_identical(outs[0]["mean"], ref.mean.values, "mean vs first_look")


def test_kernel_time_counts_launches_past_the_pending_limit():
    """The timing events of launches nobody queries are recycled past 256 pending pairs (folded into running
    totals), and the query still counts every launch since the previous one."""
    import ctypes as ct
    from cubokit_amd import _lib as L
    shape, dtype, params = synthetic.bandwidth_u16(scan=(4, 4))
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    masks = np.stack([C.DetectorMask.make(C.DetectorGeometry.circle(64, 64, 14), res.shape)])
    req = L.ReduceRequest(masks=masks.ctypes.data, n_masks=1, want_mean_max=1, want_com=1, com_weights=None)
    h = ct.c_void_p()
    assert L.lib().cubo_hip_reduce_begin(res.handle, ct.byref(req), ct.byref(h)) == 0
    res.kernel_time()
    v = L.ReduceView()
    for passes in (300, 3):
        for _ in range(passes):
            assert L.lib().cubo_hip_reduce_reset(h) == 0
            assert L.lib().cubo_hip_reduce_rows(h, 0, 4) == 0
            assert L.lib().cubo_hip_reduce_view(h, ct.byref(v)) == 0
        ms, launches = res.kernel_time()
        assert launches == passes and ms > 0
    L.lib().cubo_hip_reduce_release(h)


def test_clone_copies_the_cube_on_the_device():
    shape, dtype, params = synthetic.bandwidth_u16(scan=(6, 5))
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    twin = res.clone()
    assert twin.byte_count == res.byte_count
    assert np.array_equal(twin.read_scan_tile(range(0, 6)), res.read_scan_tile(range(0, 6)))
    odd = C.HIPResidentCube(np.arange(3 * 1 * 3 * 3, dtype=np.float32).reshape(3, 1, 3, 3))   # 108 bytes: a tail
    assert np.array_equal(odd.clone().read_scan_tile(range(0, 3)), odd.read_scan_tile(range(0, 3)))


def _fp64_special_cube(rows=7, cols=6, qr=20, qc=28):
    """A Float32 cube whose 560-pixel patterns end in a partial 512-pixel slice, with +-inf, NaN and -0 pixels in
    some patterns and fractional values elsewhere."""
    n = qr * qc
    p = np.arange(rows * cols)[:, None]
    i = np.arange(n)[None, :]
    v = (np.cos((p * 13 + i * 5).astype(np.float64)) * 40 + (i % 7) * 0.37).astype(np.float32).reshape(-1)
    v[3 * n + 5] = np.inf
    v[9 * n + 530] = -np.inf
    v[11 * n + 17] = np.nan
    v[20 * n + 559] = -0.0
    v[20 * n:21 * n] = -0.0
    return v.reshape(rows, cols, qr, qc)


@pytest.mark.parametrize("case", ["u16_fractional_2GiB", "f32_small_workload", "f32_special_tail"])
def test_both_fp64_kernels_agree_with_each_other_and_the_oracle(monkeypatch, case):
    """The vectorised FP64 kernel (16-byte loads, masks in registers, arithmetic unit-weight CoM rows) and the
    one-pixel-per-lane kernel on the same requests: images, mean and CoM sums within 1e-12 of each other
    (summation orders differ) and within 1e-10 of the oracle, the max bit for bit, special values the same. Cases:
    the 2 GiB uint16 cube with fractional masks, the 20x20x512x512 Float32 workload, and a Float32 cube with
    non-finite pixels whose patterns end in a partial slice."""
    if case == "u16_fractional_2GiB":
        shape, dtype, params = synthetic.bandwidth_u16()
        res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
        res.fill_synthetic(params)
        geos = [C.DetectorGeometry.circle(64, 64, 14), C.DetectorGeometry.annulus(64, 64, 20, 60)]
        masks = np.stack([0.75 * C.DetectorMask.make(geos[0], res.shape), C.DetectorMask.make(geos[1], res.shape)])
        rows = range(100, 102)                                        # oracle rows
    elif case == "f32_small_workload":
        shape, dtype, params = synthetic.small_f32()
        res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
        res.fill_synthetic(params)
        g = [C.DetectorGeometry.circle(256, 256, 32), C.DetectorGeometry.annulus(256, 256, 40, 200)]
        masks = np.stack([C.DetectorMask.make(x, res.shape) for x in g] +
                         [np.random.default_rng(1).uniform(-1, 1, res.shape.detector_count).astype(np.float32)])
        rows = range(0, shape[0])
    else:
        cube = _fp64_special_cube()
        res = _resident(cube)
        rng = np.random.default_rng(2)
        masks = rng.uniform(-1, 1, (4, res.shape.detector_count)).astype(np.float32)
        rows = range(0, cube.shape[0])
    masks = np.ascontiguousarray(masks, np.float32)
    assert C.reduce_path(res, masks, want_com=True) == 2
    runs = {}
    for kernel in ("vector", "scalar"):
        monkeypatch.setenv("CUBO_HIP_F64_KERNEL", kernel)
        print(f"case: FP64 kernels, {case}, {kernel}", flush=True)
        runs[kernel] = C.first_look(res, [C.DetectorGeometry.mask(m) for m in masks])
    a, b = runs["vector"], runs["scalar"]
    for i in range(len(masks)):
        _same_special(a.images[i].values, b.images[i].values, 1e-12, f"image {i}")
    _same_special(a.mean.values, b.mean.values, 1e-12, "mean")
    assert_identical(a.max.values, b.max.values, "max")
    for f in ("intensity", "moment_row", "moment_column", "row", "column"):
        _same_special(getattr(a.com, f).values, getattr(b.com, f).values, 1e-12, f"com {f}")
    host = res.read_scan_tile(rows).astype(np.float32)
    sl = slice(rows.start * res.shape.scan_columns, rows.stop * res.shape.scan_columns)
    for i, m in enumerate(masks):
        _same_special(a.images[i].values[sl], O.virtual_image(host, m), 1e-10, f"image {i} vs oracle")
    ref = O.com(host)
    for f in ("intensity", "moment_row", "moment_column"):
        _same_special(getattr(a.com, f).values[sl], ref[f], 1e-10, f"com {f} vs oracle")
    if rows.stop - rows.start == res.shape.scan_rows:
        om, ox = O.mean_max(host)
        _same_special(a.mean.values, om, 1e-10, "mean vs oracle")
        assert_identical(a.max.values, ox, "max vs oracle")
