"""Ingestion into HBM on the device: .npy and raw files (cubo_hip_cube_load_*), host arrays through the pinned feed
(cubo_hip_cube_feed_rows) and streamed cubes whose tiles are reduced as they arrive.

Bars: every element that lands in HBM is the bits of NumPy's astype(np.float32) (uint16 stays uint16), as the
reference's lazy cube must equal the file (CuboCoreTests.swift:79-130); a streamed cube gives the bits of the same
reduction over the resident cube with the same tiling.
"""
import numpy as np
import pytest

import cubokit_amd as C
from tests.support import assert_identical, golden

pytestmark = pytest.mark.gpu

FIXTURES = ["digits", "gaussian", "counts16", "counts32", "thirds", "negative", "boundary", "shift"]


def _want(arr):
    return arr if arr.dtype == np.uint16 else arr.astype(np.float32)


def test_every_fixture_loads_as_the_file():
    """CuboCoreTests.swift:79-91 on the device: load_npy, a host array through the feed, the last pattern."""
    for name in FIXTURES:
        print(f"case: fixture {name}", flush=True)
        arr = np.load(golden(name + ".npy"), allow_pickle=False)
        cube = C.HIPResidentCube.load_npy(golden(name + ".npy"))
        assert cube.shape.extents == list(arr.shape)
        got = cube.read_scan_tile(range(0, arr.shape[0]))
        np.testing.assert_array_equal(got.view(np.uint8), _want(arr).view(np.uint8))
        fed = C.HIPResidentCube(arr, tiling=C.ScanTiling.rows(1))
        np.testing.assert_array_equal(fed.read_scan_tile(range(0, arr.shape[0])).view(np.uint8),
                                      _want(arr).view(np.uint8))
        last = cube.read_pattern(arr.shape[0] - 1, arr.shape[1] - 1)
        np.testing.assert_array_equal(last, _want(arr)[-1, -1])


def test_float64_rounds_to_nearest():
    """CuboCoreTests.swift:93-105: k/3 in float64 becomes the nearest Float32."""
    d = np.load(golden("thirds.npy"), allow_pickle=False)
    got = C.HIPResidentCube.load_npy(golden("thirds.npy")).read_scan_tile(range(0, d.shape[0])).astype(np.float64)
    ulp = np.spacing(np.abs(got.astype(np.float32))).astype(np.float64)
    assert np.all(np.abs(got - d) <= ulp / 2)
    assert np.any(got != d)


def test_every_element_type_converts(tmp_path):
    """CuboCoreTests.swift:106-130: exact up to 2^24, nearest even above."""
    cases = [(np.uint8, [0, 1, 127, 128, 255]), (np.int8, [-128, -1, 0, 1, 127]),
             (np.uint16, [0, 1, 32767, 32768, 65535]), (np.int16, [-32768, -1, 0, 1, 32767]),
             (np.uint32, [0, 60000, 16_777_216, 16_777_217, 4_294_967_295]),
             (np.int32, [-2_147_483_648, -16_777_217, -1, 0, 2_147_483_647]),
             (np.uint64, [0, 16_777_217, (1 << 40) + 1, (1 << 64) - 1]),
             (np.int64, [-(1 << 63), -(1 << 40) - 1, -1, 0, (1 << 63) - 1]),
             (np.float32, [-1.5, 0, np.finfo(np.float32).smallest_subnormal, 3.4e38]),
             (np.float64, [1 / 3, -2.5e-40, 1e300, -0.0])]
    for dt, raw in cases:
        print(f"case: element type {np.dtype(dt).name}", flush=True)
        a = np.array(raw, dtype=dt).reshape(1, 1, 1, -1)
        path = str(tmp_path / f"{np.dtype(dt).name}.npy")
        np.save(path, a)
        cube = C.HIPResidentCube.load_npy(path)
        want = a if dt == np.uint16 else a.astype(np.float32)
        np.testing.assert_array_equal(cube.read_scan_tile(range(0, 1)).view(np.uint8), want.view(np.uint8))
        fed = C.HIPResidentCube(a)
        np.testing.assert_array_equal(fed.read_scan_tile(range(0, 1)).view(np.uint8), want.view(np.uint8))


def test_raw_file_with_an_offset(tmp_path):
    rng = np.random.default_rng(2)
    a = rng.integers(0, 65536, size=(5, 4, 6, 8)).astype(np.uint16)
    path = str(tmp_path / "raw.bin")
    with open(path, "wb") as f:
        f.write(b"\x00" * 24)
        f.write(a.tobytes())
    cube = C.HIPResidentCube.load_file(path, 24, "uint16", C.CubeShape(5, 4, 6, 8))
    np.testing.assert_array_equal(cube.read_scan_tile(range(0, 5)), a)


def _three_tiles():
    rng = np.random.default_rng(5)
    return rng.poisson(30.0, size=(11, 9, 16, 24)).astype(np.uint16)


@pytest.mark.parametrize("dtype", [np.uint16, np.float32, np.float64])
def test_streamed_reduction_gives_the_resident_bits(dtype):
    arr = _three_tiles().astype(dtype)
    if dtype != np.uint16:
        arr[3, 2, 5, 7] = np.nan
        arr[4, 1] *= -0.5
    tiling = C.ScanTiling.rows(3)
    geos = [C.DetectorGeometry.circle(8, 12, 5), C.DetectorGeometry.annulus(8, 12, 3, 9)]
    res = C.HIPResidentCube(arr, tiling=tiling)
    a = C.first_look(res, geos, tiling=tiling)
    print(f"case: streamed first look {np.dtype(dtype).name}", flush=True)
    b = C.first_look(arr, geos, tiling=tiling)                       # a host array: a streamed cube
    for x, y, what in ((a.images[0], b.images[0], "image 0"), (a.images[1], b.images[1], "image 1"),
                       (a.mean, b.mean, "mean"), (a.max, b.max, "max"), (a.com.intensity, b.com.intensity, "S"),
                       (a.com.row, b.com.row, "row")):
        assert_identical(np.nan_to_num(y.values, nan=12345.0), np.nan_to_num(x.values, nan=12345.0), what)
    mask = np.arange(99) % 3 != 0
    for stat in (C.DiffractionStatistic.mean, C.DiffractionStatistic.max):
        p1 = C.VirtualDiffraction.pattern(res, stat, C.ScanSelection.mask(mask), tiling=tiling)
        p2 = C.VirtualDiffraction.pattern(arr, stat, C.ScanSelection.mask(mask), tiling=tiling)
        assert_identical(np.nan_to_num(p2.values, nan=1.0), np.nan_to_num(p1.values, nan=1.0), f"pattern {stat}")


def test_a_memory_mapped_file_streams(tmp_path):
    arr = _three_tiles()
    path = str(tmp_path / "cube.npy")
    np.save(path, arr)
    mapped = np.load(path, mmap_mode="r")
    tiling = C.ScanTiling.rows(4)
    a = C.VirtualDiffraction.mean_and_max(C.HIPResidentCube.load_npy(path), tiling=tiling)
    b = C.VirtualDiffraction.mean_and_max(mapped, tiling=tiling)
    assert_identical(b[0].values, a[0].values, "mean")
    assert_identical(b[1].values, a[1].values, "max")


def test_streamed_cube_refusals():
    import ctypes as ct
    from cubokit_amd import _lib as L
    arr = _three_tiles()
    s = C.HIPStreamedCube(arr, C.ScanTiling.rows(3))
    out = np.zeros((1,) + arr.shape[1:], np.uint16)
    assert L.lib().cubo_hip_cube_read_rows(s.handle, 0, 1, out.ctypes.data) == L.ERR_UNSUPPORTED
    m = np.ones((1, 16 * 24), np.float32)
    h = ct.c_void_p()
    assert L.lib().cubo_hip_stack_plan_create(s.handle, m.ctypes.data, 1, 0, ct.byref(h)) == L.ERR_UNSUPPORTED
    big = np.ascontiguousarray(arr[:4])
    assert L.lib().cubo_hip_cube_feed_rows(s.handle, 0, 4, big.ctypes.data, 4) == L.ERR_INVALID_ARGUMENT
    req = L.ReduceRequest(masks=None, n_masks=0, want_mean_max=1, want_com=0, com_weights=None)
    r = ct.c_void_p()
    assert L.lib().cubo_hip_reduce_begin(s.handle, ct.byref(req), ct.byref(r)) == 0
    assert L.lib().cubo_hip_reduce_rows(r, 0, 3) == L.ERR_INVALID_ARGUMENT      # rows not fed yet
    L.lib().cubo_hip_reduce_cancel(r)
    s.release()


def test_full_size_npy_round_trip(tmp_path):
    """A 256x256x128x128 uint16 file (2 GiB) into HBM by load_npy and, as a host array, through the feed."""
    from cubokit_amd import synthetic
    shape, dtype, params = synthetic.bandwidth_u16()
    src = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    src.fill_synthetic(params)
    host = src.read_scan_tile(range(0, shape[0]))
    path = str(tmp_path / "big.npy")
    np.save(path, host)
    print("case: full-size load_npy", flush=True)
    loaded = C.HIPResidentCube.load_npy(path)
    a = C.VirtualDiffraction.mean_and_max(src, tiling=C.ScanTiling.rows(256))
    b = C.VirtualDiffraction.mean_and_max(loaded, tiling=C.ScanTiling.rows(256))
    assert_identical(b[0].values, a[0].values, "mean after load")
    assert_identical(b[1].values, a[1].values, "max after load")
    print("case: full-size streamed", flush=True)
    c = C.VirtualDiffraction.mean_and_max(host, tiling=C.ScanTiling.rows(16))
    d = C.VirtualDiffraction.mean_and_max(src, tiling=C.ScanTiling.rows(16))
    assert_identical(c[0].values, d[0].values, "streamed mean")
    assert_identical(c[1].values, d[1].values, "streamed max")
