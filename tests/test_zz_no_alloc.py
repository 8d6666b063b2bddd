"""cubo_hip_reduce_noalloc (the fused reduction with its scratch in a static device array) against
cubo_hip_reduce, through the C ABI: the same bits on two calls in a row, the second on a smaller cube, on the
exact integer kernel and the FP64 kernel; a request that needs more than the static scratch is refused before any
device work, and the allocating entry point still serves it."""
import ctypes as ct

import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import _lib as L
from cubokit_amd import synthetic
from tests.support import assert_identical, load_cube

pytestmark = pytest.mark.gpu

KEYS = ("images", "mean", "max", "com_intensity", "com_moment_row", "com_moment_column", "com_row", "com_column")


def _run(entry, cube, masks, want_mean_max=True, want_com=True, com_weights=None):
    s = cube.shape
    m = np.ascontiguousarray(masks, np.float32).reshape(-1, s.detector_count)
    cw = None if com_weights is None else np.ascontiguousarray(com_weights, np.float32)
    req = L.ReduceRequest(masks=m.ctypes.data if m.shape[0] else None, n_masks=m.shape[0],
                          want_mean_max=int(want_mean_max), want_com=int(want_com),
                          com_weights=None if cw is None else cw.ctypes.data)
    sizes = {"images": m.shape[0] * s.scan_count, "mean": s.detector_count, "max": s.detector_count}
    out = {k: np.full(sizes.get(k, s.scan_count), 7.25) for k in KEYS}
    res = L.ReduceResult(**{k: v.ctypes.data for k, v in out.items()})
    st = getattr(L.lib(), entry)(cube.handle, ct.byref(req), ct.byref(res))
    return st, out


def _same(cube, masks, what, **kw):
    print(f"case: no-alloc reduction, {what}, cube {cube.shape.extents}", flush=True)
    st, got = _run("cubo_hip_reduce_noalloc", cube, masks, **kw)
    assert st == 0, L.last_error()
    st, want = _run("cubo_hip_reduce", cube, masks, **kw)
    assert st == 0, L.last_error()
    for k in KEYS:
        assert_identical(got[k], want[k], f"{what}: {k}")
    return got


def test_noalloc_reduction_matches_the_allocating_one():
    # 1. the 256x256x128x128 uint16 workload (2 GiB) on the exact integer kernel: BF and ADF, mean, max, CoM
    shape, dtype, params = synthetic.bandwidth_u16()
    big = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    big.fill_synthetic(params)
    bf = C.DetectorMask.make(C.DetectorGeometry.circle(64, 64, 14), big.shape)
    adf = C.DetectorMask.make(C.DetectorGeometry.annulus(64, 64, 20, 60), big.shape)
    assert C.reduce_path(big, np.stack([bf, adf]), want_com=True) == 1
    got = _same(big, np.stack([bf, adf]), "integer kernel, 2 GiB")
    assert got["max"].max() > 400
    # 2. a smaller cube right after, with fewer groups and other rows: the scratch keeps the last call's contents
    small = C.HIPResidentCube(load_cube("counts16"))
    m = C.DetectorMask.make(C.DetectorGeometry.circle(5.5, 5.5, 3.6), small.shape)
    _same(small, m[None], "integer kernel, counts16")
    # 3. the FP64 kernel: fractional masks and CoM weights on a Float32 cube, then on the uint16 cube
    f32 = C.HIPResidentCube(load_cube("gaussian"))
    qn = f32.shape.detector_count
    rng = np.random.default_rng(5)
    fm = rng.uniform(-1, 1, (3, qn)).astype(np.float32)
    _same(f32, fm, "FP64 kernel, Float32 cube", com_weights=rng.uniform(0, 2, qn))
    _same(big, (bf * 0.5)[None], "FP64 kernel, 2 GiB uint16, fractional mask", want_com=False)
    # 4. only some outputs requested; the rest untouched
    _same(small, m[None], "images only", want_mean_max=False, want_com=False)
    st, out = _run("cubo_hip_reduce_noalloc", small, np.zeros((0, small.shape.detector_count)), want_com=False)
    assert st == 0 and np.all(out["com_row"] == 7.25) and np.all(out["images"] == 7.25)


def test_noalloc_reduction_refuses_what_exceeds_its_scratch():
    """Twelve fractional masks plus the CoM rows on the 2 GiB cube run on the FP64 kernel, whose row partials
    ([16 slices][65536 positions][16 rows] Doubles) need more than the static scratch."""
    shape, dtype, params = synthetic.bandwidth_u16()
    big = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    big.fill_synthetic(params)
    masks = np.random.default_rng(2).uniform(0, 1, (12, big.shape.detector_count)).astype(np.float32)
    st, out = _run("cubo_hip_reduce_noalloc", big, masks)
    assert st == L.ERR_EXCEEDS_BUFFER_LIMIT and "MiB" in L.last_error()
    assert np.all(out["mean"] == 7.25)
    st, out = _run("cubo_hip_reduce", big, masks)
    assert st == 0 and np.all(np.isfinite(out["images"]))
    st, _ = _run("cubo_hip_reduce_noalloc", big, masks[:1], want_com=False)              # one mask fits
    assert st == 0, L.last_error()
