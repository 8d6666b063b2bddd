"""GPU parity of the shifted mask stack (MaskShift.bilinear per position) and its adjoint.

Against the oracle's FP64 restatement (oc_vi_bilinear, oc_vi_bilinear_grad_w, oc_vi_bilinear_grad_origin),
against central finite differences of the oracle's FP64 forward (h = 1e-3), and, end to end, against central
differences of the GPU forward itself.

Tolerances (the products are exact, the sums FP32 within 256 pixels, FP64 across):
- Y within 1e-6 of sum |x w| per output (the shifted |mask|'s image);
- dL/dmasks within 1e-6 of the same sum taken with |dL/dY|;
- dL/dorigin within 2e-6 of sum_k |dL/dY| * (the corner contractions with |mask|), per position: the origin
  adjoint is a difference of neighbouring corner sums, so its error is relative to their size.
"""
import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cube(scan, det, counts, seed):
    p = synthetic.params(seed, (det / 2, det / 2), det / 8, counts, edge=0.5, background=2.0,
                         rings=[(det / 4, 15.0, 1.5)], waves=[(1.5, 0.0, 1 / 8, 0.0, 0.0), (0.0, 1.5, 0.0, 1 / 8, 0.0)])
    res = C.HIPResidentCube(shape=C.CubeShape.of((scan[0], scan[1], det, det)), dtype=np.uint16)
    res.fill_synthetic(p)
    return res


def _masks(det, k, seed, signed=False):
    rng = np.random.default_rng(seed)
    m = rng.uniform(-1 if signed else 0, 1, (k, det * det)).astype(np.float32)
    m[0] = 1.0
    return m


def _origins(n, seed, spread=3.0, avoid_edges=False):
    rng = np.random.default_rng(seed)
    if not avoid_edges:
        return 40.0 + rng.uniform(-spread, spread, n), 25.0 + rng.uniform(-spread, spread, n)
    # shifts +-e with fractional parts in [0.2, 0.8]: their mean is 0 up to rounding, so every shift stays
    # inside one bilinear cell under a small step
    assert n % 2 == 0
    out = []
    for base in (40.0, 25.0):
        e = rng.integers(0, int(spread), n // 2) + rng.uniform(0.2, 0.8, n // 2)
        out.append(base + np.concatenate([e, -e]))
    return out[0], out[1]


def _host(res):
    return res.read_scan_tile(range(0, res.shape.scan_rows))


# 32 columns: every corner's weights loaded per chunk; 64, 128 and 256 columns (1, 2 and 4 chunks per row): the
# lower corners' weights taken from the LDS ring of the upper corners' chunks
@pytest.mark.parametrize("det", [32, 64, 128, 256])
def test_forward_matches_the_oracle_per_mask(det):
    res = _cube((6, 7), det, 300.0, 1)
    masks = _masks(det, 5, 2, signed=True)
    orow, ocol = _origins(42, 3)
    y = C.ShiftedStack(res, masks).forward(orow, ocol)
    host = _host(res)
    for k in range(masks.shape[0]):
        ref = O.vi_bilinear(host, masks[k].astype(np.float64), orow, ocol)
        cond = O.vi_bilinear(host, np.abs(masks[k]).astype(np.float64), orow, ocol)
        err = np.max(np.abs(y[:, k] - ref) / cond)
        assert err <= 1e-6, f"mask {k}: {err:.3e} of sum |x w|"


def test_equal_origins_give_the_plain_stack_bit_for_bit():
    res = _cube((5, 9), 64, 500.0, 4)
    masks = _masks(64, 17, 5)
    n = 45
    y = C.ShiftedStack(res, masks).forward(np.full(n, 63.5), np.full(n, 61.25))
    assert np.array_equal(y, C.MaskStack(res, masks, C.MaskStack.MFMA_SPLIT).run())


@pytest.mark.parametrize("det,counts", [(32, 400.0), (24, 400.0), (96, 400.0), (160, 400.0), (64, 400.0),
                                        (128, 400.0), (256, 400.0), (192, 400.0), (128, 30000.0), (96, 30000.0)])
def test_backward_matches_the_oracle(det, counts):
    """Rows of 32, 96 and 160 px: the matrix-core mask adjoint over per-position gathers (bands of three output rows
    for 32 and 96, of one for 160; 96 px rows stage with 9 waves, not a divisor of the batch); 24 px rows: the
    vector-ALU one. Rows of 64, 128, 192 and 256 px: the mask adjoint over batches of one column shift staged
    transposed (bands of 8, 4, 2 and 2 rows), and the origin adjoint's lower corners from the LDS weight ring.
    Counts up to 30000: the data's high parts (counts of 2048 and more) on both mask-adjoint kernels."""
    res = _cube((9, 11), det, counts, 6)        # 99 positions in many shift classes: partial tiles
    if counts > 2048:
        assert res.read_scan_tile(range(0, 9)).max() >= 2048
    masks = _masks(det, 6, 7, signed=True)
    orow, ocol = _origins(99, 8)
    rng = np.random.default_rng(9)
    dy = rng.normal(size=(99, 6))
    st = C.ShiftedStack(res, masks)
    dm, dr, dc = st.backward(orow, ocol, dy)
    host = _host(res)
    gr_ref, gc_ref = np.zeros(99), np.zeros(99)
    scale = np.zeros(99)
    for k in range(6):
        ref = O.vi_bilinear_grad_w(host, orow, ocol, dy[:, k])
        cond = O.vi_bilinear_grad_w(host, orow, ocol, np.abs(dy[:, k]))
        err = np.max(np.abs(dm[k] - ref) / np.maximum(cond, 1e-300))
        assert err <= 1e-6, f"dL/dmask {k}: {err:.3e} of sum |dy x|"
        g = O.vi_bilinear_grad_origin(host, masks[k].astype(np.float64), orow, ocol, dy[:, k])
        gr_ref += g[0]
        gc_ref += g[1]
        scale += np.abs(dy[:, k]) * O.vi_bilinear(host, np.abs(masks[k]).astype(np.float64), orow, ocol)
    # the mean removal spreads every position's error over all of them
    tol = 2e-6 * (scale + scale.mean())
    assert np.all(np.abs(dr - gr_ref) <= tol), f"dL/dorigin_row: {np.max(np.abs(dr - gr_ref) / tol):.3f} of tol"
    assert np.all(np.abs(dc - gc_ref) <= tol), f"dL/dorigin_col: {np.max(np.abs(dc - gc_ref) / tol):.3f} of tol"
    # only one of the two adjoints, and the class preparation reused from the previous call
    dm2, dr2, _ = st.backward(orow, ocol, dy, origin=False)
    assert dr2 is None and np.array_equal(dm2, dm)
    dm3, dr3, dc3 = st.backward(orow, ocol, dy, masks=False)
    assert dm3 is None and np.array_equal(dr3, dr) and np.array_equal(dc3, dc)


@pytest.mark.parametrize("det,counts", [(64, 400.0), (128, 400.0), (128, 30000.0), (32, 400.0), (256, 400.0)])
def test_origin_adjoint_from_the_forwards_jacobian_rows(det, counts):
    """A bilinear forward keeps dY/d(shift) for its origin map: the next backward on that map takes dL/dorigin from
    them, with no pass over the cube. Both routes against the oracle (rows of 32 pixels: every corner's weights
    loaded; 64, 128 and 256: the weight ring), and the cases that must take the pass: another origin map, a rounded
    forward in between, the cube written since the forward, the option switched off."""
    import ctypes as ct
    from cubokit_amd import _lib as L
    res = _cube((9, 11), det, counts, 31)
    masks = _masks(det, 6, 32, signed=True)
    orow, ocol = _origins(99, 33)
    dy = np.random.default_rng(34).normal(size=(99, 6))
    host = _host(res)
    gr_ref, gc_ref, scale = np.zeros(99), np.zeros(99), np.zeros(99)
    for k in range(6):
        g = O.vi_bilinear_grad_origin(host, masks[k].astype(np.float64), orow, ocol, dy[:, k])
        gr_ref += g[0]
        gc_ref += g[1]
        scale += np.abs(dy[:, k]) * O.vi_bilinear(host, np.abs(masks[k]).astype(np.float64), orow, ocol)
    tol = 2e-6 * (scale + scale.mean())

    def check(dr, dc, what):
        assert np.all(np.abs(dr - gr_ref) <= tol), f"{what} row: {np.max(np.abs(dr - gr_ref) / tol):.3f} of tol"
        assert np.all(np.abs(dc - gc_ref) <= tol), f"{what} column: {np.max(np.abs(dc - gc_ref) / tol):.3f} of tol"

    kept = True
    st = C.ShiftedStack(res, masks, keep_jacobian=True)
    y = st.forward(orow, ocol)
    for k in (0, 5):
        ref = O.vi_bilinear(host, masks[k].astype(np.float64), orow, ocol)
        cond = O.vi_bilinear(host, np.abs(masks[k]).astype(np.float64), orow, ocol)
        assert np.max(np.abs(y[:, k] - ref) / cond) <= 1e-6
    st.jacobian_uses()
    dm, dr, dc = st.backward(orow, ocol, dy)
    assert st.jacobian_uses() == (1 if kept else 0)
    check(dr, dc, "from the forward" if kept else "pass")
    # a plan with no forward takes the pass; both routes agree within the tolerance
    fresh = C.ShiftedStack(res, masks)
    dm2, dr2, dc2 = fresh.backward(orow, ocol, dy)
    assert fresh.jacobian_uses() == 0 and np.array_equal(dm2, dm)
    check(dr2, dc2, "pass")
    # the kept rows serve every backward on the map until it changes
    _, dr3, _ = st.backward(orow, ocol, dy, masks=False)
    assert st.jacobian_uses() == (1 if kept else 0) and np.array_equal(dr3, dr)
    # another map: the pass
    orow2 = orow + 0.37
    st.backward(orow2, ocol, dy, masks=False)
    assert st.jacobian_uses() == 0
    # a rounded forward on the map, then the bilinear backward: the pass
    st.forward(orow, ocol)
    st.forward(orow, ocol, rounded=True)
    st.backward(orow, ocol, dy, masks=False)
    assert st.jacobian_uses() == 0
    # the cube written after the forward (the same rows back): the pass, with the same gradients
    st.forward(orow, ocol)
    row = np.ascontiguousarray(host[2:3])
    assert L.lib().cubo_hip_cube_upload_rows(res.handle, 2, 1, row.ctypes.data_as(ct.c_void_p)) == 0
    _, dr4, dc4 = st.backward(orow, ocol, dy, masks=False)
    assert st.jacobian_uses() == 0 and np.array_equal(dr4, dr2) and np.array_equal(dc4, dc2)
    # switched off: the forward keeps nothing (and gives the same Y)
    st.keep_jacobian(False)
    assert np.array_equal(st.forward(orow, ocol), y)
    st.backward(orow, ocol, dy, masks=False)
    assert st.jacobian_uses() == 0


def test_gradients_against_central_differences_of_the_fp64_forward():
    res = _cube((4, 5), 32, 400.0, 10)
    masks = _masks(32, 3, 11)
    orow, ocol = _origins(20, 12, avoid_edges=True)
    rng = np.random.default_rng(13)
    dy = rng.normal(size=(20, 3))
    host = _host(res)

    def loss(m, r, c):
        return sum(float(np.dot(dy[:, k], O.vi_bilinear(host, m[k].astype(np.float64), r, c))) for k in range(3))

    dm, dr, dc = C.ShiftedStack(res, masks).backward(orow, ocol, dy)
    h = 1e-3
    for trial in range(3):
        e = rng.normal(size=masks.shape)
        fd = (loss(masks + h * e, orow, ocol) - loss(masks - h * e, orow, ocol)) / (2 * h)
        an = float(np.sum(dm * e))
        assert abs(fd - an) <= 1e-6 * (abs(fd) + np.sum(np.abs(dm * e))), f"masks direction {trial}: {fd} vs {an}"
        vr, vc = rng.normal(size=20), rng.normal(size=20)
        fd = (loss(masks, orow + h * vr, ocol + h * vc) - loss(masks, orow - h * vr, ocol - h * vc)) / (2 * h)
        an = float(np.dot(dr, vr) + np.dot(dc, vc))
        assert abs(fd - an) <= 1e-5 * (abs(fd) + np.abs(dr * vr).sum() + np.abs(dc * vc).sum()), \
            f"origin direction {trial}: {fd} vs {an}"


def test_gpu_forward_and_adjoint_agree_end_to_end():
    """Central differences of the GPU forward against the GPU adjoint (h = 1e-3). The forward's FP32 partial
    sums limit the difference quotient to about 1e-7 of sum |x w| / h."""
    res = _cube((6, 6), 32, 500.0, 14)
    masks = _masks(32, 4, 15)
    orow, ocol = _origins(36, 16, avoid_edges=True)
    rng = np.random.default_rng(17)
    dy = rng.normal(size=(36, 4))
    st = C.ShiftedStack(res, masks)
    dm, dr, dc = st.backward(orow, ocol, dy)
    h = 1e-3
    vr, vc = rng.normal(size=36), rng.normal(size=36)
    lp = float(np.sum(dy * st.forward(orow + h * vr, ocol + h * vc)))
    lm = float(np.sum(dy * st.forward(orow - h * vr, ocol - h * vc)))
    fd = (lp - lm) / (2 * h)
    an = float(np.dot(dr, vr) + np.dot(dc, vc))
    mag = float(np.sum(np.abs(dy) * np.abs(st.forward(orow, ocol))))
    assert abs(fd - an) <= 1e-3 * mag, f"{fd} vs {an} (scale {mag:.3e})"
    e = rng.normal(size=masks.shape)
    mp, mm = (masks + h * e).astype(np.float32), (masks - h * e).astype(np.float32)
    lp = float(np.sum(dy * C.ShiftedStack(res, mp).forward(orow, ocol)))
    lm = float(np.sum(dy * C.ShiftedStack(res, mm).forward(orow, ocol)))
    fd = (lp - lm) / (2 * h)
    an = float(np.sum(dm * (mp.astype(np.float64) - mm) / (2 * h)))
    assert abs(fd - an) <= 1e-3 * mag, f"masks: {fd} vs {an} (scale {mag:.3e})"


def test_large_classes_span_several_tiles():
    res = _cube((24, 30), 64, 300.0, 18)          # 720 positions, origins within one cell: 1-4 classes
    masks = _masks(64, 9, 19)
    n = 720
    rng = np.random.default_rng(20)
    orow, ocol = 30.0 + rng.uniform(0, 0.45, n), 30.0 + rng.uniform(0, 0.45, n)
    st = C.ShiftedStack(res, masks)
    y = st.forward(orow, ocol)
    host = _host(res)
    for k in (0, 4, 8):
        ref = O.vi_bilinear(host, masks[k].astype(np.float64), orow, ocol)
        cond = O.vi_bilinear(host, np.abs(masks[k]).astype(np.float64), orow, ocol)
        assert np.max(np.abs(y[:, k] - ref) / cond) <= 1e-6
    dy = rng.normal(size=(n, 9))
    dm, _, _ = st.backward(orow, ocol, dy, origin=False)
    ref = O.vi_bilinear_grad_w(host, orow, ocol, dy[:, 4])
    cond = O.vi_bilinear_grad_w(host, orow, ocol, np.abs(dy[:, 4]))
    assert np.max(np.abs(dm[4] - ref) / np.maximum(cond, 1e-300)) <= 1e-6


def test_split_tiles_match_the_oracle():
    """A scan whose whole tiles (256 positions of one shift class) fill more than one round of the device, so each
    class's partial tile is split into chunk ranges run by separate blocks and added afterwards: 320 x 256
    positions, origins in a 2 x 2 pixel square (four classes of about 20,480 positions, each with a partial tile).
    Y on every position and the origin adjoint against the oracle."""
    scan, det, n = (320, 256), 64, 320 * 256
    res = _cube(scan, det, 300.0, 22)
    masks = _masks(det, 5, 23, signed=True)
    rng = np.random.default_rng(24)
    orow, ocol = 30.0 + rng.uniform(-1, 1, n), 30.0 + rng.uniform(-1, 1, n)
    # the shift classes as the host forms them: whole tiles and partial ones
    fr, fc = np.floor(orow - orow.mean()), np.floor(ocol - ocol.mean())
    _, counts = np.unique(fr * 1000 + fc, return_counts=True)
    assert np.sum(counts // 256) >= C.device_info(0).compute_units and np.any(counts % 256)
    st = C.ShiftedStack(res, masks, keep_jacobian=True)
    y = st.forward(orow, ocol)
    host = _host(res)
    for k in range(masks.shape[0]):
        ref = O.vi_bilinear(host, masks[k].astype(np.float64), orow, ocol)
        cond = O.vi_bilinear(host, np.abs(masks[k]).astype(np.float64), orow, ocol)
        err = np.max(np.abs(y[:, k] - ref) / cond)
        assert err <= 1e-6, f"mask {k}: {err:.3e} of sum |x w|"
    dy = rng.normal(size=(n, masks.shape[0]))
    st.jacobian_uses()
    _, dr, dc = st.backward(orow, ocol, dy, masks=False)
    assert st.jacobian_uses() == 1                  # the split tiles' Jacobian rows added as their Y
    gr_ref, gc_ref, scale = np.zeros(n), np.zeros(n), np.zeros(n)
    for k in range(masks.shape[0]):
        g = O.vi_bilinear_grad_origin(host, masks[k].astype(np.float64), orow, ocol, dy[:, k])
        gr_ref += g[0]
        gc_ref += g[1]
        scale += np.abs(dy[:, k]) * O.vi_bilinear(host, np.abs(masks[k]).astype(np.float64), orow, ocol)
    tol = 2e-6 * (scale + scale.mean())
    assert np.all(np.abs(dr - gr_ref) <= tol), f"dL/dorigin_row: {np.max(np.abs(dr - gr_ref) / tol):.3f} of tol"
    assert np.all(np.abs(dc - gc_ref) <= tol), f"dL/dorigin_col: {np.max(np.abs(dc - gc_ref) / tol):.3f} of tol"


def test_refusals():
    res = _cube((2, 2), 32, 100.0, 21)
    with pytest.raises(C.CuboVirtualError):
        C.ShiftedStack(res, np.ones((33, 1024), np.float32))
    # a Float32 cube has the forward pass (vector ALUs) but not the adjoint
    f32 = C.HIPResidentCube(np.ones((2, 2, 32, 32), np.float32))
    st32 = C.ShiftedStack(f32, np.ones((1, 1024), np.float32))
    st32.forward(np.zeros(4), np.zeros(4))
    with pytest.raises(C.CuboVirtualError):
        st32.backward(np.zeros(4), np.zeros(4), np.ones((4, 1)))
    st = C.ShiftedStack(res, np.ones((1, 1024), np.float32))
    with pytest.raises(C.CuboVirtualError):
        st.forward(np.array([1.0, np.nan, 2.0, 3.0]), np.zeros(4))


@pytest.mark.parametrize("dtype,det", [(np.float32, (16, 20)), (np.uint16, (15, 13))])
def test_vector_path_matches_the_oracle(dtype, det):
    """Cubes the matrix-core path does not take (Float32; a pattern length not a multiple of 8) run the forward
    pass on the vector ALUs in Double: within 1e-12 of sum |x w| of the oracle, bilinear and rounded."""
    rng = np.random.default_rng(31)
    qy, qx = det
    if dtype == np.float32:
        arr = rng.normal(50.0, 20.0, size=(5, 6, qy, qx)).astype(np.float32)
    else:
        arr = rng.integers(0, 65536, size=(5, 6, qy, qx)).astype(np.uint16)
    res = C.HIPResidentCube(arr)
    masks = rng.uniform(-1, 1, (3, qy * qx)).astype(np.float32)
    orow, ocol = rng.normal(7, 5, 30), rng.normal(9, 6, 30)
    st = C.ShiftedStack(res, masks)
    for rounded in (False, True):
        y = st.forward(orow, ocol, rounded=rounded)
        f = O.vi_rounded if rounded else O.vi_bilinear
        for k in range(3):
            ref = f(arr, masks[k].astype(np.float64), orow, ocol)
            cond = f(np.abs(arr.astype(np.float32)), np.abs(masks[k]).astype(np.float64), orow, ocol)
            assert np.max(np.abs(y[:, k] - ref) / cond) <= 1e-12, (rounded, k)
