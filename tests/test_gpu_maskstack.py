"""GPU parity of the mask stack Y = X W against the oracle's FP64 contraction (oc_mask_stack).

Tolerance: relative 1e-6 per output against FP64 with the reference's floor (1e-6 * max|want| of each mask
column), for the split-FP16 matrix-core method and for the plain VALU method.
"""
import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import synthetic
from oracle import oracle as O
from tests.support import assert_relative

pytestmark = pytest.mark.gpu


def _check_columns(y, ref, tol, msg):
    worst = 0.0
    for k in range(ref.shape[1]):
        worst = max(worst, assert_relative(y[:, k], ref[:, k], tol, f"{msg} column {k}"))
    return worst


def _cube(scan, det, counts, seed, dtype=np.uint16):
    shape = (scan[0], scan[1], det, det)
    p = synthetic.params(seed, (det / 2, det / 2), det / 10, counts, edge=0.5, background=3.0,
                         rings=[(det / 4, 20.0, 1.5)], waves=[(2.0, 0.0, 1 / 16, 0.0, 0.0), (0.0, 2.0, 0.0, 1 / 16, 0.0)])
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(p)
    return res


@pytest.mark.parametrize("method", [C.MaskStack.MFMA_SPLIT, C.MaskStack.PLAIN])
def test_stack_of_64_masks_within_1e6_of_fp64(method):
    res = _cube((12, 10), 64, 500.0, 11)
    masks = synthetic.stack_masks(64, 64)
    st = C.MaskStack(res, masks, method)
    assert st.method == method
    y = st.run()
    host = res.read_scan_tile(range(0, 12))
    ref = O.mask_stack(host, masks.astype(np.float64))
    w = _check_columns(y, ref, 1e-6, f"method {method}")
    print(f"method {method}: worst relative error against FP64 {w:.3e}")


def test_counts_above_2048_take_the_high_part():
    res = _cube((6, 7), 64, 30000.0, 12)
    host = res.read_scan_tile(range(0, 6))
    assert host.max() >= 2048
    masks = synthetic.stack_masks(64, 64, n_sectors=5, n_dense=8)
    y = C.MaskStack(res, masks, C.MaskStack.MFMA_SPLIT).run()
    _check_columns(y, O.mask_stack(host, masks.astype(np.float64)), 1e-6, "high counts")


@pytest.mark.parametrize("k", [1, 5, 33, 100, 200])
def test_odd_mask_counts_and_ragged_tiles(k):
    res = _cube((5, 7), 96, 300.0, 13)                 # 35 positions: a partial 128-position tile
    rng = np.random.default_rng(k)
    masks = rng.uniform(-1, 1, (k, 96 * 96)).astype(np.float32)
    host = res.read_scan_tile(range(0, 5))
    ref = O.mask_stack(host, masks.astype(np.float64))
    # signed weights cancel, so the error is bounded against the condition sum |x w| (counts are >= 0)
    cond = O.mask_stack(host, np.abs(masks).astype(np.float64))
    for method in (C.MaskStack.MFMA_SPLIT, C.MaskStack.PLAIN):
        y = C.MaskStack(res, masks, method).run()
        err = np.max(np.abs(y - ref) / cond)
        assert err <= 1e-6, f"k={k} method {method}: error {err:.3e} of sum |x w|"


def test_float32_cube_uses_the_plain_method():
    res = _cube((4, 4), 64, 200.0, 14, dtype=np.float32)
    masks = synthetic.stack_masks(64, 64, n_sectors=3, n_dense=4)
    st = C.MaskStack(res, masks)
    assert st.method == C.MaskStack.PLAIN
    with pytest.raises(C.CuboVirtualError):
        C.MaskStack(res, masks, C.MaskStack.MFMA_SPLIT)
    _check_columns(st.run(), O.mask_stack(res.read_scan_tile(range(0, 4)), masks.astype(np.float64)), 1e-6, "f32")


def test_stack_columns_equal_single_images_and_com():
    """Each stack column is a virtual image; {1, q_row, q_col} give the centre of mass (division in FP64)."""
    res = _cube((8, 8), 128, 500.0, 15)
    masks = synthetic.stack_masks(128, 128)
    y = C.MaskStack(res, masks).run()
    com = C.CenterOfMass.measure(res)
    got = C.center_of_mass_from_stack(res.shape, y)
    assert_relative(got.intensity.values, com.intensity.values, 1e-6, "S")
    assert np.max(np.abs(got.row.values - com.row.values)) <= 1e-4
    sector = masks[5]
    img = C.VirtualImaging.images(res, sector[None])[0]
    assert_relative(y[:, 5], img.values, 1e-6, "sector image")


def test_full_size_stack_on_a_sample_of_rows():
    """The 256x256x128x128 uint16 cube with K = 64: rows 0-7 against the FP64 oracle."""
    shape, dtype, params = synthetic.bandwidth_u16()
    res = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    res.fill_synthetic(params)
    masks = synthetic.stack_masks(128, 128)
    y = C.MaskStack(res, masks, C.MaskStack.MFMA_SPLIT).run()
    host = res.read_scan_tile(range(0, 8))
    ref = O.mask_stack(host, masks.astype(np.float64))
    w = _check_columns(y[: ref.shape[0]], ref, 1e-6, "full size")
    print(f"full-size K=64 split-FP16 worst relative error {w:.3e}")
