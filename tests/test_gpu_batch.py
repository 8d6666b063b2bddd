"""GPU parity of batched mask stacks: many uneven uint16 cubes in one persistent launch, replayed from a hipGraph.

Each item must give exactly the bits of the same cube's single split-FP16 mask stack (same operands, same
accumulation order per position), and lie within 1e-6 of sum |x w| of the oracle's FP64 contraction
(oc_mask_stack). Graph replays with other device work between them must repeat the first result bit for bit.
"""
import ctypes as ct

import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import _lib as L
from cubokit_amd import synthetic
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (scan rows, scan columns, detector side, masks): uneven scans (partial 128- and 256-position tiles, a single
# position), uneven detectors (one 64-pixel chunk, a partial chunk, many chunks), 1..64 masks.
ITEMS = [((5, 7), 64, 17), ((12, 10), 96, 64), ((1, 1), 32, 1), ((16, 20), 128, 40), ((9, 11), 36, 33),
         ((3, 40), 44, 8), ((20, 20), 64, 64)]


def _cube(scan, det, counts, seed):
    p = synthetic.params(seed, (det / 2, det / 2), det / 10, counts, edge=0.5, background=3.0,
                         rings=[(det / 4, 20.0, 1.5)], waves=[(2.0, 0.0, 1 / 16, 0.0, 0.0)])
    res = C.HIPResidentCube(shape=C.CubeShape.of((scan[0], scan[1], det, det)), dtype=np.uint16)
    res.fill_synthetic(p)
    return res


def _batch_items(items=ITEMS, seed=0):
    out = []
    for i, (scan, det, k) in enumerate(items):
        res = _cube(scan, det, 400.0 if i % 2 else 4000.0, 100 + seed + i)
        rng = np.random.default_rng(1000 + seed + i)
        masks = rng.uniform(-1, 1, (k, det * det)).astype(np.float32)
        masks[0] = 1.0                                       # a plain sum column
        out.append((res, masks))
    return out


def _check_against_oracle(items, ys):
    for j, ((res, masks), y) in enumerate(zip(items, ys)):
        host = res.read_scan_tile(range(0, res.shape.scan_rows))
        ref = O.mask_stack(host, masks.astype(np.float64))
        cond = O.mask_stack(host, np.abs(masks).astype(np.float64))
        with np.errstate(invalid="ignore", divide="ignore"):
            err = np.where(cond > 0, np.abs(y - ref) / cond, np.abs(y - ref))
        assert err.max() <= 1e-6, f"item {j}: error {err.max():.3e} of sum |x w|"


def test_batch_matches_single_stacks_bit_for_bit_and_the_oracle():
    items = _batch_items()
    b = C.StackBatch(items)
    ys = b.run()
    for j, ((res, masks), y) in enumerate(zip(items, ys)):
        single = C.MaskStack(res, masks, C.MaskStack.MFMA_SPLIT).run()
        assert y.shape == single.shape
        assert np.array_equal(y, single), f"item {j}: batch differs from the single stack"
    _check_against_oracle(items, ys)
    ms, runs = b.kernel_time()
    assert runs == 1 and ms > 0


def test_graph_replays_repeat_the_first_result():
    items = _batch_items(seed=7)
    b = C.StackBatch(items)
    first = b.run()
    other = _cube((8, 8), 64, 300.0, 55)
    for rep in range(4):
        # eager device work of the program between replays, on other streams and buffers
        C.MaskStack(other, synthetic.stack_masks(64, 64, n_sectors=3, n_dense=4)).run()
        C.first_look(other)
        again = b.run()
        for j, (x, y) in enumerate(zip(first, again)):
            assert np.array_equal(x, y), f"replay {rep} item {j} differs"
    b.run(copy_out=False)
    assert b.kernel_time()[1] == 6


def test_view_reads_the_results_in_place():
    """cubo_hip_batch_view: an error before the first run; afterwards each cube's results in the pinned buffer,
    the same bits as the copy out, and read-only."""
    items = _batch_items(seed=11)
    b = C.StackBatch(items)
    with pytest.raises(C.CuboVirtualError):
        b.view()
    copied = b.run()
    viewed = b.run(view=True)
    for j, (x, y) in enumerate(zip(copied, viewed)):
        assert y.shape == x.shape and np.array_equal(x, y), f"item {j}: view differs from the copy"
        assert not y.flags.writeable
    ptr = ct.POINTER(ct.c_double)()
    assert L.lib().cubo_hip_batch_view(b._h, len(items), ct.byref(ptr)) != 0       # no such item
    assert L.lib().cubo_hip_batch_view(b._h, -1, ct.byref(ptr)) != 0


def test_tuned_shape_gives_identical_results():
    items = _batch_items(seed=3)
    b = C.StackBatch(items)
    before = b.run()
    shape = b.tune()
    assert shape["waves"] in (4, 8) and shape["blocks_per_cu"] >= 1
    after = b.run()
    for j, (x, y) in enumerate(zip(before, after)):
        assert np.array_equal(x, y), f"item {j} differs after tuning to {shape}"
    # a second batch of the same shapes starts from the tuned shape and still agrees
    b2 = C.StackBatch(items)
    for j, (x, y) in enumerate(zip(before, b2.run())):
        assert np.array_equal(x, y), f"item {j} differs in a batch built after tuning"


def test_batch_rejects_what_the_matrix_core_method_cannot_run():
    ok = _cube((2, 2), 32, 100.0, 1)
    f32 = C.HIPResidentCube(shape=C.CubeShape.of((2, 2, 32, 32)), dtype=np.float32)
    odd = C.HIPResidentCube(shape=C.CubeShape.of((2, 2, 9, 9)), dtype=np.uint16)
    m = np.ones((1, 32 * 32), np.float32)
    with pytest.raises(C.CuboVirtualError):
        C.StackBatch([(ok, m), (f32, m)])
    with pytest.raises(C.CuboVirtualError):
        C.StackBatch([(odd, np.ones((1, 81), np.float32))])
    with pytest.raises(C.CuboVirtualError):
        C.StackBatch([(ok, np.ones((65, 32 * 32), np.float32))])
    bad = m.copy()
    bad[0, 3] = np.nan
    with pytest.raises(C.CuboVirtualError):
        C.StackBatch([(ok, bad)])
    with pytest.raises(C.CuboError):
        C.StackBatch([])


def test_view_keeps_the_batch_alive():
    """The in-place views hold the batch (and its pinned buffer) through their base: a batch dropped right after
    run(view=True) still leaves readable, unchanged results."""
    import gc
    items = _batch_items(seed=13)
    copied = C.StackBatch(items).run()
    views = C.StackBatch(items).run(view=True)             # the batch object itself is not kept here
    gc.collect()
    for j, (x, v) in enumerate(zip(copied, views)):
        base = v
        while base is not None and not hasattr(base, "_batch"):
            base = getattr(base, "base", None)
        assert base is not None and isinstance(base._batch, C.StackBatch), f"item {j}: no batch in the base chain"
        assert np.array_equal(x, v), f"item {j}: the view changed after the batch was dropped"


def test_screening_batch_of_48_cubes():
    """The 48-cube screening workload as specified: scans {32^2, 64^2, 96x128, 128^2, 256^2}, detectors
    {64^2, 96^2, 128^2, 256^2} (two 256^4 cubes of 8 GiB; 47.4 GiB in all), 8..64 masks each, after an
    availability check of the whole batch. Every item equals its own cube's single split-FP16 stack bit for bit, and
    eight sampled positions of every cube are within 1e-6 of the oracle's FP64 contraction (relative to sum x w,
    the masks being non-negative)."""
    specs = synthetic.batch_items()
    assert len(specs) == 48 and max(s[2] for s, _, _ in specs) == 256 and max(s[0] for s, _, _ in specs) == 256
    assert C.StackBatch.availability([s for s, _, _ in specs]) is None, "the batch does not fit the device"
    items = []
    for j, (shape, p, k) in enumerate(specs):
        c = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=np.uint16)
        c.fill_synthetic(p)
        items.append((c, np.ascontiguousarray(synthetic.stack_masks(shape[2], shape[3])[:k])))
    print(f"case: 48-cube batch, {synthetic.batch_bytes(specs) / 2**30:.1f} GiB", flush=True)
    b = C.StackBatch(items)
    b.tune()
    ys = b.run()
    rng = np.random.default_rng(48)
    for j, ((c, masks), y) in enumerate(zip(items, ys)):
        print(f"case: batch item {j} {c.shape.extents} k={masks.shape[0]}", flush=True)
        single = C.MaskStack(c, masks, C.MaskStack.MFMA_SPLIT).run()
        assert np.array_equal(y, single), f"item {j}: batch differs from the single stack"
        rows = sorted(set(int(r) for r in rng.integers(0, c.shape.scan_rows, 2)))
        for r in rows:
            host = c.read_scan_tile(range(r, r + 1)).reshape(c.shape.scan_columns, -1)
            cols = rng.choice(c.shape.scan_columns, size=min(4, c.shape.scan_columns), replace=False)
            sub = host[cols].astype(np.float32)
            ref = O.mask_stack(sub.reshape(len(cols), 1, c.shape.detector_rows, c.shape.detector_columns),
                               masks.astype(np.float64))
            got = y[r * c.shape.scan_columns + cols]
            assert np.all(masks >= 0)
            err = np.abs(got - ref) / np.maximum(ref, np.finfo(np.float64).tiny)
            assert err.max() <= 1e-6, f"item {j} row {r}: error {err.max():.3e}"
    b.release()
    for c, _ in items:
        c.release()
