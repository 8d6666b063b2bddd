"""Benchmark of the fused first-processing pass on one MI355X.

Workload: the 256x256x128x128 uint16 synthetic cube (2 GiB, seeded Poisson counts: disk r=12 at 500 counts/px
with a +-2 px sinusoidal deflection, Gaussian rings at q=30 and 45 px of amplitude 20). One step is one fused
pass over the HBM-resident cube yielding the BF (r=14) and ADF (20-60) virtual images, the mean and maximum
patterns and the centre of mass, with the Double results copied to host memory.

Prints ONE JSON line. `value` is cube bytes per second of whole steps; `roofline.achieved` is the cube bytes
of one fused-kernel launch over that kernel's average duration from HIP events on the cube's stream.
"""
from __future__ import annotations

import argparse
import ctypes as ct
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "fused VI+COM cube GB/s (% HBM peak); BF16-split vs FP32 speedup at ≤1e-6 rel err"
ROOT = os.path.dirname(os.path.abspath(__file__))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per fused-kernel launch measured with rocprofv3 counters (profiles/traffic.json), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("fused_u16_int_bandwidth_u16", {}).get("hbm_bytes_per_launch")


def mfma_utilisation(kernel_ms: float):
    """Matrix-core utilisation of the K = 64 split-FP16 stack: the FP16 matrix flops one pass executes, counted by
    rocprofv3 (SQ_INSTS_VALU_MFMA_MOPS_F16 x 512 per dispatch, profiles/traffic.json; fixed for this workload),
    over this run's event-timed kernel duration, against MI355X's dense FP16 peak (2.5 PFLOP/s)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        e = json.load(f).get("mask_stack_mfma<2, 8>_bandwidth_u16_k64", {})
    if "mfma_executed_flop" not in e or kernel_ms <= 0:
        return None
    tflops = e["mfma_executed_flop"] / (kernel_ms / 1e3) / 1e12
    return {"tflops": round(tflops, 1), "frac_of_dense_fp16_peak": round(tflops / 2500.0, 4),
            "executed_gflop_per_pass": round(e["mfma_executed_flop"] / 1e9, 1),
            "source": "executed flops from rocprofv3 counters (SQ_INSTS_VALU_MFMA_MOPS_F16), kernel time from "
                      "HIP events in this run"}


def host_cpu():
    """(threads the baseline uses, nproc, CPU model). The threads are every core this process may run on, capped by
    OMP_NUM_THREADS where the host sets it (the GPU machines give one GPU's process a 16-core share)."""
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = nproc
    cap = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(usable, int(cap)) if cap.isdigit() and int(cap) > 0 else usable
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, nproc, usable, model


def cpu_baseline(cube, bf, adf, seconds: float):
    """The oracle's restatement of CuboKit's .cpu pass, on a bounded sample of the same cube."""
    from oracle import oracle as O
    rows = 16
    sample = cube.read_scan_tile(range(0, rows))
    threads, nproc, usable, model = host_cpu()
    nbytes = sample.nbytes
    passes, t0 = 0, time.perf_counter()
    while True:
        O.cpu_first_processing_u16(sample, 128, 128, sample.shape[1], bf, adf, threads=threads)
        passes += 1
        el = time.perf_counter() - t0
        if el >= seconds or passes >= 500:
            break
    return {"value": round(nbytes * passes / el / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "nproc": nproc, "affinity_cores": usable, "cpu_model": model,
            "sample": f"{rows} of 256 scan rows ({nbytes / 2**20:.0f} MiB of uint16), {passes} passes of the oracle's "
                      f".cpu restatement (VI serial, mean+max and CoM on {threads} threads), {el:.1f} s"}


def stack_error_vs_fp64(cube, masks, ys, rows=(0, 97, 255)):
    """C3's error of every output against the FP64 oracle on sampled scan rows: per mask column the largest
    |y - y64| / y64 (the masks are non-negative), for each method; and the centre of mass from {1, q_row, q_col}
    (the FP64 division of the stack's columns 1, 2 by 0) against the oracle's, in pixels."""
    from oracle import oracle as O
    s = cube.shape
    host = np.concatenate([cube.read_scan_tile(range(r, r + 1)) for r in rows])
    ref = O.mask_stack(host, masks.astype(np.float64))
    pos = np.concatenate([np.arange(r * s.scan_columns, (r + 1) * s.scan_columns) for r in rows])
    out = {"sample": f"scan rows {list(rows)} ({pos.size} positions), every one of the {masks.shape[0]} outputs"}
    tiny = np.finfo(np.float64).tiny
    for name, y in ys.items():
        err = np.max(np.abs(y[pos] - ref) / np.maximum(np.abs(ref), tiny), axis=0)
        out[name] = {"max": float(err.max()), "per_output": [float(f"{e:.3g}") for e in err]}
        cr, cc = y[pos, 1] / y[pos, 0], y[pos, 2] / y[pos, 0]
        out[name]["com_px_vs_fp64"] = float(max(np.max(np.abs(cr - ref[:, 1] / ref[:, 0])),
                                                np.max(np.abs(cc - ref[:, 2] / ref[:, 0]))))
    return out


def batch_section(C):
    """48 seeded uneven uint16 cubes (47.4 GiB, up to 256x256x256x256) with 8..64 masks each: one tuned graph replay
    of the persistent batch kernel against a loop of single-cube split-FP16 stacks and a loop of plain stacks (each
    with its results copied to host memory), and the replay with the results read in place in its pinned buffer (no
    copy). GB/s are cube bytes per second of wall time. Per-cube latency: the device time of each cube's own launch
    in the loops (HIP events on the cube's stream) against the batch's device time per cube."""
    from cubokit_amd import synthetic
    specs = synthetic.batch_items()
    why = C.StackBatch.availability([s for s, _, _ in specs])
    if why is not None:
        raise RuntimeError(f"the 48-cube batch does not fit the device: {why.message}")
    items = []
    for shape, p, k in specs:
        c = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=np.uint16)
        c.fill_synthetic(p)
        items.append((c, np.ascontiguousarray(synthetic.stack_masks(shape[2], shape[3])[:k])))
    total = sum(c.byte_count for c, _ in items)
    log(f"batch: {len(items)} cubes, {total / 2**30:.1f} GiB, {sum(m.shape[0] for _, m in items)} masks")
    b = C.StackBatch(items)
    shape = b.tune()
    ys = b.run()
    reps = 10
    b.kernel_time()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run(ys)
    graph_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        views = b.run(view=True)                        # the results read in place in pinned host memory
    view_s = (time.perf_counter() - t0) / reps
    views_match = all(np.array_equal(v, y) for v, y in zip(views, ys))
    gms, gn = b.kernel_time()
    res = {"cubes": len(items), "cube_bytes": total, "masks": int(sum(m.shape[0] for _, m in items)),
           "largest_cube": list(max((c.shape.extents for c, _ in items), key=lambda e: int(np.prod(e)))),
           "tuned": shape, "graph_ms": round(graph_s * 1e3, 3), "graph_device_ms": round(gms / gn, 3),
           "graph_gbs": round(total / graph_s / 1e9, 1), "graph_view_ms": round(view_s * 1e3, 3),
           "graph_view_gbs": round(total / view_s / 1e9, 1), "view_matches_copy": views_match,
           "per_cube_device_ms_batched": round(gms / gn / len(items), 4)}
    for name, method, r in (("mfma_loop", C.MaskStack.MFMA_SPLIT, 5), ("plain_loop", C.MaskStack.PLAIN, 2)):
        plans = [C.MaskStack(c, m, method) for c, m in items]
        outs = [p.run() for p in plans]
        for c, _ in items:
            c.kernel_time()                             # drop the warm-up launches
        t0 = time.perf_counter()
        for _ in range(r):
            for p, y in zip(plans, outs):
                p.run(y)
        el = (time.perf_counter() - t0) / r
        per_cube = []
        for c, _ in items:
            ms, n = c.kernel_time()
            per_cube.append(ms / r)                     # every launch of the cube's passes, per run
        for p in plans:
            p.release()
        res[name + "_ms"] = round(el * 1e3, 3)
        res[name + "_gbs"] = round(total / el / 1e9, 1)
        res[name + "_per_cube_device_ms"] = {"mean": round(float(np.mean(per_cube)), 4),
                                             "median": round(float(np.median(per_cube)), 4),
                                             "max": round(float(np.max(per_cube)), 4),
                                             "sum": round(float(np.sum(per_cube)), 3)}
        if method == C.MaskStack.MFMA_SPLIT:
            res["identical_to_single_stacks"] = all(np.array_equal(x, y) for x, y in zip(ys, outs))
    res["speedup_vs_mfma_loop"] = round(res["mfma_loop_ms"] / res["graph_ms"], 2)
    res["speedup_vs_plain_loop"] = round(res["plain_loop_ms"] / res["graph_ms"], 2)
    log(f"batch: graph {res['graph_ms']} ms ({res['graph_gbs']} GB/s), single-stack loop {res['mfma_loop_ms']} ms, "
        f"plain loop {res['plain_loop_ms']} ms; per cube {res['per_cube_device_ms_batched']} ms batched vs "
        f"{res['plain_loop_per_cube_device_ms']['mean']} ms plain; tuned {shape}")
    b.release()
    for c, _ in items:
        c.release()
    return res


def diff_section(C):
    """The 512x512x128x128 uint16 differentiation workload (8 GiB): origins from one fused CoM pass, the K = 16
    + CoM shifted stack forward, 20 fixed-step DPC iterations on the 512x512 centre-of-mass field and their
    adjoint, then the stack's adjoint with respect to the masks (re-reading the cube) and the origins (from the
    forward's kept Jacobian rows; also timed as a pass of its own, for a plan with no forward), and
    CuboPhase's plane reference and rotation solve on the same field. Kernel times from HIP events; GB/s are cube
    bytes per kernel second; the phase calls are timed whole (host arrays in and out)."""
    from cubokit_amd import synthetic
    from cubokit_amd.phase import DPC
    shape, dtype, params = synthetic.differentiation_u16()
    cube = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    cube.fill_synthetic(params)
    nb = cube.byte_count
    npos = shape[0] * shape[1]
    log(f"differentiation: cube {shape} {nb / 2**30:.1f} GiB")
    fl = C.first_look(cube, [], com=True)
    orow, ocol = fl.com.row.values.copy(), fl.com.column.values.copy()
    masks = synthetic.differentiation_masks()
    st = C.ShiftedStack(cube, masks)
    y = st.forward(orow, ocol)                              # prepares the shift classes
    st.kernel_time()
    for _ in range(3):
        y = st.forward(orow, ocol)
    ms, _ = st.kernel_time()
    plain_fwd_ms = ms[0] / 3
    st.keep_jacobian(True)                                 # the forward of a differentiation step
    y = st.forward(orow, ocol)
    st.kernel_time()
    t0 = time.perf_counter()
    for _ in range(3):
        y = st.forward(orow, ocol)
    fwd_wall = (time.perf_counter() - t0) / 3
    ms, _ = st.kernel_time()
    fwd_ms = ms[0] / 3
    S, Mr, Mc = y[:, 16], y[:, 17], y[:, 18]
    comr, comc = Mr / S, Mc / S
    fx, fy = comr - comr.mean(), comc - comc.mean()
    dpc = DPC(shape[0], shape[1])
    rng = np.random.default_rng(5)
    g_phase = rng.normal(size=npos)
    dpc.chain_backward(fx, fy, 20, g_phase)                 # warm-up
    dpc.kernel_time()
    t0 = time.perf_counter()
    phase, errors, gfx, gfy = dpc.chain_backward(fx, fy, 20, g_phase)
    dpc_wall = time.perf_counter() - t0
    dpc_ms, _ = dpc.kernel_time()
    # through the mean removal and the quotients com = M / S
    gcr, gcc = gfx - gfx.mean(), gfy - gfy.mean()
    dy = np.zeros((npos, 19))
    dy[:, :16] = rng.normal(size=(npos, 16))
    dy[:, 17], dy[:, 18] = gcr / S, gcc / S
    dy[:, 16] = -(gcr * Mr + gcc * Mc) / (S * S)
    # the backward after the forward on the same origin map: dL/dorigin from the forward's kept Jacobian rows, the
    # mask adjoint re-reading the cube
    st.backward(orow, ocol, dy)                            # warm-up
    st.kernel_time()
    st.jacobian_uses()
    t0 = time.perf_counter()
    for _ in range(2):
        dm, dr, dc = st.backward(orow, ocol, dy)
    bwd_wall = (time.perf_counter() - t0) / 2
    ms, calls = st.kernel_time()
    kept = st.jacobian_uses() == calls
    org_ms, msk_ms = ms[1] / calls, ms[2] / calls
    # the origin adjoint of a plan with no forward: its own pass over the cube
    st2 = C.ShiftedStack(cube, masks)
    st2.backward(orow, ocol, dy, masks=False)              # warm-up (prepares the shift classes)
    st2.kernel_time()
    for _ in range(2):
        st2.backward(orow, ocol, dy, masks=False)
    ms2, calls2 = st2.kernel_time()
    org_pass_ms = ms2[1] / calls2
    st2.release()
    # CuboPhase's reference plane and rotation on the same 512x512 centre-of-mass field (host arrays in and out,
    # so the call times include the copies over the host link)
    from cubokit_amd.phase import normalize, rotation_solve
    _, _, nx, ny = normalize(comr, comc, shape[0], shape[1])         # warm-up
    t0 = time.perf_counter()
    for _ in range(5):
        _, _, nx, ny = normalize(comr, comc, shape[0], shape[1])
    norm_ms = (time.perf_counter() - t0) / 5 * 1e3
    rot = rotation_solve(nx, ny, shape[0], shape[1])                 # warm-up
    t0 = time.perf_counter()
    for _ in range(5):
        rot = rotation_solve(nx, ny, shape[0], shape[1])
    rot_ms = (time.perf_counter() - t0) / 5 * 1e3
    ok = bool(np.all(np.isfinite(dm)) and np.all(np.isfinite(dr)) and np.all(np.isfinite(phase)))
    res = {"workload": "512x512x128x128 uint16 (8 GiB), disk r=12, 8-sinusoid +-2 px deflection; K=16 masks + CoM "
                       "rows, bilinear per-position origin; 20 fixed-step DPC iterations on the 512x512 field",
           "cube_bytes": nb, "forward_kernel_ms": round(fwd_ms, 3), "forward_gbs": round(nb / fwd_ms / 1e6, 1),
           "forward_keeps_jacobian": True,
           "plain_forward_kernel_ms": round(plain_fwd_ms, 3), "plain_forward_gbs": round(nb / plain_fwd_ms / 1e6, 1),
           "origin_adjoint_ms": round(org_ms, 3),
           "origin_adjoint_from": "the forward's kept Jacobian rows (no pass over the cube)" if kept else "a pass",
           "origin_adjoint_pass_ms": round(org_pass_ms, 3), "origin_adjoint_pass_gbs": round(nb / org_pass_ms / 1e6, 1),
           "mask_adjoint_ms": round(msk_ms, 3), "mask_adjoint_gbs": round(nb / msk_ms / 1e6, 1),
           "forward_and_backward_device_ms": round(fwd_ms + org_ms + msk_ms, 3),
           "dpc_20_steps_and_adjoint_device_ms": round(dpc_ms, 3), "dpc_20_steps_and_adjoint_wall_ms": round(dpc_wall * 1e3, 3),
           "forward_call_wall_ms": round(fwd_wall * 1e3, 3), "backward_call_wall_ms": round(bwd_wall * 1e3, 3),
           "dpc_final_error": float(errors[-1]), "plane_normalize_call_ms": round(norm_ms, 3),
           "rotation_solve_call_ms": round(rot_ms, 3), "rotation_degrees": rot["degrees"],
           "rotation_transposed": rot["transposed"], "finite": ok}
    log(f"differentiation: forward {fwd_ms:.3f} ms, origin adjoint {org_ms:.3f} ms (kept rows: {kept}; as a pass "
        f"{org_pass_ms:.3f} ms), mask adjoint {msk_ms:.3f} ms, "
        f"DPC 20 steps + adjoint {dpc_ms:.3f} ms device ({dpc_wall * 1e3:.1f} ms wall); plane reference "
        f"{norm_ms:.3f} ms, rotation (358 scores) {rot_ms:.3f} ms per call")
    st.release()
    dpc.release()
    cube.release()
    return res


def f64_section(C, cube16, bf, adf):
    """The FP64 fused kernel (fused_f64), the path of every Float32 cube and of fractional masks: the
    256x256x128x128 workload as Float32 (4 GiB, the same synthetic counts) with BF, ADF, mean, max and CoM, and the
    uint16 cube with fractional masks (3/4 of the BF disk, the ADF annulus) and the same outputs. One launch per
    pass; GB/s = cube bytes over the kernel's event-timed average."""
    from cubokit_amd import _lib as L
    from cubokit_amd import synthetic
    shape, _, params = synthetic.bandwidth_u16()
    f32 = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=np.float32)
    f32.fill_synthetic(params)
    lib = L.lib()
    out = {}
    for name, cube, masks in (("f32_cube", f32, np.stack([bf, adf])), ("u16_fractional", cube16,
                                                                       np.stack([0.75 * bf, adf]))):
        masks = np.ascontiguousarray(masks, np.float32)
        assert C.reduce_path(cube, masks, want_com=True) == 2
        req = L.ReduceRequest(masks=masks.ctypes.data, n_masks=2, want_mean_max=1, want_com=1, com_weights=None)
        h = ct.c_void_p()
        view = L.ReduceView()
        if lib.cubo_hip_reduce_begin(cube.handle, ct.byref(req), ct.byref(h)):
            raise RuntimeError(L.last_error())
        for _ in range(2):
            lib.cubo_hip_reduce_reset(h)
            lib.cubo_hip_reduce_rows(h, 0, cube.shape.scan_rows)
        lib.cubo_hip_reduce_view(h, ct.byref(view))
        cube.kernel_time()
        n = 10
        t0 = time.perf_counter()
        for _ in range(n):
            if lib.cubo_hip_reduce_reset(h) or lib.cubo_hip_reduce_rows(h, 0, cube.shape.scan_rows):
                raise RuntimeError(L.last_error())
        if lib.cubo_hip_reduce_view(h, ct.byref(view)):
            raise RuntimeError(L.last_error())
        el = (time.perf_counter() - t0) / n
        kms, nl = cube.kernel_time()
        lib.cubo_hip_reduce_release(h)
        ms = kms / max(1, nl)
        nb = cube.byte_count
        out[name] = {"cube_bytes": nb, "kernel_ms": round(ms, 4), "kernel_gbs": round(nb / ms / 1e6, 1),
                     "frac_of_hbm_peak": round(nb / ms / 1e6 / HBM_PEAK_GBS, 4), "pass_ms": round(el * 1e3, 4)}
        log(f"fused_f64 {name}: kernel {ms:.3f} ms ({nb / ms / 1e6:.0f} GB/s), pass {el * 1e3:.3f} ms")
    f32.release()
    return out


def pattern_section(C, cube):
    """Selected-area patterns on the bandwidth cube: a disk of half the scan positions (exact integer kernel), every
    position weighted (FP64 kernel), the disk moved by the rounded centre of mass, and by the Fourier shift (FP64
    FFTs of every pattern; the mean inverts one summed spectrum per block, the maximum every pattern). GB/s over
    the bytes of the patterns each pass reads."""
    s = cube.shape
    npos = s.scan_count
    r, c = np.divmod(np.arange(npos), s.scan_columns)
    disk = (r - s.scan_rows / 2) ** 2 + (c - s.scan_columns / 2) ** 2 < 0.5 * npos / np.pi
    weights = 0.5 + (r + c) / (s.scan_rows + s.scan_columns)
    com = C.CenterOfMass.measure(cube)
    shift = C.PatternShift.rounded(C.OriginMap(com.row.values, com.column.values))
    pbytes = s.detector_count * 2
    cases = (("mask_mean", C.DiffractionStatistic.mean, C.ScanSelection.mask(disk), None, int(disk.sum())),
             ("mask_max", C.DiffractionStatistic.max, C.ScanSelection.mask(disk), None, int(disk.sum())),
             ("weights_mean", C.DiffractionStatistic.mean, C.ScanSelection.weights(weights), None, npos),
             ("shifted_mask_mean", C.DiffractionStatistic.mean, C.ScanSelection.mask(disk), shift, int(disk.sum())),
             ("fourier_mask_mean", C.DiffractionStatistic.mean, C.ScanSelection.mask(disk),
              C.PatternShift.fourier_shift(C.OriginMap(com.row.values, com.column.values)), int(disk.sum())),
             ("fourier_mask_max", C.DiffractionStatistic.max, C.ScanSelection.mask(disk),
              C.PatternShift.fourier_shift(C.OriginMap(com.row.values, com.column.values)), int(disk.sum())))
    out = {"selected_positions": int(disk.sum())}
    for name, stat, scan, sh, n in cases:
        one = C.ScanTiling.rows(s.scan_rows)                            # the whole scan as one tile: one launch
        C.VirtualDiffraction.pattern(cube, stat, scan, shift=sh, tiling=one)        # warm-up
        cube.kernel_time()
        for _ in range(5):
            C.VirtualDiffraction.pattern(cube, stat, scan, shift=sh, tiling=one)
        tms, nl = cube.kernel_time()
        ms = tms / max(1, nl)
        out[name + "_ms"] = round(ms, 4)
        out[name + "_gbs"] = round(n * pbytes / (ms / 1e3) / 1e9, 1)
        log(f"pattern {name}: {ms:.3f} ms ({n * pbytes / (ms / 1e3) / 1e9:.0f} GB/s over the selected patterns)")
    return out


def ingest_section(C, cube):
    """Host to HBM on the bandwidth cube (2 GiB of uint16): an .npy file in the page cache loaded by load_npy, the
    same array fed from host memory into a resident cube, and a mean-and-max pass over the host array streamed in
    16-row tiles (PCIe included; the copies overlap the reduction)."""
    import tempfile
    s = cube.shape
    host = cube.read_scan_tile(range(0, s.scan_rows))
    nbytes = host.nbytes
    out = {"bytes": int(nbytes)}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cube.npy")
        np.save(path, host)
        C.HIPResidentCube.load_npy(path).release()                     # warm the page cache and the path
        t0 = time.perf_counter()
        loaded = C.HIPResidentCube.load_npy(path)
        el = time.perf_counter() - t0
        loaded.release()
    out["load_npy_gbs"] = round(nbytes / el / 1e9, 2)
    t0 = time.perf_counter()
    fed = C.HIPResidentCube(host, tiling=C.ScanTiling.rows(16))
    fed.read_scan_tile(range(0, 1))                                     # waits for the copies
    el = time.perf_counter() - t0
    fed.release()
    out["feed_resident_gbs"] = round(nbytes / el / 1e9, 2)
    C.VirtualDiffraction.mean_and_max(host, tiling=C.ScanTiling.rows(16))
    t0 = time.perf_counter()
    C.VirtualDiffraction.mean_and_max(host, tiling=C.ScanTiling.rows(16))
    el = time.perf_counter() - t0
    out["streamed_mean_max_gbs"] = round(nbytes / el / 1e9, 2)
    log(f"ingest: load_npy {out['load_npy_gbs']} GB/s, feed {out['feed_resident_gbs']} GB/s, streamed mean+max "
        f"{out['streamed_mean_max_gbs']} GB/s (PCIe included)")
    return out


def origins_section(C, cube):
    """Origins of every position of the bandwidth cube: py4DSTEM's get_origin (Double Gaussian filter, argmax,
    centre of mass) and the iterated window (Float32), one launch each over the whole scan."""
    s = cube.shape
    mx = C.VirtualDiffraction.max(cube)
    radius = C.ProbeSize.estimate(mx).radius
    one = C.ScanTiling.rows(s.scan_rows)
    out = {"radius": radius}
    for name, run in (("get_origin", lambda: C.OriginMeasurement.measure(cube, radius, tiling=one)),
                      ("iterated_window", lambda: C.OriginMeasurement.measure_with(cube, radius, tiling=one))):
        run()
        cube.kernel_time()
        for _ in range(3):
            run()
        ms, n = cube.kernel_time()
        out[name + "_ms"] = round(ms / max(1, n), 3)
        out[name + "_positions_per_s"] = round(s.scan_count / (ms / max(1, n) / 1e3), 0)
        log(f"origins {name}: {ms / max(1, n):.3f} ms for {s.scan_count} positions")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="also run one device copy of 2 GiB (read 2 GiB, write 2 GiB) for counter calibration")
    ap.add_argument("--no-stack", action="store_true", help="skip the K = 64 mask-stack comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-cube batched-stack comparison")
    ap.add_argument("--no-diff", action="store_true", help="skip the 8 GiB differentiation workload")
    ap.add_argument("--no-pattern", action="store_true", help="skip the selected-area pattern timings")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 fused-kernel timings")
    ap.add_argument("--no-ingest", action="store_true", help="skip the host-to-HBM ingestion timings")
    ap.add_argument("--no-origins", action="store_true", help="skip the origin timings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the workload runs on one MI355X "
              "and no multi-GPU path is built).", file=sys.stderr)
        return 2

    import cubokit_amd as C
    from cubokit_amd import _lib as L
    from cubokit_amd import synthetic

    info = C.device_info(0)
    log(f"device {info.name.decode()} {info.arch.decode()} CUs={info.compute_units} HBM={info.hbm_bytes / 2**30:.0f} GiB")
    shape, dtype, params = synthetic.bandwidth_u16()
    cube = C.HIPResidentCube(shape=C.CubeShape.of(shape), dtype=dtype)
    cube.fill_synthetic(params)
    cube_bytes = cube.byte_count
    log(f"cube {shape} uint16 {cube_bytes / 2**30:.2f} GiB synthesised")

    s = cube.shape
    bf = C.DetectorMask.make(C.DetectorGeometry.circle(64, 64, 14), s)
    adf = C.DetectorMask.make(C.DetectorGeometry.annulus(64, 64, 20, 60), s)
    masks = np.ascontiguousarray(np.stack([bf, adf]))
    assert C.reduce_path(cube, masks, want_com=True) == 1
    req = L.ReduceRequest(masks=masks.ctypes.data, n_masks=2, want_mean_max=1, want_com=1, com_weights=None)
    npos, ndet = s.scan_count, s.detector_count
    out = {"images": np.empty((2, npos)), "mean": np.empty(ndet), "max": np.empty(ndet)}
    for k in ("com_intensity", "com_moment_row", "com_moment_column", "com_row", "com_column"):
        out[k] = np.empty(npos)
    res = L.ReduceResult(**{k: v.ctypes.data for k, v in out.items()})
    lib = L.lib()
    h = ct.c_void_p()
    if lib.cubo_hip_reduce_begin(cube.handle, ct.byref(req), ct.byref(h)) != 0:
        raise RuntimeError(L.last_error())

    view = L.ReduceView()

    def step():
        # One pass: reset, every row as one tile: the fused kernel and the column totals, which write every result
        # into (mapped, pinned) host memory. Passes are queued back to back on the cube's stream; the host waits
        # once, after the last (a serving loop that reads a pass's results waits for it with cubo_hip_reduce_view).
        if lib.cubo_hip_reduce_reset(h) or lib.cubo_hip_reduce_rows(h, 0, s.scan_rows):
            raise RuntimeError(L.last_error())

    def wait():
        if lib.cubo_hip_reduce_view(h, ct.byref(view)):
            raise RuntimeError(L.last_error())

    for _ in range(args.warmup):
        step()
    wait()
    cube.kernel_time()                                  # drop the warm-up launches
    log(f"warmup {args.warmup} done; timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    wait()                                              # every pass complete, its results in host memory
    el = time.perf_counter() - t0
    kms, launches = cube.kernel_time()
    if lib.cubo_hip_reduce_results(h, ct.byref(res)):
        raise RuntimeError(L.last_error())
    # For reference: the same passes with the host waiting for each one's results before queueing the next.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        wait()
    sync_ms = (time.perf_counter() - t0) / args.steps * 1e3
    cube.kernel_time()
    lib.cubo_hip_reduce_release(h)
    ms_per_step = el / args.steps * 1e3
    value = cube_bytes / (el / args.steps) / 1e9
    kernel_ms = kms / max(1, launches)
    achieved = cube_bytes / (kernel_ms / 1e3) / 1e9
    log(f"step {ms_per_step:.3f} ms ({value:.0f} GB/s); fused kernel {kernel_ms:.3f} ms ({achieved:.0f} GB/s)")
    assert out["max"].max() > 400 and np.all(np.isfinite(out["images"]))

    if args.calibration_copy:
        twin = cube.clone()                              # one copy kernel: 2 GiB read, 2 GiB written
        twin.release()
        log("calibration copy of 2 GiB done")

    # The K = 64 mask stack on the same cube: split-FP16 matrix cores against the plain VALU kernel.
    stack_masks = synthetic.stack_masks(128, 128)
    stack = {}
    ys = {}
    for name, method in (() if args.no_stack else
                         (("plain", C.MaskStack.PLAIN), ("mfma_split_fp16", C.MaskStack.MFMA_SPLIT))):
        st = C.MaskStack(cube, stack_masks, method)
        y = st.run()                                    # warm-up
        cube.kernel_time()
        reps = 3 if name == "plain" else 10
        for _ in range(reps):
            st.run(y)
        tms, nl = cube.kernel_time()
        st.release()
        ys[name] = y
        stack[name + "_ms"] = round(tms / nl, 4)
        log(f"mask stack {name}: {tms / nl:.3f} ms per pass")
    if ys:
        stack["speedup"] = round(stack["plain_ms"] / stack["mfma_split_fp16_ms"], 2)
        den = np.maximum(np.abs(ys["plain"]), 1e-6 * np.abs(ys["plain"]).max(axis=0))
        stack["max_rel_diff_vs_plain"] = float(np.max(np.abs(ys["mfma_split_fp16"] - ys["plain"]) / den))
        stack["error_vs_fp64"] = stack_error_vs_fp64(cube, stack_masks, ys)
        log(f"mask stack error vs FP64: plain {stack['error_vs_fp64']['plain']['max']:.2e}, split-FP16 "
            f"{stack['error_vs_fp64']['mfma_split_fp16']['max']:.2e}")
        stack["k"] = int(stack_masks.shape[0])
        stack["mfma_gbs"] = round(cube_bytes / (stack["mfma_split_fp16_ms"] / 1e3) / 1e9, 1)
        stack["mfma_utilisation"] = mfma_utilisation(stack["mfma_split_fp16_ms"])

    fused64 = None if args.no_f64 else f64_section(C, cube, bf, adf)
    patterns = None if args.no_pattern else pattern_section(C, cube)
    ingest = None if args.no_ingest else ingest_section(C, cube)
    origins = None if args.no_origins else origins_section(C, cube)
    batch = None if args.no_batch else batch_section(C)
    diff = None if args.no_diff else diff_section(C)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(cube, bf, adf, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']} GB/s on {cpu['cores']} threads")
    traffic = traffic_from_profiles()
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": "bandwidth_u16: 256x256x128x128 uint16 Poisson cube (2 GiB), one fused pass: "
                               "BF r=14 + ADF 20-60 images, mean, max, centre of mass",
                   "cube_bytes": cube_bytes, "pct_hbm_peak_whole_step": round(100 * value / HBM_PEAK_GBS, 2),
                   "fused_kernel_ms": round(kernel_ms, 4), "kernel_path": "exact integer (fused_u16_int)",
                   "steps_queued": "back to back on the cube's stream, one host wait after the last",
                   "ms_per_step_waiting_each": round(sync_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "mask_stack_k64": stack,
        "fused_f64": fused64,
        "selected_area": patterns,
        "ingest": ingest,
        "origins": origins,
        "batch_48": batch,
        "differentiation": diff,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
