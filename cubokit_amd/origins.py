"""Mirror of CuboVirtual's origins over the C ABI (include/cubo_hip.h, "origins").

- ProbeSize.estimate (Sources/CuboVirtual/ProbeSize.swift:25-58), on the host in the library;
- OriginMeasurement.measure (OriginMeasurement.swift:44-95, py4DSTEM's get_origin) and the iterated-window
  estimator (IteratedWindowOrigin.swift:183-241), on the device, tile by tile with the reference's cancellation and
  progress.
"""
from __future__ import annotations

import ctypes as ct
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .virtual import (CancellationError, ComputePath, CubeShape, CuboError, CuboVirtualError, DetectorImage,
                      HIPResidentCube, IsCancelled, OriginMap, Progress, ScanImage, ScanTiling, VirtualDiffraction,
                      _check, _require_hip, _resident)


@dataclass(frozen=True)
class ProbeEstimate:
    radius: float
    center_row: float
    center_column: float
    threshold: float


class ProbeSize:
    @staticmethod
    def estimate(pattern: DetectorImage, lower_threshold: float = 0.01, upper_threshold: float = 0.99,
                 count: int = 100) -> ProbeEstimate:
        v = np.ascontiguousarray(pattern.values, np.float64)
        out = L.ProbeEstimate()
        _check(L.lib().cubo_hip_probe_size(v.ctypes.data, pattern.detector_rows, pattern.detector_columns,
                                           float(lower_threshold), float(upper_threshold), int(count), ct.byref(out)),
               "probe size")
        return ProbeEstimate(out.radius, out.center_row, out.center_column, out.threshold)


@dataclass(frozen=True)
class OriginEstimator:
    """OriginEstimator (IteratedWindowOrigin.swift:14-27): py4DSTEM's get_origin or the iterated window."""
    kind: str
    passes: int = 4
    minimum_margin: float = 1.5
    binned_coarse: bool = True

    @staticmethod
    def py4dstem() -> "OriginEstimator":
        return OriginEstimator("py4DSTEM")

    @staticmethod
    def iterated_window(passes: int = 4, minimum_margin: float = 1.5, binned_coarse: bool = True) -> "OriginEstimator":
        return OriginEstimator("iteratedWindow", passes, minimum_margin, binned_coarse)


def _measure(source, request: L.OriginRequest, tiling: ScanTiling, progress: Progress,
             is_cancelled: IsCancelled) -> OriginMap:
    tiling = tiling or ScanTiling.automatic
    if is_cancelled is not None and is_cancelled():
        raise CancellationError()
    cube, temporary, feed = _resident(source, tiling)
    try:
        shape = cube.shape
        rows, cols = np.zeros(shape.scan_count), np.zeros(shape.scan_count)
        for rg in tiling.ranges(shape):
            if is_cancelled is not None and is_cancelled():
                raise CancellationError()
            if feed is not None:
                feed(rg)
            _check(L.lib().cubo_hip_origins(cube.handle, ct.byref(request), rg.start, len(rg), rows.ctypes.data,
                                            cols.ctypes.data), "measuring origins")
            if progress is not None:
                progress(rg.stop / shape.scan_rows)
        img = lambda v: ScanImage(shape.scan_rows, shape.scan_columns, v, ComputePath.hip)   # noqa: E731
        return OriginMap(img(rows), img(cols))
    finally:
        if temporary:
            cube.release()


class OriginMeasurement:
    @staticmethod
    def measure(source, radius: Optional[float] = None, radius_scale: float = 1.2, scan_mask=None,
                fast_center: bool = False, tiling: ScanTiling = None, progress: Progress = None,
                is_cancelled: IsCancelled = None):
        """py4DSTEM's get_origin: (OriginMap, radius); the radius defaults to the probe size of the maximum pattern."""
        shape = source.shape if isinstance(source, HIPResidentCube) else None
        if radius is None:
            mx = VirtualDiffraction.max(source, tiling=tiling)
            radius = ProbeSize.estimate(mx).radius
        take = None
        if scan_mask is not None:
            take = np.ascontiguousarray(np.asarray(scan_mask, bool).reshape(-1).astype(np.uint8))
            n = (shape.scan_count if shape is not None else int(np.prod(np.asarray(source).shape[:2])))
            if take.size != n:
                raise CuboVirtualError("maskLengthMismatch", f"expected {n}, received {take.size}")
        req = L.OriginRequest(estimator=L.ORIGIN_PY4DSTEM, radius=float(radius), radius_scale=float(radius_scale),
                              scan_mask=None if take is None else take.ctypes.data, fast_center=int(fast_center),
                              passes=0, minimum_margin=0.0, binned_coarse=0)
        return _measure(source, req, tiling, progress, is_cancelled), float(radius)

    @staticmethod
    def measure_with(source, radius: float, radius_scale: float = 1.2,
                     estimator: OriginEstimator = OriginEstimator.iterated_window(),
                     path: ComputePath = ComputePath.hip, tiling: ScanTiling = None, progress: Progress = None,
                     is_cancelled: IsCancelled = None) -> OriginMap:
        """measure(of:radius:radiusScale:estimator:path:...) (IteratedWindowOrigin.swift:183-241)."""
        _require_hip(path)
        if estimator.kind == "py4DSTEM":
            return OriginMeasurement.measure(source, radius, radius_scale, tiling=tiling, progress=progress,
                                             is_cancelled=is_cancelled)[0]
        req = L.OriginRequest(estimator=L.ORIGIN_ITERATED_WINDOW, radius=float(radius),
                              radius_scale=float(radius_scale), scan_mask=None, fast_center=0,
                              passes=int(estimator.passes), minimum_margin=float(estimator.minimum_margin),
                              binned_coarse=int(estimator.binned_coarse))
        return _measure(source, req, tiling, progress, is_cancelled)


def sensitivity_sample_indices(total: int, limit: int, pattern_bytes: int, budget_bytes: int) -> list:
    """OriginMeasurement.sensitivitySampleIndices (IteratedWindowOrigin.swift:243-252): every n-th position, n the
    smallest step that keeps the count at or under `limit` and, when pattern_bytes > 0, under
    budget_bytes / pattern_bytes (at least one position)."""
    if total <= 0 or limit <= 0:
        return []
    by_bytes = max(1, budget_bytes // pattern_bytes) if pattern_bytes > 0 else limit
    bound = min(limit, by_bytes)
    step = (total + bound - 1) // bound
    return list(range(0, total, step))


def window_sensitivity(source, radius: float, shipped_scale: float = 1.2, wide_scale: float = 2.5,
                       sample_limit: int = 200, sample_budget_bytes: int = 256 << 20,
                       path: ComputePath = ComputePath.hip) -> Optional[float]:
    """OriginMeasurement.windowSensitivity (IteratedWindowOrigin.swift:254-290): the median, over a strided sample of
    scan positions, of the distance between the iterated-window origins at radiusScale = shipped_scale and at
    wide_scale (both with the estimator's defaults), in Float32 as the estimator computes it; None when the radius or
    a scale is not positive and finite, or no distance is finite. The sampled patterns are gathered into a one-row
    cube on the device and both measurements run there; the distances and their median are host arithmetic, as in
    the reference."""
    _require_hip(path)
    resident = isinstance(source, HIPResidentCube)
    arr = None if resident else np.asarray(source)
    if arr is not None and arr.ndim != 4:
        raise CuboError(f"invalid shape {list(arr.shape)}")
    shape = source.shape if resident else CubeShape.of(arr.shape)
    idx = sensitivity_sample_indices(shape.scan_count, sample_limit, shape.detector_count * 4, sample_budget_bytes)
    if not idx or not all(math.isfinite(v) and v > 0 for v in (radius, shipped_scale, wide_scale)):
        return None
    # readPattern of each sampled position (on the device for a resident cube), gathered into a 1 x n scan
    pats = source.read_patterns(idx) if resident else arr.reshape(shape.scan_count, shape.detector_rows,
                                                                   shape.detector_columns)[idx]
    sample = HIPResidentCube(np.ascontiguousarray(pats).reshape(1, len(idx), shape.detector_rows,
                                                                shape.detector_columns))
    try:
        shipped = OriginMeasurement.measure_with(sample, radius, shipped_scale)
        wide = OriginMeasurement.measure_with(sample, radius, wide_scale)
    finally:
        sample.release()
    f = np.float32
    dr = wide.row.astype(f) - shipped.row.astype(f)
    dc = wide.column.astype(f) - shipped.column.astype(f)
    d = np.sqrt(dr * dr + dc * dc, dtype=f)
    d = np.sort(d[np.isfinite(d)])
    if d.size == 0:
        return None
    m = d.size // 2
    return float(d[m] if d.size % 2 else (d[m - 1] + d[m]) / f(2))


def gaussian_filter(cube: HIPResidentCube, row: int, column: int, sigma: float) -> np.ndarray:
    """GaussianFilter(sigma).apply of one resident pattern, on the device."""
    out = np.empty(cube.shape.detector_count)
    _check(L.lib().cubo_hip_gaussian_filter_pattern(cube.handle, row, column, float(sigma), out.ctypes.data), "filter")
    return out.reshape(cube.shape.detector_rows, cube.shape.detector_columns)


__all__ = ["OriginEstimator", "OriginMeasurement", "ProbeEstimate", "ProbeSize", "gaussian_filter",
           "sensitivity_sample_indices", "window_sensitivity"]

OriginMeasurement.window_sensitivity = staticmethod(window_sensitivity)
