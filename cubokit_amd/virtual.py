"""Host-side mirror of CuboKit's CuboVirtual API for the `.hip` compute path.

Names, argument meanings and refusals follow the Swift reference:
- CubeShape and ScanTiling: Sources/CuboCore/CubeShape.swift:11-58, Sources/CuboCore/CubeSource.swift:85-112;
- DetectorGeometry and DetectorMask.make: Sources/CuboVirtual/DetectorGeometry.swift:23-98;
- ComputePath, ScanImage, DetectorImage, CuboVirtualError: Sources/CuboVirtual/Results.swift:13-97;
- VirtualImaging.image, VirtualDiffraction.mean/max/meanAndMax: Sources/CuboVirtual/VirtualImaging.swift:20-66;
- CenterOfMass.measure and CenterOfMassMaps: Sources/CuboVirtual/CenterOfMass.swift:14-47;
- ResidentCube (here HIPResidentCube): Sources/CuboVirtual/ResidentCube.swift:31-128;
- progress / isCancelled: Sources/CuboVirtual/Interruption.swift:10-25.

Every reduction runs on the GPU through the C ABI (include/cubo_hip.h); this module only prepares masks,
walks the scan tiles in order (asking `is_cancelled` before each, reporting `progress` after each) and wraps
the Double results.
"""
from __future__ import annotations

import ctypes as ct
import enum
import math
import os
from dataclasses import dataclass
from typing import Callable, Optional, Sequence, Union

import numpy as np

from . import _lib as L

# ------------------------------------------------------------------------------------------------
# Errors
# ------------------------------------------------------------------------------------------------


class CuboError(Exception):
    """CuboCore errors (Sources/CuboCore/CuboError.swift); `case` names the reference's case (cannotOpen, notNPY,
    fortranOrderNotSupported, bigEndianNotSupported, unsupportedElementType, wrongRank, fileTooShort,
    misalignedOffset, invalidShape) where the library reports one."""

    def __init__(self, message: str = "", case: str = ""):
        super().__init__(message)
        self.case = case


class CuboVirtualError(Exception):
    """CuboVirtualError (Results.swift:65-97) plus the `.hip` cases that replace `.noMetalDevice` / `.metal`."""

    def __init__(self, case: str, detail: str = ""):
        super().__init__(f"{case}: {detail}" if detail else case)
        self.case = case
        self.detail = detail


class CancellationError(Exception):
    """Thrown when `is_cancelled` answers True before a tile (Interruption.swift:15-18)."""


class ResidentCubeError(Exception):
    """ResidentCubeError.unavailable(reason) (ResidentCube.swift:14-21)."""

    def __init__(self, reason: "UnavailableReason"):
        super().__init__(reason.message)
        self.reason = reason


def _check(status: int, what: str) -> None:
    if status == L.OK:
        return
    msg = L.last_error()
    if status == L.ERR_NO_DEVICE:
        raise CuboVirtualError("noHIPDevice", msg or what)
    if status == L.ERR_INVALID_ARGUMENT:
        raise CuboVirtualError("invalidInput", f"{what}: {msg}")
    if status == L.ERR_CANCELLED:
        raise CancellationError(f"{what}: {msg}")
    if status == L.ERR_FILE:
        raise CuboError(f"{what}: {msg}", case=msg.split(":")[0].strip())
    raise CuboVirtualError("hip", f"{what}: {msg} (status {status})")


# ------------------------------------------------------------------------------------------------
# Shape and tiling
# ------------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class CubeShape:
    scan_rows: int
    scan_columns: int
    detector_rows: int
    detector_columns: int

    def __post_init__(self):
        ext = self.extents
        if any(int(e) <= 0 for e in ext) or math.prod(ext) * 8 >= 2 ** 63:
            raise CuboError(f"invalid shape {ext}")

    @staticmethod
    def of(extents: Sequence[int]) -> "CubeShape":
        if len(extents) != 4:
            raise CuboError(f"invalid shape {list(extents)}")
        return CubeShape(*[int(e) for e in extents])

    @property
    def extents(self):
        return [self.scan_rows, self.scan_columns, self.detector_rows, self.detector_columns]

    @property
    def scan_count(self):
        return self.scan_rows * self.scan_columns

    @property
    def detector_count(self):
        return self.detector_rows * self.detector_columns

    @property
    def element_count(self):
        return self.scan_count * self.detector_count

    @property
    def elements_per_scan_row(self):
        return self.scan_columns * self.detector_count


class ScanTiling:
    """`.automatic` (largest row count whose Float32 tile is at most 64 MiB) or `.rows(n)`."""

    automatic_tile_bytes = 64 * 1024 * 1024

    def __init__(self, rows: Optional[int] = None):
        self._rows = rows

    @staticmethod
    def rows(n: int) -> "ScanTiling":
        return ScanTiling(n)

    def rows_per_tile(self, shape: CubeShape) -> int:
        if self._rows is None:
            row_bytes = shape.elements_per_scan_row * 4
            return max(1, min(shape.scan_rows, self.automatic_tile_bytes // max(1, row_bytes)))
        if self._rows <= 0:
            raise CuboError(f"invalid tiling {self._rows}")
        return min(self._rows, shape.scan_rows)

    def ranges(self, shape: CubeShape):
        step = self.rows_per_tile(shape)
        return [range(r, min(shape.scan_rows, r + step)) for r in range(0, shape.scan_rows, step)]

    def __repr__(self):
        return "ScanTiling.automatic" if self._rows is None else f"ScanTiling.rows({self._rows})"


ScanTiling.automatic = ScanTiling()

# ------------------------------------------------------------------------------------------------
# Results
# ------------------------------------------------------------------------------------------------


class ComputePath(str, enum.Enum):
    """Where a reduction ran. `.cpu` and `.metal` are CuboKit's Swift paths; this package is the `.hip` path."""
    cpu = "cpu"
    metal = "metal"
    hip = "hip"


@dataclass
class ScanImage:
    scan_rows: int
    scan_columns: int
    values: np.ndarray
    path: ComputePath

    def __post_init__(self):
        if self.values.size != self.scan_rows * self.scan_columns:
            raise ValueError("ScanImage: wrong number of values")

    def __getitem__(self, rc):
        return self.values[rc[0] * self.scan_columns + rc[1]]


@dataclass
class DetectorImage:
    detector_rows: int
    detector_columns: int
    values: np.ndarray
    path: ComputePath

    def __getitem__(self, rc):
        return self.values[rc[0] * self.detector_columns + rc[1]]


@dataclass
class CenterOfMassMaps:
    row: ScanImage
    column: ScanImage
    intensity: ScanImage
    moment_row: ScanImage
    moment_column: ScanImage


# ------------------------------------------------------------------------------------------------
# Detector geometry (py4DSTEM make_detector, evaluated in Double)
# ------------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class DetectorGeometry:
    kind: str
    params: tuple

    @staticmethod
    def point(row: int, column: int):
        return DetectorGeometry("point", (int(row), int(column)))

    @staticmethod
    def circle(center_row: float, center_column: float, radius: float):
        return DetectorGeometry("circle", (float(center_row), float(center_column), float(radius)))

    @staticmethod
    def annulus(center_row: float, center_column: float, inner: float, outer: float):
        return DetectorGeometry("annulus", (float(center_row), float(center_column), float(inner), float(outer)))

    @staticmethod
    def rectangle(row_min: float, row_max: float, column_min: float, column_max: float):
        return DetectorGeometry("rectangle", (float(row_min), float(row_max), float(column_min), float(column_max)))

    @staticmethod
    def mask(weights):
        return DetectorGeometry("mask", (np.asarray(weights, dtype=np.float32).reshape(-1),))


class DetectorMask:
    @staticmethod
    def make(geometry: DetectorGeometry, shape: CubeShape = None, detector_rows: int = None,
             detector_columns: int = None) -> np.ndarray:
        """The Float32 mask of `geometry`, row-major (DetectorGeometry.swift:43-98)."""
        qy = shape.detector_rows if shape is not None else detector_rows
        qx = shape.detector_columns if shape is not None else detector_columns
        if qy is None or qx is None or qy <= 0 or qx <= 0:
            raise CuboError(f"invalid shape {[qy, qx]}")
        m = np.zeros(qy * qx, np.float32)
        k, p = geometry.kind, geometry.params
        if k == "point":
            r, c = p
            if not (0 <= r < qy and 0 <= c < qx):
                raise CuboVirtualError("pointOutsideDetector", f"({r}, {c})")
            m[r * qx + c] = 1
        elif k in ("circle", "annulus"):
            if not all(math.isfinite(v) for v in p):
                raise CuboVirtualError("nonFiniteGeometry")
            cr, cc = p[0], p[1]
            di = np.arange(qy, dtype=np.float64) - cr
            dj = np.arange(qx, dtype=np.float64) - cc
            d2 = (di * di)[:, None] + (dj * dj)[None, :]
            if k == "circle":
                if p[2] < 0:
                    raise CuboVirtualError("negativeRadius", str(p[2]))
                sel = d2 < p[2] * p[2]
            else:
                inner, outer = p[2], p[3]
                if inner < 0:
                    raise CuboVirtualError("negativeRadius", str(inner))
                if not outer > inner:
                    raise CuboVirtualError("innerNotBelowOuter", f"{inner} {outer}")
                sel = (d2 > inner * inner) & (d2 < outer * outer)
            m[sel.reshape(-1)] = 1
        elif k == "rectangle":
            if not all(math.isfinite(v) for v in p):
                raise CuboVirtualError("nonFiniteGeometry")
            rb = [float(np.round(v)) for v in p]          # numpy round: half to even
            if any(v < 0 for v in rb):
                raise CuboVirtualError("negativeRectangleBound")
            r0, r1 = int(min(rb[0], qy)), int(min(rb[1], qy))
            c0, c1 = int(min(rb[2], qx)), int(min(rb[3], qx))
            if r0 < r1 and c0 < c1:
                m.reshape(qy, qx)[r0:r1, c0:c1] = 1
        elif k == "mask":
            w = p[0]
            if w.size != qy * qx:
                raise CuboVirtualError("maskLengthMismatch", f"expected {qy * qx}, received {w.size}")
            if not np.all(np.isfinite(w)):
                raise CuboVirtualError("nonFiniteGeometry")
            m = w.astype(np.float32).copy()
        else:
            raise ValueError(k)
        return m


# ------------------------------------------------------------------------------------------------
# Devices and the resident cube
# ------------------------------------------------------------------------------------------------


def device_count() -> int:
    n = ct.c_int(0)
    _check(L.lib().cubo_hip_device_count(ct.byref(n)), "device count")
    return n.value


def hip_is_available() -> bool:
    """Whether this machine can run the `.hip` path (VirtualImaging.swift:69-72's metalIsAvailable)."""
    try:
        return device_count() > 0
    except (L.LibraryMissing, OSError):
        return False


def device_info(device: int = 0) -> L.DeviceInfo:
    info = L.DeviceInfo()
    _check(L.lib().cubo_hip_get_device_info(device, ct.byref(info)), "device info")
    return info


@dataclass(frozen=True)
class UnavailableReason:
    kind: str                  # "noHIPDevice" | "exceedsMemoryCap" | "exceedsBufferLimit"
    bytes: int = 0
    cap: int = 0

    @property
    def message(self) -> str:
        if self.kind == "noHIPDevice":
            return "This machine has no HIP device, so the cube cannot be held resident."
        if self.kind == "exceedsMemoryCap":
            return (f"The cube is {self.bytes / 1e9:.1f} GB in device memory; a resident cube may use at most "
                    f"{self.cap / 1e9:.1f} GB of HBM.")
        return f"The cube is {self.bytes / 1e9:.1f} GB; that much HBM is not free."


_DTYPES = {np.dtype(np.uint16): L.U16, np.dtype(np.float32): L.F32}


def native_dtype(dtype) -> np.dtype:
    """uint16 stays native; every other element type is held as Float32, the reference's read type."""
    return np.dtype(np.uint16) if np.dtype(dtype) == np.uint16 else np.dtype(np.float32)


class HIPResidentCube:
    """A cube held in one HBM buffer in its native element type (uint16) or Float32."""

    memory_cap_fraction = 0.9

    def __init__(self, loading=None, *, shape: CubeShape = None, dtype=np.uint16, tiling: ScanTiling = None,
                 device: int = 0):
        if loading is not None:
            arr = np.asarray(loading)
            if arr.ndim != 4:
                raise CuboError(f"invalid shape {list(arr.shape)}")
            shape = CubeShape.of(arr.shape)
            dtype = native_dtype(arr.dtype)
        if shape is None:
            raise CuboError("a shape or an array to load is needed")
        self.shape = shape
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise CuboError(f"unsupported resident dtype {self.dtype}")
        self.device = device
        avail = HIPResidentCube.availability(shape, self.dtype, device=device)
        if avail is not None and avail.kind != "noHIPDevice":
            raise ResidentCubeError(avail)
        h = ct.c_void_p()
        ext = (ct.c_int64 * 4)(*shape.extents)
        _check(L.lib().cubo_hip_cube_create(device, ext, _DTYPES[self.dtype], ct.byref(h)), "allocating the cube")
        self._h = h
        if loading is not None:
            # through the pinned double-buffered feed; other element types are converted on the device
            for rg in (tiling or ScanTiling.automatic).ranges(shape):
                _feed(self._h, arr, rg)

    @staticmethod
    def _adopt(h, shape: "CubeShape", dtype, device: int) -> "HIPResidentCube":
        c = HIPResidentCube.__new__(HIPResidentCube)
        c.shape, c.dtype, c.device, c._h = shape, np.dtype(dtype), device, h
        return c

    @staticmethod
    def load_npy(path: str, device: int = 0) -> "HIPResidentCube":
        """MappedCube(npy:) loaded into HBM (NPYFile.swift:28-71, MappedCube.swift:130-136): uint16 stays uint16,
        every other element type becomes Float32 on the device."""
        hd = NPYFile.header(path)
        if len(hd.shape) != 4:
            raise CuboError(f"wrongRank: expected 4, received {len(hd.shape)}", case="wrongRank")
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_cube_load_npy(device, os.fsencode(path), ct.byref(h)), f"loading {path}")
        return HIPResidentCube._adopt(h, CubeShape.of(hd.shape), native_dtype(hd.element_type), device)

    @staticmethod
    def load_file(path: str, data_offset: int, element_type: str, shape: "CubeShape", device: int = 0):
        """MappedCube(url:dataOffset:elementType:shape:) loaded into HBM (MappedCube.swift:98-127)."""
        ext = (ct.c_int64 * 4)(*shape.extents)
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_cube_load_file(device, os.fsencode(path), int(data_offset),
                                               L.ELEMENT_TYPES.index(element_type), ext, ct.byref(h)),
               f"loading {path}")
        return HIPResidentCube._adopt(h, shape, native_dtype(element_type), device)

    @staticmethod
    def availability(shape: CubeShape, dtype=np.uint16, hbm_bytes: int = 0, device: int = 0):
        """None when available, else the UnavailableReason (checks the cap, then the device, then the buffer)."""
        reason, b, cap = ct.c_int(), ct.c_uint64(), ct.c_uint64()
        ext = (ct.c_int64 * 4)(*shape.extents)
        _check(L.lib().cubo_hip_availability(device, ext, _DTYPES[np.dtype(dtype)], hbm_bytes, ct.byref(reason),
                                             ct.byref(b), ct.byref(cap)), "availability")
        kind = {L.AVAILABLE: None, L.UNAVAILABLE_NO_DEVICE: "noHIPDevice",
                L.UNAVAILABLE_EXCEEDS_MEMORY_CAP: "exceedsMemoryCap",
                L.UNAVAILABLE_EXCEEDS_BUFFER_LIMIT: "exceedsBufferLimit"}[reason.value]
        return None if kind is None else UnavailableReason(kind, b.value, cap.value)

    @staticmethod
    def fits(shape: CubeShape, dtype=np.uint16) -> bool:
        return HIPResidentCube.availability(shape, dtype) is None

    @property
    def handle(self):
        if self._h is None:
            raise CuboVirtualError("hip", "the cube has been released")
        return self._h

    @property
    def byte_count(self) -> int:
        b = ct.c_uint64()
        _check(L.lib().cubo_hip_cube_byte_count(self.handle, ct.byref(b)), "byte count")
        return b.value

    def read_scan_tile(self, rows: range) -> np.ndarray:
        out = np.empty((len(rows), self.shape.scan_columns, self.shape.detector_rows, self.shape.detector_columns),
                       self.dtype)
        _check(L.lib().cubo_hip_cube_read_rows(self.handle, rows.start, len(rows), out.ctypes.data_as(ct.c_void_p)),
               "reading rows")
        return out

    def read_pattern(self, row: int, column: int) -> np.ndarray:
        if not (0 <= row < self.shape.scan_rows and 0 <= column < self.shape.scan_columns):
            raise CuboError(f"position ({row}, {column}) outside the scan")
        return self.read_patterns([row * self.shape.scan_columns + column])[0]

    def read_patterns(self, indices) -> np.ndarray:
        """readPattern for several scan positions (row * scanColumns + column): (n, qy, qx) in the cube's dtype."""
        idx = np.ascontiguousarray(indices, np.int64).reshape(-1)
        out = np.empty((idx.size, self.shape.detector_rows, self.shape.detector_columns), self.dtype)
        _check(L.lib().cubo_hip_cube_read_patterns(self.handle, idx.ctypes.data, idx.size, out.ctypes.data),
               "reading patterns")
        return out

    def clone(self) -> "HIPResidentCube":
        """A second resident cube with the same contents, copied on the device."""
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_cube_clone(self.handle, ct.byref(h)), "cloning the cube")
        c = HIPResidentCube.__new__(HIPResidentCube)
        c.shape, c.dtype, c.device, c._h = self.shape, self.dtype, self.device, h
        return c

    def fill_synthetic(self, params: L.SynthParams) -> None:
        _check(L.lib().cubo_hip_cube_fill_synthetic(self.handle, ct.byref(params)), "synthesising the cube")

    def kernel_time(self):
        """(total ms, launches) of the fused streaming kernel since the last query (HIP events)."""
        t, n = ct.c_double(), ct.c_int64()
        _check(L.lib().cubo_hip_cube_kernel_time(self.handle, ct.byref(t), ct.byref(n)), "kernel time")
        return t.value, n.value

    def release(self):
        if getattr(self, "_h", None) is not None:
            L.lib().cubo_hip_cube_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()


Source = Union[HIPResidentCube, np.ndarray]


@dataclass(frozen=True)
class NPYHeader:
    """What an .npy header says (NPYFile.swift:11-22)."""
    major_version: int
    minor_version: int
    element_type: str
    shape: list
    data_offset: int


class NPYFile:
    @staticmethod
    def header(path: str) -> NPYHeader:
        """NPYFile.header(of:) (NPYFile.swift:28-71), parsed by the library."""
        h = L.NpyHeader()
        _check(L.lib().cubo_hip_npy_read_header(os.fsencode(path), ct.byref(h)), f"reading the header of {path}")
        return NPYHeader(h.major_version, h.minor_version, L.ELEMENT_TYPES[h.element_type],
                         [int(h.shape[k]) for k in range(h.rank)], int(h.data_offset))


def _element_type(dtype: np.dtype) -> Optional[int]:
    """The stored element type of a native little-endian NumPy dtype, or None."""
    dt = np.dtype(dtype)
    if dt.byteorder == ">" or dt.name not in L.ELEMENT_TYPES:
        return None
    return L.ELEMENT_TYPES.index(dt.name)


def _feed(h, arr: np.ndarray, rg: range) -> None:
    """Scan rows `rg` of `arr` to the cube `h` through the library's pinned feed."""
    t = _element_type(arr.dtype)
    tile = arr[rg.start:rg.stop]
    if t is None:                           # float16, bool, big-endian: Float32 on the host first
        tile, t = np.asarray(tile, np.float32), L.ELEMENT_TYPES.index("float32")
    tile = np.ascontiguousarray(tile)
    _check(L.lib().cubo_hip_cube_feed_rows(h, rg.start, len(rg), tile.ctypes.data_as(ct.c_void_p), t), "feeding rows")


class HIPStreamedCube:
    """A cube fed tile by tile from host memory (a NumPy array or np.memmap of a file) into a two-slot device ring:
    the staging case of the reference's TileFeed (MetalReducers.swift:494-573), with the copies double-buffered
    through pinned memory and overlapped with the reductions."""

    def __init__(self, source: np.ndarray, tiling: ScanTiling, device: int = 0):
        self.source = source
        self.shape = CubeShape.of(source.shape)
        self.dtype = native_dtype(source.dtype)
        self.device = device
        self.tile_rows = tiling.rows_per_tile(self.shape)
        ext = (ct.c_int64 * 4)(*self.shape.extents)
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_cube_create_streamed(device, ext, _DTYPES[self.dtype], self.tile_rows, ct.byref(h)),
               "creating the streamed cube")
        self._h = h

    @property
    def handle(self):
        return self._h

    def feed(self, rg: range) -> None:
        _feed(self._h, self.source, rg)

    def release(self):
        if getattr(self, "_h", None) is not None:
            L.lib().cubo_hip_cube_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _resident(source: Source, tiling: ScanTiling):
    """(cube, temporary, feed): a resident cube as it is, or a host array as a streamed cube whose tiles `feed`
    copies to the device before each tile's reduction (the TileFeed staging case, MetalReducers.swift:504-573)."""
    if isinstance(source, HIPResidentCube):
        return source, False, None
    arr = np.asarray(source)
    if arr.ndim != 4:
        raise CuboError(f"invalid shape {list(arr.shape)}")
    cube = HIPStreamedCube(arr, tiling)
    return cube, True, cube.feed


# ------------------------------------------------------------------------------------------------
# The reductions
# ------------------------------------------------------------------------------------------------

Progress = Optional[Callable[[float], None]]
IsCancelled = Optional[Callable[[], bool]]


def _require_hip(path: ComputePath):
    if ComputePath(path) != ComputePath.hip:
        raise CuboVirtualError("unsupportedPath", f"this package implements the .hip path, not .{ComputePath(path).value}")


def _reduce(source: Source, *, masks: Optional[np.ndarray], want_cols: bool, want_com: bool,
            com_weights: Optional[np.ndarray], tiling: ScanTiling, progress: Progress, is_cancelled: IsCancelled):
    tiling = tiling or ScanTiling.automatic
    if is_cancelled is not None and is_cancelled():
        raise CancellationError()
    cube, temporary, feed = _resident(source, tiling)
    try:
        shape = cube.shape
        n, npos = shape.detector_count, shape.scan_count
        m = None if masks is None else np.ascontiguousarray(masks, np.float32).reshape(-1, n)
        cw = None if com_weights is None else np.ascontiguousarray(com_weights, np.float32).reshape(n)
        req = L.ReduceRequest(masks=None if m is None else m.ctypes.data, n_masks=0 if m is None else m.shape[0],
                              want_mean_max=int(want_cols), want_com=int(want_com),
                              com_weights=None if cw is None else cw.ctypes.data)
        out = {}
        res = L.ReduceResult()
        if m is not None:
            out["images"] = np.empty((m.shape[0], npos))
            res.images = out["images"].ctypes.data
        if want_cols:
            out["mean"], out["max"] = np.empty(n), np.empty(n)
            res.mean, res.max = out["mean"].ctypes.data, out["max"].ctypes.data
        if want_com:
            for k in ("com_intensity", "com_moment_row", "com_moment_column", "com_row", "com_column"):
                out[k] = np.empty(npos)
                setattr(res, k, out[k].ctypes.data)
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_reduce_begin(cube.handle, ct.byref(req), ct.byref(h)), "starting the reduction")
        try:
            for rg in tiling.ranges(shape):
                if is_cancelled is not None and is_cancelled():
                    raise CancellationError()
                if feed is not None:
                    feed(rg)
                _check(L.lib().cubo_hip_reduce_rows(h, rg.start, len(rg)), "reducing rows")
                if progress is not None:
                    progress(rg.stop / shape.scan_rows)
        except BaseException:
            L.lib().cubo_hip_reduce_cancel(h)
            raise
        _check(L.lib().cubo_hip_reduce_end(h, ct.byref(res)), "finishing the reduction")
        return shape, out
    finally:
        if temporary:
            cube.release()


def reduce_path(source: HIPResidentCube, masks=None, want_com=False, com_weights=None) -> int:
    """1 = exact integer kernel, 2 = FP64 kernel, for this request on this cube."""
    n = source.shape.detector_count
    m = None if masks is None else np.ascontiguousarray(masks, np.float32).reshape(-1, n)
    cw = None if com_weights is None else np.ascontiguousarray(com_weights, np.float32)
    req = L.ReduceRequest(masks=None if m is None else m.ctypes.data, n_masks=0 if m is None else m.shape[0],
                          want_mean_max=1, want_com=int(want_com), com_weights=None if cw is None else cw.ctypes.data)
    p = ct.c_int()
    _check(L.lib().cubo_hip_reduce_path(source.handle, ct.byref(req), ct.byref(p)), "path")
    return p.value


def _shape_of(source: Source) -> CubeShape:
    return source.shape if isinstance(source, HIPResidentCube) else CubeShape.of(np.asarray(source).shape)


class VirtualImaging:
    @staticmethod
    def image(source: Source, geometry: DetectorGeometry, path: ComputePath = ComputePath.hip,
              tiling: ScanTiling = None, progress: Progress = None, is_cancelled: IsCancelled = None) -> ScanImage:
        """For every scan position, the sum over the detector of pattern x mask (VirtualImaging.swift:20-31)."""
        _require_hip(path)
        shape = _shape_of(source)
        mask = DetectorMask.make(geometry, shape)
        return VirtualImaging.images(source, mask[None], tiling=tiling, progress=progress,
                                     is_cancelled=is_cancelled)[0]

    @staticmethod
    def images(source: Source, masks: np.ndarray, tiling: ScanTiling = None, progress: Progress = None,
               is_cancelled: IsCancelled = None):
        """Several masks in one pass over the cube (at most 16 weight rows per pass)."""
        shape, out = _reduce(source, masks=masks, want_cols=False, want_com=False, com_weights=None, tiling=tiling,
                             progress=progress, is_cancelled=is_cancelled)
        return [ScanImage(shape.scan_rows, shape.scan_columns, v, ComputePath.hip) for v in out["images"]]


class MaskShift:
    """How the detector mask follows the beam from position to position (ShiftedImaging.swift:14-21)."""

    def __init__(self, kind: str, origin: Optional["OriginMap"] = None):
        self.kind, self.origin = kind, origin

    @staticmethod
    def none() -> "MaskShift":
        return MaskShift("none")

    @staticmethod
    def rounded(origin: "OriginMap") -> "MaskShift":
        return MaskShift("rounded", origin)

    @staticmethod
    def bilinear(origin: "OriginMap") -> "MaskShift":
        return MaskShift("bilinear", origin)


def _transform(g: DetectorGeometry, origin, pixel_size) -> DetectorGeometry:
    """get_calibrated_detector_geometry (ShiftedImaging.swift:77-97): the origin added, then every coordinate divided
    by the pixel size; a point truncates toward zero."""
    if origin is not None and pixel_size is not None:
        raise CuboVirtualError("centeredAndCalibrated")
    if pixel_size is not None and not (pixel_size > 0 and math.isfinite(pixel_size)):
        raise CuboVirtualError("invalidPixelSize", str(pixel_size))
    r0, c0 = (float(origin[0]), float(origin[1])) if origin is not None else (0.0, 0.0)
    sc = float(pixel_size) if pixel_size is not None else 1.0
    p = g.params
    if g.kind == "mask":
        return g
    if g.kind == "point":
        fr, fc = (p[0] + r0) / sc, (p[1] + c0) / sc
        if not (math.isfinite(fr) and math.isfinite(fc)) or abs(math.trunc(fr)) >= 1e15 or abs(math.trunc(fc)) >= 1e15:
            raise CuboVirtualError("nonFiniteGeometry")
        return DetectorGeometry.point(math.trunc(fr), math.trunc(fc))
    if g.kind == "circle":
        return DetectorGeometry.circle((p[0] + r0) / sc, (p[1] + c0) / sc, p[2] / sc)
    if g.kind == "annulus":
        return DetectorGeometry.annulus((p[0] + r0) / sc, (p[1] + c0) / sc, p[2] / sc, p[3] / sc)
    return DetectorGeometry.rectangle((p[0] + r0) / sc, (p[1] + r0) / sc, (p[2] + c0) / sc, (p[3] + c0) / sc)


def image_centered(source: Source, geometry: DetectorGeometry, centered_on=None, pixel_size: Optional[float] = None,
                   shift: Optional[MaskShift] = None, tiling: ScanTiling = None) -> ScanImage:
    """VirtualImaging.image(of:geometry:centeredOn:pixelSize:shift:) (ShiftedImaging.swift:34-75): the geometry
    centred on `centered_on` (normally the origin map's mean) or calibrated by `pixel_size`, then the mask fixed
    (.none) or moved per position by the rounded or bilinear shift of the origin map, on the matrix cores."""
    shape = _shape_of(source)
    g = _transform(geometry, centered_on, pixel_size)
    if shift is None or shift.kind == "none":
        return VirtualImaging.image(source, g, tiling=tiling)
    origin = shift.origin
    if origin.row.size != shape.scan_count:
        raise CuboVirtualError("originMapMismatch", f"the origin map has {origin.row.size} positions, the scan "
                               f"{shape.scan_count}")
    if not (np.all(np.isfinite(origin.row)) and np.all(np.isfinite(origin.column))):
        raise CuboVirtualError("originMapMismatch", "the origin map has non-finite values")
    mask = DetectorMask.make(g, shape)
    cube, temporary = (source, False) if isinstance(source, HIPResidentCube) else (HIPResidentCube(source), True)
    try:
        st = ShiftedStack(cube, mask[None])
        try:
            y = st.forward(origin.row, origin.column, rounded=shift.kind == "rounded")
        finally:
            st.release()
    finally:
        if temporary:
            cube.release()
    return ScanImage(shape.scan_rows, shape.scan_columns, y[:, 0].copy(), ComputePath.hip)


class VirtualDiffraction:
    @staticmethod
    def mean(source: Source, path: ComputePath = ComputePath.hip, tiling: ScanTiling = None,
             progress: Progress = None, is_cancelled: IsCancelled = None) -> DetectorImage:
        return VirtualDiffraction.mean_and_max(source, path, tiling, progress, is_cancelled)[0]

    @staticmethod
    def max(source: Source, path: ComputePath = ComputePath.hip, tiling: ScanTiling = None,
            progress: Progress = None, is_cancelled: IsCancelled = None) -> DetectorImage:
        return VirtualDiffraction.mean_and_max(source, path, tiling, progress, is_cancelled)[1]

    @staticmethod
    def mean_and_max(source: Source, path: ComputePath = ComputePath.hip, tiling: ScanTiling = None,
                     progress: Progress = None, is_cancelled: IsCancelled = None):
        """Both patterns in one pass over the cube (VirtualImaging.swift:56-66)."""
        _require_hip(path)
        shape, out = _reduce(source, masks=None, want_cols=True, want_com=False, com_weights=None, tiling=tiling,
                             progress=progress, is_cancelled=is_cancelled)
        qy, qx = shape.detector_rows, shape.detector_columns
        return (DetectorImage(qy, qx, out["mean"], ComputePath.hip), DetectorImage(qy, qx, out["max"], ComputePath.hip))

    @staticmethod
    def hip_is_available() -> bool:
        return hip_is_available()

    @staticmethod
    def pattern(source: Source, statistic: "DiffractionStatistic", scan: "ScanSelection" = None,
                shift: "PatternShift" = None, path: ComputePath = ComputePath.hip, tiling: ScanTiling = None,
                median_working_bytes: int = 256 << 20, progress: Progress = None,
                is_cancelled: IsCancelled = None) -> DetectorImage:
        """A virtual diffraction pattern over a selection of positions, optionally centre-shifted
        (SelectedDiffraction.swift:46-108). The `.hip` path runs the mean and the max over `.all`, `.mask` and
        `.weights`, each pattern optionally moved by PatternShift.rounded or .fourier; the median is refused, as the
        reference's `.metal` path refuses it."""
        _require_hip(path)
        statistic = DiffractionStatistic(statistic)
        scan = scan if scan is not None else ScanSelection.all()
        if median_working_bytes <= 0:
            raise CuboVirtualError("invalidInput", "the median's working memory must be positive")
        shape = _shape_of(source)
        take, weights = scan.arrays(shape)
        if statistic == DiffractionStatistic.median:
            raise CuboVirtualError("invalidInput", "the .hip path covers the mean and the max over a selection")
        if scan.kind == "all" and shift is None:
            # VirtualDiffraction.mean / max (SelectedDiffraction.swift:89-92)
            mm = VirtualDiffraction.mean_and_max(source, path, tiling, progress, is_cancelled)
            return mm[0] if statistic == DiffractionStatistic.mean else mm[1]
        return _pattern(source, statistic, take, weights, shift, L.TIES_FIRST, tiling, progress, is_cancelled)

    @staticmethod
    def mean_and_max_shifted(source: Source, shift: "PatternShift", tiling: ScanTiling = None,
                             progress: Progress = None, is_cancelled: IsCancelled = None):
        """meanAndMax(of:shift:) (ShiftedImaging.swift:125-167): every pattern moved by the map's mean minus its
        origin (rounded, or through the Fourier shift theorem); the maximum keeps the later of equal values. Two
        passes, one per statistic."""
        half = None if progress is None else (lambda f: progress(0.5 * f))
        mean = _pattern(source, DiffractionStatistic.mean, None, None, shift, L.TIES_LATER, tiling, half, is_cancelled)
        rest = None if progress is None else (lambda f: progress(0.5 + 0.5 * f))
        mx = _pattern(source, DiffractionStatistic.max, None, None, shift, L.TIES_LATER, tiling, rest, is_cancelled)
        return mean, mx


class DiffractionStatistic(str, enum.Enum):
    """How the patterns of a selection are combined (SelectedDiffraction.swift:14-17)."""
    mean = "mean"
    max = "max"
    median = "median"


class ScanSelection:
    """Which scan positions a virtual diffraction pattern combines (SelectedDiffraction.swift:19-27)."""

    def __init__(self, kind: str, values=None):
        self.kind = kind
        self.values = values

    @staticmethod
    def all() -> "ScanSelection":
        return ScanSelection("all")

    @staticmethod
    def mask(flags) -> "ScanSelection":
        return ScanSelection("mask", np.asarray(flags, bool).reshape(-1))

    @staticmethod
    def weights(values) -> "ScanSelection":
        return ScanSelection("weights", np.asarray(values, np.float64).reshape(-1))

    def arrays(self, shape: CubeShape):
        """(take, weights) for the C ABI, with the reference's refusals (SelectedDiffraction.swift:64-79)."""
        if self.kind == "all":
            return None, None
        if self.values.size != shape.scan_count:
            raise CuboVirtualError("maskLengthMismatch", f"expected {shape.scan_count}, received {self.values.size}")
        if self.kind == "mask":
            if not self.values.any():
                raise CuboVirtualError("invalidInput", "the selection holds no scan position")
            return np.ascontiguousarray(self.values.astype(np.uint8)), None
        if not np.all(np.isfinite(self.values)):
            raise CuboVirtualError("invalidInput", "the weights must be finite")
        return None, np.ascontiguousarray(self.values)


class OriginMap:
    """An origin per scan position (OriginMeasurement.swift:14-40): row and column maps (flattened in scan order) and
    their NumPy means."""

    def __init__(self, row, column):
        r, c = np.asarray(getattr(row, "values", row), np.float64), np.asarray(getattr(column, "values", column), np.float64)
        if r.shape != c.shape:
            raise CuboVirtualError("originMapMismatch", "the row and column maps differ in shape")
        self.row, self.column = np.ascontiguousarray(r.reshape(-1)), np.ascontiguousarray(c.reshape(-1))
        self.scan_shape = (getattr(row, "scan_rows", None), getattr(row, "scan_columns", None))

    @property
    def mean_row(self) -> float:
        return float(np.mean(self.row))          # NumPy's pairwise mean (NumPySum.mean)

    @property
    def mean_column(self) -> float:
        return float(np.mean(self.column))

    def replacing_nan_with_mean(self) -> "OriginMap":
        """py4DSTEM's remove_NaN (OriginMeasurement.swift:30-39)."""
        def filled(m):
            keep = m[~np.isnan(m)]
            fill = float(np.mean(keep)) if keep.size else np.nan
            return np.where(np.isnan(m), fill, m)
        return OriginMap(filled(self.row), filled(self.column))


class PatternShift:
    """How each pattern is moved before the mean and maximum (ShiftedImaging.swift:23-29)."""

    def __init__(self, origin: OriginMap, fourier: bool):
        self.origin, self.fourier = origin, fourier

    @staticmethod
    def rounded(origin: OriginMap) -> "PatternShift":
        return PatternShift(origin, False)

    @staticmethod
    def fourier_shift(origin: OriginMap) -> "PatternShift":
        return PatternShift(origin, True)


def _pattern(source: Source, statistic: DiffractionStatistic, take, weights, shift: Optional[PatternShift], ties: int,
             tiling: ScanTiling, progress: Progress, is_cancelled: IsCancelled) -> DetectorImage:
    tiling = tiling or ScanTiling.automatic
    if is_cancelled is not None and is_cancelled():
        raise CancellationError()
    cube, temporary, feed = _resident(source, tiling)
    try:
        shape = cube.shape
        if shift is not None and shift.origin.row.size != shape.scan_count:
            raise CuboVirtualError("originMapMismatch", f"the origin map has {shift.origin.row.size} positions, "
                                   f"the scan {shape.scan_count}")
        req = L.PatternRequest(statistic=L.PATTERN_MEAN if statistic == DiffractionStatistic.mean else L.PATTERN_MAX,
                               take=None if take is None else take.ctypes.data,
                               weights=None if weights is None else weights.ctypes.data,
                               origin_row=None if shift is None else shift.origin.row.ctypes.data,
                               origin_col=None if shift is None else shift.origin.column.ctypes.data,
                               max_ties=ties, shift_kind=L.SHIFT_FOURIER if shift is not None and shift.fourier
                               else L.SHIFT_ROUNDED)
        out = np.empty(shape.detector_count)
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_pattern_begin(cube.handle, ct.byref(req), ct.byref(h)), "starting the pattern")
        try:
            for rg in tiling.ranges(shape):
                if is_cancelled is not None and is_cancelled():
                    raise CancellationError()
                if feed is not None:
                    feed(rg)
                _check(L.lib().cubo_hip_pattern_rows(h, rg.start, len(rg)), "pattern rows")
                if progress is not None:
                    progress(rg.stop / shape.scan_rows)
        except BaseException:
            L.lib().cubo_hip_pattern_cancel(h)
            raise
        _check(L.lib().cubo_hip_pattern_end(h, out.ctypes.data), "finishing the pattern")
        return DetectorImage(shape.detector_rows, shape.detector_columns, out, ComputePath.hip)
    finally:
        if temporary:
            cube.release()


def pattern_path(source: "HIPResidentCube", take=None, weights=None) -> int:
    """1 = exact integer kernel, 2 = FP64 kernel, for a selected pattern on this cube."""
    t = None if take is None else np.ascontiguousarray(np.asarray(take, bool).astype(np.uint8))
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    req = L.PatternRequest(statistic=L.PATTERN_MEAN, take=None if t is None else t.ctypes.data,
                           weights=None if w is None else w.ctypes.data, origin_row=None, origin_col=None,
                           max_ties=L.TIES_FIRST)
    p = ct.c_int()
    _check(L.lib().cubo_hip_pattern_path(source.handle, ct.byref(req), ct.byref(p)), "path")
    return p.value


class CenterOfMass:
    @staticmethod
    def measure(source: Source, mask: Optional[DetectorGeometry] = None, path: ComputePath = ComputePath.hip,
                tiling: ScanTiling = None, progress: Progress = None,
                is_cancelled: IsCancelled = None) -> CenterOfMassMaps:
        """The centre of mass of every pattern, weighted by `mask` (CenterOfMass.swift:27-47)."""
        _require_hip(path)
        shape = _shape_of(source)
        w = None if mask is None else DetectorMask.make(mask, shape)
        shape, out = _reduce(source, masks=None, want_cols=False, want_com=True, com_weights=w, tiling=tiling,
                             progress=progress, is_cancelled=is_cancelled)
        return _com_maps(shape, out)


def _com_maps(shape: CubeShape, out: dict) -> CenterOfMassMaps:
    def img(v):
        return ScanImage(shape.scan_rows, shape.scan_columns, v, ComputePath.hip)
    return CenterOfMassMaps(row=img(out["com_row"]), column=img(out["com_column"]), intensity=img(out["com_intensity"]),
                            moment_row=img(out["com_moment_row"]), moment_column=img(out["com_moment_column"]))


@dataclass
class FirstLook:
    """Everything one fused pass yields: virtual images, mean and max pattern, centre of mass."""
    images: list
    mean: DetectorImage
    max: DetectorImage
    com: Optional[CenterOfMassMaps]


def first_look(source: Source, geometries: Sequence[DetectorGeometry] = (), com: bool = True,
               com_mask: Optional[DetectorGeometry] = None, tiling: ScanTiling = None, progress: Progress = None,
               is_cancelled: IsCancelled = None) -> FirstLook:
    """Virtual images of `geometries`, the mean and max pattern and the centre of mass in ONE pass."""
    shape = _shape_of(source)
    masks = np.stack([DetectorMask.make(g, shape) for g in geometries]) if geometries else None
    w = None if com_mask is None else DetectorMask.make(com_mask, shape)
    shape, out = _reduce(source, masks=masks, want_cols=True, want_com=com, com_weights=w, tiling=tiling,
                         progress=progress, is_cancelled=is_cancelled)
    qy, qx = shape.detector_rows, shape.detector_columns
    images = [ScanImage(shape.scan_rows, shape.scan_columns, v, ComputePath.hip) for v in out.get("images", [])]
    return FirstLook(images, DetectorImage(qy, qx, out["mean"], ComputePath.hip),
                     DetectorImage(qy, qx, out["max"], ComputePath.hip), _com_maps(shape, out) if com else None)


class MaskStack:
    """Y = X W for a stack of weight rows, in one pass (include/cubo_hip.h, "mask stack")."""

    AUTO, PLAIN, MFMA_SPLIT = L.STACK_AUTO, L.STACK_PLAIN, L.STACK_MFMA_SPLIT

    def __init__(self, cube: HIPResidentCube, masks: np.ndarray, method: int = L.STACK_AUTO):
        self.cube = cube
        self.masks = np.ascontiguousarray(masks, np.float32).reshape(-1, cube.shape.detector_count)
        self.k = self.masks.shape[0]
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_stack_plan_create(cube.handle, self.masks.ctypes.data, self.k, method, ct.byref(h)),
               "preparing the mask stack")
        self._h = h
        m = ct.c_int32()
        _check(L.lib().cubo_hip_stack_plan_method(h, ct.byref(m)), "stack method")
        self.method = m.value

    def run(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """(scanCount, K) Double."""
        y = out if out is not None else np.empty((self.cube.shape.scan_count, self.k))
        _check(L.lib().cubo_hip_stack_plan_run(self._h, y.ctypes.data), "running the mask stack")
        return y

    def release(self):
        if getattr(self, "_h", None) is not None:
            L.lib().cubo_hip_stack_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class StackBatch:
    """Mask stacks over many resident uint16 cubes in one persistent launch (include/cubo_hip.h, "batches").

    items: (HIPResidentCube, masks (k, detectorCount)) pairs, k in 1..64, all cubes on one device. The
    launch is captured once in a hipGraph; `run` replays it. `tune` times the launch shapes on the device
    and keeps the fastest for every batch of the same shapes in this process."""

    def __init__(self, items: Sequence[tuple]):
        if not items:
            raise CuboError("an empty batch")
        self.cubes = [c for c, _ in items]
        self.masks = [np.ascontiguousarray(m, np.float32).reshape(-1, c.shape.detector_count) for c, m in items]
        self.k = [m.shape[0] for m in self.masks]
        n = len(items)
        cubes = (ct.c_void_p * n)(*[c.handle.value if isinstance(c.handle, ct.c_void_p) else c.handle
                                     for c in self.cubes])
        masks = (ct.c_void_p * n)(*[m.ctypes.data for m in self.masks])
        ks = (ct.c_int32 * n)(*self.k)
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_batch_create(n, cubes, masks, ks, ct.byref(h)), "preparing the batch")
        self._h = h
        self.shape = None

    @staticmethod
    def availability(shapes: Sequence, dtype=np.uint16, hbm_bytes: int = 0, device: int = 0):
        """ResidentCube.availability for the cubes of a batch held together: None when their summed bytes fit the
        memory cap and the free HBM, else the UnavailableReason (of the summed bytes). shapes: CubeShapes or
        4-tuples."""
        total = sum(int(np.prod(getattr(s, "extents", s), dtype=np.int64)) for s in shapes)
        return HIPResidentCube.availability(CubeShape(1, 1, 1, max(1, total)), dtype, hbm_bytes, device)

    def tune(self) -> dict:
        w, b, o = ct.c_int32(), ct.c_int32(), ct.c_int32()
        _check(L.lib().cubo_hip_batch_tune(self._h, ct.byref(w), ct.byref(b), ct.byref(o)), "tuning the batch")
        self.shape = {"waves": w.value, "blocks_per_cu": b.value,
                      "order": ("given", "largest detector first", "interleaved")[o.value]}
        return self.shape

    def run(self, out: Optional[Sequence[np.ndarray]] = None, copy_out: bool = True, view: bool = False) -> list:
        """[(scanCount_i, k_i) Double] per item; copy_out=False replays the graph without the host copy.
        view=True returns read-only arrays over the batch's pinned result buffer (no copy); they keep the batch
        alive, and the next run overwrites them."""
        if view or not copy_out:
            _check(L.lib().cubo_hip_batch_run(self._h, None), "running the batch")
            return self.view() if view else []
        ys = list(out) if out is not None else [np.empty((c.shape.scan_count, k)) for c, k in zip(self.cubes, self.k)]
        ptrs = (ct.c_void_p * len(ys))(*[y.ctypes.data for y in ys])
        _check(L.lib().cubo_hip_batch_run(self._h, ptrs), "running the batch")
        return ys

    def view(self) -> list:
        """The last run's results in place: read-only (scanCount_i, k_i) arrays over the pinned buffer. Each array
        holds the batch alive through its base (a ctypes array over the buffer that keeps a reference to this
        batch), so the buffer outlives every view; the values change at the next run."""
        out = []
        for i, (c, k) in enumerate(zip(self.cubes, self.k)):
            p = ct.POINTER(ct.c_double)()
            _check(L.lib().cubo_hip_batch_view(self._h, i, ct.byref(p)), "viewing the batch results")
            n = c.shape.scan_count * k
            raw = (ct.c_double * n).from_address(ct.addressof(p.contents))
            raw._batch = self
            a = np.frombuffer(raw, dtype=np.float64).reshape(c.shape.scan_count, k)
            a.flags.writeable = False
            out.append(a)
        return out

    def kernel_time(self):
        """(total device milliseconds, replays) since the previous call."""
        t, n = ct.c_double(), ct.c_int64()
        _check(L.lib().cubo_hip_batch_time(self._h, ct.byref(t), ct.byref(n)), "batch time")
        return t.value, n.value

    def release(self):
        if getattr(self, "_h", None) is not None:
            L.lib().cubo_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class ShiftedStack:
    """The mask stack with every position's masks moved by MaskShift.bilinear, and its adjoint
    (include/cubo_hip.h, "shifted stack"; Sources/CuboVirtual/ShiftedImaging.swift:43-73).

    forward(origin_row, origin_col) -> Y (scanCount, K); backward(...) -> (dL/dmasks (K, detectorCount),
    dL/dorigin_row, dL/dorigin_col) for a given dL/dY, re-reading the resident cube. With keep_jacobian=True the
    bilinear forward keeps the Jacobian rows dY/d(shift) on the device, so a backward on the same origin map (cube
    unchanged) reads the cube only for dL/dmasks."""

    def __init__(self, cube: HIPResidentCube, masks: np.ndarray, keep_jacobian: bool = False):
        self.cube = cube
        self.masks = np.ascontiguousarray(masks, np.float32).reshape(-1, cube.shape.detector_count)
        self.k = self.masks.shape[0]
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_shift_plan_create(cube.handle, self.masks.ctypes.data, self.k, ct.byref(h)),
               "preparing the shifted stack")
        self._h = h
        if keep_jacobian:
            self.keep_jacobian(True)

    def keep_jacobian(self, keep: bool) -> None:
        """Whether the bilinear forward keeps dY/d(shift) for the next backward on its origin map."""
        _check(L.lib().cubo_hip_shift_keep_jacobian(self._h, 1 if keep else 0), "keeping the Jacobian rows")

    def _origins(self, origin_row, origin_col):
        n = self.cube.shape.scan_count
        r = np.ascontiguousarray(origin_row, np.float64).reshape(n)
        c = np.ascontiguousarray(origin_col, np.float64).reshape(n)
        return r, c

    def forward(self, origin_row, origin_col, rounded: bool = False) -> np.ndarray:
        """MaskShift.bilinear, or with rounded=True MaskShift.rounded (one roll, no blend)."""
        r, c = self._origins(origin_row, origin_col)
        y = np.empty((self.cube.shape.scan_count, self.k))
        f = L.lib().cubo_hip_shift_forward_rounded if rounded else L.lib().cubo_hip_shift_forward
        _check(f(self._h, r.ctypes.data, c.ctypes.data, y.ctypes.data), "shifted stack")
        return y

    def backward(self, origin_row, origin_col, dy, masks: bool = True, origin: bool = True):
        r, c = self._origins(origin_row, origin_col)
        n = self.cube.shape.scan_count
        d = np.ascontiguousarray(dy, np.float64).reshape(n, self.k)
        dm = np.empty((self.k, self.cube.shape.detector_count)) if masks else None
        dr = np.empty(n) if origin else None
        dc = np.empty(n) if origin else None
        ptr = (lambda a: a.ctypes.data if a is not None else None)
        _check(L.lib().cubo_hip_shift_backward(self._h, r.ctypes.data, c.ctypes.data, d.ctypes.data, ptr(dm), ptr(dr),
                                               ptr(dc)), "shifted stack adjoint")
        return dm, dr, dc

    def kernel_time(self):
        """([forward, origin adjoint, mask adjoint] device ms, backward calls) since the previous call."""
        ms = (ct.c_double * 3)()
        n = ct.c_int64()
        _check(L.lib().cubo_hip_shift_kernel_time(self._h, ms, ct.byref(n)), "shift kernel time")
        return list(ms), n.value

    def jacobian_uses(self) -> int:
        """Backward calls since the previous call whose dL/dorigin came from the forward's kept Jacobian rows."""
        n = ct.c_int64()
        _check(L.lib().cubo_hip_shift_jacobian_uses(self._h, ct.byref(n)), "shift Jacobian uses")
        return n.value

    def release(self):
        if getattr(self, "_h", None) is not None:
            L.lib().cubo_hip_shift_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def center_of_mass_from_stack(cube_shape: CubeShape, y: np.ndarray, columns=(0, 1, 2)) -> CenterOfMassMaps:
    """The centre of mass from the stack's {w, w*q_row, w*q_col} columns; the division in FP64."""
    s, mr, mc = (np.ascontiguousarray(y[:, c], np.float64) for c in columns)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = {"com_intensity": s, "com_moment_row": mr, "com_moment_column": mc, "com_row": mr / s,
               "com_column": mc / s}
    return _com_maps(cube_shape, out)
