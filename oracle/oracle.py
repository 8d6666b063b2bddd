"""ctypes binding of the oracle library (oracle/_build/libcubo_oracle.so).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline
leg of ``bench.py``, always as the checker or the CPU column, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libcubo_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_i64 = ct.c_int64


class OracleError(RuntimeError):
    pass


def build() -> None:
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ct.CDLL(LIB_PATH)
    sig = {
        "oc_mask_point": (ct.c_int, [ct.c_int, ct.c_int, ct.c_int, ct.c_int, _fp]),
        "oc_mask_circle": (ct.c_int, [ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, _fp]),
        "oc_mask_annulus": (ct.c_int, [ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_double, _fp]),
        "oc_mask_rectangle": (ct.c_int, [ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_double, _fp]),
        "oc_virtual_image": (None, [_fp, _i64, _i64, _fp, _dp]),
        "oc_mean_max": (None, [_fp, _i64, _i64, _dp, _dp]),
        "oc_com_sums": (None, [_fp, _i64, ct.c_int, ct.c_int, ct.c_void_p, _dp, _dp, _dp]),
        "oc_mask_stack": (None, [_fp, _i64, _i64, _dp, ct.c_int, _dp]),
        "oc_mask_stack_grad_w": (None, [_fp, _i64, _i64, _dp, ct.c_int, _dp]),
        "oc_vi_bilinear": (ct.c_int, [_fp, _i64, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp]),
        "oc_vi_rounded": (ct.c_int, [_fp, _i64, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp]),
        "oc_vi_bilinear_grad_w": (None, [_fp, _i64, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp]),
        "oc_vi_bilinear_grad_origin": (ct.c_int, [_fp, _i64, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "oc_pattern": (ct.c_int, [_fp, _i64, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                 ct.c_void_p, ct.c_int, ct.c_int, _dp]),
        "oc_gaussian_weights": (ct.c_int, [ct.c_double, _dp, ct.c_int]),
        "oc_gaussian_filter": (None, [_dp, ct.c_int, ct.c_int, _dp, ct.c_int, _dp]),
        "oc_origin_py4dstem": (ct.c_int, [_fp, _i64, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_void_p,
                                         ct.c_int, _dp, _dp]),
        "oc_probe_size": (ct.c_int, [_dp, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_int, _dp]),
        "oc_origin_iterated": (None, [_fp, _i64, ct.c_int, ct.c_int, ct.c_float, ct.c_float, ct.c_int, ct.c_float,
                                     ct.c_int, _dp, _dp]),
        "oc_pairwise_sum": (ct.c_double, [_dp, _i64]),
        "oc_fft2d": (None, [_dp, _dp, ct.c_int, ct.c_int, ct.c_int]),
        "oc_plane_fit": (ct.c_int, [_dp, ct.c_int, ct.c_int, ct.c_double, _dp]),
        "oc_normalize": (ct.c_int, [_dp, _dp, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_void_p,
                                   ct.c_void_p, _dp, _dp, _dp, _dp]),
        "oc_rotate": (None, [_dp, _dp, _i64, ct.c_double, ct.c_int, _dp, _dp]),
        "oc_curl": (ct.c_double, [_dp, _dp, ct.c_int, ct.c_int]),
        "oc_divergence": (ct.c_double, [_dp, _dp, ct.c_int, ct.c_int]),
        "oc_rotation_curve": (ct.c_int, [_dp, _dp, ct.c_int, ct.c_int, _dp, ct.c_int, ct.c_int, ct.c_int, _dp]),
        "oc_rotation_solve": (ct.c_int, [_dp, _dp, ct.c_int, ct.c_int, ct.c_void_p, ct.c_int, ct.c_int,
                                        ct.c_void_p, ct.c_int, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int),
                                        ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), _dp, _dp, _dp, _dp]),
        "oc_dpc_reconstruct": (ct.c_int, [_dp, _dp, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_int,
                                         ct.c_double, ct.c_double, ct.c_int, _dp, _dp, _ip,
                                         ct.POINTER(ct.c_int), ct.POINTER(ct.c_double)]),
        "oc_dpc_step": (ct.c_int, [_dp, _dp, _dp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                  ct.c_double, _dp, ct.POINTER(ct.c_double)]),
        "oc_dpc_step_vjp": (ct.c_int, [_dp, _dp, _dp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                                      ct.c_double, _dp, ct.c_double, _dp, _dp, _dp]),
        "oc_cpu_first_processing_u16": (None, [_u16p, _i64, ct.c_int, ct.c_int, _i64, _i64, _fp, _fp, ct.c_int,
                                               _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _check(st: int, what: str) -> None:
    if st != 0:
        raise OracleError(f"{what}: oracle status {st}")


def _f32(cube) -> np.ndarray:
    """The reference converts every element type to Float32 on read (MappedCube.swift:56-83)."""
    return np.ascontiguousarray(np.asarray(cube).astype(np.float32, copy=False))


def _flat(cube):
    a = _f32(cube)
    ry, rx, qy, qx = a.shape
    return a.reshape(ry * rx, qy * qx), ry, rx, qy, qx


# --- masks --------------------------------------------------------------------------------------
def mask(kind: str, qy: int, qx: int, **g) -> np.ndarray:
    out = np.zeros(qy * qx, np.float32)
    L = lib()
    if kind == "point":
        st = L.oc_mask_point(qy, qx, int(g["row"]), int(g["column"]), out)
    elif kind == "circle":
        st = L.oc_mask_circle(qy, qx, g["centerRow"], g["centerColumn"], g["radius"], out)
    elif kind == "annulus":
        st = L.oc_mask_annulus(qy, qx, g["centerRow"], g["centerColumn"], g["inner"], g["outer"], out)
    elif kind == "rectangle":
        st = L.oc_mask_rectangle(qy, qx, g["rowMin"], g["rowMax"], g["columnMin"], g["columnMax"], out)
    elif kind == "mask":
        w = np.asarray(g["weights"], np.float32).reshape(-1)
        if w.size != qy * qx:
            raise OracleError("mask length mismatch")
        if not np.all(np.isfinite(w)):
            raise OracleError("non-finite mask")
        return w.copy()
    else:
        raise ValueError(kind)
    _check(st, f"mask {kind}")
    return out


# --- reductions ---------------------------------------------------------------------------------
def virtual_image(cube, m) -> np.ndarray:
    x, ry, rx, qy, qx = _flat(cube)
    out = np.zeros(ry * rx)
    lib().oc_virtual_image(x, x.shape[0], x.shape[1], np.ascontiguousarray(m, np.float32), out)
    return out


def mean_max(cube):
    x, ry, rx, qy, qx = _flat(cube)
    mean = np.zeros(qy * qx)
    mx = np.zeros(qy * qx)
    lib().oc_mean_max(x, x.shape[0], x.shape[1], mean, mx)
    return mean, mx


def pattern(cube, statistic: str, take=None, weights=None, origin=None, ties: str = "first",
            fourier: bool = False) -> np.ndarray:
    """VirtualDiffraction.pattern / meanAndMax(of:shift:) for 'mean' or 'max' (SelectedDiffraction.swift:110-179,
    225-271; ShiftedImaging.swift:125-167). take: flags per position; weights: Doubles per position; origin:
    (row map, column map) for PatternShift.rounded; ties 'first' (np.max) or 'later' (Swift.max)."""
    x, ry, rx, qy, qx = _flat(cube)
    npos = ry * rx
    keep = []

    def ptr(a, dt):
        if a is None:
            return None
        arr = np.ascontiguousarray(np.asarray(a).reshape(-1), dt)
        if arr.size != npos:
            raise OracleError("selection length mismatch")
        keep.append(arr)
        return arr.ctypes.data
    t = ptr(None if take is None else np.asarray(take, bool).astype(np.uint8), np.uint8)
    w = ptr(weights, np.float64)
    orow = ocol = None
    if origin is not None:
        orow, ocol = ptr(origin[0], np.float64), ptr(origin[1], np.float64)
    out = np.zeros(qy * qx)
    st = lib().oc_pattern(x, npos, qy, qx, 0 if statistic == "mean" else 1, t, w, orow, ocol,
                          1 if ties == "later" else 0, int(fourier), out)
    _check(st, f"pattern {statistic}")
    return out


def gaussian_filter(pattern, sigma: float) -> np.ndarray:
    """scipy.ndimage.gaussian_filter(pattern, sigma, mode='nearest') in the reference's order."""
    p = np.ascontiguousarray(pattern, np.float64)
    qy, qx = p.shape
    w = np.zeros(2 * int(4.0 * sigma + 0.5) + 1)
    R = lib().oc_gaussian_weights(float(sigma), w, w.size)
    out = np.zeros(qy * qx)
    lib().oc_gaussian_filter(p.ravel(), qy, qx, w, R, out)
    return out.reshape(qy, qx)


def origin_py4dstem(cube, radius: float, radius_scale: float = 1.2, take=None, fast_center: bool = False):
    """OriginMeasurement.measure (py4DSTEM get_origin): (row map, column map) per position."""
    x, ry, rx, qy, qx = _flat(cube)
    r, c = np.zeros(ry * rx), np.zeros(ry * rx)
    t = None if take is None else np.ascontiguousarray(np.asarray(take, bool).astype(np.uint8))
    _check(lib().oc_origin_py4dstem(x, ry * rx, qy, qx, float(radius), float(radius_scale),
                                    None if t is None else t.ctypes.data, int(fast_center), r, c), "origin")
    return r, c


def probe_size(pattern, lower=0.01, upper=0.99, count=100):
    """ProbeSize.estimate: (radius, centre row, centre column, threshold)."""
    p = np.ascontiguousarray(pattern, np.float64)
    out = np.zeros(4)
    _check(lib().oc_probe_size(p.ravel(), p.shape[0], p.shape[1], lower, upper, count, out), "probe size")
    return tuple(out)


def origin_iterated(cube, radius: float, radius_scale: float = 1.2, passes: int = 4, margin: float = 1.5,
                    binned: bool = True):
    """The iterated-window estimator in Float32: (row map, column map)."""
    x, ry, rx, qy, qx = _flat(cube)
    r, c = np.zeros(ry * rx), np.zeros(ry * rx)
    lib().oc_origin_iterated(x, ry * rx, qy, qx, float(radius), float(radius_scale), int(passes), float(margin),
                             int(binned), r, c)
    return r, c


def com_sums(cube, weights=None):
    x, ry, rx, qy, qx = _flat(cube)
    n = ry * rx
    s, r, c = np.zeros(n), np.zeros(n), np.zeros(n)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    lib().oc_com_sums(x, n, qy, qx, None if w is None else w.ctypes.data_as(ct.c_void_p), s, r, c)
    return s, r, c


def com(cube, weights=None):
    s, r, c = com_sums(cube, weights)
    with np.errstate(invalid="ignore", divide="ignore"):
        return dict(intensity=s, moment_row=r, moment_column=c, row=r / s, column=c / s)


def mask_stack(cube, w) -> np.ndarray:
    x, ry, rx, qy, qx = _flat(cube)
    w = np.ascontiguousarray(w, np.float64).reshape(-1, qy * qx)
    y = np.zeros((ry * rx, w.shape[0]))
    lib().oc_mask_stack(x, x.shape[0], x.shape[1], w, w.shape[0], y)
    return y


def mask_stack_grad_w(cube, dy) -> np.ndarray:
    x, ry, rx, qy, qx = _flat(cube)
    dy = np.ascontiguousarray(dy, np.float64).reshape(ry * rx, -1)
    dw = np.zeros((dy.shape[1], qy * qx))
    lib().oc_mask_stack_grad_w(x, x.shape[0], x.shape[1], dy, dy.shape[1], dw)
    return dw


def vi_bilinear(cube, m, origin_r, origin_c) -> np.ndarray:
    x, ry, rx, qy, qx = _flat(cube)
    out = np.zeros(ry * rx)
    _check(lib().oc_vi_bilinear(x, x.shape[0], qy, qx, np.ascontiguousarray(m, np.float64),
                                np.ascontiguousarray(origin_r, np.float64),
                                np.ascontiguousarray(origin_c, np.float64), out), "vi_bilinear")
    return out


def vi_rounded(cube, m, origin_r, origin_c) -> np.ndarray:
    x, ry, rx, qy, qx = _flat(cube)
    out = np.zeros(ry * rx)
    _check(lib().oc_vi_rounded(x, x.shape[0], qy, qx, np.ascontiguousarray(m, np.float64),
                               np.ascontiguousarray(origin_r, np.float64),
                               np.ascontiguousarray(origin_c, np.float64), out), "vi_rounded")
    return out


def vi_bilinear_grad_w(cube, origin_r, origin_c, dy) -> np.ndarray:
    x, ry, rx, qy, qx = _flat(cube)
    dw = np.zeros(qy * qx)
    lib().oc_vi_bilinear_grad_w(x, x.shape[0], qy, qx, np.ascontiguousarray(origin_r, np.float64),
                                np.ascontiguousarray(origin_c, np.float64), np.ascontiguousarray(dy, np.float64), dw)
    return dw


def vi_bilinear_grad_origin(cube, m, origin_r, origin_c, dy):
    x, ry, rx, qy, qx = _flat(cube)
    gr, gc = np.zeros(ry * rx), np.zeros(ry * rx)
    _check(lib().oc_vi_bilinear_grad_origin(x, x.shape[0], qy, qx, np.ascontiguousarray(m, np.float64),
                                            np.ascontiguousarray(origin_r, np.float64),
                                            np.ascontiguousarray(origin_c, np.float64),
                                            np.ascontiguousarray(dy, np.float64), gr, gc), "grad_origin")
    return gr, gc


# --- numerics helpers ---------------------------------------------------------------------------
def pairwise_sum(a) -> float:
    a = np.ascontiguousarray(a, np.float64)
    return lib().oc_pairwise_sum(a, a.size)


def fft2(re, im=None, inverse=False):
    re = np.ascontiguousarray(np.array(re, np.float64))
    im = np.zeros_like(re) if im is None else np.ascontiguousarray(np.array(im, np.float64))
    lib().oc_fft2d(re, im, re.shape[0], re.shape[1], int(inverse))
    return re, im


# --- CuboPhase chain ----------------------------------------------------------------------------
_REF = {"none": 0, "plane": 1, "constant": 2, "given": 3}


def plane_fit(m, R, C, tol=1e-12):
    out = np.zeros(R * C)
    _check(lib().oc_plane_fit(np.ascontiguousarray(m, np.float64), R, C, tol, out), "plane_fit")
    return out


def normalize(mx, my, R, C, reference="plane", sampling=1.0, tol=1e-12, given=None):
    """reference "given" takes given = (row field, column field), each R*C."""
    fx, fy, nx, ny = (np.zeros(R * C) for _ in range(4))
    g = [None, None] if given is None else [np.ascontiguousarray(v, np.float64).reshape(R * C) for v in given]
    _check(lib().oc_normalize(np.ascontiguousarray(mx, np.float64), np.ascontiguousarray(my, np.float64), R, C,
                              _REF[reference], sampling, tol, None if g[0] is None else g[0].ctypes.data,
                              None if g[1] is None else g[1].ctypes.data, fx, fy, nx, ny), "normalize")
    return fx, fy, nx, ny


DEFAULT_ANGLES = np.arange(179, dtype=np.float64) - 89.0


def rotation_solve(nx, ny, R, C, angles=DEFAULT_ANGLES, maximize=False, force_degrees=None, force_transpose=None):
    """RotationSolver.solve; curves not evaluated come back empty."""
    angles = np.ascontiguousarray(angles, np.float64)
    deg = ct.c_double()
    t, n0, n1 = ct.c_int(), ct.c_int(), ct.c_int()
    m = max(angles.size, 1)
    curve, curve_t = np.zeros(m), np.zeros(m)
    cx, cy = np.zeros(R * C), np.zeros(R * C)
    forced = None if force_degrees is None else ct.byref(ct.c_double(float(force_degrees)))
    ft = -1 if force_transpose is None else int(bool(force_transpose))
    _check(lib().oc_rotation_solve(np.ascontiguousarray(nx, np.float64), np.ascontiguousarray(ny, np.float64), R, C,
                                   angles.ctypes.data, angles.size, int(maximize), forced, ft, ct.byref(deg),
                                   ct.byref(t), ct.byref(n0), ct.byref(n1), curve, curve_t, cx, cy), "rotation")
    return dict(degrees=deg.value, transposed=bool(t.value), curve=curve[: n0.value],
                curve_transposed=curve_t[: n1.value], cx=cx, cy=cy)


def rotation_curve(nx, ny, R, C, angles=DEFAULT_ANGLES, transpose=False, maximize=False):
    angles = np.ascontiguousarray(angles, np.float64)
    out = np.zeros(angles.size)
    _check(lib().oc_rotation_curve(np.ascontiguousarray(nx, np.float64), np.ascontiguousarray(ny, np.float64), R, C,
                                   angles, angles.size, int(transpose), int(maximize), out), "rotation_curve")
    return out


def dpc_reconstruct(fx, fy, R, C, sampling=(1.0, 1.0), padding=2.0, max_iter=64, step=0.5, stop=1e-6,
                    backtrack=True):
    phase = np.zeros(R * C)
    errors = np.zeros(max(max_iter, 1))
    acc = np.zeros(max(max_iter, 1), np.int32)
    n = ct.c_int()
    fs = ct.c_double()
    _check(lib().oc_dpc_reconstruct(np.ascontiguousarray(fx, np.float64), np.ascontiguousarray(fy, np.float64), R, C,
                                    sampling[0], sampling[1], padding, max_iter, step, stop, int(backtrack), phase,
                                    errors, acc, ct.byref(n), ct.byref(fs)), "dpc")
    accepted = [bool(a) for a in acc[: n.value]]
    return dict(phase=phase, errors=errors[: sum(accepted)], accepted=accepted, final_step=fs.value)


def dpc_step(phi, fx, fy, R, C, Rp, Cp, sampling=(1.0, 1.0), step=0.5):
    out = np.zeros(Rp * Cp)
    err = ct.c_double()
    _check(lib().oc_dpc_step(np.ascontiguousarray(phi, np.float64), np.ascontiguousarray(fx, np.float64),
                             np.ascontiguousarray(fy, np.float64), R, C, Rp, Cp, sampling[0], sampling[1], step, out,
                             ct.byref(err)), "dpc_step")
    return out, err.value


def dpc_step_vjp(phi, fx, fy, R, C, Rp, Cp, g_phi_out, g_err, sampling=(1.0, 1.0), step=0.5):
    g_phi, g_fx, g_fy = np.zeros(Rp * Cp), np.zeros(R * C), np.zeros(R * C)
    _check(lib().oc_dpc_step_vjp(np.ascontiguousarray(phi, np.float64), np.ascontiguousarray(fx, np.float64),
                                 np.ascontiguousarray(fy, np.float64), R, C, Rp, Cp, sampling[0], sampling[1], step,
                                 np.ascontiguousarray(g_phi_out, np.float64), float(g_err), g_phi, g_fx, g_fy),
           "dpc_step_vjp")
    return g_phi, g_fx, g_fy


def cpu_first_processing_u16(cube_u16: np.ndarray, qy: int, qx: int, positions_per_row: int, bf, adf,
                             threads: int, rows_per_tile: int = 0):
    """The reference's .cpu pass (VI x2, mean, max, CoM) over a uint16 cube, timed as the CPU baseline.
    rows_per_tile 0 is ScanTiling.automatic (the most scan rows whose Float32 tile fits 64 MiB)."""
    x = np.ascontiguousarray(cube_u16, np.uint16).reshape(-1)
    ndet = qy * qx
    npos = x.size // ndet
    row_bytes = positions_per_row * ndet * 4
    if rows_per_tile <= 0:
        rows_per_tile = max(1, min(npos // positions_per_row, (64 << 20) // max(1, row_bytes)))
    out = [np.zeros(npos), np.zeros(npos), np.zeros(ndet), np.zeros(ndet), np.zeros(npos), np.zeros(npos),
           np.zeros(npos)]
    lib().oc_cpu_first_processing_u16(x, npos, qy, qx, rows_per_tile, positions_per_row,
                                      np.ascontiguousarray(bf, np.float32), np.ascontiguousarray(adf, np.float32),
                                      threads, *out)
    return dict(vi_bf=out[0], vi_adf=out[1], mean=out[2], max=out[3], intensity=out[4], moment_row=out[5],
                moment_column=out[6])
