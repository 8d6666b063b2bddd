"""Mirror of CuboPhase's DPC over the C ABI (include/cubo_hip.h, "iterative DPC").

DPC.reconstruct (Sources/CuboPhase/DPC.swift:26-102) on the device in Double, plus the adjoints of one
iteration and of a fixed-step chain of iterations, which the reference does not have.
"""
from __future__ import annotations

import ctypes as ct
import threading
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .virtual import CancellationError, CuboError, _check


ANGLES = np.arange(179, dtype=np.float64) - 89.0      # RotationSolver's default angles, -89 ... 89 degrees
_REFERENCE = {"none": 0, "plane": 1, "constant": 2, "given": 3}


def normalize(mx, my, rows: int, cols: int, reference: str = "plane", sampling: float = 1.0, tol: float = 1e-12,
              given=None, device: int = 0):
    """CenterOfMassReference.normalize: (reference row, reference column, normalised row, normalised column).
    reference "given" is .given(field): given = (row field, column field), each rows x cols, e.g. an OriginMap's
    row and column maps (py4DSTEM's force_com_shifts)."""
    n = rows * cols
    x = np.ascontiguousarray(mx, np.float64).reshape(n)
    y = np.ascontiguousarray(my, np.float64).reshape(n)
    g = [None, None]
    if reference == "given":
        if given is None:
            raise CuboError("the given reference needs a field", case="invalidField")
        g = [np.ascontiguousarray(getattr(v, "values", v), np.float64) for v in given]
        if any(v.size != n for v in g):                          # CenterOfMassReference.swift:46-48
            raise CuboError(f"the given reference has {g[0].size} and {g[1].size} entries, the field {rows} x {cols}",
                            case="invalidField")
        g = [v.reshape(n) for v in g]
    fx, fy, nx, ny = (np.zeros(n) for _ in range(4))
    _check(L.lib().cubo_hip_phase_normalize(device, x.ctypes.data, y.ctypes.data, rows, cols, _REFERENCE[reference],
                                            float(sampling), float(tol), None if g[0] is None else g[0].ctypes.data,
                                            None if g[1] is None else g[1].ctypes.data, fx.ctypes.data,
                                            fy.ctypes.data, nx.ctypes.data, ny.ctypes.data), "normalize")
    return fx, fy, nx, ny


def _cancellable(is_cancelled, call):
    """Runs call(flag pointer or None). The reference's isCancelled closure becomes the C ABI's flag: it is asked
    before the call, then polled from a helper thread while the library works (ctypes releases the GIL), which sets
    the flag the library reads."""
    if is_cancelled is None:
        return call(None)
    if is_cancelled():
        raise CancellationError()
    flag = ct.c_int32(0)
    stop = threading.Event()

    def watch():
        while not stop.wait(0.0002):
            if is_cancelled():
                flag.value = 1
                return

    th = threading.Thread(target=watch, daemon=True)
    th.start()
    try:
        return call(ct.byref(flag))
    finally:
        stop.set()
        th.join()


def rotation_solve(nx, ny, rows: int, cols: int, angles=ANGLES, maximize: bool = False,
                   force_degrees: Optional[float] = None, force_transpose: Optional[bool] = None,
                   is_cancelled=None, device: int = 0) -> dict:
    """RotationSolver.solve: the chosen angle and orientation, the curves that were evaluated (empty otherwise) and
    the corrected field. force_degrees / force_transpose are forceRotationDegrees / forceTranspose; is_cancelled a
    callable (CancellationError when it answers True)."""
    n = rows * cols
    x = np.ascontiguousarray(nx, np.float64).reshape(n)
    y = np.ascontiguousarray(ny, np.float64).reshape(n)
    a = np.ascontiguousarray(angles, np.float64)
    deg, t, n0, n1 = ct.c_double(), ct.c_int32(), ct.c_int32(), ct.c_int32()
    m = 1 if force_degrees is not None else max(a.size, 1)
    curve, curve_t, cx, cy = np.zeros(m), np.zeros(m), np.zeros(n), np.zeros(n)
    forced = None if force_degrees is None else ct.c_double(float(force_degrees))
    ft = -1 if force_transpose is None else int(bool(force_transpose))
    _cancellable(is_cancelled, lambda flag: _check(L.lib().cubo_hip_phase_rotation(
        device, x.ctypes.data, y.ctypes.data, rows, cols, a.ctypes.data, a.size, int(maximize),
        None if forced is None else ct.addressof(forced), ft, flag, ct.byref(deg), ct.byref(t), ct.byref(n0),
        ct.byref(n1), curve.ctypes.data, curve_t.ctypes.data, cx.ctypes.data, cy.ctypes.data), "rotation"))
    return dict(degrees=deg.value, radians=deg.value * (np.pi / 180), transposed=bool(t.value),
                angles=(np.array([deg.value]) if force_degrees is not None and force_transpose is None
                        else (a if force_degrees is None else np.zeros(0))),
                curve=curve[: n0.value], curve_transposed=curve_t[: n1.value], cx=cx, cy=cy)


def rotation_curve(nx, ny, rows: int, cols: int, angles=ANGLES, maximize: bool = False, transpose: bool = False,
                   is_cancelled=None, device: int = 0) -> np.ndarray:
    """RotationSolver.curve (RotationSolver.swift:36-44): the criterion over the angles for one orientation."""
    s = rotation_solve(nx, ny, rows, cols, angles, maximize, force_transpose=transpose, is_cancelled=is_cancelled,
                       device=device)
    return s["curve_transposed"] if transpose else s["curve"]


class DPC:
    """One scan shape, padding and sampling: DPC(rows, cols, padding=2, sampling=(1, 1))."""

    def __init__(self, rows: int, cols: int, padding: float = 2.0, sampling: Sequence[float] = (1.0, 1.0),
                 device: int = 0):
        self.rows, self.cols = int(rows), int(cols)
        h = ct.c_void_p()
        _check(L.lib().cubo_hip_dpc_create(device, self.rows, self.cols, float(padding), float(sampling[0]),
                                           float(sampling[1]), ct.byref(h)), "preparing DPC")
        self._h = h
        r, c = ct.c_int32(), ct.c_int32()
        _check(L.lib().cubo_hip_dpc_padded_shape(h, ct.byref(r), ct.byref(c)), "DPC shape")
        self.padded_shape = (r.value, c.value)

    def _field(self, fx, fy):
        n = self.rows * self.cols
        x = np.ascontiguousarray(fx, np.float64).reshape(n)
        y = np.ascontiguousarray(fy, np.float64).reshape(n)
        return x, y

    def reconstruct(self, fx, fy, max_iter: int = 64, step: float = 0.5, stop: float = 1e-6,
                    backtrack: bool = True) -> dict:
        """phase (rows x cols, flattened), errors of the accepted iterations, the accepted flags, final step."""
        x, y = self._field(fx, fy)
        phase = np.zeros(self.rows * self.cols)
        errors = np.zeros(max(max_iter, 1))
        acc = np.zeros(max(max_iter, 1), np.int32)
        n, fs = ct.c_int32(), ct.c_double()
        _check(L.lib().cubo_hip_dpc_reconstruct(self._h, x.ctypes.data, y.ctypes.data, int(max_iter), float(step),
                                                float(stop), int(backtrack), phase.ctypes.data, errors.ctypes.data,
                                                acc.ctypes.data, ct.byref(n), ct.byref(fs)), "DPC")
        accepted = [bool(a) for a in acc[: n.value]]
        return dict(phase=phase, errors=errors[: sum(accepted)], accepted=accepted, final_step=fs.value)

    def step(self, phi, fx, fy, step: float = 0.5):
        """One accepted iteration on the padded grid: (phi_out, error of phi)."""
        x, y = self._field(fx, fy)
        p = np.ascontiguousarray(phi, np.float64).reshape(-1)
        out = np.zeros_like(p)
        err = ct.c_double()
        _check(L.lib().cubo_hip_dpc_step(self._h, p.ctypes.data, x.ctypes.data, y.ctypes.data, float(step),
                                         out.ctypes.data, ct.byref(err)), "DPC step")
        return out, err.value

    def step_backward(self, phi, fx, fy, g_phi_out, g_error: float = 0.0, step: float = 0.5):
        """(dL/dphi, dL/dfx, dL/dfy) of one iteration."""
        x, y = self._field(fx, fy)
        p = np.ascontiguousarray(phi, np.float64).reshape(-1)
        g = np.ascontiguousarray(g_phi_out, np.float64).reshape(-1)
        g_phi, g_fx, g_fy = np.zeros_like(p), np.zeros_like(x), np.zeros_like(y)
        _check(L.lib().cubo_hip_dpc_step_backward(self._h, p.ctypes.data, x.ctypes.data, y.ctypes.data, float(step),
                                                  g.ctypes.data, float(g_error), g_phi.ctypes.data, g_fx.ctypes.data,
                                                  g_fy.ctypes.data), "DPC step adjoint")
        return g_phi, g_fx, g_fy

    def chain_backward(self, fx, fy, nsteps: int, g_phase, g_errors: Optional[np.ndarray] = None, step: float = 0.5):
        """nsteps fixed-step iterations from phase 0 and their adjoint: (phase, errors, dL/dfx, dL/dfy)."""
        x, y = self._field(fx, fy)
        gp = np.ascontiguousarray(g_phase, np.float64).reshape(self.rows * self.cols)
        ge = None if g_errors is None else np.ascontiguousarray(g_errors, np.float64).reshape(nsteps)
        phase = np.zeros(self.rows * self.cols)
        errors = np.zeros(nsteps)
        g_fx, g_fy = np.zeros_like(x), np.zeros_like(y)
        _check(L.lib().cubo_hip_dpc_chain_backward(self._h, x.ctypes.data, y.ctypes.data, int(nsteps), float(step),
                                                   gp.ctypes.data, ge.ctypes.data if ge is not None else None,
                                                   phase.ctypes.data, errors.ctypes.data, g_fx.ctypes.data,
                                                   g_fy.ctypes.data), "DPC chain adjoint")
        return phase, errors, g_fx, g_fy

    def kernel_time(self):
        ms, n = ct.c_double(), ct.c_int64()
        _check(L.lib().cubo_hip_dpc_time(self._h, ct.byref(ms), ct.byref(n)), "DPC time")
        return ms.value, n.value

    def release(self):
        if getattr(self, "_h", None) is not None:
            L.lib().cubo_hip_dpc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


__all__ = ["ANGLES", "DPC", "CuboError", "normalize", "rotation_curve", "rotation_solve"]
