"""CPU-only checks of the boundary and the host logic: the library loads, exports every symbol the header
declares, and the host mirror's masks, tiling, shape checks and availability follow the reference."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import _lib as L
from tests.support import geometry_kwargs, virtual_pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "cubo_hip.h")).read()
    return sorted(set(re.findall(r"^\s*int\s+(cubo_hip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "the ctypes table and the header disagree"


def test_no_device_here_is_reported_not_faked():
    if C.device_count() > 0:
        pytest.skip("a GPU is present")
    assert not C.hip_is_available()
    with pytest.raises(C.CuboVirtualError) as e:
        C.HIPResidentCube(np.zeros((1, 1, 8, 8), np.uint16))
    assert e.value.case == "noHIPDevice"


def test_availability_checks_the_cap_before_the_device():
    """ResidentCubeAvailabilityTests.swift:21-37: an injected memory size; the cap is inclusive."""
    shape = C.CubeShape(256, 256, 128, 128)            # 2 GiB of uint16
    r = C.HIPResidentCube.availability(shape, np.uint16, hbm_bytes=2 << 30)
    assert r is not None and r.kind == "exceedsMemoryCap" and r.bytes == 2 << 30
    assert "GB" in r.message
    cap_exact = int(np.ceil((2 << 30) / 0.9))
    r = C.HIPResidentCube.availability(shape, np.uint16, hbm_bytes=cap_exact)
    assert r is None or r.kind != "exceedsMemoryCap"
    r = C.HIPResidentCube.availability(shape, np.float32, hbm_bytes=cap_exact)
    assert r is not None and r.kind == "exceedsMemoryCap" and r.bytes == 4 << 30


def test_masks_are_make_detectors():
    for c in virtual_pins()["cubes"]:
        qy, qx = c["shape"][2], c["shape"][3]
        for g in c["geometries"]:
            kind, kw = geometry_kwargs(g["geometry"])
            geo = {"point": lambda: C.DetectorGeometry.point(kw["row"], kw["column"]),
                   "circle": lambda: C.DetectorGeometry.circle(kw["centerRow"], kw["centerColumn"], kw["radius"]),
                   "annulus": lambda: C.DetectorGeometry.annulus(kw["centerRow"], kw["centerColumn"], kw["inner"],
                                                                 kw["outer"]),
                   "rectangle": lambda: C.DetectorGeometry.rectangle(kw["rowMin"], kw["rowMax"], kw["columnMin"],
                                                                     kw["columnMax"]),
                   "mask": lambda: C.DetectorGeometry.mask(kw["weights"])}[kind]()
            m = C.DetectorMask.make(geo, detector_rows=qy, detector_columns=qx)
            assert np.nonzero(m)[0].tolist() == g["mask_pixels"], f"{c['name']}/{g['name']}"


def test_mask_refusals():
    with pytest.raises(C.CuboVirtualError):
        C.DetectorMask.make(C.DetectorGeometry.annulus(4, 4, 3, 3), detector_rows=8, detector_columns=8)
    with pytest.raises(C.CuboVirtualError):
        C.DetectorMask.make(C.DetectorGeometry.circle(4, 4, -1), detector_rows=8, detector_columns=8)
    with pytest.raises(C.CuboVirtualError):
        C.DetectorMask.make(C.DetectorGeometry.point(8, 0), detector_rows=8, detector_columns=8)
    with pytest.raises(C.CuboVirtualError):
        C.DetectorMask.make(C.DetectorGeometry.rectangle(-1, 2, 0, 2), detector_rows=8, detector_columns=8)
    with pytest.raises(C.CuboVirtualError):
        C.DetectorMask.make(C.DetectorGeometry.mask(np.ones(3)), detector_rows=8, detector_columns=8)


def test_tiling_follows_the_reference():
    s = C.CubeShape(256, 256, 128, 128)               # f32 row = 16 MiB: 4 rows per 64 MiB tile
    assert C.ScanTiling.automatic.rows_per_tile(s) == 4
    assert len(C.ScanTiling.automatic.ranges(s)) == 64
    assert [len(r) for r in C.ScanTiling.rows(3).ranges(C.CubeShape(7, 2, 4, 4))] == [3, 3, 1]
    with pytest.raises(C.CuboError):
        C.ScanTiling.rows(0).rows_per_tile(s)
    with pytest.raises(C.CuboError):
        C.CubeShape(0, 1, 1, 1)


def test_non_hip_paths_are_refused_not_emulated():
    with pytest.raises(C.CuboVirtualError):
        C.VirtualImaging.image(np.zeros((1, 1, 4, 4), np.float32), C.DetectorGeometry.point(0, 0),
                               path=C.ComputePath.cpu)


def test_phase_calls_refuse_before_any_device_call():
    """The reference's refusals come before the device is touched, so they hold here without a GPU: a degeneracy
    tolerance that is negative or not finite (CenterOfMassReference.swift:30-32), a given reference of the wrong size
    (:46-48), non-finite angles (invalidAngles, RotationSolver.swift:67, 74-76), a non-finite field or a scan under
    3 x 3 (:128-135); a cancel flag already set gives CancellationError."""
    from cubokit_amd import phase as P
    x = np.arange(20, dtype=np.float64)
    for tol in (-1e-12, np.nan, np.inf):
        with pytest.raises(C.CuboVirtualError) as e:
            P.normalize(x, x, 4, 5, "plane", tol=tol)
        assert e.value.case == "invalidInput"
    with pytest.raises(C.CuboError) as e:
        P.normalize(x, x, 4, 5, "given", given=(x[:19], x[:19]))
    assert e.value.case == "invalidField"
    for kw in (dict(angles=np.array([0.0, np.nan])), dict(angles=np.array([np.inf])), dict(angles=np.zeros(0)),
               dict(force_degrees=np.nan), dict(force_degrees=np.inf, force_transpose=True)):
        with pytest.raises(C.CuboVirtualError) as e:
            P.rotation_solve(x, x, 4, 5, **kw)
        assert e.value.case == "invalidInput", kw
    bad = x.copy()
    bad[3] = np.inf
    with pytest.raises(C.CuboVirtualError):
        P.rotation_solve(bad, x, 4, 5, force_degrees=3.0, force_transpose=False)
    with pytest.raises(C.CuboVirtualError):
        P.rotation_solve(x[:6], x[:6], 2, 3)
    with pytest.raises(C.CancellationError):
        P.rotation_solve(x, x, 4, 5, is_cancelled=lambda: True)
    # through the C ABI: a forced angle and orientation (nothing scored) with the flag already set
    flag, forced = ct.c_int32(1), ct.c_double(1.0)
    deg, t, n0, n1 = ct.c_double(), ct.c_int32(), ct.c_int32(), ct.c_int32()
    cx, cy = np.zeros(20), np.zeros(20)
    st = L.lib().cubo_hip_phase_rotation(0, x.ctypes.data, x.ctypes.data, 4, 5, None, 0, 0, ct.addressof(forced), 0,
                                         ct.addressof(flag), ct.byref(deg), ct.byref(t), ct.byref(n0), ct.byref(n1),
                                         None, None, cx.ctypes.data, cy.ctypes.data)
    assert st == L.ERR_CANCELLED


def test_window_sensitivity_sample_indices():
    """OriginMeasurement.sensitivitySampleIndices (IteratedWindowOrigin.swift:243-252) by hand: 20 positions with a
    limit of 7 take every third (IteratedWindowOriginTests.swift:77); a byte budget of five patterns takes every
    twentieth of 100; a budget below one pattern still takes one; nothing to sample gives nothing."""
    from cubokit_amd.origins import sensitivity_sample_indices as s
    assert s(20, 7, 0, 0) == [0, 3, 6, 9, 12, 15, 18]
    assert s(20, 7, 1600 * 4, 256 << 20) == [0, 3, 6, 9, 12, 15, 18]
    assert s(100, 200, 1000, 5000) == [0, 20, 40, 60, 80]
    assert s(100, 200, 1000, 10) == [0]
    assert s(0, 7, 4, 100) == [] and s(5, 0, 4, 100) == []
    assert s(65536, 200, 16384 * 4, 256 << 20) == list(range(0, 65536, 328))
