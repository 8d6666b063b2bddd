"""CPU-only checks of the boundary and the host logic: the library loads, exports every symbol the header
declares, and the host mirror's masks, tiling, shape checks and availability follow the reference."""
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
