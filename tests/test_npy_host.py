"""The .npy reader of the library (cubo_hip_npy_read_header) and the file checks of cubo_hip_cube_load_file /
cubo_hip_cube_load_npy, which all run before any device call, so on the CPU. They follow CuboCoreTests.swift:150-268:
the header of NumPy's own files, version 2.0 and 3.0 headers (one longer than 64 KiB), and the refusals by case."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import cubokit_amd as C
from cubokit_amd import _lib as L
from tests.support import golden

FIXTURES = ["digits", "gaussian", "counts16", "counts32", "thirds", "negative", "boundary", "shift"]


def _numpy_header(path):
    with open(path, "rb") as f:
        version = np.lib.format.read_magic(f)
        shape, fortran, dtype = (np.lib.format.read_array_header_1_0(f) if version == (1, 0)
                                 else np.lib.format.read_array_header_2_0(f))
        return version, shape, fortran, dtype, f.tell()


def _raw_npy(path, text: str, payload: int, version=1):
    """CuboCoreTests.swift:252-261: a header dict padded with spaces to 64 bytes and a newline, then payload bytes."""
    pre = 10 if version == 1 else 12
    pad = (64 - (pre + len(text.encode()) + 1) % 64) % 64
    header = (text + " " * pad + "\n").encode()
    data = b"\x93NUMPY" + bytes([version, 0])
    data += len(header).to_bytes(2 if version == 1 else 4, "little") + header + bytes(payload)
    with open(path, "wb") as f:
        f.write(data)


def _refusal(fn):
    with pytest.raises(C.CuboError) as e:
        fn()
    return e.value


def test_header_of_every_fixture_is_numpys():
    for name in FIXTURES:
        path = golden(name + ".npy")
        version, shape, fortran, dtype, offset = _numpy_header(path)
        h = C.NPYFile.header(path)
        assert (h.major_version, h.minor_version) == version
        assert h.shape == list(shape) and not fortran
        assert h.element_type == dtype.name and h.data_offset == offset and offset % 64 == 0


@pytest.mark.parametrize("version", [(1, 0), (2, 0), (3, 0)])
def test_numpy_written_versions(tmp_path, version):
    arr = np.arange(2 * 3 * 4 * 5, dtype=np.int16).reshape(2, 3, 4, 5)
    path = str(tmp_path / f"v{version[0]}.npy")
    with open(path, "wb") as f:
        np.lib.format.write_array(f, arr, version=version)
    h = C.NPYFile.header(path)
    assert (h.major_version, h.minor_version) == version
    assert h.element_type == "int16" and h.shape == [2, 3, 4, 5]
    assert h.data_offset == _numpy_header(path)[4]


def test_version_2_header_longer_than_64_kib(tmp_path):
    """CuboCoreTests.swift:170-202: the third length byte is not zero."""
    path = str(tmp_path / "long.npy")
    text = "{'descr': '<u2', 'fortran_order': False, 'shape': (1, 2, 1, 3), }" + " " * 70_000
    _raw_npy(path, text, 0, version=2)
    with open(path, "ab") as f:
        f.write(np.array([7, 8, 9, 50000, 11, 12], "<u2").tobytes())
    h = C.NPYFile.header(path)
    m = int.from_bytes(open(path, "rb").read(12)[8:12], "little")
    assert m > 65535 and h.data_offset == 12 + m and h.shape == [1, 2, 1, 3]


def test_element_types_and_orders(tmp_path):
    for descr, name in [("<f4", "float32"), ("<f8", "float64"), ("|u1", "uint8"), ("|i1", "int8"), ("<u2", "uint16"),
                        ("<i2", "int16"), ("<u4", "uint32"), ("<i4", "int32"), ("<u8", "uint64"), ("<i8", "int64"),
                        ("=f4", "float32"), (">u1", "uint8"), (">i1", "int8")]:
        path = str(tmp_path / "t.npy")
        _raw_npy(path, f"{{'descr': '{descr}', 'fortran_order': False, 'shape': (1, 1, 2, 2), }}", 64)
        assert C.NPYFile.header(path).element_type == name, descr


@pytest.mark.parametrize("descr,fortran,shape,case", [
    (">f4", "False", "(1, 1, 2, 2)", "bigEndianNotSupported"),
    ("<c8", "False", "(1, 1, 2, 2)", "unsupportedElementType"),
    ("<f4", "True", "(1, 1, 2, 2)", "fortranOrderNotSupported"),
])
def test_header_refusals(tmp_path, descr, fortran, shape, case):
    """CuboCoreTests.swift:204-216."""
    path = str(tmp_path / "r.npy")
    _raw_npy(path, f"{{'descr': '{descr}', 'fortran_order': {fortran}, 'shape': {shape}, }}", 16)
    assert _refusal(lambda: C.NPYFile.header(path)).case == case


def _load_npy_status(path):
    h = ct.c_void_p()
    st = L.lib().cubo_hip_cube_load_npy(0, os.fsencode(path), ct.byref(h))
    if st == 0:
        L.lib().cubo_hip_cube_destroy(h)
    return st, L.last_error()


def test_load_refusals_come_before_the_device(tmp_path):
    """CuboCoreTests.swift:204-238: wrong rank, a file one byte short, an unknown key, no magic, a misaligned offset
    and an absent file, each refused by its case, with or without a GPU."""
    p = str(tmp_path / "rank.npy")
    _raw_npy(p, "{'descr': '<f4', 'fortran_order': False, 'shape': (2, 2, 2), }", 32)
    st, msg = _load_npy_status(p)
    assert st == L.ERR_FILE and msg.startswith("wrongRank")
    p = str(tmp_path / "short.npy")
    _raw_npy(p, "{'descr': '<f4', 'fortran_order': False, 'shape': (1, 1, 2, 2), }", 15)
    st, msg = _load_npy_status(p)
    assert st == L.ERR_FILE and msg.startswith("fileTooShort")
    needed, holds = map(int, re.findall(r"needs (\d+) bytes, holds (\d+)", msg)[0])
    assert needed - holds == 1
    p = str(tmp_path / "key.npy")
    _raw_npy(p, "{'descr': '<f4', 'fortran_order': False, 'shape': (1, 1, 2, 2), 'x': 1, }", 16)
    st, msg = _load_npy_status(p)
    assert st == L.ERR_FILE and msg == "notNPY: unknown key 'x'"
    p = str(tmp_path / "magic.npy")
    with open(p, "wb") as f:
        f.write(b"PNG not numpy at all")
    st, msg = _load_npy_status(p)
    assert st == L.ERR_FILE and msg == "notNPY: no NUMPY magic"
    e = _refusal(lambda: C.HIPResidentCube.load_file(golden("digits.npy"), 2, "float32", C.CubeShape(2, 3, 5, 7)))
    assert e.case == "misalignedOffset"
    e = _refusal(lambda: C.NPYFile.header(str(tmp_path / "absent.npy")))
    assert e.case == "cannotOpen"


@pytest.mark.parametrize("shape", ["(4294967296, 4294967296, 4294967296, 4294967296)", "(9223372036854775807, 2)",
                                   "(3000000000, 3000000000)", "(1152921504606846976,)"])
def test_overflowing_shapes_are_refused(tmp_path, shape):
    """CuboCoreTests.swift:240-249."""
    p = str(tmp_path / "overflow.npy")
    _raw_npy(p, f"{{'descr': '<f8', 'fortran_order': False, 'shape': {shape}, }}", 64)
    assert _refusal(lambda: C.NPYFile.header(p)).case == "notNPY"
    assert _load_npy_status(p)[0] == L.ERR_FILE


def test_dict_grammar():
    """The reference's dict parser (NPYFile.swift:158-218): quotes of either kind, 'L' suffixes, a trailing comma
    or none, keys in any order; duplicates and junk refused."""
    import tempfile
    ok = ["{\"descr\": \"<u2\", \"fortran_order\": False, \"shape\": (1, 2, 3, 4)}",
          "{'shape': (1L, 2L, 3L, 4L), 'fortran_order': False, 'descr': '<u2', }",
          "{'descr':'<u2','fortran_order':False,'shape':(1,2,3,4),}"]
    bad = ["{'descr': '<u2', 'descr': '<u2', 'fortran_order': False, 'shape': (1, 2, 3, 4), }",
           "{'descr': '<u2', 'fortran_order': false, 'shape': (1, 2, 3, 4), }",
           "{'descr': '<u2', 'fortran_order': False, 'shape': (1, x, 3, 4), }",
           "{'descr': '<u2', 'fortran_order': False, 'shape': (1, 2, 3, 4) 'y' }",
           "['descr', '<u2']",
           "{'descr': '<u2', 'fortran_order': False}"]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.npy")
        for text in ok:
            _raw_npy(p, text, 48)
            assert C.NPYFile.header(p).shape == [1, 2, 3, 4], text
        for text in bad:
            _raw_npy(p, text, 48)
            assert _refusal(lambda: C.NPYFile.header(p)).case == "notNPY", text
