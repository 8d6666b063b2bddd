"""ctypes binding of libcubo_hip.so, the C ABI declared in include/cubo_hip.h.

The library is built in-tree by cubokit_amd/build.py. Loading fails loudly when it is missing: there is no
fallback implementation behind this package.
"""
from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcubo_hip.so")

OK = 0
ERR_NO_DEVICE = -1
ERR_OUT_OF_MEMORY = -2
ERR_EXCEEDS_CAP = -3
ERR_INVALID_ARGUMENT = -4
ERR_HIP = -5
ERR_UNSUPPORTED = -6
ERR_INCOMPLETE = -7
ERR_EXCEEDS_BUFFER_LIMIT = -8
ERR_FILE = -9
ERR_CANCELLED = -10

U16 = 1
F32 = 2

STACK_AUTO = 0
STACK_PLAIN = 1
STACK_MFMA_SPLIT = 2

AVAILABLE = 0
UNAVAILABLE_NO_DEVICE = 1
UNAVAILABLE_EXCEEDS_MEMORY_CAP = 2
UNAVAILABLE_EXCEEDS_BUFFER_LIMIT = 3


class DeviceInfo(ct.Structure):
    _fields_ = [("hbm_bytes", ct.c_uint64), ("free_bytes", ct.c_uint64), ("max_alloc_bytes", ct.c_uint64),
                ("compute_units", ct.c_int32), ("wavefront_size", ct.c_int32), ("name", ct.c_char * 64),
                ("arch", ct.c_char * 32)]


class ReduceRequest(ct.Structure):
    _fields_ = [("masks", ct.c_void_p), ("n_masks", ct.c_int32), ("want_mean_max", ct.c_int32),
                ("want_com", ct.c_int32), ("com_weights", ct.c_void_p)]


class ReduceResult(ct.Structure):
    _fields_ = [(n, ct.c_void_p) for n in ("images", "mean", "max", "com_intensity", "com_moment_row",
                                           "com_moment_column", "com_row", "com_column")]


class ReduceView(ct.Structure):
    _fields_ = [(n, ct.POINTER(ct.c_double)) for n in ("images", "mean", "max", "com_intensity", "com_moment_row",
                                                       "com_moment_column", "com_row", "com_column")]


# stored element types, in ElementType's order
ELEMENT_TYPES = ["float32", "float64", "uint8", "int8", "uint16", "int16", "uint32", "int32", "uint64", "int64"]


class NpyHeader(ct.Structure):
    _fields_ = [("major_version", ct.c_int32), ("minor_version", ct.c_int32), ("element_type", ct.c_int32),
                ("rank", ct.c_int32), ("shape", ct.c_int64 * 32), ("data_offset", ct.c_int64)]


ORIGIN_PY4DSTEM = 0
ORIGIN_ITERATED_WINDOW = 1


class ProbeEstimate(ct.Structure):
    _fields_ = [("radius", ct.c_double), ("center_row", ct.c_double), ("center_column", ct.c_double),
                ("threshold", ct.c_double)]


class OriginRequest(ct.Structure):
    _fields_ = [("estimator", ct.c_int32), ("radius", ct.c_double), ("radius_scale", ct.c_double),
                ("scan_mask", ct.c_void_p), ("fast_center", ct.c_int32), ("passes", ct.c_int32),
                ("minimum_margin", ct.c_double), ("binned_coarse", ct.c_int32)]


PATTERN_MEAN = 0
PATTERN_MAX = 1
TIES_FIRST = 0
TIES_LATER = 1


class PatternRequest(ct.Structure):
    _fields_ = [("statistic", ct.c_int32), ("take", ct.c_void_p), ("weights", ct.c_void_p),
                ("origin_row", ct.c_void_p), ("origin_col", ct.c_void_p), ("max_ties", ct.c_int32),
                ("shift_kind", ct.c_int32)]


SHIFT_ROUNDED = 0
SHIFT_FOURIER = 1


class SynthParams(ct.Structure):
    _fields_ = [("seed", ct.c_uint64), ("center_row", ct.c_double), ("center_column", ct.c_double),
                ("radius", ct.c_double), ("edge", ct.c_double), ("counts", ct.c_double),
                ("background", ct.c_double), ("n_rings", ct.c_int32), ("ring_q", ct.c_double * 4),
                ("ring_amplitude", ct.c_double * 4), ("ring_sigma", ct.c_double * 4), ("n_waves", ct.c_int32),
                ("wave_row_amp", ct.c_double * 8), ("wave_col_amp", ct.c_double * 8), ("wave_fr", ct.c_double * 8),
                ("wave_fc", ct.c_double * 8), ("wave_phase", ct.c_double * 8), ("poisson", ct.c_int32)]


# name -> (restype, argtypes); every symbol include/cubo_hip.h declares.
SIGNATURES = {
    "cubo_hip_device_count": (ct.c_int, [ct.POINTER(ct.c_int)]),
    "cubo_hip_get_device_info": (ct.c_int, [ct.c_int, ct.POINTER(DeviceInfo)]),
    "cubo_hip_last_error": (ct.c_int, [ct.c_char_p, ct.c_size_t]),
    "cubo_hip_availability": (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int64), ct.c_int, ct.c_uint64,
                                         ct.POINTER(ct.c_int), ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_uint64)]),
    "cubo_hip_cube_create": (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int64), ct.c_int, ct.POINTER(ct.c_void_p)]),
    "cubo_hip_cube_upload_rows": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_void_p]),
    "cubo_hip_cube_read_rows": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_void_p]),
    "cubo_hip_cube_read_patterns": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p]),
    "cubo_hip_cube_destroy": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_cube_clone": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_void_p)]),
    "cubo_hip_cube_byte_count": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_uint64)]),
    "cubo_hip_reduce_begin": (ct.c_int, [ct.c_void_p, ct.POINTER(ReduceRequest), ct.POINTER(ct.c_void_p)]),
    "cubo_hip_reduce_rows": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64]),
    "cubo_hip_reduce_end": (ct.c_int, [ct.c_void_p, ct.POINTER(ReduceResult)]),
    "cubo_hip_reduce_cancel": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_reduce_results": (ct.c_int, [ct.c_void_p, ct.POINTER(ReduceResult)]),
    "cubo_hip_reduce_reset": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_reduce_view": (ct.c_int, [ct.c_void_p, ct.POINTER(ReduceView)]),
    "cubo_hip_reduce_release": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_reduce": (ct.c_int, [ct.c_void_p, ct.POINTER(ReduceRequest), ct.POINTER(ReduceResult)]),
    "cubo_hip_reduce_noalloc": (ct.c_int, [ct.c_void_p, ct.POINTER(ReduceRequest), ct.POINTER(ReduceResult)]),
    "cubo_hip_reduce_path": (ct.c_int, [ct.c_void_p, ct.POINTER(ReduceRequest), ct.POINTER(ct.c_int)]),
    "cubo_hip_cube_kernel_time": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64)]),
    "cubo_hip_npy_read_header": (ct.c_int, [ct.c_char_p, ct.POINTER(NpyHeader)]),
    "cubo_hip_cube_load_file": (ct.c_int, [ct.c_int, ct.c_char_p, ct.c_int64, ct.c_int32, ct.POINTER(ct.c_int64),
                                           ct.POINTER(ct.c_void_p)]),
    "cubo_hip_cube_load_npy": (ct.c_int, [ct.c_int, ct.c_char_p, ct.POINTER(ct.c_void_p)]),
    "cubo_hip_cube_create_streamed": (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int64), ct.c_int, ct.c_int64,
                                                 ct.POINTER(ct.c_void_p)]),
    "cubo_hip_cube_feed_rows": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_void_p, ct.c_int32]),
    "cubo_hip_pattern_begin": (ct.c_int, [ct.c_void_p, ct.POINTER(PatternRequest), ct.POINTER(ct.c_void_p)]),
    "cubo_hip_pattern_rows": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64]),
    "cubo_hip_pattern_results": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "cubo_hip_pattern_reset": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_pattern_release": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_pattern_end": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "cubo_hip_pattern_cancel": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_pattern": (ct.c_int, [ct.c_void_p, ct.POINTER(PatternRequest), ct.c_void_p]),
    "cubo_hip_pattern_path": (ct.c_int, [ct.c_void_p, ct.POINTER(PatternRequest), ct.POINTER(ct.c_int)]),
    "cubo_hip_probe_size": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_double, ct.c_double, ct.c_int32,
                                       ct.POINTER(ProbeEstimate)]),
    "cubo_hip_origins": (ct.c_int, [ct.c_void_p, ct.POINTER(OriginRequest), ct.c_int64, ct.c_int64, ct.c_void_p,
                                    ct.c_void_p]),
    "cubo_hip_gaussian_filter_pattern": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_double, ct.c_void_p]),
    "cubo_hip_stack_plan_create": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_int32,
                                              ct.POINTER(ct.c_void_p)]),
    "cubo_hip_stack_plan_method": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32)]),
    "cubo_hip_stack_plan_run": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "cubo_hip_stack_plan_destroy": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_mask_stack": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_void_p]),
    "cubo_hip_batch_create": (ct.c_int, [ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_void_p)]),
    "cubo_hip_batch_tune": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32),
                                       ct.POINTER(ct.c_int32)]),
    "cubo_hip_batch_run": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "cubo_hip_batch_view": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.POINTER(ct.POINTER(ct.c_double))]),
    "cubo_hip_batch_time": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64)]),
    "cubo_hip_batch_destroy": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_shift_plan_create": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.POINTER(ct.c_void_p)]),
    "cubo_hip_shift_forward": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "cubo_hip_shift_forward_rounded": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "cubo_hip_shift_backward": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                           ct.c_void_p, ct.c_void_p]),
    "cubo_hip_shift_kernel_time": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int64)]),
    "cubo_hip_shift_jacobian_uses": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int64)]),
    "cubo_hip_shift_keep_jacobian": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "cubo_hip_shift_plan_destroy": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_phase_normalize": (ct.c_int, [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_int32,
                                            ct.c_double, ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                            ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "cubo_hip_phase_rotation": (ct.c_int, [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_void_p,
                                           ct.c_int32, ct.c_int32, ct.c_void_p, ct.c_int32, ct.c_void_p,
                                           ct.POINTER(ct.c_double), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32),
                                           ct.POINTER(ct.c_int32), ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                           ct.c_void_p]),
    "cubo_hip_dpc_create": (ct.c_int, [ct.c_int, ct.c_int32, ct.c_int32, ct.c_double, ct.c_double, ct.c_double,
                                       ct.POINTER(ct.c_void_p)]),
    "cubo_hip_dpc_padded_shape": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32)]),
    "cubo_hip_dpc_reconstruct": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_double, ct.c_double,
                                            ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int32),
                                            ct.POINTER(ct.c_double)]),
    "cubo_hip_dpc_step": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_void_p,
                                     ct.POINTER(ct.c_double)]),
    "cubo_hip_dpc_step_backward": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double,
                                              ct.c_void_p, ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "cubo_hip_dpc_chain_backward": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_double,
                                               ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                               ct.c_void_p]),
    "cubo_hip_dpc_time": (ct.c_int, [ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64)]),
    "cubo_hip_dpc_destroy": (ct.c_int, [ct.c_void_p]),
    "cubo_hip_cube_fill_synthetic": (ct.c_int, [ct.c_void_p, ct.POINTER(SynthParams)]),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """The loaded library. Raises LibraryMissing when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is not built; run `python cubokit_amd/build.py` "
                                 "(the HIP path has no fallback)")
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    buf = ct.create_string_buffer(512)
    lib().cubo_hip_last_error(buf, 512)
    return buf.value.decode(errors="replace")
