"""ctypes binding of libgt355.so (declared in include/gt355.h).

This is the same binding a Python caller of the C ABI would write; the C++ headers under
gridtools_amd/include bind the same symbols natively.  Every call is checked: a nonzero status raises
RuntimeError with gt355_last_error(), the reference's error convention (common/cuda_util.hpp:20-35).
The library must exist: there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgt355.so")

F64 = 0
F32 = 1
NATIVE = 0
WIDEN = 1

c_i64_3 = ctypes.c_int64 * 3
c_i32_3 = ctypes.c_int32 * 3
c_halo = (ctypes.c_int32 * 5) * 3


class Field(ctypes.Structure):
    """gt355_field: origin, strides, bounds (the SID triple, sid/concept.hpp:36-120)."""

    _fields_ = [
        ("origin", ctypes.c_void_p),
        ("stride", ctypes.c_int64 * 3),
        ("lower", ctypes.c_int32 * 3),
        ("upper", ctypes.c_int32 * 3),
        ("dtype", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Layout(ctypes.Structure):
    _fields_ = [
        ("lengths", ctypes.c_int64 * 3),
        ("halos", ctypes.c_int64 * 3),
        ("padded", ctypes.c_int64 * 3),
        ("stride", ctypes.c_int64 * 3),
        ("elem_size", ctypes.c_int64),
        ("alloc_bytes", ctypes.c_int64),
    ]


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("cu_count", ctypes.c_int32),
        ("xcd_count", ctypes.c_int32),
        ("hbm_bytes", ctypes.c_int64),
        ("lds_per_cu", ctypes.c_int64),
        ("l2_bytes", ctypes.c_int64),
        ("clock_khz", ctypes.c_int32),
        ("mem_clock_khz", ctypes.c_int32),
        ("mem_bus_width", ctypes.c_int32),
        ("wavefront", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_FP = ctypes.POINTER(Field)

# name -> argtypes (every entry point returns int unless listed in _RESTYPES)
SIGNATURES = {
    "gt355_init": [_I],
    "gt355_finalize": [],
    "gt355_tmp_acquire": [ctypes.c_size_t, _P, ctypes.POINTER(_P)],
    "gt355_tmp_release": [_P, _P],
    "gt355_tmp_release_n": [_P, _I, _P],
    "gt355_stream_set_current": [_P],
    "gt355_stream_current": [ctypes.POINTER(_P)],
    "gt355_debug_sync": [_I],
    "gt355_tune_env_refresh": [],
    "gt355_last_error": [],
    "gt355_version": [],
    "gt355_get_device_info": [ctypes.POINTER(DeviceInfo)],
    "gt355_storage_layout": [ctypes.c_int64, c_i64_3, c_i64_3, ctypes.POINTER(Layout)],
    "gt355_storage_base": [ctypes.POINTER(Layout), _P, ctypes.POINTER(_P)],
    "gt355_malloc": [ctypes.c_size_t, ctypes.POINTER(_P)],
    "gt355_free": [_P],
    "gt355_memcpy_h2d": [_P, _P, ctypes.c_size_t],
    "gt355_memcpy_d2h": [_P, _P, ctypes.c_size_t],
    "gt355_memcpy_d2d": [_P, _P, ctypes.c_size_t, _P],
    "gt355_memset": [_P, _I, ctypes.c_size_t, _P],
    "gt355_synchronize": [],
    "gt355_stream_synchronize": [_P],
    "gt355_fill_uniform": [_I, _P, c_i64_3, c_i64_3, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                           ctypes.c_double, _P],
    "gt355_hdiff": [_I, _FP, _FP, _FP, _I, _I, _I, _P],
    "gt355_laplacian": [_I, _FP, _FP, _I, _I, _I, _P],
    "gt355_copy": [_FP, _FP, _I, _I, _I, _P],
    "gt355_hdiff_vadv": [_I, _FP, _FP, _FP, _FP, _FP, _FP, _FP, ctypes.c_double, _I, _I, _I, _P],
    "gt355_convert": [_I, _P, c_i64_3, _I, _P, c_i64_3, c_i64_3, _P],
    "gt355_transform_layout": [_I, _P, _P, _I, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                               ctypes.POINTER(ctypes.c_int64), _P],
    "gt355_vadv": [_I, _FP, _FP, _FP, _FP, _FP, ctypes.c_double, _I, _I, _I, _P],
    "gt355_vadv_noalloc": [_I, _FP, _FP, _FP, _FP, _FP, ctypes.c_double, _I, _I, _I, _P],
    "gt355_simple_hori_diff": [_I, _FP, _FP, _FP, _FP, _FP, _I, _I, _I, _P],
    "gt355_prepare_tracers": [_I, _FP, _FP, _I, _FP, _I, _I, _I, _P],
    "gt355_copy_fields": [_FP, _FP, _I, _I, _I, _I, _P],
    "gt355_extended_4d": [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), _FP, c_i64_3, _FP,
                          c_i64_3, _FP, ctypes.c_int64, _I, _I, _I, _P],
    "gt355_bc_apply": [_I, _I, ctypes.POINTER(_P), ctypes.POINTER(_P), _I, ctypes.c_double, c_halo, c_i64_3,
                       ctypes.c_uint32, _P],
    "gt355_halo_fill": [_I, ctypes.POINTER(_P), _I, c_halo, c_i64_3, c_i32_3, _P],
    "gt355_halo_buffer_size": [_I, _I, c_halo, ctypes.c_uint32, ctypes.POINTER(ctypes.c_size_t)],
    "gt355_halo_pack": [_I, ctypes.POINTER(_P), _I, c_halo, c_i64_3, ctypes.c_uint32, _P, _P],
    "gt355_halo_unpack": [_I, ctypes.POINTER(_P), _I, c_halo, c_i64_3, ctypes.c_uint32, _P, _P],
    "gt355_halo_exchange_local": [_I, ctypes.POINTER(_P), _I, c_halo, c_i64_3, ctypes.c_uint32, _P, _P],
    "gt355_stream_create": [ctypes.POINTER(_P)],
    "gt355_stream_destroy": [_P],
    "gt355_event_create": [ctypes.POINTER(_P)],
    "gt355_event_destroy": [_P],
    "gt355_event_record": [_P, _P],
    "gt355_event_synchronize": [_P],
    "gt355_event_elapsed_ms": [_P, _P, ctypes.POINTER(ctypes.c_float)],
    "gt355_mall_flush": [_P],
    "gt355_range_push": [ctypes.c_char_p],
    "gt355_range_pop": [],
    "gt355_tune_hdiff": [_I, _FP, _FP, _FP, _I, _I, _I, _I, _P, ctypes.POINTER(ctypes.c_int32),
                         ctypes.POINTER(ctypes.c_double)],
    "gt355_tune_save": [ctypes.c_char_p],
    "gt355_tune_load": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)],
    "gt355_tune_clear": [],
    "gt355_graph_begin": [_P],
    "gt355_graph_end": [_P, ctypes.POINTER(_P)],
    "gt355_graph_launch": [_P, _P],
    "gt355_graph_destroy": [_P],
    "gt355_batch_create": [_I, _P, _I, _I, ctypes.POINTER(_P)],
    "gt355_batch_run": [_P, ctypes.c_double, _P],
    "gt355_batch_info": [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "gt355_batch_destroy": [_P],
    "gt355_batch_tune": [_P, _I, _P, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double)],
    "gt355_batch_get_config": [_P, ctypes.POINTER(ctypes.c_int32)],
    "gt355_batch_set_config": [_P, ctypes.POINTER(ctypes.c_int32)],
    "gt355_tune_lookup": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), _I, ctypes.POINTER(ctypes.c_int32),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)],
    "gt355_tune_store": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), _I, ctypes.POINTER(ctypes.c_int32),
                         ctypes.c_double],
    "gt355_batch_begin": [_P],
    "gt355_batch_flush": [_P],
    "gt355_batch_end": [_P],
}
_RESTYPES = {"gt355_last_error": ctypes.c_char_p, "gt355_version": ctypes.c_char_p}

_lib = None


def lib():
    """Loads libgt355.so (once) and declares every signature.  Raises if the library was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gridtools_amd.build`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = handle
    return _lib


def last_error() -> str:
    return lib().gt355_last_error().decode()


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise RuntimeError(f"{what or 'gt355'} failed ({rc}): {last_error()}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
