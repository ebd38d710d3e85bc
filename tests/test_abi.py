"""CPU checks of the C ABI: the library loads, exports every entry point include/gt355.h declares, and the
pure host-side entry points (layout arithmetic, argument validation) behave like the reference storage."""
import ctypes
import os
import re

import numpy as np
import pytest

from gridtools_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(gt355_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.lib()
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for every entry point
    assert names <= set(_abi.SIGNATURES), sorted(names - set(_abi.SIGNATURES))


def test_version_string():
    assert b"gfx950" in _abi.lib().gt355_version()


@pytest.mark.parametrize("elem,lengths,halos,padded_i", [
    (8, (262, 262, 80), (3, 3, 0), 272),   # vadv FP64 256^2 + 3-halo: 262 -> 272 (multiple of 16 doubles)
    (8, (68, 68, 60), (2, 2, 0), 80),
    (4, (1542, 1542, 80), (3, 3, 0), 1568),  # FP32: multiple of 32 floats
    (4, (32, 5, 3), (0, 0, 0), 32),
])
def test_storage_layout_matches_reference_padding(elem, lengths, halos, padded_i):
    """storage/info.hpp:38-85: i padded to 128/gcd(sizeof T,128) elements; layout_map<2,1,0> strides."""
    lay = _abi.Layout()
    _abi.call("gt355_storage_layout", elem, _abi.c_i64_3(*lengths), _abi.c_i64_3(*halos), ctypes.byref(lay))
    assert lay.padded[0] == padded_i
    assert tuple(lay.stride) == (1, padded_i, padded_i * lengths[1])
    assert lay.alloc_bytes == (padded_i * lengths[1] * lengths[2] + 128 // elem) * elem


def test_storage_base_aligns_first_interior_point():
    """storage/data_store.hpp:63-72: the first interior point lands on a 128-byte boundary."""
    lay = _abi.Layout()
    _abi.call("gt355_storage_layout", 8, _abi.c_i64_3(262, 262, 80), _abi.c_i64_3(3, 3, 0), ctypes.byref(lay))
    for raw in (0x10000, 0x10008, 0x10078, 0x123458):
        base = ctypes.c_void_p()
        _abi.call("gt355_storage_base", ctypes.byref(lay), ctypes.c_void_p(raw), ctypes.byref(base))
        first = base.value + (3 + 3 * lay.stride[1]) * 8
        assert first % 128 == 0
        assert raw <= base.value < raw + 128


def test_bad_arguments_are_rejected_without_a_device():
    lib = _abi.lib()
    lay = _abi.Layout()
    assert lib.gt355_storage_layout(3, _abi.c_i64_3(4, 4, 4), _abi.c_i64_3(0, 0, 0), ctypes.byref(lay)) != 0
    assert "element size" in _abi.last_error()
    assert lib.gt355_storage_layout(8, _abi.c_i64_3(4, 4, 4), _abi.c_i64_3(2, 0, 0), ctypes.byref(lay)) != 0
    # stencil entry points validate before touching the device
    f = _abi.Field()
    assert lib.gt355_hdiff(0, None, ctypes.byref(f), ctypes.byref(f), 4, 4, 4, None) != 0
    assert lib.gt355_vadv(0, ctypes.byref(f), ctypes.byref(f), ctypes.byref(f), ctypes.byref(f), ctypes.byref(f),
                          ctypes.c_double(0.15), 4, 4, 1, None) != 0
    assert "nk = 1" in _abi.last_error()


def test_hdiff_bounds_check_rejects_small_halo():
    """A field whose allocation cannot hold the 2-point rim hdiff reads is rejected before launch (the role of
    the reference's debug field-bounds check, stencil/frontend/run.hpp:215-232)."""
    lib = _abi.lib()
    f = _abi.Field()
    f.origin = 0x100000
    f.stride = (ctypes.c_int64 * 3)(1, 80, 80 * 68)
    f.lower = (ctypes.c_int32 * 3)(-1, -1, 0)  # only a 1-point halo
    f.upper = (ctypes.c_int32 * 3)(65, 65, 60)
    f.dtype = _abi.F64
    rc = lib.gt355_hdiff(0, ctypes.byref(f), ctypes.byref(f), ctypes.byref(f), 64, 64, 60, None)
    assert rc == -1
    assert "outside bounds" in _abi.last_error()


def test_tune_table_save_load_roundtrip(tmp_path):
    """The autotuner's table persists as text and reloads (no device needed)."""
    from gridtools_amd import tune

    tune.clear()
    src = tmp_path / "in.txt"
    src.write_text("hdiff 8 8 256 256 80 0 64 8 5 24.900\nhdiff 4 8 256 256 80 1 16 4 0 17.100\n")
    assert tune.load(src) == 2
    out = tmp_path / "out.txt"
    tune.save(out)
    lines = sorted(out.read_text().strip().splitlines())
    assert lines == sorted(["hdiff 8 8 256 256 80 0 64 8 5 24.900", "hdiff 4 8 256 256 80 1 16 4 0 17.100"])
    tune.clear()
    tune.save(out)
    assert out.read_text() == ""


def test_batch_create_validates_domains_without_a_device():
    """gt355_batch_create rejects bad domains (k depth beyond the on-chip limit, undersized halos) before it
    touches the device."""
    from gridtools_amd.batch import BatchDomain

    arr = (BatchDomain * 1)()
    d = arr[0]
    for n in ("in_", "coeff", "lap", "u_stage", "utens_stage", "wcon", "u_pos", "utens"):
        f = getattr(d, n)
        f.origin = 0x100000
        f.stride = (ctypes.c_int64 * 3)(1, 96, 96 * 72)
        f.lower = (ctypes.c_int32 * 3)(-3, -3, 0)
        f.upper = (ctypes.c_int32 * 3)(67, 67, 150)
        f.dtype = _abi.F32
    d.ni, d.nj, d.nk = 64, 64, 145
    h = ctypes.c_void_p()
    lib = _abi.lib()
    assert lib.gt355_batch_create(_abi.WIDEN, arr, 1, 2, ctypes.byref(h)) == -1
    assert "145" in _abi.last_error()
    d.nk = 60
    d.in_.lower = (ctypes.c_int32 * 3)(-1, -1, 0)
    assert lib.gt355_batch_create(_abi.WIDEN, arr, 1, 2, ctypes.byref(h)) == -1
    assert "outside bounds" in _abi.last_error()
    assert lib.gt355_batch_create(_abi.WIDEN, arr, 1, 0, ctypes.byref(h)) == -1
    assert not h.value


def test_batch_entry_points_reject_null_handles_without_a_device():
    lib = _abi.lib()
    n, g = ctypes.c_int(), ctypes.c_int()
    assert lib.gt355_batch_run(None, ctypes.c_double(0.15), None) == -1
    assert lib.gt355_batch_info(None, ctypes.byref(n), ctypes.byref(g)) == -1
    w, us = ctypes.c_int(), ctypes.c_double()
    assert lib.gt355_batch_tune(None, 3, None, ctypes.byref(w), ctypes.byref(us)) == -1
    assert lib.gt355_batch_destroy(None) == 0
    assert "null batch" in _abi.last_error()


def test_tune_hdiff_rejects_bad_arguments_without_a_device():
    lib = _abi.lib()
    f = _abi.Field()
    best = (ctypes.c_int32 * 4)()
    us = ctypes.c_double()
    assert lib.gt355_tune_hdiff(0, None, ctypes.byref(f), ctypes.byref(f), 8, 8, 8, 5, None, best,
                                ctypes.byref(us)) == -1


def test_range_markers_need_no_device():
    from gridtools_amd.timing import Range

    with Range("outer"):
        with Range("inner"):
            pass
    assert _abi.lib().gt355_range_push(None) == -1


def _field(origin, strides, lower, upper, dtype=None):
    f = _abi.Field()
    f.origin = origin
    f.stride = (ctypes.c_int64 * 3)(*strides)
    f.lower = (ctypes.c_int32 * 3)(*lower)
    f.upper = (ctypes.c_int32 * 3)(*upper)
    f.dtype = _abi.F64 if dtype is None else dtype
    return f


def test_cartesian_entry_points_validate_without_a_device():
    """simple_hori_diff / prepare_tracers / copy_fields reject bad fields before any launch: a masked (stride 0)
    dimension is accepted only on read-only metric / input fields, the metric box is checked in its unmasked
    dimension, members must share a layout, and a written field's index map must be injective."""
    lib = _abi.lib()
    ni, nj, nk = 64, 32, 4
    full = _field(0x100000, (1, 80, 80 * 40), (-2, -2, 0), (66, 34, nk))
    jonly = _field(0x200000, (0, 1, 0), (-(1 << 30), -2, -(1 << 30)), (1 << 30, 34, 1 << 30))
    short = _field(0x200000, (0, 1, 0), (-(1 << 30), 0, -(1 << 30)), (1 << 30, 34, 1 << 30))  # no row -1
    b = ctypes.byref
    # a j-only metric passes validation; a null origin elsewhere then fails, so nothing launches
    nul = _field(0, (1, 80, 80 * 40), (-2, -2, 0), (66, 34, nk))
    assert lib.gt355_simple_hori_diff(0, b(nul), b(full), b(jonly), b(jonly), b(full), ni, nj, nk, None) == -1
    assert "null origin" in _abi.last_error()
    assert lib.gt355_simple_hori_diff(0, b(full), b(full), b(short), b(jonly), b(full), ni, nj, nk, None) == -1
    assert "crlato" in _abi.last_error() and "outside bounds" in _abi.last_error()
    # precision form needs FP32 storage
    assert lib.gt355_simple_hori_diff(1, b(full), b(full), b(jonly), b(jonly), b(full), ni, nj, nk, None) == -1
    # prepare_tracers: members with different layouts
    other = _field(0x300000, (1, 96, 96 * 40), (-2, -2, 0), (66, 34, nk))
    data = (_abi.Field * 2)(full, other)
    nnow = (_abi.Field * 2)(full, full)
    assert lib.gt355_prepare_tracers(0, data, nnow, 2, b(full), ni, nj, nk, None) == -1
    assert "layout" in _abi.last_error()
    # copy_fields: a masked output, then an output whose i stride overlaps its rows
    outs = (_abi.Field * 1)(jonly)
    ins = (_abi.Field * 1)(full)
    assert lib.gt355_copy_fields(outs, ins, 1, ni, nj, nk, None) == -1
    assert "masked" in _abi.last_error()
    overlap = _field(0x400000, (2, 80, 80 * 40), (0, 0, 0), (ni, nj, nk))
    outs = (_abi.Field * 1)(overlap)
    assert lib.gt355_copy_fields(outs, ins, 1, ni, nj, nk, None) == -1
    assert "overlap" in _abi.last_error()
    # empty calls are no-ops
    assert lib.gt355_copy_fields(None, None, 0, ni, nj, nk, None) == 0
    assert lib.gt355_prepare_tracers(0, None, None, 0, b(full), ni, nj, nk, None) == 0


def test_extended_4d_validates_without_a_device():
    lib = _abi.lib()
    ni, nj, nk = 8, 4, 2
    sp = _field(0x100000, (1, 16, 64), (0, 0, 0), (ni, nj, nk))
    tab = (ctypes.c_double * 16)(*([1.0] * 16))
    vol = 64 * nk
    ok3 = _abi.c_i64_3(vol, 2 * vol, 4 * vol)
    # result slices that overlap each other are rejected
    bad3 = _abi.c_i64_3(vol, vol, 4 * vol)
    rc = lib.gt355_extended_4d(tab, tab, ctypes.byref(sp), bad3, ctypes.byref(sp), ok3, ctypes.byref(sp),
                               ctypes.c_int64(vol), ni, nj, nk, None)
    assert rc == -1 and "overlap" in _abi.last_error()
    # a quadrature stride inside the spatial box is rejected
    rc = lib.gt355_extended_4d(tab, tab, ctypes.byref(sp), ok3, ctypes.byref(sp), ok3, ctypes.byref(sp),
                               ctypes.c_int64(8), ni, nj, nk, None)
    assert rc == -1 and "quadrature" in _abi.last_error()
    assert lib.gt355_extended_4d(None, tab, ctypes.byref(sp), ok3, ctypes.byref(sp), ok3, ctypes.byref(sp),
                                 ctypes.c_int64(vol), ni, nj, nk, None) == -1


def test_transform_layout_validates_without_a_device():
    lib = _abi.lib()
    arr = ctypes.c_int64 * 2
    # overlapping destination (two dimensions with one stride), bad element size, too many dimensions
    assert lib.gt355_transform_layout(8, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), 2, arr(4, 5), arr(1, 1),
                                      arr(1, 4), None) == -1
    assert "overlap" in _abi.last_error()
    assert lib.gt355_transform_layout(3, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), 2, arr(4, 5), arr(1, 4),
                                      arr(1, 4), None) == -1
    assert lib.gt355_transform_layout(8, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), 9, arr(4, 5), arr(1, 4),
                                      arr(1, 4), None) == -1
    # an empty box is a no-op
    assert lib.gt355_transform_layout(8, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), 2, arr(0, 5), arr(1, 4),
                                      arr(1, 4), None) == 0


def test_deferred_scope_bookkeeping_without_a_device():
    """gt355_batch_begin / flush / end: one scope per stream and thread; an empty scope launches nothing."""
    lib = _abi.lib()
    assert lib.gt355_batch_flush(None) == 0  # no scope open: nothing to run
    assert lib.gt355_batch_end(None) == -1
    assert "no scope" in _abi.last_error()
    assert lib.gt355_batch_begin(None) == 0
    assert lib.gt355_batch_begin(None) == -1
    assert "already open" in _abi.last_error()
    other = ctypes.c_void_p(0x1234)  # a scope is kept per stream handle; this one is never launched on
    assert lib.gt355_batch_begin(other) == 0
    assert lib.gt355_batch_flush(None) == 0
    assert lib.gt355_batch_end(other) == 0
    assert lib.gt355_batch_end(None) == 0
    assert lib.gt355_batch_end(None) == -1


def test_tune_table_lookup_and_store_by_key_without_a_device(tmp_path):
    """gt355_tune_lookup / gt355_tune_store: the tuning table by kernel key; a stored hdiff entry must be an
    instantiated configuration and survives save / load."""
    from gridtools_amd import tune

    tune.clear()
    shape = (4, 8, 256, 256, 80)
    assert tune.lookup("hdiff", shape) is None
    tune.store("hdiff", shape, [2, 8, 4, 0], 13.5)
    assert tune.lookup("hdiff", shape) == ([2, 8, 4, 0], 13.5)
    with pytest.raises(RuntimeError, match="not instantiated"):
        tune.store("hdiff", shape, [2, 7, 4, 0])
    with pytest.raises(RuntimeError, match="unknown kernel key"):
        tune.lookup("vadv", shape)
    bshape = (0x1234abcd5678, 4, 8, 3)
    tune.store("batch", bshape, [16, 4, 2, 89], 2130.0)
    out = tmp_path / "t.txt"
    tune.save(out)
    tune.clear()
    assert tune.lookup("batch", bshape) is None
    assert tune.load(out) == 2
    assert tune.lookup("batch", bshape) == ([16, 4, 2, 89], 2130.0)
    assert tune.lookup("hdiff", shape) == ([2, 8, 4, 0], 13.5)
    tune.clear()
