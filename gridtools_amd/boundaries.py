"""Boundary conditions and local halo handling on the device.

``make_boundary(halos, bc).apply(*fields)`` mirrors boundaries/boundary.hpp:57-79 for the reference's
predefined functors zero_boundary / value_boundary / copy_boundary (zero.hpp, value.hpp, copy.hpp), with the
direction predicate given as a 27-bit mask (the pre-evaluated predicate of boundaries/apply_gpu.hpp:52-83).
``halo_fill`` is the single-rank periodic exchange of GCL plus a zero-gradient boundary, and
``halo_pack``/``halo_unpack`` produce and consume the per-direction message buffers.
"""
import ctypes

from . import _abi
from .grid import halo_descriptor

ZERO, VALUE, COPY = 0, 1, 2
NONE, ZERO_GRADIENT, PERIODIC = 0, 1, 2

ALL_DIRECTIONS = 0x7FFFFFF & ~(1 << 13)


def direction_bit(di, dj, dk):
    return (di + 1) * 9 + (dj + 1) * 3 + (dk + 1)


def direction_mask(pred):
    """Pre-evaluates a predicate pred(di, dj, dk) -> bool over the 26 directions (apply_gpu.hpp:52-83)."""
    m = 0
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                if (di, dj, dk) != (0, 0, 0) and pred(di, dj, dk):
                    m |= 1 << direction_bit(di, dj, dk)
    return m


def _halo_arg(halos):
    if len(halos) != 3:
        raise ValueError("three halo descriptors are needed")
    rows = []
    for h in halos:
        if not isinstance(h, halo_descriptor):
            h = halo_descriptor(*h)
        rows.append((ctypes.c_int32 * 5)(*h.as_tuple()))
    return _abi.c_halo(*rows)


def _ptrs(fields):
    return (ctypes.c_void_p * len(fields))(*[f.base for f in fields])


def _common(fields):
    if not fields:
        raise ValueError("at least one field is needed")
    dt = fields[0].code
    st = fields[0].strides
    for f in fields:
        if f.code != dt or f.strides != st:
            raise ValueError("all fields of a boundary call must share dtype and layout")
    return dt, st


class boundary:
    def __init__(self, halos, kind, value=0.0, dir_mask=ALL_DIRECTIONS):
        self.halos = halos
        self.kind = kind
        self.value = float(value)
        self.dir_mask = dir_mask

    def apply(self, *fields, src=None, stream=None):
        """Applies the condition to ``fields`` (kind COPY copies from the fields in ``src``)."""
        dt, st = _common(fields)
        srcp = None
        if self.kind == COPY:
            if src is None or len(src) != len(fields):
                raise ValueError("copy boundary needs one source per field")
            _common(list(fields) + list(src))
            srcp = _ptrs(src)
        _abi.call("gt355_bc_apply", self.kind, dt, _ptrs(fields), srcp, len(fields), ctypes.c_double(self.value),
                  _halo_arg(self.halos), _abi.c_i64_3(*st), ctypes.c_uint32(self.dir_mask), stream)


def make_boundary(halos, kind, value=0.0, predicate=None):
    mask = ALL_DIRECTIONS if predicate is None else direction_mask(predicate)
    return boundary(halos, kind, value, mask)


def halo_fill(fields, halos, modes, stream=None):
    """Fills the halos of ``fields``: modes[d] in {NONE, ZERO_GRADIENT, PERIODIC} per dimension."""
    dt, st = _common(fields)
    _abi.call("gt355_halo_fill", dt, _ptrs(fields), len(fields), _halo_arg(halos), _abi.c_i64_3(*st),
              _abi.c_i32_3(*modes), stream)


def halo_buffer_size(dtype_code, nfields, halos, dir_mask=ALL_DIRECTIONS):
    n = ctypes.c_size_t()
    _abi.call("gt355_halo_buffer_size", dtype_code, nfields, _halo_arg(halos), ctypes.c_uint32(dir_mask),
              ctypes.byref(n))
    return n.value


def halo_pack(fields, halos, buffer_ptr, dir_mask=ALL_DIRECTIONS, stream=None):
    dt, st = _common(fields)
    _abi.call("gt355_halo_pack", dt, _ptrs(fields), len(fields), _halo_arg(halos), _abi.c_i64_3(*st),
              ctypes.c_uint32(dir_mask), ctypes.c_void_p(buffer_ptr), stream)


def halo_unpack(fields, halos, buffer_ptr, dir_mask=ALL_DIRECTIONS, stream=None):
    dt, st = _common(fields)
    _abi.call("gt355_halo_unpack", dt, _ptrs(fields), len(fields), _halo_arg(halos), _abi.c_i64_3(*st),
              ctypes.c_uint32(dir_mask), ctypes.c_void_p(buffer_ptr), stream)


def halo_exchange_local(fields, halos, buffer_ptr, dir_mask=ALL_DIRECTIONS, stream=None):
    """The single-rank periodic exchange: pack the slabs of dir_mask's directions into the buffer
    (halo_buffer_size bytes), then unpack every halo region d from the slab packed for -d."""
    dt, st = _common(fields)
    _abi.call("gt355_halo_exchange_local", dt, _ptrs(fields), len(fields), _halo_arg(halos), _abi.c_i64_3(*st),
              ctypes.c_uint32(dir_mask), ctypes.c_void_p(buffer_ptr), stream)
