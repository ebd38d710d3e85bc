"""The hot-path stencils, called through the C ABI.

Each function takes DataStores and a Grid like ``stencil::run(spec, backend, grid, fields...)``
(stencil/frontend/run.hpp:241-250): fields are shifted to the grid origin (stencil/core/backend.hpp:28-34) and
handed to libgt355 with their allocation bounds, which the library checks against the stencil's extents.
Launches are asynchronous on ``stream`` (None = null stream).
"""
import ctypes

from . import _abi
from .grid import Grid

NATIVE = _abi.NATIVE
WIDEN = _abi.WIDEN


def _sid(ds, grid: Grid):
    return ds.sid(grid.origin)


def hdiff(out, in_, coeff, grid: Grid, precision=NATIVE, stream=None):
    """Horizontal diffusion (tests/regression/horizontal_diffusion.cpp:35-107)."""
    o, i, c = _sid(out, grid), _sid(in_, grid), _sid(coeff, grid)
    _abi.call("gt355_hdiff", precision, ctypes.byref(o), ctypes.byref(i), ctypes.byref(c), grid.i_size,
              grid.j_size, grid.k_size, stream)


def copy(out, in_, grid: Grid, stream=None):
    """copy_stencil (tests/regression/copy_stencil.cpp:23-44): out = in over the interior."""
    o, i = _sid(out, grid), _sid(in_, grid)
    _abi.call("gt355_copy", ctypes.byref(o), ctypes.byref(i), grid.i_size, grid.j_size, grid.k_size, stream)


def laplacian(out, in_, grid: Grid, precision=NATIVE, stream=None):
    """Laplacian (tests/regression/laplacian.cpp:20-38)."""
    o, i = _sid(out, grid), _sid(in_, grid)
    _abi.call("gt355_laplacian", precision, ctypes.byref(o), ctypes.byref(i), grid.i_size, grid.j_size,
              grid.k_size, stream)


def hdiff_vadv(in_, coeff, u_stage, utens_stage, wcon, u_pos, utens, dtr_stage, grid: Grid, precision=NATIVE,
               stream=None):
    """The hdiff -> vadv composition: u_stage = hdiff(in, coeff), then vertical advection of u_stage."""
    s = [_sid(f, grid) for f in (in_, coeff, u_stage, utens_stage, wcon, u_pos, utens)]
    _abi.call("gt355_hdiff_vadv", precision, *[ctypes.byref(x) for x in s], ctypes.c_double(dtr_stage),
              grid.i_size, grid.j_size, grid.k_size, stream)


def convert(dst, src, stream=None):
    """dst = src elementwise over the whole field (same lengths; the element type may differ)."""
    if dst.lengths != src.lengths:
        raise ValueError(f"lengths differ: {dst.lengths} vs {src.lengths}")
    _abi.call("gt355_convert", dst.code, ctypes.c_void_p(dst.base), _abi.c_i64_3(*dst.strides), src.code,
              ctypes.c_void_p(src.base), _abi.c_i64_3(*src.strides), _abi.c_i64_3(*src.lengths), stream)


def vadv(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, grid: Grid, precision=NATIVE, stream=None):
    """Vertical advection dycore (tests/regression/vertical_advection_dycore.cpp:32-150).  utens_stage is
    read and overwritten."""
    s = [_sid(f, grid) for f in (utens_stage, u_stage, wcon, u_pos, utens)]
    _abi.call("gt355_vadv", precision, *[ctypes.byref(x) for x in s], ctypes.c_double(dtr_stage), grid.i_size,
              grid.j_size, grid.k_size, stream)
