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


def vadv(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage, grid: Grid, precision=NATIVE, stream=None):
    """Vertical advection dycore (tests/regression/vertical_advection_dycore.cpp:32-150).  utens_stage is
    read and overwritten."""
    s = [_sid(f, grid) for f in (utens_stage, u_stage, wcon, u_pos, utens)]
    _abi.call("gt355_vadv", precision, *[ctypes.byref(x) for x in s], ctypes.c_double(dtr_stage), grid.i_size,
              grid.j_size, grid.k_size, stream)
