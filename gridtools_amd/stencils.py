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


def simple_hori_diff(out, in_, crlato, crlatu, coeff, grid: Grid, precision=NATIVE, stream=None):
    """simple_hori_diff (tests/regression/simple_hori_diff.cpp:25-85): wlap into an ij-cached lap, then divflux.
    crlato / crlatu are usually j-only stores (builder().selector(0, 1, 0))."""
    s = [_sid(f, grid) for f in (out, in_, crlato, crlatu, coeff)]
    _abi.call("gt355_simple_hori_diff", precision, *[ctypes.byref(x) for x in s], grid.i_size, grid.j_size,
              grid.k_size, stream)


def _field_array(fields, grid):
    arr = (_abi.Field * len(fields))()
    for n, f in enumerate(fields):
        arr[n] = _sid(f, grid)
    return arr


def prepare_tracers(data, data_nnow, rho, grid: Grid, precision=NATIVE, stream=None):
    """expandable_run of prepare_tracers (tests/regression/advection_pdbott_prepare_tracers.cpp:23-57):
    data[m] = rho * data_nnow[m] for every member of the two equally long lists, up to 32 members per launch."""
    if len(data) != len(data_nnow):
        raise ValueError(f"expandable_run: vector arguments differ in size ({len(data)} vs {len(data_nnow)})")
    d, n, r = _field_array(data, grid), _field_array(data_nnow, grid), _sid(rho, grid)
    _abi.call("gt355_prepare_tracers", precision, d, n, len(data), ctypes.byref(r), grid.i_size, grid.j_size,
              grid.k_size, stream)


def copy_fields(out, in_, grid: Grid, stream=None):
    """copy_stencil_tuple (tests/regression/copy_stencil_tuple.cpp:23-100): out[f] = in[f] for equally long lists
    of stores (or of gt355_field descriptors, e.g. the members of an array-of-structs buffer)."""
    if len(out) != len(in_):
        raise ValueError(f"copy_fields: {len(out)} outputs for {len(in_)} inputs")

    def arr(fields):
        a = (_abi.Field * len(fields))()
        for n, f in enumerate(fields):
            a[n] = f if isinstance(f, _abi.Field) else _sid(f, grid)
        return a

    _abi.call("gt355_copy_fields", arr(out), arr(in_), len(out), grid.i_size, grid.j_size, grid.k_size, stream)


def extended_4d(result, f, jac, phi, psi, grid: Grid, stream=None):
    """extended_4D (tests/regression/extended_4D.cpp:58-135): result and f are stores with extra dimensions
    (2, 2, 2), jac with (2,); phi / psi are the global parameters as 2x2x2x2 tables [x, y, z, q] (a scalar is the
    reference's constant elemental)."""
    import numpy as np

    if tuple(result.extra) != (2, 2, 2) or tuple(f.extra) != (2, 2, 2) or tuple(jac.extra) != (2,):
        raise ValueError("result and f need extra dimensions (2, 2, 2), jac needs (2,)")

    def table(t):
        a = np.broadcast_to(np.asarray(t, dtype=np.float64), (2, 2, 2, 2))
        return (ctypes.c_double * 16)(*np.ascontiguousarray(a).reshape(16))

    r, fs, js = _sid(result, grid), _sid(f, grid), _sid(jac, grid)
    _abi.call("gt355_extended_4d", table(phi), table(psi), ctypes.byref(r), _abi.c_i64_3(*result.extra_strides),
              ctypes.byref(fs), _abi.c_i64_3(*f.extra_strides), ctypes.byref(js), ctypes.c_int64(jac.extra_strides[0]),
              grid.i_size, grid.j_size, grid.k_size, stream)
