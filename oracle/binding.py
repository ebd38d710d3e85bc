"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
never by the gridtools_amd package.  Arrays are numpy arrays in the device layout [k, j, i_padded] (as
returned by DataStore.padded_numpy()) or any C-contiguous [k, j, i] array; the origin of the interior and the
strides are derived from the halo.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_PD = ctypes.c_ssize_t

_SIGS = {
    "gto_fill_uniform_f64": [_P, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, ctypes.c_double],
    "gto_hdiff_f64": [_P, _P, _P, _I, _I, _I, _PD, _PD, _I],
    "gto_hdiff_f32": [_P, _P, _P, _I, _I, _I, _PD, _PD, _I],
    "gto_hdiff_mixed": [_P, _P, _P, _I, _I, _I, _PD, _PD, _I],
    "gto_lap_f64": [_P, _P, _I, _I, _I, _PD, _PD],
    "gto_lap_f32": [_P, _P, _I, _I, _I, _PD, _PD],
    "gto_vadv_f64": [_P, _P, _P, _P, _P, ctypes.c_double, _I, _I, _I, _PD, _PD, _PD, _PD, _I],
    "gto_vadv_f32": [_P, _P, _P, _P, _P, ctypes.c_float, _I, _I, _I, _PD, _PD, _PD, _PD, _I],
    "gto_vadv_mixed": [_P, _P, _P, _P, _P, ctypes.c_float, _I, _I, _I, _PD, _PD, _PD, _PD, _I],
    "gto_tridiagonal_f64": [_P, _P, _P, _P, _P, _I, _I, _I, _PD, _PD],
    "gto_transform_layout": [_I, _P, _P, _I, _P, _P, _P],
    "gto_hdiff_kfirst_f64": [_P, _P, _P, _I, _I, _I, _PD, _PD, _I],
    "gto_vadv_kfirst_f64": [_P, _P, _P, _P, _P, ctypes.c_double, _I, _I, _I, _PD, _PD, _PD, _PD, _I],
    "gto_shd_f64": [_P, _P, _P, _P, _PD, _PD, _PD, _P, _I, _I, _I, _PD, _PD],
    "gto_shd_f32": [_P, _P, _P, _P, _PD, _PD, _PD, _P, _I, _I, _I, _PD, _PD],
    "gto_shd_mixed": [_P, _P, _P, _P, _PD, _PD, _PD, _P, _I, _I, _I, _PD, _PD],
    "gto_extended_4d_f64": [_P, _P, _P, _P, _P, _PD, _P, _P, _I, _I, _I, _PD, _PD],
    "gto_extended_4d_f32": [_P, _P, _P, _P, _P, _PD, _P, _P, _I, _I, _I, _PD, _PD],
    "gto_prepare_tracers_f64": [_P, _P, _P, _I, _I, _I, _PD, _PD],
    "gto_prepare_tracers_f32": [_P, _P, _P, _I, _I, _I, _PD, _PD],
    "gto_bc_f64": [_I, _P, _P, ctypes.c_double, _P, _PD, _PD, ctypes.c_uint32],
    "gto_bc_f32": [_I, _P, _P, ctypes.c_float, _P, _PD, _PD, ctypes.c_uint32],
    "gto_halo_fill_f64": [_P, _P, _PD, _PD, _P],
    "gto_halo_fill_f32": [_P, _P, _PD, _PD, _P],
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int if name.startswith("gto_vadv") else None
        _lib = h
    return _lib


class View:
    """A [k, j, i] C-contiguous array with an interior origin at (hk, hj, hi)."""

    def __init__(self, arr, halo_i, halo_j, halo_k=0):
        if not arr.flags.c_contiguous:
            raise ValueError("oracle arrays must be C-contiguous [k, j, i]")
        self.arr = arr
        self.sj = arr.shape[2]
        self.sk = arr.shape[1] * arr.shape[2]
        self.off = halo_k * self.sk + halo_j * self.sj + halo_i
        self.halo = (halo_i, halo_j, halo_k)

    @property
    def ptr(self):
        return self.arr.ctypes.data + self.off * self.arr.itemsize


def _kind(arrs):
    dt = {a.arr.dtype for a in arrs}
    if len(dt) != 1:
        raise TypeError("mixed dtypes")
    return dt.pop()


def hdiff(out: View, in_: View, coeff: View, ni, nj, nk, mode="native", threads=1):
    """mode: native (compute in storage type), mixed (FP32 storage, FP64 compute) or kfirst (FP64, the cpu_kfirst
    traversal)."""
    dt = _kind([out, in_, coeff])
    assert out.sj == in_.sj == coeff.sj and out.sk == in_.sk == coeff.sk
    fn = {("float64", "native"): "gto_hdiff_f64", ("float32", "native"): "gto_hdiff_f32",
          ("float32", "mixed"): "gto_hdiff_mixed", ("float64", "kfirst"): "gto_hdiff_kfirst_f64"}[(dt.name, mode)]
    getattr(lib(), fn)(out.ptr, in_.ptr, coeff.ptr, ni, nj, nk, out.sj, out.sk, threads)


def laplacian(out: View, in_: View, ni, nj, nk):
    """laplacian.cpp:20-29 in the storage type (float64 or float32; nothing else is accepted)."""
    dt = _kind([out, in_])
    assert out.sj == in_.sj and out.sk == in_.sk
    fn = {"float64": "gto_lap_f64", "float32": "gto_lap_f32"}[dt.name]
    getattr(lib(), fn)(out.ptr, in_.ptr, ni, nj, nk, out.sj, out.sk)


def vadv(uts: View, us: View, wc: View, up: View, ut: View, dtr, ni, nj, nk, mode="native", threads=1):
    dt = _kind([uts, us, wc, up, ut])
    fn = {("float64", "native"): "gto_vadv_f64", ("float32", "native"): "gto_vadv_f32",
          ("float32", "mixed"): "gto_vadv_mixed", ("float64", "kfirst"): "gto_vadv_kfirst_f64"}[(dt.name, mode)]
    rc = getattr(lib(), fn)(uts.ptr, us.ptr, wc.ptr, up.ptr, ut.ptr, dtr, ni, nj, nk, uts.sj, uts.sk, wc.sj, wc.sk,
                            threads)
    if rc != 0:
        raise ValueError(f"{fn} rejected nk={nk}")


def simple_hori_diff(out: View, in_: View, crlato, crlatu, coeff: View, ni, nj, nk, mode="native"):
    """simple_hori_diff.cpp:25-80.  crlato / crlatu: either 1-D C-contiguous arrays over the storage j index (the
    j-only fields of selector<0,1,0>; element halo_j + j belongs to grid row j, rows -1 .. nj are read) or Views
    of full fields laid out like ``in_`` (read over i in [-1, ni], j in [-1, nj])."""
    dt = _kind([out, in_, coeff])
    assert out.sj == in_.sj == coeff.sj and out.sk == in_.sk == coeff.sk
    hj = in_.halo[1]
    ptrs = []
    for c in (crlato, crlatu):
        if isinstance(c, View):
            if c.arr.dtype != dt or c.sj != in_.sj or c.sk != in_.sk or min(c.halo[:2]) < 1:
                raise TypeError("a 3-D metric field must be laid out like in and have a 1-point halo")
            ptrs.append(c.ptr)
            strides = (1, c.sj, c.sk)
        else:
            if c.dtype != dt or c.ndim != 1 or not c.flags.c_contiguous or c.shape[0] < hj + nj + 1 or hj < 1:
                raise TypeError("crlato / crlatu: 1-D arrays of the field dtype covering rows -1 .. nj")
            ptrs.append(c.ctypes.data + hj * c.itemsize)
            strides = (0, 1, 0)
    if isinstance(crlato, View) != isinstance(crlatu, View):
        raise TypeError("crlato and crlatu must be both j-only or both 3-D")
    fn = {("float64", "native"): "gto_shd_f64", ("float32", "native"): "gto_shd_f32",
          ("float32", "mixed"): "gto_shd_mixed"}[(dt.name, mode)]
    getattr(lib(), fn)(out.ptr, in_.ptr, ptrs[0], ptrs[1], *strides, coeff.ptr, ni, nj, nk, out.sj, out.sk)


def prepare_tracers(data: View, data_nnow: View, rho: View, ni, nj, nk):
    """advection_pdbott_prepare_tracers.cpp:23-34 for one member: data = rho * data_nnow."""
    dt = _kind([data, data_nnow, rho])
    assert data.sj == data_nnow.sj == rho.sj and data.sk == data_nnow.sk == rho.sk
    fn = {"float64": "gto_prepare_tracers_f64", "float32": "gto_prepare_tracers_f32"}[dt.name]
    getattr(lib(), fn)(data.ptr, data_nnow.ptr, rho.ptr, ni, nj, nk, data.sj, data.sk)


def extended_4d(result, f, jac, phi, psi, ni, nj, nk):
    """extended_4D.cpp:58-91.  result / f: C-contiguous [c, b, a, k, j, i] arrays (the extra dimensions slowest, as
    the gpu layout orders them), jac: [q, k, j, i]; all share the spatial shape.  phi / psi: 2x2x2x2 tables
    [x, y, z, q] (float64: the reference's elemental returns double)."""
    dt = result.dtype
    if f.dtype != dt or jac.dtype != dt or dt not in (np.float64, np.float32):
        raise TypeError("result, f and jac must share a float64 or float32 dtype")
    for a in (result, f, jac):
        if not a.flags.c_contiguous or a.shape[-3:] != result.shape[-3:]:
            raise TypeError("C-contiguous arrays with one spatial shape")
    if result.shape[:3] != (2, 2, 2) or f.shape[:3] != (2, 2, 2) or jac.shape[0] != 2:
        raise TypeError("result / f have extra dimensions 2x2x2, jac has 2 quadrature points")
    nz, ny, nx = result.shape[-3:]
    sj, sk, vol = nx, nx * ny, nx * ny * nz
    ext = np.array([vol, 2 * vol, 4 * vol], dtype=np.int64)  # strides of the extra dimensions (a, b, c)
    tphi = np.ascontiguousarray(np.asarray(phi, dtype=np.float64).reshape(16))
    tpsi = np.ascontiguousarray(np.asarray(psi, dtype=np.float64).reshape(16))
    fn = "gto_extended_4d_f64" if dt == np.float64 else "gto_extended_4d_f32"
    getattr(lib(), fn)(result.ctypes.data, ext.ctypes.data, f.ctypes.data, ext.ctypes.data, jac.ctypes.data, vol,
                       tphi.ctypes.data, tpsi.ctypes.data, ni, nj, nk, sj, sk)


def transform_layout(dst, src, dims, dst_strides, src_strides):
    """layout_transformation/cpu.hpp:20-55 on numpy buffers: dst / src are C-contiguous arrays of one dtype whose
    flat element n is index n of the strided layouts (strides in elements)."""
    if dst.dtype != src.dtype or not dst.flags.c_contiguous or not src.flags.c_contiguous:
        raise TypeError("dst and src: C-contiguous arrays of one dtype")
    n = len(dims)
    d, ds, ss = (np.ascontiguousarray(np.asarray(x, dtype=np.int64)) for x in (dims, dst_strides, src_strides))
    if ds.size != n or ss.size != n:
        raise ValueError("dims and strides differ in length")
    lib().gto_transform_layout(dst.itemsize, dst.ctypes.data, src.ctypes.data, n, d.ctypes.data, ds.ctypes.data,
                               ss.ctypes.data)


def tridiagonal(out: View, inf: View, diag: View, sup: View, rhs: View, ni, nj, nk):
    lib().gto_tridiagonal_f64(out.ptr, inf.ptr, diag.ptr, sup.ptr, rhs.ptr, ni, nj, nk, out.sj, out.sk)


def _halo_arr(halos):
    return np.ascontiguousarray(np.array(halos, dtype=np.int32).reshape(3, 5))


def _full(arr, like=None):
    """A full [k, j, i] float64/float32 C-contiguous array (and the same dtype and shape as ``like``)."""
    if arr.dtype not in (np.float64, np.float32) or not arr.flags.c_contiguous or arr.ndim != 3:
        raise TypeError("oracle arrays are C-contiguous 3-D float64 or float32")
    if like is not None and (arr.dtype != like.dtype or arr.shape != like.shape):
        raise TypeError("src must match dst in dtype and shape")
    return arr


def bc(kind, dst, src, value, halos, dir_mask):
    """dst/src: full [k, j, i] arrays (storage index (0,0,0) at element 0)."""
    _full(dst)
    if src is not None:
        _full(src, dst)
    h = _halo_arr(halos)
    sj, sk = dst.shape[2], dst.shape[1] * dst.shape[2]
    srcp = src.ctypes.data if src is not None else None
    if dst.dtype == np.float64:
        lib().gto_bc_f64(kind, dst.ctypes.data, srcp, value, h.ctypes.data, sj, sk, dir_mask)
    else:
        lib().gto_bc_f32(kind, dst.ctypes.data, srcp, value, h.ctypes.data, sj, sk, dir_mask)


def halo_fill(f, halos, modes):
    _full(f)
    h = _halo_arr(halos)
    m = np.ascontiguousarray(np.array(modes, dtype=np.int32))
    sj, sk = f.shape[2], f.shape[1] * f.shape[2]
    if f.dtype == np.float64:
        lib().gto_halo_fill_f64(f.ctypes.data, h.ctypes.data, sj, sk, m.ctypes.data)
    else:
        lib().gto_halo_fill_f32(f.ctypes.data, h.ctypes.data, sj, sk, m.ctypes.data)


def fill_uniform(n, seed, field_id, a, b):
    out = np.empty(n, dtype=np.float64)
    lib().gto_fill_uniform_f64(out.ctypes.data, n, seed, field_id, a, b)
    return out
