"""Device storage with the reference's gpu storage semantics.

Mirrors ``storage::builder<Traits>.type<T>().dimensions(..).halos(..).initializer(f)/value(v).build()``
(storage/builder.hpp:168-326) and ``data_store`` (storage/data_store.hpp:33-222) for 3-D fields:
i-innermost layout with the i length padded to 128 bytes, first interior point 128-byte aligned
(storage/gpu.hpp:34-105, storage/info.hpp:38-85, storage/data_store.hpp:63-72).  The layout arithmetic is
done by the library (gt355_storage_layout / gt355_storage_base) so that Python, C++ and the kernels agree.
"""
import ctypes

import numpy as np

from . import _abi

_DTYPES = {np.dtype(np.float64): _abi.F64, np.dtype(np.float32): _abi.F32}


def dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt not in _DTYPES:
        raise TypeError(f"unsupported storage type {dt}; use float64 or float32")
    return _DTYPES[dt]


class DataStore:
    """A 3-D field in device memory (lengths include the halo)."""

    def __init__(self, dtype, lengths, halos=(0, 0, 0), name="", selector=(1, 1, 1)):
        self.dtype = np.dtype(dtype)
        self.code = dtype_code(self.dtype)
        lengths = tuple(int(x) for x in lengths)
        if len(lengths) < 3 or len(halos) != 3 or len(selector) != 3:
            raise ValueError("DataStore has three spatial dimensions (plus optional extra data dimensions)")
        # extra data dimensions follow k in memory (the gpu layout_map<N-1, ..., 0>, storage/gpu.hpp:27-40)
        self.lengths = lengths[:3]
        self.extra = lengths[3:]
        self.halos = tuple(int(x) for x in halos)
        self.name = name
        # selector (storage::builder::selector, storage/builder.hpp:168-326): a masked dimension holds one value
        # for every index -- it is allocated with length 1 and has stride 0 in the SID
        self.selector = tuple(bool(x) for x in selector)
        if self.extra and not all(self.selector):
            raise ValueError("masked dimensions and extra data dimensions are not combined")
        self.alloc_lengths = tuple(n if m else 1 for n, m in zip(self.lengths, self.selector))
        alloc_halos = tuple(h if m else 0 for h, m in zip(self.halos, self.selector))
        lay = _abi.Layout()
        _abi.call("gt355_storage_layout", self.dtype.itemsize, _abi.c_i64_3(*self.alloc_lengths),
                  _abi.c_i64_3(*alloc_halos), ctypes.byref(lay))
        self.layout = lay
        self.alloc_strides = tuple(int(x) for x in lay.stride)
        self.strides = tuple(st if m else 0 for st, m in zip(self.alloc_strides, self.selector))
        self.padded = tuple(int(x) for x in lay.padded)
        span = self.alloc_strides[2] * self.padded[2]  # elements of one spatial slice
        self.extra_strides = tuple(int(span * np.prod(self.extra[:d], dtype=np.int64)) for d in range(len(self.extra)))
        nslices = int(np.prod(self.extra, dtype=np.int64))
        alloc_bytes = lay.alloc_bytes + (nslices - 1) * span * self.dtype.itemsize
        raw = ctypes.c_void_p()
        _abi.call("gt355_malloc", ctypes.c_size_t(alloc_bytes), ctypes.byref(raw))
        self._raw = raw
        # padding and alignment slack are zeroed once so that no copy of the allocation ever carries garbage
        _abi.call("gt355_memset", raw, 0, ctypes.c_size_t(alloc_bytes), None)
        base = ctypes.c_void_p()
        _abi.call("gt355_storage_base", ctypes.byref(lay), raw, ctypes.byref(base))
        self.base = base.value  # address of storage index (0, 0, 0)
        self.nelem = span * nslices  # elements from base covered by the layout

    @property
    def masked(self):
        return not all(self.selector)

    # ------------------------------------------------------------------ lifetime
    def free(self):
        if getattr(self, "_raw", None) is not None and self._raw.value:
            _abi.lib().gt355_free(self._raw)
            self._raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ------------------------------------------------------------------ addressing
    def address(self, i, j, k, *extra):
        """Device address of storage index (i, j, k, extra...)."""
        off = i * self.strides[0] + j * self.strides[1] + k * self.strides[2]
        off += sum(e * s for e, s in zip(extra, self.extra_strides))
        return self.base + off * self.dtype.itemsize

    def sid(self, origin=None, extra=()):
        """gt355_field whose origin is storage index ``origin`` (default: the first interior point) of the spatial
        slice ``extra``, with the allocation box as bounds -- the SID after core::shift_origin
        (stencil/core/backend.hpp:28-34).  The structure is built once per (origin, extra) and shared: treat it as
        read-only (a launch's host cost is mostly building these)."""
        o = self.halos if origin is None else tuple(origin)
        key = (o, tuple(extra))
        cache = self.__dict__.setdefault("_sids", {})
        f = cache.get(key)
        if f is None:
            f = cache[key] = self._make_sid(o, extra)
        return f

    def _make_sid(self, o, extra):
        f = _abi.Field()
        f.origin = self.address(*o, *extra)
        f.stride = (ctypes.c_int64 * 3)(*self.strides)
        # a masked dimension can be indexed anywhere (stride 0)
        f.lower = (ctypes.c_int32 * 3)(*[-x if m else -(1 << 30) for x, m in zip(o, self.selector)])
        f.upper = (ctypes.c_int32 * 3)(*[n - x if m else (1 << 30) for n, x, m in zip(self.lengths, o,
                                                                                      self.selector)])
        f.dtype = self.code
        return f

    # ------------------------------------------------------------------ host mirror
    def _host_buffer(self):
        return np.zeros(tuple(reversed(self.extra)) + (self.padded[2], self.padded[1], self.alloc_strides[1]),
                        dtype=self.dtype)

    def to_numpy(self):
        """Host copy of the whole field indexed [i, j, k, extra...] (the reference's host_view,
        data_store.hpp:140-146); masked dimensions are broadcast to the field's lengths."""
        buf = self._host_buffer()
        _abi.call("gt355_memcpy_d2h", buf.ctypes.data, ctypes.c_void_p(self.base), ctypes.c_size_t(buf.nbytes))
        ni, nj, nk = self.alloc_lengths
        arr = np.ascontiguousarray(buf[..., :nk, :nj, :ni].transpose(*range(buf.ndim - 1, -1, -1)))
        return np.broadcast_to(arr, self.lengths).copy() if self.masked else arr

    def padded_numpy(self):
        """Host copy in the device layout: array [k, j, i_padded] (for passing to CPU code with the same strides);
        a masked dimension has length 1 here."""
        buf = self._host_buffer()
        _abi.call("gt355_memcpy_d2h", buf.ctypes.data, ctypes.c_void_p(self.base), ctypes.c_size_t(buf.nbytes))
        return buf

    def from_numpy(self, arr):
        """Uploads an [i, j, k] array of the field's lengths (storage_update_target, storage/gpu.hpp:83-90); a
        masked dimension takes index 0 of the array (an array of the allocation's lengths is accepted too)."""
        arr = np.asarray(arr, dtype=self.dtype)
        if arr.shape != self.lengths + self.extra and arr.shape != self.alloc_lengths + self.extra:
            raise ValueError(f"shape {arr.shape} != field lengths {self.lengths + self.extra}")
        if arr.shape[:3] != self.alloc_lengths:
            arr = arr[tuple(slice(None) if m else slice(0, 1) for m in self.selector)]
        buf = self._host_buffer()
        ni, nj, nk = self.alloc_lengths
        buf[..., :nk, :nj, :ni] = arr.transpose(*range(arr.ndim - 1, -1, -1))
        _abi.call("gt355_memcpy_h2d", ctypes.c_void_p(self.base), buf.ctypes.data, ctypes.c_size_t(buf.nbytes))
        return self

    def fill_uniform(self, seed, field_id, a, b, stream=None):
        """Synthetic input: SplitMix64 -> U(a,b) over the whole field, halo included (see inputs.py); extra
        dimensions continue the logical order after k (k is never padded, so they merge with it)."""
        n0, n1, n2 = self.alloc_lengths
        n2 *= int(np.prod(self.extra, dtype=np.int64))
        _abi.call("gt355_fill_uniform", self.code, ctypes.c_void_p(self.base), _abi.c_i64_3(n0, n1, n2),
                  _abi.c_i64_3(*self.alloc_strides), ctypes.c_uint64(seed), ctypes.c_uint64(field_id),
                  ctypes.c_double(a), ctypes.c_double(b), stream)
        return self

    def fill_value(self, v):
        return self.from_numpy(np.full(self.alloc_lengths + self.extra, v, dtype=self.dtype))


class builder:
    """storage::builder mirror: builder().type(np.float64).dimensions(ni, nj, nk).halos(2, 2, 0).value(0).build()."""

    def __init__(self):
        self._dtype = np.float64
        self._dims = None
        self._halos = (0, 0, 0)
        self._init = None
        self._name = ""
        self._selector = (1, 1, 1)

    def _copy(self, **kw):
        b = builder()
        b.__dict__.update(self.__dict__)
        b.__dict__.update(kw)
        return b

    def type(self, dtype):
        return self._copy(_dtype=dtype)

    def dimensions(self, *dims):
        return self._copy(_dims=tuple(dims))

    def halos(self, *halos):
        return self._copy(_halos=tuple(halos))

    def name(self, name):
        return self._copy(_name=name)

    def selector(self, *mask):
        """selector<I, J, K> (storage/builder.hpp:168-326): dimensions with mask 0 hold one value for every index
        (selector(0, 1, 0) is a j-only field)."""
        if len(mask) != 3:
            raise ValueError("selector takes three flags")
        return self._copy(_selector=tuple(int(bool(m)) for m in mask))

    def value(self, v):
        return self._copy(_init=("value", v))

    def initializer(self, fn):
        """fn(i, j, k) over index arrays of the full allocation, halo included (storage/builder.hpp:67-84)."""
        return self._copy(_init=("fn", fn))

    def uniform(self, seed, field_id, a, b):
        return self._copy(_init=("uniform", (seed, field_id, a, b)))

    def build(self):
        if self._dims is None:
            raise ValueError("dimensions() not set")
        ds = DataStore(self._dtype, self._dims, self._halos, self._name, self._selector)
        kind = self._init[0] if self._init else "value"
        if kind == "value":
            ds.fill_value(self._init[1] if self._init else 0)
        elif kind == "fn":
            # fn(i, j, k, extra...); evaluated at index 0 of every masked dimension
            shape = ds.alloc_lengths + ds.extra
            idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
            ds.from_numpy(np.broadcast_to(self._init[1](*idx), shape))
        else:
            ds.fill_uniform(*self._init[1])
            _abi.call("gt355_synchronize")
        return ds
