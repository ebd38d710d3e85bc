"""transform_layout (include/gridtools/layout_transformation.hpp:64-75): copy an N-dimensional box between two
strided layouts on the device (gt355_transform_layout)."""
import ctypes

from . import _abi


def transform_layout(dst, src, dims, dst_strides, src_strides, elem_size=8, stream=None):
    """dst / src: device addresses (int); dims, dst_strides, src_strides: equally long sequences, strides in
    elements.  Elements of dst outside the index map are not touched."""
    n = len(dims)
    if len(dst_strides) != n or len(src_strides) != n:
        raise ValueError("dims, dst_strides and src_strides must have the same length")
    arr = ctypes.c_int64 * n
    _abi.call("gt355_transform_layout", elem_size, ctypes.c_void_p(dst), ctypes.c_void_p(src), n, arr(*dims),
              arr(*dst_strides), arr(*src_strides), stream)
