"""Pins the CPU oracle's transform_layout with the reference's own unit tests
(tests/unit_tests/layout_transformation/test_layout_transformation.cpp:52-120): reversed 2-D, 3-D and 4-D layouts
with index-coded values, and a 1-D copy into stride 2 that must leave the other elements untouched; plus the
regression test's (0,1,2) -> (2,1,0) storage layouts (tests/regression/layout_transformation.cpp:30-40)."""
import itertools

import numpy as np
import pytest


@pytest.mark.parametrize("shape", [(4, 5), (4, 5, 6), (4, 5, 6, 7)])
def test_reverse_layout(oracle, shape):
    n = len(shape)
    src = np.zeros(shape)  # C order: last dimension fastest
    for idx in itertools.product(*[range(x) for x in shape]):
        src[idx] = sum(10 ** (n - 1 - d) * i for d, i in enumerate(idx)) * 10 ** (4 - n)
    dst = np.full(shape[::-1], -1.0)
    src_strides = [int(np.prod(shape[d + 1:])) for d in range(n)]
    dst_strides = [int(np.prod(shape[:d])) for d in range(n)]
    oracle.transform_layout(dst, src, shape, dst_strides, src_strides)
    assert np.array_equal(dst, src.transpose(*range(n - 1, -1, -1)))


def test_1d_into_stride_2(oracle):
    src = np.arange(4, dtype=np.float64)
    dst = np.full((4, 2), -1.0)
    oracle.transform_layout(dst, src, [4], [2], [1])
    assert np.array_equal(dst[:, 0], src)
    assert np.all(dst[:, 1] == -1)  # the non-indexable elements are not touched


def test_regression_storage_layouts(oracle):
    """layout<0,1,2> (k-fastest... i slowest) source to layout<2,1,0> (i fastest) destination, value i + j + k."""
    ni, nj, nk = 12, 33, 61
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    src = np.ascontiguousarray((i + j + k).astype(np.float32))  # [i, j, k]: k fastest
    dst = np.zeros((nk, nj, ni), np.float32)                     # [k, j, i]: i fastest
    oracle.transform_layout(dst, src, [ni, nj, nk], [1, ni, ni * nj], [nj * nk, nk, 1])
    assert np.array_equal(dst.transpose(2, 1, 0), src)
