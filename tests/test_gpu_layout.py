"""GPU transform_layout (gt355_transform_layout) against the oracle restatement of transform_cpu_loop and the
reference's known answers (tests/unit_tests/layout_transformation/test_layout_transformation.cpp:52-120,
tests/regression/layout_transformation.cpp:30-40): reversed layouts in 2-5 dimensions (the LDS-tiled transpose
path), copies whose fastest dimension agrees (the flat path), a 1-D copy into stride 2 that leaves the gaps
untouched, every element size, a storage <-> dense exchange with halo and padding at 256x256x80, and the 16-byte
access kernel for 4- and 8-byte elements with partial tiles."""
import ctypes
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DevBuf:
    def __init__(self, arr):
        from gridtools_amd import _abi

        self._abi = _abi
        self.nbytes = arr.nbytes
        self.p = ctypes.c_void_p()
        _abi.call("gt355_malloc", ctypes.c_size_t(max(self.nbytes, 16)), ctypes.byref(self.p))
        _abi.call("gt355_memcpy_h2d", self.p, arr.ctypes.data, ctypes.c_size_t(self.nbytes))

    def get(self, like):
        out = np.empty_like(like)
        self._abi.call("gt355_memcpy_d2h", out.ctypes.data, self.p, ctypes.c_size_t(self.nbytes))
        return out

    def __del__(self):
        try:
            self._abi.lib().gt355_free(self.p)
        except Exception:
            pass


def run_both(oracle, dst0, src, dims, ds, ss):
    """The device transform and the oracle on copies of dst0; returns both results."""
    from gridtools_amd import layout

    d_src, d_dst = DevBuf(src), DevBuf(dst0)
    layout.transform_layout(d_dst.p.value, d_src.p.value, dims, ds, ss, elem_size=src.itemsize)
    got = d_dst.get(dst0)
    want = dst0.copy()
    oracle.transform_layout(want, src, dims, ds, ss)
    return got, want


@pytest.mark.parametrize("shape", [(4, 5), (4, 5, 6), (4, 5, 6, 7), (70, 3, 130, 2, 5), (256, 96), (96, 300)])
def test_reverse_layout(gpu, oracle, shape):
    """Tiles 64 x 16 (b extents 4, 70), 64 x 64 (b extent 256) and 64 x 32 (b extent 96)."""
    print(f"[case] transform_layout reverse {shape}", flush=True)
    n = len(shape)
    src = np.zeros(shape)
    src[...] = np.arange(src.size, dtype=np.float64).reshape(shape)
    dst0 = np.full(shape[::-1], -1.0)
    ss = [int(np.prod(shape[d + 1:])) for d in range(n)]
    ds = [int(np.prod(shape[:d])) for d in range(n)]
    got, want = run_both(oracle, dst0, src, shape, ds, ss)
    assert np.array_equal(got, want)
    assert np.array_equal(got, src.transpose(*range(n - 1, -1, -1)))


@pytest.mark.parametrize("vec", ["1", "0"])
@pytest.mark.parametrize("dtype", [np.float32, np.int32, np.float64, np.int64])
def test_reverse_layout_16_byte_accesses(gpu, oracle, monkeypatch, dtype, vec):
    """A reversed 132 x 4 x 264 layout: unit stride along the source's and the destination's fastest dimensions and
    every other stride a multiple of 16 bytes, so the transpose moves 16 bytes per access (vec "1"; "0" forces one
    element per access).  Partial tiles along both: 264 = 4 x 64 + 8 along a, 132 = 8 x 16 + 4 along b."""
    monkeypatch.setenv("GT355_TUNE_layout_vec", vec)
    shape = (132, 4, 264)
    print(f"[case] transform_layout 16-byte accesses {np.dtype(dtype).name} vec={vec} {shape}", flush=True)
    n = len(shape)
    src = (np.arange(int(np.prod(shape))) % 1000003).astype(dtype).reshape(shape)
    dst0 = np.full(shape[::-1], -1, dtype=dtype)
    ss = [int(np.prod(shape[d + 1:])) for d in range(n)]
    ds = [int(np.prod(shape[:d])) for d in range(n)]
    got, want = run_both(oracle, dst0, src, shape, ds, ss)
    assert np.array_equal(got, want)
    assert np.array_equal(got, src.transpose(2, 1, 0))


def test_1d_into_stride_2(gpu, oracle):
    src = np.arange(4, dtype=np.float64)
    dst0 = np.full((4, 2), -1.0)
    got, want = run_both(oracle, dst0, src, [4], [2], [1])
    assert np.array_equal(got, want)
    assert np.array_equal(got[:, 0], src) and np.all(got[:, 1] == -1)


@pytest.mark.parametrize("dtype", [np.uint8, np.float16, np.float32, np.float64, np.complex128])
def test_element_sizes(gpu, oracle, dtype):
    print(f"[case] transform_layout element {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(3)
    shape = (37, 29, 5)
    src = (rng.uniform(0, 100, size=shape)).astype(dtype)
    dst0 = np.zeros(shape[::-1], dtype=dtype)
    got, want = run_both(oracle, dst0, src, shape, [1, 37, 37 * 29], [29 * 5, 5, 1])
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_same_fastest_dimension_flat_path(gpu, oracle):
    """Both layouts i-fastest with different padding and a permuted k/extra order: the flat path."""
    ni, nj, nk, ne = 300, 7, 9, 3
    rng = np.random.default_rng(4)
    src = rng.uniform(-1, 1, size=(ne, nk, nj, 304))       # [e, k, j, i_padded]
    dst0 = np.full((nk, ne, nj, 320), 7.0)                 # [k, e, j, i_padded]
    dims = [ni, nj, nk, ne]
    ss = [1, 304, 304 * nj, 304 * nj * nk]
    ds = [1, 320, 320 * nj * ne, 320 * nj]
    got, want = run_both(oracle, dst0, src, dims, ds, ss)
    assert np.array_equal(got, want)
    assert np.all(got[..., ni:] == 7.0)


def test_storage_to_dense_fortran_exchange(gpu, oracle):
    """A C/Fortran caller's dense array [k][j][i] (Fortran a(i, j, k)) into a padded 256x256x80 storage interior with
    a 3-point halo and back -- the data exchange of the reference's bindings."""
    print("[case] transform_layout storage <-> dense 256x256x80", flush=True)
    ni, nj, nk, h = 256, 256, 80, 3
    b = gpu.builder().type(np.float64).dimensions(ni + 2 * h, nj + 2 * h, nk).halos(h, h, 0)
    st = b.value(0).build()
    rng = np.random.default_rng(5)
    dense = rng.uniform(-1, 1, size=(nk, nj, ni))
    d_dense = DevBuf(dense)
    from gridtools_amd import layout

    layout.transform_layout(st.address(h, h, 0), d_dense.p.value, [ni, nj, nk], list(st.strides), [1, ni, ni * nj])
    got = st.to_numpy()
    assert np.array_equal(got[h:h + ni, h:h + nj, :], dense.transpose(2, 1, 0))
    assert not got[:h].any() and not got[:, :h].any()
    # back into a k-fastest (C-order a[i][j][k]) dense buffer: a transpose through LDS tiles
    back0 = np.zeros((ni, nj, nk))
    d_back = DevBuf(back0)
    layout.transform_layout(d_back.p.value, st.address(h, h, 0), [ni, nj, nk], [nj * nk, nk, 1], list(st.strides))
    assert np.array_equal(d_back.get(back0), dense.transpose(2, 1, 0))
