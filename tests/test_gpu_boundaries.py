"""GPU parity of the boundary kernels (libgt355 through the C ABI) against the CPU oracle: the predefined
boundary functors with direction predicates (boundaries/apply_gpu.hpp, zero.hpp, value.hpp, copy.hpp), the
local halo fill (periodic / zero-gradient), and pack -> periodic self-exchange -> unpack round trips
(gcl m_pack*/m_unpack* kernels, descriptors_manual_gpu.hpp:342-512).  Integer-free copies: bit-exact."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def halos3(ni, nj, nk, hi, hj, hk):
    def hd(h, n):
        return (h, h, h, h + n - 1, n + 2 * h)
    return [hd(hi, ni), hd(hj, nj), hd(hk, nk)]


def field(gt, dtype, lengths, seed, fid):
    return gt.builder().type(dtype).dimensions(*lengths).halos(0, 0, 0).uniform(seed, fid, -1.0, 1.0).build()


def box(ds, arr):
    """The logical box of a padded [k, j, i_padded] copy (padding columns excluded)."""
    ni, nj, nk = ds.lengths
    return arr[:nk, :nj, :ni]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_bc_apply_matches_oracle(gpu, oracle, dtype, kind):
    from gridtools_amd import boundaries

    ni, nj, nk = 33, 12, 7
    hal = halos3(ni, nj, nk, 3, 2, 1)
    lengths = (ni + 6, nj + 4, nk + 2)
    dst = [field(gpu, dtype, lengths, 1, 10 + f) for f in range(2)]
    src = [field(gpu, dtype, lengths, 2, 20 + f) for f in range(2)]
    pred = lambda di, dj, dk: not (dj == 1 and di == -1)  # noqa: E731  an asymmetric predicate
    bc = boundaries.make_boundary(hal, kind, value=0.75, predicate=pred)
    exp = [d.padded_numpy() for d in dst]
    for e, s in zip(exp, src):
        oracle.bc(kind, e, s.padded_numpy() if kind == 2 else None, 0.75, hal, bc.dir_mask)
    bc.apply(*dst, src=src if kind == 2 else None)
    for d, e in zip(dst, exp):
        assert np.array_equal(box(d, d.padded_numpy()), box(d, e))


@pytest.mark.parametrize("modes", [(2, 1, 0), (1, 1, 0), (2, 2, 0), (2, 0, 0), (1, 2, 2)])
def test_halo_fill_matches_oracle(gpu, oracle, modes):
    from gridtools_amd import boundaries

    ni, nj, nk = 40, 17, 6
    hk = 1 if modes[2] else 0
    hal = halos3(ni, nj, nk, 3, 3, hk)
    lengths = (ni + 6, nj + 6, nk + 2 * hk)
    fs = [field(gpu, np.float32, lengths, 3, 30 + f) for f in range(3)]
    exp = [f.padded_numpy() for f in fs]
    for e in exp:
        oracle.halo_fill(e, hal, modes)
    boundaries.halo_fill(fs, hal, modes)
    for f, e in zip(fs, exp):
        assert np.array_equal(box(f, f.padded_numpy()), box(f, e))


def test_pack_self_exchange_unpack_is_periodic_fill(gpu, oracle):
    """On one rank with a periodic process grid (gcl/low_level/proc_grids_3D.hpp:83-88) a halo exchange sends
    each direction's slab to itself: unpack(direction d) receives pack(direction -d)."""
    from gridtools_amd import _abi, boundaries

    ni, nj, nk = 24, 19, 5
    hal = halos3(ni, nj, nk, 3, 3, 0)
    lengths = (ni + 6, nj + 6, nk)
    fs = [field(gpu, np.float64, lengths, 5, 50 + f) for f in range(2)]
    exp = [f.padded_numpy() for f in fs]
    for e in exp:
        oracle.halo_fill(e, hal, (2, 2, 0))
    mask = boundaries.direction_mask(lambda di, dj, dk: dk == 0)
    nbytes = boundaries.halo_buffer_size(_abi.F64, 2, hal, mask)
    assert nbytes == 2 * 8 * (2 * 3 * nj * nk + 2 * 3 * ni * nk + 4 * 9 * nk)
    send, recv = ctypes.c_void_p(), ctypes.c_void_p()
    _abi.call("gt355_malloc", ctypes.c_size_t(nbytes), ctypes.byref(send))
    _abi.call("gt355_malloc", ctypes.c_size_t(nbytes), ctypes.byref(recv))
    try:
        boundaries.halo_pack(fs, hal, send.value, mask)
        host = np.zeros(nbytes // 8)
        _abi.call("gt355_memcpy_d2h", host.ctypes.data, send, ctypes.c_size_t(nbytes))
        # segment table in direction-bit order (symmetric halos: pack and unpack sizes coincide)
        dirs = [(di, dj, 0) for di in (-1, 0, 1) for dj in (-1, 0, 1) if (di, dj) != (0, 0)]
        dirs.sort(key=lambda d: boundaries.direction_bit(*d))
        size = {}
        for d in dirs:
            size[d] = 2 * (ni if d[0] == 0 else 3) * (nj if d[1] == 0 else 3) * nk
        off, o = {}, 0
        for d in dirs:
            off[d] = o
            o += size[d]
        recv_host = np.zeros_like(host)
        for d in dirs:
            opp = (-d[0], -d[1], 0)
            recv_host[off[d]:off[d] + size[d]] = host[off[opp]:off[opp] + size[opp]]
        _abi.call("gt355_memcpy_h2d", recv, recv_host.ctypes.data, ctypes.c_size_t(nbytes))
        boundaries.halo_unpack(fs, hal, recv.value, mask)
        for f, e in zip(fs, exp):
            assert np.array_equal(box(f, f.padded_numpy()), box(f, e))
    finally:
        _abi.lib().gt355_free(send)
        _abi.lib().gt355_free(recv)


@pytest.mark.parametrize("dtype,ni", [(np.float32, 40), (np.float64, 36), (np.float32, 41)])
@pytest.mark.parametrize("gap", ["1", "0"])
def test_halo_fill_row_gap_runs(gpu, oracle, monkeypatch, dtype, ni, gap):
    """The halo fill with the i faces written as whole-sector row-gap runs (fields of the library's layout, ni times
    the element size a multiple of 32 bytes: 40 FP32, 36 FP64) equals the per-point fill (41 FP32, and
    GT355_TUNE_halo_gap=0) and the oracle; the row padding stays zero."""
    from gridtools_amd import _abi, boundaries

    monkeypatch.setenv("GT355_TUNE_halo_gap", gap)
    _abi.call("gt355_tune_env_refresh")
    nj, nk = 13, 7
    for modes in [(2, 1, 0), (1, 2, 0), (2, 0, 0)]:
        print(f"[case] halo fill {np.dtype(dtype).name} ni={ni} modes={modes} gap={gap}", flush=True)
        hal = halos3(ni, nj, nk, 3, 3, 0)
        lengths = (ni + 6, nj + 6, nk)
        # the library's layout for a halo of 3: the first interior point aligned
        fs = [gpu.builder().type(dtype).dimensions(*lengths).halos(3, 3, 0).uniform(3, 70 + f, -1.0, 1.0).build()
              for f in range(2)]
        exp = [f.padded_numpy() for f in fs]
        for e in exp:
            oracle.halo_fill(e, hal, modes)
        boundaries.halo_fill(fs, hal, modes)
        for f, e in zip(fs, exp):
            got = f.padded_numpy()
            assert np.array_equal(box(f, got), box(f, e)), modes
            assert not got[:, :, lengths[0]:].any(), "row padding changed"
    monkeypatch.delenv("GT355_TUNE_halo_gap")
    _abi.call("gt355_tune_env_refresh")
