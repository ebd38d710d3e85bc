"""gt355_vadv_noalloc, the entry point that allocates nothing: called through the C ABI, its results equal
gt355_vadv's bit for bit on two calls in a row, the second on a smaller input (the static scratch keeps the first
call's contents), on the deep-column path (nk > 80, ccol/dcol in the static scratch, in row chunks when a domain's
ccol/dcol exceed it) and on the on-chip path; a row whose ccol/dcol exceed the static scratch is refused before any
launch."""
import ctypes

import numpy as np
import pytest

from test_gpu_stencils import vadv_case

pytestmark = pytest.mark.gpu

NAMES = ("utens_stage", "u_stage", "wcon", "u_pos", "utens")


def _call(entry, precision, f, grid, ni, nj, nk):
    from gridtools_amd import _abi

    sids = [f[n].sid(grid.origin) for n in NAMES]
    args = [ctypes.byref(s) for s in sids]
    return getattr(_abi.lib(), entry)(precision, *args, ctypes.c_double(0.15), ni, nj, nk, None)


# (storage type, precision, first shape, second (smaller) shape)
CASES = [
    (np.float64, 0, (96, 40, 100), (64, 24, 90)),      # deep columns, one chunk
    (np.float64, 0, (256, 160, 137), (200, 100, 120)),  # 89.5 MB of ccol/dcol: two chunks of rows, then one
    (np.float32, 1, (130, 70, 97), (65, 9, 85)),        # precision form (FP64 ccol/dcol)
    (np.float32, 0, (70, 33, 110), (33, 17, 81)),       # FP32 storage and compute
    (np.float64, 0, (100, 30, 80), (40, 20, 33)),       # on-chip path (no scratch used)
]


@pytest.mark.parametrize("dtype,precision,first,second", CASES)
def test_noalloc_equals_vadv(gpu, dtype, precision, first, second):
    from gridtools_amd import _abi

    for ni, nj, nk in (first, second):
        print(f"[case] vadv_noalloc {np.dtype(dtype).name} precision={precision} {ni}x{nj}x{nk}", flush=True)
        a, grid = vadv_case(gpu, ni, nj, nk, dtype)
        b, _ = vadv_case(gpu, ni, nj, nk, dtype)
        assert _call("gt355_vadv", precision, a, grid, ni, nj, nk) == 0, _abi.last_error()
        assert _call("gt355_vadv_noalloc", precision, b, grid, ni, nj, nk) == 0, _abi.last_error()
        _abi.call("gt355_synchronize")
        ref = a["utens_stage"].padded_numpy()
        got = b["utens_stage"].padded_numpy()
        # the whole allocation, halos included: equal interiors and nothing written outside them
        assert np.array_equal(got, ref)


def test_noalloc_refuses_a_row_beyond_the_static_scratch(gpu):
    """sizeof(double) * ni * (2 nk - 1) = 8 * 32768 * 273 bytes (71.6 MB) for one row is more than the 64 MiB
    static scratch: refused with a status and a message, nothing launched (the output keeps its values)."""
    from gridtools_amd import _abi

    ni, nj, nk = 32768, 1, 137
    f, grid = vadv_case(gpu, ni, nj, nk, np.float64, h=1)
    before = f["utens_stage"].padded_numpy()
    rc = _call("gt355_vadv_noalloc", 0, f, grid, ni, nj, nk)
    assert rc == -1
    assert "static scratch" in _abi.last_error()
    _abi.call("gt355_synchronize")
    assert np.array_equal(f["utens_stage"].padded_numpy(), before)
