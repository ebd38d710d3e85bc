"""The runtime's pools and modes through the C ABI: the temporary pool's stream rules, the deep-column scratch
policy under graph capture, the debug launch mode and the current stream.  Host-side checks of returned pointers and
status codes (no kernel is made to fail)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _acq(nbytes, stream):
    from gridtools_amd import _abi

    p = ctypes.c_void_p()
    _abi.call("gt355_tmp_acquire", ctypes.c_size_t(nbytes), stream, ctypes.byref(p))
    return p.value


def test_tmp_pool_reuse_rules(gpu):
    from gridtools_amd import _abi
    from gridtools_amd.timing import Stream

    a, b = Stream(), Stream()
    p = _acq(1 << 20, a.handle)
    assert p
    q = _acq(1 << 20, a.handle)  # p is in use: a second block
    assert q and q != p
    _abi.call("gt355_tmp_release", ctypes.c_void_p(p), a.handle)
    # the same stream takes the released block at once (stream order separates its users)
    assert _acq(1 << 20, a.handle) == p
    _abi.call("gt355_tmp_release", ctypes.c_void_p(p), a.handle)
    # another stream takes it once the work queued before the release has completed
    a.synchronize()
    assert _acq(1 << 20, b.handle) == p
    # a block more than twice as large as the request is not handed out
    _abi.call("gt355_tmp_release", ctypes.c_void_p(p), b.handle)
    b.synchronize()
    r = _acq(1 << 18, b.handle)
    assert r != p
    _abi.call("gt355_tmp_release", ctypes.c_void_p(q), a.handle)
    _abi.call("gt355_tmp_release", ctypes.c_void_p(r), b.handle)
    # releasing an unknown pointer is an error, not a crash
    assert _abi.lib().gt355_tmp_release(ctypes.c_void_p(0x1234), None) != 0
    assert "not acquired" in _abi.last_error()
    _abi.call("gt355_synchronize")


def test_tmp_pool_release_n(gpu):
    """gt355_tmp_release_n: several blocks back to the pool with one event; a list holding an unknown or repeated
    pointer releases nothing; null entries are skipped."""
    from gridtools_amd import _abi
    from gridtools_amd.timing import Stream

    a, b = Stream(), Stream()
    nb = (3 << 20) + 12288  # a size no other test asks for
    p, q = _acq(nb, a.handle), _acq(nb, a.handle)
    assert p and q and p != q
    arr = (ctypes.c_void_p * 3)(p, None, q)
    # an unknown pointer in the list: an error, and p and q stay in use
    bad = (ctypes.c_void_p * 2)(p, 0x1234)
    assert _abi.lib().gt355_tmp_release_n(bad, 2, a.handle) != 0
    assert "not acquired" in _abi.last_error()
    twice = (ctypes.c_void_p * 2)(p, p)
    assert _abi.lib().gt355_tmp_release_n(twice, 2, a.handle) != 0
    r = _acq(nb, a.handle)  # p and q still in use: a third block
    assert r not in (p, q)
    _abi.call("gt355_tmp_release_n", arr, 3, a.handle)
    # both back: the same stream takes them at once, another only after the release's work completed
    a.synchronize()
    got = {_acq(nb, b.handle), _acq(nb, b.handle)}
    assert got == {p, q}
    arr2 = (ctypes.c_void_p * 3)(p, q, r)
    _abi.call("gt355_tmp_release_n", arr2, 3, b.handle)
    _abi.call("gt355_tmp_release_n", arr2, 0, b.handle)  # an empty list
    _abi.call("gt355_synchronize")


def test_deep_scratch_growth_refused_during_capture(gpu, oracle):
    """The nk > 80 vadv keeps ccol/dcol in a per-stream scratch buffer.  Growing it inside a capture is refused with
    a status (a graph must not reference an allocation made for it); after a run of the shape outside capture, the
    capture succeeds and its replay equals the eager result."""
    from gridtools_amd import _abi, stencils
    from gridtools_amd.timestep import capture
    from gridtools_amd.timing import Stream
    from test_gpu_stencils import oracle_vadv, vadv_case

    st = Stream()
    ni, nj, nk = 48, 40, 97
    f, grid = vadv_case(gpu, ni, nj, nk, np.float64)
    host0 = f["utens_stage"].padded_numpy()
    sids = [f[n].sid(grid.origin) for n in ("utens_stage", "u_stage", "wcon", "u_pos", "utens")]
    _abi.call("gt355_graph_begin", st.handle)
    try:
        rc = _abi.lib().gt355_vadv(0, *sids, ctypes.c_double(0.15), ni, nj, nk, st.handle)
    finally:
        g = ctypes.c_void_p()
        _abi.call("gt355_graph_end", st.handle, ctypes.byref(g))
        if g.value:
            _abi.lib().gt355_graph_destroy(g)
    assert rc == -4 and "outside capture" in _abi.last_error()
    # eager run sizes the stream's scratch; a capture of the same shape then works
    ref = oracle_vadv(oracle, f, 0.15, 3, ni, nj, nk, "native")
    stencils.vadv(f["utens_stage"], f["u_stage"], f["wcon"], f["u_pos"], f["utens"], 0.15, grid,
                  stream=st.handle)
    st.synchronize()
    assert np.array_equal(f["utens_stage"].padded_numpy()[:nk, 3:3 + nj, 3:3 + ni], ref)
    f["utens_stage"].from_numpy(np.ascontiguousarray(host0.transpose(2, 1, 0)[:ni + 6, :nj + 6, :nk]))
    graph = capture(st, lambda s: stencils.vadv(f["utens_stage"], f["u_stage"], f["wcon"], f["u_pos"],
                                                f["utens"], 0.15, grid, stream=s.handle))
    graph.launch(st)
    st.synchronize()
    assert np.array_equal(f["utens_stage"].padded_numpy()[:nk, 3:3 + nj, 3:3 + ni], ref)


def test_debug_sync_mode_and_current_stream(gpu, oracle):
    from gridtools_amd import _abi, stencils
    from gridtools_amd.timing import Stream
    from test_gpu_stencils import hdiff_case, oracle_hdiff

    st = Stream()
    _abi.call("gt355_stream_set_current", st.handle)
    cur = ctypes.c_void_p()
    _abi.call("gt355_stream_current", ctypes.byref(cur))
    assert cur.value == st.handle.value
    _abi.call("gt355_stream_set_current", None)
    _abi.call("gt355_debug_sync", 1)
    try:
        inp, coeff, out, grid = hdiff_case(gpu, 40, 24, 7, np.float64)
        stencils.hdiff(out, inp, coeff, grid)
        got = out.padded_numpy()[:7, 2:26, 2:42]
        assert np.array_equal(got, oracle_hdiff(oracle, inp, coeff, 2, 40, 24, 7, "native"))
    finally:
        _abi.call("gt355_debug_sync", 0)
