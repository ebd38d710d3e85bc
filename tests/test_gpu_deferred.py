"""GPU parity of the deferred-run scope (gt355_batch_begin / gt355_batch_end): hdiff, vadv and hdiff_vadv calls on
uneven domains, recorded in a scope and run as persistent two-phase batches, equal the same calls run one after
another, bit for bit.  The program is two dependent timesteps (the second reads what the first wrote and
overwrites what the first read), so its calls spread over four launches; a domain deeper than the vertical phase
holds (150 levels) runs its vertical advection on gt355_vadv at its place in the order; a repeated scope runs from
the kept batches; a captured scope replays like eager calls.  The single-call entry points themselves are
pinned against the oracle by the other GPU tests."""
import numpy as np
import pytest

from test_gpu_batch import make_members

pytestmark = pytest.mark.gpu

SIZES = [(23, 11, 43), (70, 9, 81), (130, 37, 2), (64, 4, 137), (96, 33, 150), (1, 1, 3)]
FIELDS = ("in_", "u_stage", "utens_stage", "lap")


def program(stencils, members, precision, stream=None):
    """Step 1: u_stage = hdiff(in), utens_stage = vadv(.., u_stage, ..); step 2: the roles of in and u_stage
    swapped, as one hdiff_vadv call."""
    for m in members:
        g = m["grid"]
        stencils.hdiff(m["u_stage"], m["in_"], m["coeff"], g, precision=precision, stream=stream)
        stencils.vadv(m["utens_stage"], m["u_stage"], m["wcon"], m["u_pos"], m["utens"], 0.15, g,
                      precision=precision, stream=stream)
    for m in members:
        g = m["grid"]
        stencils.hdiff_vadv(m["u_stage"], m["coeff"], m["in_"], m["utens_stage"], m["wcon"], m["u_pos"], m["utens"],
                            0.15, g, precision=precision, stream=stream)


def bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def same_state(a, b, label):
    for d, (ma, mb) in enumerate(zip(a, b)):
        for n in FIELDS:
            x, y = ma[n].padded_numpy(), mb[n].padded_numpy()
            assert np.array_equal(bits(x), bits(y)), (label, d, n)


@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1), (np.float32, 0)])
def test_deferred_scope_equals_calls_in_order(gpu, dtype, precision):
    from gridtools_amd import stencils
    from gridtools_amd.batch import deferred

    print(f"[case] deferred scope {np.dtype(dtype).name} precision={precision} {SIZES}", flush=True)
    a = make_members(gpu, SIZES, dtype)
    b = make_members(gpu, SIZES, dtype)
    with deferred():
        program(stencils, a, precision)
    program(stencils, b, precision)
    gpu.synchronize()
    same_state(a, b, "first scope")
    # the same scope again: its batches come from the thread's cache
    with deferred():
        program(stencils, a, precision)
    program(stencils, b, precision)
    gpu.synchronize()
    same_state(a, b, "repeated scope")


def test_deferred_scope_flush_keeps_order(gpu):
    """An explicit flush and a library call that launches on the scope's stream (gt355_copy) run the recorded calls
    first."""
    from gridtools_amd import stencils
    from gridtools_amd.batch import deferred

    sizes = [(40, 30, 50), (90, 17, 96)]
    a = make_members(gpu, sizes, np.float32)
    b = make_members(gpu, sizes, np.float32)
    with deferred() as sc:
        program(stencils, a[:1], 1)
        sc.flush()
        program(stencils, a[1:], 1)
        # reads u_stage of the second domain, which a recorded hdiff writes: the copy must see the new values
        m = a[1]
        stencils.copy(m["lap"], m["u_stage"], m["grid"])
    program(stencils, b, 1)
    m = b[1]
    stencils.copy(m["lap"], m["u_stage"], m["grid"])
    gpu.synchronize()
    same_state(a, b, "flush")
    assert np.array_equal(bits(a[1]["lap"].padded_numpy()), bits(b[1]["lap"].padded_numpy()))


def test_deferred_scope_in_a_graph(gpu):
    """A scope run once eagerly (building its batches), then captured: every replay equals an eager twin."""
    from gridtools_amd import stencils
    from gridtools_amd.batch import deferred
    from gridtools_amd.timestep import capture
    from gridtools_amd.timing import Stream

    sizes = [(40, 30, 50), (90, 17, 96), (64, 64, 41)]
    a = make_members(gpu, sizes, np.float32)
    b = make_members(gpu, sizes, np.float32)
    st = Stream()
    with deferred(st):
        program(stencils, a, 1, st.handle)
    program(stencils, b, 1)

    def record(s):
        with deferred(s):
            program(stencils, a, 1, s.handle)

    g = capture(st, record)
    for r in range(3):
        g.launch(st)
        program(stencils, b, 1)
        st.synchronize()
        gpu.synchronize()
        same_state(a, b, f"replay {r}")


@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1)])
def test_deferred_single_calls_and_flushes_in_order(gpu, dtype, precision):
    """Scopes flushed after every independent hdiff / vadv pair (each slot holds one call, which runs on its own
    single-domain entry point) and a dependent pair after them, eager and captured: the results equal the calls in
    order, and so do three replays of the captured scope."""
    from gridtools_amd import stencils
    from gridtools_amd.batch import deferred
    from gridtools_amd.timestep import capture
    from gridtools_amd.timing import Stream

    sizes = [(256, 256, 80), (100, 60, 96), (77, 130, 33)]
    a = make_members(gpu, sizes, dtype)
    b = make_members(gpu, sizes, dtype)
    print(f"[case] independent phases {np.dtype(dtype).name} precision={precision}", flush=True)

    def prog(ms, stream=None):
        # hdiff of domain d with vadv of domain d+1 (independent), one pair per scope flush
        n = len(ms)
        for d in range(n):
            h, v = ms[d], ms[(d + 1) % n]
            stencils.hdiff(h["u_stage"], h["in_"], h["coeff"], h["grid"], precision=precision, stream=stream)
            stencils.vadv(v["utens_stage"], v["in_"], v["wcon"], v["u_pos"], v["utens"], 0.15, v["grid"],
                          precision=precision, stream=stream)
            yield
        # dependent: vadv reads the u_stage the hdiff writes
        m = ms[0]
        stencils.hdiff(m["lap"], m["in_"], m["coeff"], m["grid"], precision=precision, stream=stream)
        stencils.vadv(m["utens_stage"], m["lap"], m["wcon"], m["u_pos"], m["utens"], 0.15, m["grid"],
                      precision=precision, stream=stream)

    def scoped(ms, st=None):
        handle = st.handle if st is not None else None
        with deferred(st) as sc:
            for _ in prog(ms, handle):
                sc.flush()

    def direct(ms):
        for _ in prog(ms):
            pass

    st = Stream()
    scoped(a, st)
    direct(b)
    st.synchronize()
    gpu.synchronize()
    same_state(a, b, "eager")
    g = capture(st, lambda s: scoped(a, s))
    for r in range(3):
        g.launch(st)
        direct(b)
        st.synchronize()
        gpu.synchronize()
        same_state(a, b, f"replay {r}")
