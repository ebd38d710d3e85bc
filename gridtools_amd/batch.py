"""The batching form: many independent domains of uneven sizes in one persistent launch (gt355_batch_*).

The reference runs one grid per ``stencil::run`` (stencil/frontend/run.hpp:34-60); its only batching,
``expandable_run`` (stencil/frontend/expandable_run.hpp:80-183), replicates a spec over fields that share one
grid.  Here each member domain has its own ni x nj x nk, and one launch computes, for every domain and level,

    lap = laplacian(in); u_stage = hdiff(in, coeff); utens_stage = vadv(utens_stage, u_stage, wcon, u_pos, utens)

with u_stage passed to the vertical sweep in the compute type (as gt355_hdiff_vadv does).

    from gridtools_amd.batch import Batch
    b = Batch([dict(in_=..., coeff=..., lap=..., u_stage=..., utens_stage=..., wcon=..., u_pos=..., utens=...,
                    grid=grid), ...], precision=WIDEN)
    b.run(0.15, stream)

``BatchedTimesteps`` is the seeded multi-domain workload with its timestep loop (u_stage becomes the next step's
in) and graph capture.
"""
import ctypes

import numpy as np

from . import _abi, stencils
from .grid import halo_descriptor, make_grid
from .inputs import FIELD_IDS
from .storage import builder
from .timestep import capture

FIELDS = ("in_", "coeff", "lap", "u_stage", "utens_stage", "wcon", "u_pos", "utens")


class BatchDomain(ctypes.Structure):
    """gt355_batch_domain."""

    _fields_ = [(n, _abi.Field) for n in FIELDS] + [
        ("ni", ctypes.c_int32), ("nj", ctypes.c_int32), ("nk", ctypes.c_int32), ("phases", ctypes.c_int32)]


class Batch:
    """A work table over ``members`` (dicts of the eight DataStores of FIELDS plus ``grid``) and its launch.

    A member may carry ``phases``: 0 (both phases), 1 (the stencil phase alone: u_stage = hdiff(in, coeff)) or 2 (the
    vertical phase alone: vadv reading u_stage); fields its phases do not use, and ``lap``, may then be None."""

    def __init__(self, members, precision=stencils.WIDEN, wg_per_cu=2):
        self.members = list(members)  # keeps the stores alive while the table points at them
        arr = (BatchDomain * len(self.members))()
        for d, m in enumerate(self.members):
            g = m["grid"]
            for n in FIELDS:
                if m.get(n) is not None:
                    setattr(arr[d], n, m[n].sid(g.origin))
            arr[d].ni, arr[d].nj, arr[d].nk = g.i_size, g.j_size, g.k_size
            arr[d].phases = m.get("phases", 0)
        h = ctypes.c_void_p()
        _abi.call("gt355_batch_create", precision, arr, len(self.members), wg_per_cu, ctypes.byref(h))
        self.handle = h
        self.precision = precision

    def run(self, dtr_stage, stream=None):
        _abi.call("gt355_batch_run", self.handle, ctypes.c_double(dtr_stage), stream)

    def tune(self, reps=5, stream=None):
        """Picks the persistent grid size (workgroups per CU) by timing; returns (wg_per_cu, mean microseconds).
        The runs rewrite the outputs and utens_stage, as run() does."""
        w, us = ctypes.c_int(), ctypes.c_double()
        _abi.call("gt355_batch_tune", self.handle, reps, stream, ctypes.byref(w), ctypes.byref(us))
        return w.value, us.value

    def config(self):
        """The launch configuration {form, hjb, hw, w, deep_nk}: the form, the stencil phase's rows per item and
        workgroups per CU, the persistent grid setting, the depth from which the vertical phase keeps 96 levels in
        registers."""
        c = (ctypes.c_int32 * 5)()
        _abi.call("gt355_batch_get_config", self.handle, c)
        return dict(zip(("form", "hjb", "hw", "w", "deep_nk"), list(c)))

    def set_config(self, **kw):
        """Applies a launch configuration (keys of config(); the form cannot change)."""
        cur = self.config()
        cur.update(kw)
        c = (ctypes.c_int32 * 5)(*[cur[k] for k in ("form", "hjb", "hw", "w", "deep_nk")])
        _abi.call("gt355_batch_set_config", self.handle, c)

    def info(self):
        """(work items, persistent workgroups)."""
        n, g = ctypes.c_int(), ctypes.c_int()
        _abi.call("gt355_batch_info", self.handle, ctypes.byref(n), ctypes.byref(g))
        return n.value, g.value

    def close(self):
        if self.handle:
            _abi.lib().gt355_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def points(self):
        return sum(m["grid"].i_size * m["grid"].j_size * m["grid"].k_size for m in self.members)


class deferred:
    """A deferred-run scope (gt355_batch_begin / gt355_batch_end) on ``stream`` (a torch stream, a raw handle or
    None): the hdiff, vadv and hdiff_vadv calls made on that stream inside the ``with`` block are recorded and run
    at its end as persistent batches, with the results of running them one after another.

        with deferred(stream):
            for d in domains:
                stencils.hdiff(...); stencils.vadv(...)
    """

    def __init__(self, stream=None):
        # the library's Stream (.handle), a torch stream (.cuda_stream) or a raw handle
        self.handle = getattr(stream, "handle", getattr(stream, "cuda_stream", stream))

    def __enter__(self):
        _abi.call("gt355_batch_begin", self.handle)
        return self

    def flush(self):
        """Runs the calls recorded so far; the scope stays open."""
        _abi.call("gt355_batch_flush", self.handle)

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None:
            _abi.call("gt355_batch_end", self.handle)
        else:
            _abi.lib().gt355_batch_end(self.handle)  # already unwinding: the scope closes, its status is dropped
        return False


def draw_sizes(seed, n=64, ij=(48, 384), k=(40, 137)):
    """n domain sizes (ni, nj, nk): ni and nj drawn independently and uniformly from {ij[0]..ij[1]}, nk from
    {k[0]..k[1]} (numpy's PCG64 generator seeded with ``seed``)."""
    rng = np.random.default_rng(seed)
    ni = rng.integers(ij[0], ij[1] + 1, size=n)
    nj = rng.integers(ij[0], ij[1] + 1, size=n)
    nk = rng.integers(k[0], k[1] + 1, size=n)
    return [(int(a), int(b), int(c)) for a, b, c in zip(ni, nj, nk)]


class BatchedTimesteps:
    """Domains of the given sizes with seeded inputs (domain d uses seed ``seed + d``; in, utens_stage, u_pos,
    utens ~ U(-1, 1), coeff ~ U(0, 0.25), wcon ~ U(-wcon_max, wcon_max) on nk+1 levels), halo ``halo`` in i and j.

    A timestep runs the batch, then u and u_stage swap roles (two work tables, one per parity, so no table is
    rewritten).  The halos of u and u_stage both hold the seeded initial values and are not updated: the interior
    evolves under fixed (Dirichlet) halo values.

    utens_stage is read and rewritten by every step, so the tendency feeds back.  wcon_max = 0.05 (the default)
    keeps the implicit vertical system diagonally dominant (|a| + |c| <= 0.025 < dtr = 0.15) and the state finite
    over many steps; with the single-stencil distribution U(-0.5, 0.5) pivots come close to zero and FP32 storage
    overflows within a few steps (a single step is well defined either way).  The kernels have no data-dependent
    control flow, so the values do not change the timing."""

    def __init__(self, sizes, dtype=np.float32, precision=stencils.WIDEN, halo=3, seed=42, dtr=0.15, wg_per_cu=2,
                 wcon_max=0.05):
        if halo < 2:
            raise ValueError("hdiff needs a halo of at least 2")
        self.sizes, self.dtr = list(sizes), dtr
        members = ([], [])
        self.domains = []
        for d, (ni, nj, nk) in enumerate(self.sizes):
            s = seed + d
            b = builder().type(dtype).dimensions(ni + 2 * halo, nj + 2 * halo, nk).halos(halo, halo, 0)
            f = {
                "u": b.uniform(s, FIELD_IDS["in"], -1.0, 1.0).build(),
                "v": b.uniform(s, FIELD_IDS["in"], -1.0, 1.0).build(),  # same halo; interior overwritten
                "coeff": b.uniform(s, FIELD_IDS["coeff"], 0.0, 0.25).build(),
                "lap": b.value(0).build(),
                "utens_stage": b.uniform(s, FIELD_IDS["utens_stage"], -1.0, 1.0).build(),
                "u_pos": b.uniform(s, FIELD_IDS["u_pos"], -1.0, 1.0).build(),
                "utens": b.uniform(s, FIELD_IDS["utens"], -1.0, 1.0).build(),
                "wcon": builder().type(dtype).dimensions(ni + 2 * halo, nj + 2 * halo, nk + 1)
                .halos(halo, halo, 0).uniform(s, FIELD_IDS["wcon"], -wcon_max, wcon_max).build(),
            }
            grid = make_grid(halo_descriptor.symmetric(halo, ni + 2 * halo),
                             halo_descriptor.symmetric(halo, nj + 2 * halo), nk)
            f["grid"] = grid
            self.domains.append(f)
            common = {n: f[n] for n in ("coeff", "lap", "utens_stage", "wcon", "u_pos", "utens")}
            members[0].append(dict(common, in_=f["u"], u_stage=f["v"], grid=grid))
            members[1].append(dict(common, in_=f["v"], u_stage=f["u"], grid=grid))
        self.batches = [Batch(m, precision, wg_per_cu) for m in members]
        self.parity = 0

    def step(self, stream=None):
        self.batches[self.parity].run(self.dtr, stream.handle if stream is not None else None)
        self.parity ^= 1

    def current(self, d):
        """The DataStore holding domain d's current state (the last step's u_stage)."""
        return self.domains[d]["u" if self.parity == 0 else "v"]

    def capture(self, nsteps, stream):
        """A graph of nsteps timesteps (even, so the replayed graph ends in the parity it started from)."""
        if nsteps % 2:
            raise ValueError("capture an even number of steps")
        p = self.parity

        def record(st):
            for _ in range(nsteps):
                self.step(st)

        g = capture(stream, record)
        self.parity = p
        return g

    def points(self):
        return sum(ni * nj * nk for ni, nj, nk in self.sizes)
