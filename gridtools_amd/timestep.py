"""Timestep sequences on one domain: the large-domain loop and hipGraph capture.

One timestep of the horizontal-diffusion / vertical-advection operator:

1. ``u_stage = hdiff(u, coeff)`` and the vertical advection of ``u_stage`` into ``utens_stage``, as one
   composition launch (gt355_hdiff_vadv);
2. boundary conditions on the halo of ``u_stage`` and ``utens_stage``: periodic in i, zero-gradient in j, either
   as one local halo fill (gt355_halo_fill) or through the halo-pack kernels (zero-gradient j, then the single-rank
   pack / periodic exchange / unpack of the i faces, gt355_halo_exchange_local);
3. ``u`` and ``u_stage`` swap roles, so the diffused field is the next step's input.

Sequences can be captured once into a hipGraph and replayed (both swap parities are recorded, so an even number of
steps is one graph).
"""
import ctypes

import numpy as np

from . import _abi, boundaries, stencils
from .grid import halo_descriptor, make_grid
from .inputs import FIELD_IDS
from .storage import builder


class Graph:
    """A captured launch sequence (gt355_graph_*)."""

    def __init__(self, handle):
        self.handle = handle

    def launch(self, stream):
        _abi.call("gt355_graph_launch", self.handle, stream.handle)

    def __del__(self):
        try:
            if self.handle:
                _abi.lib().gt355_graph_destroy(self.handle)
        except Exception:
            pass


def capture(stream, fn):
    """Records everything fn(stream) launches on ``stream`` into a graph (stream must not be the null stream)."""
    _abi.call("gt355_graph_begin", stream.handle)
    try:
        fn(stream)
    except Exception:
        g = ctypes.c_void_p()
        _abi.lib().gt355_graph_end(stream.handle, ctypes.byref(g))
        if g.value:
            _abi.lib().gt355_graph_destroy(g)
        raise
    g = ctypes.c_void_p()
    _abi.call("gt355_graph_end", stream.handle, ctypes.byref(g))
    return Graph(g)


class HdiffVadvTimestepper:
    """The operator above on an ni x nj x nk interior with a ``halo``-point ij halo (>= 2), inputs from the seeded
    generator: u, utens_stage, u_pos, utens ~ U(-1, 1), coeff ~ U(0, 0.25), wcon ~ U(-wcon_max, wcon_max) on nk+1
    levels.

    wcon_max defaults to 0.05, which keeps the implicit vertical system diagonally dominant (|a| + |c| <= wcon_max / 2
    = 0.025 < dtr = 0.15), so the tendency utens_stage, which each step feeds back into the next, accumulates
    linearly (max |utens_stage| ~ 1e2 after 20 steps on 64 x 64 x 80, all finite).  With the single-stencil
    distribution U(-0.5, 0.5) pivots come within 1e-7 of zero and the tendency grows ~1e3-fold per step: FP32 storage
    overflows after about eight steps.

    bc="fill": one halo-fill launch per step.  bc="exchange": the halo-pack route -- zero-gradient j by halo fill,
    then the periodic i faces by the single-rank pack / exchange / unpack (gt355_halo_exchange_local) over all rows
    (j halo rows included, so the corners get the wrapped, clamped values); both give the same bits."""

    def __init__(self, ni, nj, nk, dtype=np.float32, precision=stencils.WIDEN, halo=3, seed=42, dtr=0.15,
                 bc="fill", wcon_max=0.05):
        if bc not in ("fill", "exchange"):
            raise ValueError(f"bc must be 'fill' or 'exchange', not {bc!r}")
        self.bc = bc
        if halo < 2:
            raise ValueError("hdiff needs a halo of at least 2")
        self.ni, self.nj, self.nk, self.halo = ni, nj, nk, halo
        self.precision, self.dtr = precision, dtr
        b = builder().type(dtype).dimensions(ni + 2 * halo, nj + 2 * halo, nk).halos(halo, halo, 0)
        self.u = b.uniform(seed, FIELD_IDS["in"], -1.0, 1.0).build()
        self.u_stage = b.value(0).build()
        self.coeff = b.uniform(seed, FIELD_IDS["coeff"], 0.0, 0.25).build()
        self.utens_stage = b.uniform(seed, FIELD_IDS["utens_stage"], -1.0, 1.0).build()
        self.u_pos = b.uniform(seed, FIELD_IDS["u_pos"], -1.0, 1.0).build()
        self.utens = b.uniform(seed, FIELD_IDS["utens"], -1.0, 1.0).build()
        self.wcon = builder().type(dtype).dimensions(ni + 2 * halo, nj + 2 * halo, nk + 1).halos(halo, halo, 0) \
            .uniform(seed, FIELD_IDS["wcon"], -wcon_max, wcon_max).build()
        hi = halo_descriptor.symmetric(halo, ni + 2 * halo)
        hj = halo_descriptor.symmetric(halo, nj + 2 * halo)
        self.grid = make_grid(hi, hj, nk)
        self.halos = [hi, hj, halo_descriptor.of_size(nk)]
        # the initial state satisfies the boundary conditions
        boundaries.halo_fill([self.u], self.halos, (boundaries.PERIODIC, boundaries.ZERO_GRADIENT, boundaries.NONE))
        self.steps_done = 0
        self._xbuf = None
        if bc == "exchange":
            # the i faces over every row: j treated as halo-free for the exchange
            self._xhalos = [hi, halo_descriptor.of_size(nj + 2 * halo), self.halos[2]]
            self._xmask = boundaries.direction_mask(lambda di, dj, dk: di != 0 and dj == 0 and dk == 0)
            nbytes = boundaries.halo_buffer_size(self.u.code, 2, self._xhalos, self._xmask)
            buf = ctypes.c_void_p()
            _abi.call("gt355_malloc", ctypes.c_size_t(nbytes), ctypes.byref(buf))
            self._xbuf = buf

    def step(self, stream=None):
        s = stream.handle if stream is not None else None
        stencils.hdiff_vadv(self.u, self.coeff, self.u_stage, self.utens_stage, self.wcon, self.u_pos, self.utens,
                            self.dtr, self.grid, precision=self.precision, stream=s)
        if self.bc == "fill":
            boundaries.halo_fill([self.u_stage, self.utens_stage], self.halos,
                                 (boundaries.PERIODIC, boundaries.ZERO_GRADIENT, boundaries.NONE), stream=s)
        else:
            boundaries.halo_fill([self.u_stage, self.utens_stage], self.halos,
                                 (boundaries.NONE, boundaries.ZERO_GRADIENT, boundaries.NONE), stream=s)
            boundaries.halo_exchange_local([self.u_stage, self.utens_stage], self._xhalos, self._xbuf.value,
                                           self._xmask, stream=s)
        self.u, self.u_stage = self.u_stage, self.u
        self.steps_done += 1

    def capture(self, nsteps, stream):
        """A graph of nsteps timesteps (nsteps even: both swap parities are recorded, so the fields end in the
        roles they started in and the graph can be replayed)."""
        if nsteps % 2:
            raise ValueError("capture an even number of steps")

        def record(st):
            for _ in range(nsteps):
                self.step(st)

        g = capture(stream, record)
        self.steps_done -= nsteps  # capture records, it does not run
        return g

    def points(self):
        return self.ni * self.nj * self.nk

    def __del__(self):
        try:
            if self._xbuf is not None and self._xbuf.value:
                _abi.lib().gt355_free(self._xbuf)
        except Exception:
            pass
