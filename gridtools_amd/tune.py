"""The autotuner: pick the fastest hdiff configuration for a problem shape on this device and keep it.

    from gridtools_amd import tune
    cfg, us = tune.hdiff(out, in_, coeff, grid)   # times every instantiated configuration, records the best
    tune.save("gt355_tune.txt")                   # later processes: tune.load("gt355_tune.txt")
"""
import ctypes

from . import _abi
from .stencils import NATIVE

IMPL_NAMES = {0: "lds-tile", 1: "register-sweep", 2: "register-sweep-pairs"}


def hdiff(out, in_, coeff, grid, precision=NATIVE, reps=20, stream=None):
    """Returns ({impl, a, b, c}, mean microseconds) of the winner; later gt355_hdiff calls of this shape use it."""
    o, i, c = (f.sid(grid.origin) for f in (out, in_, coeff))
    best = (ctypes.c_int32 * 4)()
    us = ctypes.c_double()
    _abi.call("gt355_tune_hdiff", precision, ctypes.byref(o), ctypes.byref(i), ctypes.byref(c), grid.i_size,
              grid.j_size, grid.k_size, reps, stream, best, ctypes.byref(us))
    impl, a, b, cc = list(best)
    cfg = {"impl": IMPL_NAMES[impl], "tile": (a, b, cc)} if impl == 0 else {"impl": IMPL_NAMES[impl], "jb": a, "p": b}
    return cfg, us.value


def save(path):
    _abi.call("gt355_tune_save", str(path).encode())


def load(path):
    n = ctypes.c_int()
    _abi.call("gt355_tune_load", str(path).encode(), ctypes.byref(n))
    return n.value


def clear():
    _abi.call("gt355_tune_clear")


def lookup(kernel_key, shape):
    """The table's entry for ``kernel_key`` ("hdiff" or "batch") and ``shape`` (hdiff: storage bytes, compute
    bytes, ni, nj, nk; batch: signature, storage bytes, compute bytes, form): (cfg list of 4, microseconds), or
    None if there is none."""
    sh = (ctypes.c_int64 * len(shape))(*shape)
    cfg = (ctypes.c_int32 * 4)()
    us, found = ctypes.c_double(), ctypes.c_int()
    _abi.call("gt355_tune_lookup", kernel_key.encode(), sh, len(shape), cfg, ctypes.byref(us), ctypes.byref(found))
    return (list(cfg), us.value) if found.value else None


def store(kernel_key, shape, cfg, us=0.0):
    """Records an entry (hdiff: cfg = [impl, a, b, c], an instantiated configuration; batch: [hjb, hw, w, deep_nk])
    for the next launch of that shape."""
    sh = (ctypes.c_int64 * len(shape))(*shape)
    _abi.call("gt355_tune_store", kernel_key.encode(), sh, len(shape), (ctypes.c_int32 * 4)(*cfg), us)
