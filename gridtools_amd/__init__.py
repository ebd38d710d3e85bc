"""gridtools_amd -- an MI355X (gfx950) stencil backend with the GridTools storage / grid / stencil / boundary
interfaces, backed by the hand-written HIP kernels of libgt355.so (C ABI: include/gt355.h)."""
from . import _abi
from .grid import Grid, halo_descriptor, make_grid
from .storage import DataStore, builder
from .timing import synchronize

NATIVE = _abi.NATIVE
WIDEN = _abi.WIDEN
F64 = _abi.F64
F32 = _abi.F32


def init(device: int = 0) -> None:
    """Selects the device (gt355_init)."""
    _abi.call("gt355_init", device)


def device_info() -> dict:
    info = _abi.DeviceInfo()
    _abi.call("gt355_get_device_info", info)
    return {
        "name": info.name.decode(),
        "arch": info.arch.decode(),
        "cu_count": info.cu_count,
        "xcd_count": info.xcd_count,
        "hbm_bytes": info.hbm_bytes,
        "lds_per_cu": info.lds_per_cu,
        "l2_bytes": info.l2_bytes,
        "clock_khz": info.clock_khz,
        "mem_clock_khz": info.mem_clock_khz,
        "mem_bus_width": info.mem_bus_width,
        "wavefront": info.wavefront,
    }


__all__ = ["Grid", "halo_descriptor", "make_grid", "DataStore", "builder", "init", "device_info", "NATIVE",
           "WIDEN", "F64", "F32", "synchronize"]
