"""hipEvent timing and streams through the C ABI (the role of timer_cuda, common/timer/timer_cuda.hpp:21-54)."""
import ctypes

from . import _abi


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        _abi.call("gt355_stream_create", ctypes.byref(s))
        self.handle = s

    def synchronize(self):
        _abi.call("gt355_stream_synchronize", self.handle)

    def __del__(self):
        try:
            if self.handle:
                _abi.lib().gt355_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _abi.call("gt355_event_create", ctypes.byref(e))
        self.handle = e

    def record(self, stream=None):
        _abi.call("gt355_event_record", self.handle, stream.handle if isinstance(stream, Stream) else stream)

    def synchronize(self):
        _abi.call("gt355_event_synchronize", self.handle)

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _abi.call("gt355_event_elapsed_ms", self.handle, end.handle, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                _abi.lib().gt355_event_destroy(self.handle)
        except Exception:
            pass


def synchronize():
    _abi.call("gt355_synchronize")


def mall_flush(stream=None):
    _abi.call("gt355_mall_flush", stream.handle if isinstance(stream, Stream) else stream)


class Range:
    """with Range("name"): ... -- a named range for tracers (gt355_range_push / gt355_range_pop)."""

    def __init__(self, name):
        self.name = name.encode()

    def __enter__(self):
        _abi.call("gt355_range_push", self.name)
        return self

    def __exit__(self, *exc):
        _abi.call("gt355_range_pop")
        return False
