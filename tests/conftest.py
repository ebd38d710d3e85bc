import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgt355.so on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding

    binding.lib()
    return binding


@pytest.fixture(scope="session")
def gpu():
    """Initialises device 0 through the C ABI; a GPU test fails (not skips) if that is impossible."""
    import gridtools_amd as gt

    gt.init(0)
    info = gt.device_info()
    print(f"\n[gpu] {info['name']} {info['arch']} CUs={info['cu_count']}", flush=True)
    return gt


class _TuneAwareMonkeypatch:
    """monkeypatch whose environment changes of GT355_TUNE_* knobs reach the library, which reads them once per
    process (gt355_tune_env_refresh)."""

    def __init__(self, mp):
        self._mp = mp

    @staticmethod
    def _refresh(name):
        if str(name).startswith("GT355_TUNE_"):
            from gridtools_amd import _abi

            _abi.call("gt355_tune_env_refresh")

    def setenv(self, name, value, prepend=None):
        self._mp.setenv(name, value, prepend)
        self._refresh(name)

    def delenv(self, name, raising=True):
        self._mp.delenv(name, raising)
        self._refresh(name)

    def __getattr__(self, attr):
        return getattr(self._mp, attr)


@pytest.fixture
def monkeypatch(monkeypatch):
    wrapped = _TuneAwareMonkeypatch(monkeypatch)
    yield wrapped
    monkeypatch.undo()
    if "gridtools_amd._abi" in sys.modules:
        sys.modules["gridtools_amd._abi"].call("gt355_tune_env_refresh")


def rel_err(a, b):
    """max |a-b| / max |b| over finite entries."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.max(np.abs(a - b)) if a.size else 0.0
    m = np.max(np.abs(b)) if b.size else 0.0
    return d / m if m > 0 else d
