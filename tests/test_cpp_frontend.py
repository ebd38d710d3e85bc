"""The C++ front end (include/gt355/gridtools.hpp): its regression program, written like the reference's
tests/regression/*.cpp, built by the package build.  Host-only cases run everywhere; the device cases (hdiff,
vadv, laplacian, copy, the precision-form composition, expandable_run, the batching form, boundaries and halo fill,
each bit-exact
against the oracle or the reference's known answers) run on the GPU."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "gt355_regression")


def run(*args, timeout):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run the package build (__graft_entry__.build())")
    res = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=timeout)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "[FAIL]" not in res.stdout
    return res.stdout


def test_cpp_frontend_host_cases():
    out = run("--host", timeout=60)
    assert out.count("[ok]") == 3


@pytest.mark.gpu
def test_cpp_frontend_regression_on_device():
    out = run(timeout=240)
    assert out.count("[ok]") >= 27
