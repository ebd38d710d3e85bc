"""The generic stencil engine (include/gt355/stencil.hpp): user specs written like the reference's cartesian
regression tests -- copy, laplacian, the four-stage hdiff with ij-cached temporaries, the forward / backward vadv with
interval overloads and k-cached temporaries, the tridiagonal solve with expressions and inout fields,
simple_hori_diff with j-only metric fields, stencil functions (horizontal_diffusion_fused and the five variations of
horizontal_diffusion_functions), positional, multistage fusion through a temporary, expandable_run of a user spec --
run on the GPU and checked against the oracle bit for bit or the reference's known answers
(tests/cpp/generic.hip, built by the package build)."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "gt355_generic")


@pytest.mark.gpu
def test_generic_engine_on_device():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run the package build (__graft_entry__.build())")
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "[FAIL]" not in res.stdout
    assert res.stdout.count("[ok]") >= 24
