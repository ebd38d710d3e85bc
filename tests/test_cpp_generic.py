"""The generic stencil engine (include/gt355/stencil.hpp) on the reference's own specs (tests/cpp/reference_specs.hpp):
copy, laplacian, the four-stage hdiff with ij-cached temporaries, the forward / backward vadv with interval overloads
and k-cached temporaries, the tridiagonal solve with expressions and inout fields, simple_hori_diff with j-only metric
fields, stencil functions (horizontal_diffusion_fused and the five variations of horizontal_diffusion_functions),
positional, multistage fusion through a temporary, expandable_run, the k-cache unit tests over axis<3>, the k-parallel
tests over axis<2> / axis<3>, vertical multistages with horizontal dependencies between stages, whole_axis_access (a renamed dimension read by a 4-D accessor), the precision policy,
temporaries on two streams and user boundary functors -- run on the GPU and checked against the oracle bit for bit
or the reference's known answers (tests/cpp/generic.hip, built by the package build).

The program runs three times: as built (the reference's hdiff and vadv specs are recognised and run on the
hand-written kernels), with GT355_ENGINE_FAST=0, which keeps every spec on the generic engine's own kernels, and with
GT355_ENGINE_MSS_FUSION=0 as well, which launches each vertical multistage on its own instead of fusing consecutive
ones that pass data at the column itself (the forward / backward pairs of the vadv and tridiagonal specs)."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "gt355_generic")


@pytest.mark.gpu
@pytest.mark.parametrize("fast,fusion", [("1", "1"), ("0", "1"), ("0", "0")])
def test_generic_engine_on_device(fast, fusion):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run the package build (__graft_entry__.build())")
    print(f"[case] gt355_generic GT355_ENGINE_FAST={fast} GT355_ENGINE_MSS_FUSION={fusion}", flush=True)
    env = dict(os.environ, GT355_ENGINE_FAST=fast, GT355_ENGINE_MSS_FUSION=fusion)
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=240, env=env)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "[FAIL]" not in res.stdout
    assert res.stdout.count("[ok]") >= 50
