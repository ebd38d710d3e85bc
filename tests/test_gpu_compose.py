"""GPU parity of the hdiff -> vadv composition (u_stage = hdiff(in, coeff); vadv of u_stage) against the oracle
chain on the same seeded inputs.

* FP64 and native FP32: bit-identical to oracle hdiff followed by oracle vadv, for the fused single-launch kernel
  and for the two-launch sequence.
* Precision form (FP32 storage, FP64 compute): the fused kernel passes u_stage to the Thomas sweep in FP64, so its
  outputs equal the FP64 oracle chain on the widened FP32 inputs rounded once to FP32, bit for bit.  The two-launch
  sequence (the default dispatch) stores u_stage in FP32 between the launches.  Both are held to the bar of 1e-5
  relative (max|err|/max|ref|) against that FP64 oracle for each output field; max and RMS errors are printed.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

SEED = 7
H = 3


def make_case(gt, ni, nj, nk, dtype):
    from gridtools_amd import inputs

    b = gt.builder().type(dtype).dimensions(ni + 2 * H, nj + 2 * H, nk).halos(H, H, 0)
    f = {
        "in": b.uniform(SEED, inputs.FIELD_IDS["in"], -1.0, 1.0).build(),
        "coeff": b.uniform(SEED, inputs.FIELD_IDS["coeff"], 0.0, 0.25).build(),
        "u_stage": b.value(0).build(),
        "utens_stage": b.uniform(SEED, inputs.FIELD_IDS["utens_stage"], -1.0, 1.0).build(),
        "u_pos": b.uniform(SEED, inputs.FIELD_IDS["u_pos"], -1.0, 1.0).build(),
        "utens": b.uniform(SEED, inputs.FIELD_IDS["utens"], -1.0, 1.0).build(),
        "wcon": gt.builder().type(dtype).dimensions(ni + 2 * H, nj + 2 * H, nk + 1).halos(H, H, 0)
        .uniform(SEED, inputs.FIELD_IDS["wcon"], -0.5, 0.5).build(),
    }
    grid = gt.make_grid(gt.halo_descriptor.symmetric(H, ni + 2 * H), gt.halo_descriptor.symmetric(H, nj + 2 * H), nk)
    return f, grid


def oracle_chain(oracle, host, ni, nj, nk, dtr, hdiff_mode, vadv_mode, round_intermediate_to=None):
    """hdiff then vadv on host arrays (dict of [k, j, i_padded]); returns (u_stage, utens_stage) interiors."""
    V = oracle.View
    ust = np.zeros_like(host["in"])
    oracle.hdiff(V(ust, H, H), V(host["in"], H, H), V(host["coeff"], H, H), ni, nj, nk, mode=hdiff_mode, threads=8)
    if round_intermediate_to is not None:
        ust = ust.astype(round_intermediate_to).astype(ust.dtype)
    uts = host["utens_stage"].copy()
    oracle.vadv(V(uts, H, H), V(ust, H, H), V(host["wcon"], H, H), V(host["u_pos"], H, H), V(host["utens"], H, H),
                dtr, ni, nj, nk, mode=vadv_mode, threads=8)
    return ust[:nk, H:H + nj, H:H + ni], uts[:nk, H:H + nj, H:H + ni]


def run_gpu(f, grid, dtr, precision):
    from gridtools_amd import stencils

    stencils.hdiff_vadv(f["in"], f["coeff"], f["u_stage"], f["utens_stage"], f["wcon"], f["u_pos"], f["utens"], dtr,
                        grid, precision=precision)
    nk, nj, ni = grid.k_size, grid.j_size, grid.i_size
    return (f["u_stage"].padded_numpy()[:nk, H:H + nj, H:H + ni],
            f["utens_stage"].padded_numpy()[:nk, H:H + nj, H:H + ni])


@pytest.mark.parametrize("fused", [1, 0, 2])
@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (70, 9, 80), (130, 37, 2), (256, 256, 80), (64, 5, 17), (1, 3, 5),
                                      (127, 4, 61)])
def test_compose_fp64_bit_exact(gpu, oracle, monkeypatch, fused, ni, nj, nk):
    monkeypatch.setenv("GT355_TUNE_hv_fused", str(fused))
    print(f"[case] compose fp64 fused={fused} {ni}x{nj}x{nk}", flush=True)
    f, grid = make_case(gpu, ni, nj, nk, np.float64)
    host = {k: v.padded_numpy() for k, v in f.items()}
    ref_u, ref_t = oracle_chain(oracle, host, ni, nj, nk, 0.15, "native", "native")
    got_u, got_t = run_gpu(f, grid, 0.15, 0)
    assert rel_err(got_t, ref_t) <= 1e-12
    assert np.array_equal(got_u, ref_u)
    assert np.array_equal(got_t, ref_t)


@pytest.mark.parametrize("fused", [0, 1, 2])
@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (96, 20, 17)])
def test_compose_fp32_native_bit_exact(gpu, oracle, monkeypatch, fused, ni, nj, nk):
    monkeypatch.setenv("GT355_TUNE_hv_fused", str(fused))
    f, grid = make_case(gpu, ni, nj, nk, np.float32)
    host = {k: v.padded_numpy() for k, v in f.items()}
    dtr = float(np.float32(0.15))
    ref_u, ref_t = oracle_chain(oracle, host, ni, nj, nk, dtr, "native", "native")
    got_u, got_t = run_gpu(f, grid, 0.15, 0)
    assert np.array_equal(got_u, ref_u)
    assert np.array_equal(got_t, ref_t)


@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (512, 512, 80)])
def test_compose_precision_form(gpu, oracle, monkeypatch, ni, nj, nk):
    print(f"[case] compose precision form {ni}x{nj}x{nk}", flush=True)
    f, grid = make_case(gpu, ni, nj, nk, np.float32)
    host32 = {k: v.padded_numpy() for k, v in f.items()}
    host64 = {k: v.astype(np.float64) for k, v in host32.items()}
    dtr = float(np.float32(0.15))
    # the FP64 oracle on the same FP32-rounded inputs
    ref_u, ref_t = oracle_chain(oracle, host64, ni, nj, nk, dtr, "native", "native")
    for fused in ("1", "2"):
        # both fused kernels (LDS tile, DPP register sweep): the FP64 chain, outputs rounded once, bit for bit
        f1, _ = make_case(gpu, ni, nj, nk, np.float32)
        monkeypatch.setenv("GT355_TUNE_hv_fused", fused)
        got_u, got_t = run_gpu(f1, grid, 0.15, 1)
        assert np.array_equal(got_u, ref_u.astype(np.float32)), fused
        assert np.array_equal(got_t, ref_t.astype(np.float32)), fused
        for name, got, ref in (("u_stage", got_u, ref_u), ("utens_stage", got_t, ref_t)):
            e = rel_err(got, ref)
            rms = np.sqrt(np.mean((got.astype(np.float64) - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
            print(f"[precision] fused={fused} {name}: max rel {e:.3e}, rms rel {rms:.3e}", flush=True)
            assert e <= 1e-5
    # two-launch sequence: u_stage is rounded to FP32 storage before the ill-conditioned solve
    f2, _ = make_case(gpu, ni, nj, nk, np.float32)
    monkeypatch.setenv("GT355_TUNE_hv_fused", "0")
    got_u2, got_t2 = run_gpu(f2, grid, 0.15, 1)
    assert np.array_equal(got_u2, ref_u.astype(np.float32))
    e2 = rel_err(got_t2, ref_t)
    rms2 = np.sqrt(np.mean((got_t2.astype(np.float64) - ref_t) ** 2)) / np.sqrt(np.mean(ref_t ** 2))
    print(f"[precision] two-launch utens_stage: max rel {e2:.3e}, rms rel {rms2:.3e}", flush=True)
    assert e2 <= 1e-5


@pytest.mark.parametrize("fused", [2, 1, 0])
def test_compose_deterministic_with_partial_strips(gpu, monkeypatch, fused):
    """utens_stage is read and rewritten in place, so each column must be owned by exactly one wavefront: repeated
    runs on identical inputs, with domains whose last i-strip is partial, give identical bits."""
    monkeypatch.setenv("GT355_TUNE_hv_fused", str(fused))
    outs = []
    for _ in range(3):
        f, grid = make_case(gpu, 130, 40, 61, np.float64)
        outs.append(run_gpu(f, grid, 0.15, 0))
    for u, t in outs[1:]:
        assert np.array_equal(u, outs[0][0])
        assert np.array_equal(t, outs[0][1])
