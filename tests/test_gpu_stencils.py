"""GPU parity of the hot-path stencils: libgt355 (through the C ABI) against the CPU oracle on the same seeded
inputs.  FP64 and native FP32 must agree bit for bit (both sides evaluate the functors' operation order with
IEEE operations and no contraction); the bar written into every test is the north-star tolerance: relative
error <= 1e-12 for FP64.  The precision form (FP32 storage, FP64 compute) must equal the oracle's mixed variant
bit for bit and stay within 1e-5 relative of the FP64 oracle fed the same FP32-rounded inputs.

Sizes: the reference's typed-suite sizes (12x33x61, 23x11x43: tests/include/test_environment.hpp:69-77), the
hdiff CPU config (64x64x60, halo 2), the 256x256x80 benchmark size, and block-remainder shapes."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5
SEED = 42

HDIFF_SIZES = [(12, 33, 61), (23, 11, 43), (64, 64, 60), (1, 1, 1), (65, 9, 3), (129, 70, 17), (256, 256, 80)]
VADV_SIZES = [(12, 33, 61), (23, 11, 43), (1, 1, 2), (65, 9, 3), (70, 130, 7), (256, 256, 80)]


def hdiff_case(gt, ni, nj, nk, dtype, h=2):
    from gridtools_amd import inputs

    b = gt.builder().type(dtype).dimensions(ni + 2 * h, nj + 2 * h, nk).halos(h, h, 0)
    inp = b.uniform(SEED, inputs.FIELD_IDS["in"], -1.0, 1.0).build()
    coeff = b.uniform(SEED, inputs.FIELD_IDS["coeff"], 0.0, 0.25).build()
    out = b.value(0).build()
    grid = gt.make_grid(gt.halo_descriptor.symmetric(h, ni + 2 * h), gt.halo_descriptor.symmetric(h, nj + 2 * h), nk)
    return inp, coeff, out, grid


def oracle_hdiff(oracle, inp, coeff, h, ni, nj, nk, mode):
    a_in, a_co = inp.padded_numpy(), coeff.padded_numpy()
    if mode == "wide":
        a_in, a_co, mode = a_in.astype(np.float64), a_co.astype(np.float64), "native"
    res = np.zeros_like(a_in)
    V = oracle.View
    oracle.hdiff(V(res, h, h), V(a_in, h, h), V(a_co, h, h), ni, nj, nk, mode=mode, threads=8)
    return res[:nk, h:h + nj, h:h + ni]


@pytest.mark.parametrize("ni,nj,nk", HDIFF_SIZES)
def test_hdiff_fp64_parity(gpu, oracle, ni, nj, nk):
    from gridtools_amd import stencils

    print(f"[case] hdiff fp64 {ni}x{nj}x{nk}", flush=True)
    inp, coeff, out, grid = hdiff_case(gpu, ni, nj, nk, np.float64)
    stencils.hdiff(out, inp, coeff, grid)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    ref = oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, "native")
    assert rel_err(got, ref) <= FP64_TOL
    assert np.array_equal(got, ref), "FP64 path is expected to be bit-identical to the oracle"
    # the halo of out is untouched
    full = out.padded_numpy()
    assert np.all(full[:nk, :2, :] == 0) and np.all(full[:nk, :, :2] == 0)


@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (64, 64, 60), (256, 256, 80)])
def test_hdiff_fp32_native_parity(gpu, oracle, ni, nj, nk):
    from gridtools_amd import stencils

    print(f"[case] hdiff fp32 {ni}x{nj}x{nk}", flush=True)
    inp, coeff, out, grid = hdiff_case(gpu, ni, nj, nk, np.float32)
    stencils.hdiff(out, inp, coeff, grid, precision=stencils.NATIVE)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    ref = oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, "native")
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (256, 256, 80)])
def test_hdiff_precision_form(gpu, oracle, ni, nj, nk):
    from gridtools_amd import stencils

    print(f"[case] hdiff widen {ni}x{nj}x{nk}", flush=True)
    inp, coeff, out, grid = hdiff_case(gpu, ni, nj, nk, np.float32)
    stencils.hdiff(out, inp, coeff, grid, precision=stencils.WIDEN)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    assert np.array_equal(got, oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, "mixed"))
    ref64 = oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, "wide")
    assert rel_err(got, ref64) <= MIXED_TOL


@pytest.mark.parametrize("impl", [0, 1, 2])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 0), (np.float32, 1)])
@pytest.mark.parametrize("ni,nj,nk", [(12, 33, 5), (130, 70, 3), (256, 256, 4)])
def test_hdiff_both_kernels_all_precisions(gpu, oracle, monkeypatch, impl, dtype, precision, ni, nj, nk):
    """The LDS-tile kernel, the DPP register sweep and its column-pair form are each picked by the tuner for some
    shapes; all must agree with the oracle for every precision (selected through the GT355_TUNE_hdiff_impl knob)."""
    from gridtools_amd import stencils

    monkeypatch.setenv("GT355_TUNE_hdiff_impl", str(impl))
    print(f"[case] hdiff impl={impl} {np.dtype(dtype).name} prec={precision} {ni}x{nj}x{nk}", flush=True)
    inp, coeff, out, grid = hdiff_case(gpu, ni, nj, nk, dtype)
    stencils.hdiff(out, inp, coeff, grid, precision=precision)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    mode = "mixed" if precision == 1 else "native"
    assert np.array_equal(got, oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, mode))


@pytest.mark.parametrize("jb,p", [(4, 4), (8, 4), (16, 8), (32, 4), (32, 8)])
def test_hdiff_sweep_variants(gpu, oracle, monkeypatch, jb, p):
    from gridtools_amd import stencils

    monkeypatch.setenv("GT355_TUNE_hdiff_impl", "1")
    monkeypatch.setenv("GT355_TUNE_hdiff_jb", str(jb))
    monkeypatch.setenv("GT355_TUNE_hdiff_p", str(p))
    ni, nj, nk = 97, 75, 3
    inp, coeff, out, grid = hdiff_case(gpu, ni, nj, nk, np.float64)
    stencils.hdiff(out, inp, coeff, grid)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    assert np.array_equal(got, oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, "native"))


@pytest.mark.parametrize("jb,p", [(4, 4), (8, 4), (8, 8), (16, 4), (16, 8)])
@pytest.mark.parametrize("ni,nj", [(130, 37), (256, 20), (2, 9), (97, 11)])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1)])
def test_hdiff_pair_sweep_variants(gpu, oracle, monkeypatch, jb, p, ni, nj, dtype, precision):
    """Column-pair sweep shapes, with a partial last 128-column strip (130), exactly two strips (256), a single pair
    (2) and an odd ni (97, which the launcher hands to the one-column sweep)."""
    from gridtools_amd import stencils

    monkeypatch.setenv("GT355_TUNE_hdiff_impl", "2")
    monkeypatch.setenv("GT355_TUNE_hdiff_jb", str(jb))
    monkeypatch.setenv("GT355_TUNE_hdiff_p", str(p))
    nk = 3
    print(f"[case] hdiff pairs jb={jb} p={p} {np.dtype(dtype).name} {ni}x{nj}x{nk}", flush=True)
    inp, coeff, out, grid = hdiff_case(gpu, ni, nj, nk, dtype)
    stencils.hdiff(out, inp, coeff, grid, precision=precision)
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    mode = "mixed" if precision == 1 else "native"
    assert np.array_equal(got, oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, mode))


def vadv_case(gt, ni, nj, nk, dtype, h=3):
    from gridtools_amd import inputs

    b = gt.builder().type(dtype).dimensions(ni + 2 * h, nj + 2 * h, nk).halos(h, h, 0)
    f = {}
    for name in ("utens_stage", "u_stage", "u_pos", "utens"):
        f[name] = b.uniform(SEED, inputs.FIELD_IDS[name], -1.0, 1.0).build()
    # wcon lives on nk+1 staggered levels
    f["wcon"] = gt.builder().type(dtype).dimensions(ni + 2 * h, nj + 2 * h, nk + 1).halos(h, h, 0) \
        .uniform(SEED, inputs.FIELD_IDS["wcon"], -0.5, 0.5).build()
    grid = gt.make_grid(gt.halo_descriptor.symmetric(h, ni + 2 * h), gt.halo_descriptor.symmetric(h, nj + 2 * h), nk)
    return f, grid


def oracle_vadv(oracle, f, dtr, h, ni, nj, nk, mode):
    a = {k: v.padded_numpy() for k, v in f.items()}
    if mode == "wide":
        a = {k: v.astype(np.float64) for k, v in a.items()}
        mode = "native"
        dtr = float(np.float32(dtr))
    V = oracle.View
    res = a["utens_stage"].copy()
    oracle.vadv(V(res, h, h), V(a["u_stage"], h, h), V(a["wcon"], h, h), V(a["u_pos"], h, h), V(a["utens"], h, h),
                dtr, ni, nj, nk, mode=mode, threads=8)
    return res[:nk, h:h + nj, h:h + ni]


def run_vadv(f, grid, dtr, precision):
    from gridtools_amd import stencils

    stencils.vadv(f["utens_stage"], f["u_stage"], f["wcon"], f["u_pos"], f["utens"], dtr, grid, precision=precision)
    return f["utens_stage"].padded_numpy()


@pytest.mark.parametrize("ni,nj,nk", VADV_SIZES)
def test_vadv_fp64_parity(gpu, oracle, ni, nj, nk):
    print(f"[case] vadv fp64 {ni}x{nj}x{nk}", flush=True)
    f, grid = vadv_case(gpu, ni, nj, nk, np.float64)
    ref = oracle_vadv(oracle, f, 0.15, 3, ni, nj, nk, "native")
    got = run_vadv(f, grid, 0.15, 0)[:nk, 3:3 + nj, 3:3 + ni]
    assert rel_err(got, ref) <= FP64_TOL
    assert np.array_equal(got, ref), "FP64 path is expected to be bit-identical to the oracle"


@pytest.mark.parametrize("ni,nj,nk", [(70, 9, 81), (33, 20, 137)])
def test_vadv_deep_columns_use_scratch_kernel(gpu, oracle, ni, nj, nk):
    """nk above the largest register bucket (80) runs the kernel that keeps ccol/dcol in a pooled buffer."""
    print(f"[case] vadv fp64 deep {ni}x{nj}x{nk}", flush=True)
    f, grid = vadv_case(gpu, ni, nj, nk, np.float64)
    ref = oracle_vadv(oracle, f, 0.15, 3, ni, nj, nk, "native")
    got = run_vadv(f, grid, 0.15, 0)[:nk, 3:3 + nj, 3:3 + ni]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("upos_lds", [0, 1])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1)])
@pytest.mark.parametrize("nk", [2, 17, 43, 64, 80])
def test_vadv_register_buckets_and_upos_placement(gpu, oracle, monkeypatch, upos_lds, dtype, precision, nk):
    """Every k-depth bucket (16..80, including nk far below the bucket size) and both u_pos placements."""
    monkeypatch.setenv("GT355_TUNE_vadv_upos_lds", str(upos_lds))
    ni, nj = 67, 5
    f, grid = vadv_case(gpu, ni, nj, nk, dtype)
    dtr = 0.15 if dtype == np.float64 else float(np.float32(0.15))
    ref = oracle_vadv(oracle, f, dtr, 3, ni, nj, nk, "native" if precision == 0 else "mixed")
    got = run_vadv(f, grid, 0.15, precision)[:nk, 3:3 + nj, 3:3 + ni]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("depth", [4, 8, 12, 16])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1)])
@pytest.mark.parametrize("nk", [70, 80])
def test_vadv_load_ahead_depths(gpu, oracle, monkeypatch, depth, dtype, precision, nk):
    """Every load-ahead depth of the 80-level bucket (GT355_TUNE_vadv_d) is bit-identical to the oracle."""
    monkeypatch.setenv("GT355_TUNE_vadv_d", str(depth))
    ni, nj = 67, 5
    f, grid = vadv_case(gpu, ni, nj, nk, dtype)
    dtr = 0.15 if dtype == np.float64 else float(np.float32(0.15))
    ref = oracle_vadv(oracle, f, dtr, 3, ni, nj, nk, "native" if precision == 0 else "mixed")
    got = run_vadv(f, grid, 0.15, precision)[:nk, 3:3 + nj, 3:3 + ni]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1), (np.float32, 0)])
@pytest.mark.parametrize("ni,nj,nk", [(130, 700, 67), (256, 512, 80), (1, 2100, 72)])
def test_vadv_persistent_strips(gpu, oracle, monkeypatch, dtype, precision, ni, nj, nk):
    """Domains with at least two rounds of 64-column strips per SIMD take the persistent kernel (next strip's first
    levels prefetched during the backward sweep): bit-identical to the oracle, with a partial last strip (130), a
    one-column domain (1) and a depth below the bucket (67, 72)."""
    monkeypatch.setenv("GT355_TUNE_vadv_wide", "0")  # FP32 storage with ni % 4 == 0 would take the wide form
    print(f"[case] vadv persistent {np.dtype(dtype).name} prec={precision} {ni}x{nj}x{nk}", flush=True)
    f, grid = vadv_case(gpu, ni, nj, nk, dtype)
    dtr = 0.15 if dtype == np.float64 else float(np.float32(0.15))
    ref = oracle_vadv(oracle, f, dtr, 3, ni, nj, nk, "native" if precision == 0 else "mixed")
    got = run_vadv(f, grid, 0.15, precision)[:nk, 3:3 + nj, 3:3 + ni]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("precision", [1, 0])
@pytest.mark.parametrize("ni,nj,nk", [(130, 700, 67), (256, 512, 80), (1, 2100, 72)])
def test_vadv_persistent_strips_depth16(gpu, oracle, monkeypatch, precision, ni, nj, nk):
    """The persistent kernel with 16 levels of load-ahead (FP32 storage, GT355_TUNE_vadv_persist_d=16)."""
    monkeypatch.setenv("GT355_TUNE_vadv_persist_d", "16")
    monkeypatch.setenv("GT355_TUNE_vadv_wide", "0")
    print(f"[case] vadv persistent D=16 prec={precision} {ni}x{nj}x{nk}", flush=True)
    f, grid = vadv_case(gpu, ni, nj, nk, np.float32)
    ref = oracle_vadv(oracle, f, float(np.float32(0.15)), 3, ni, nj, nk, "native" if precision == 0 else "mixed")
    got = run_vadv(f, grid, 0.15, precision)[:nk, 3:3 + nj, 3:3 + ni]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("form", ["lds", 3, 4, 5])
@pytest.mark.parametrize("precision", [1, 0])
@pytest.mark.parametrize("ni,nj,nk", [(4, 3, 65), (60, 7, 80), (64, 5, 66), (100, 9, 77), (256, 256, 80),
                                      (132, 500, 71), (512, 40, 79)])
def test_vadv_wide_requests(gpu, oracle, monkeypatch, form, precision, ni, nj, nk):
    """The FP32-storage form with 16-byte requests (four rows per load, staged in LDS; four levels per store) for
    65..80 levels and ni % 4 == 0, both kernels: rows loaded straight into LDS slots ("lds", the default) and rows
    through a register ring of 3, 4 and 5 groups of four levels.  Bit-identical to the oracle with a partial last
    strip (4, 60, 100, 132: lanes past ni clamp their output chunk onto the last four columns), depths below the
    bucket (65, 66, 71, 77, 79: rows above the top level), one and several strips per wavefront (132x500: 1500
    strips)."""
    if form == "lds":
        monkeypatch.setenv("GT355_TUNE_vadv_wide_lds", "1")
    else:
        monkeypatch.setenv("GT355_TUNE_vadv_wide_lds", "0")
        monkeypatch.setenv("GT355_TUNE_vadv_wide_dg", str(form))
    print(f"[case] vadv wide form={form} prec={precision} {ni}x{nj}x{nk}", flush=True)
    f, grid = vadv_case(gpu, ni, nj, nk, np.float32)
    ref = oracle_vadv(oracle, f, float(np.float32(0.15)), 3, ni, nj, nk, "native" if precision == 0 else "mixed")
    got = run_vadv(f, grid, 0.15, precision)
    assert np.array_equal(got[:nk, 3:3 + nj, 3:3 + ni], ref)
    # nothing written outside the interior
    host0 = vadv_case(gpu, ni, nj, nk, np.float32)[0]["utens_stage"].padded_numpy()
    mask = np.ones(got.shape, bool)
    mask[:nk, 3:3 + nj, 3:3 + ni] = False
    assert np.array_equal(got[mask], host0[mask])


@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (256, 256, 80)])
def test_vadv_fp32_native_parity(gpu, oracle, ni, nj, nk):
    print(f"[case] vadv fp32 {ni}x{nj}x{nk}", flush=True)
    f, grid = vadv_case(gpu, ni, nj, nk, np.float32)
    dtr = float(np.float32(0.15))
    ref = oracle_vadv(oracle, f, dtr, 3, ni, nj, nk, "native")
    got = run_vadv(f, grid, 0.15, 0)[:nk, 3:3 + nj, 3:3 + ni]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("ni,nj,nk", [(23, 11, 43), (256, 256, 80)])
def test_vadv_precision_form(gpu, oracle, ni, nj, nk):
    print(f"[case] vadv widen {ni}x{nj}x{nk}", flush=True)
    f, grid = vadv_case(gpu, ni, nj, nk, np.float32)
    dtr = float(np.float32(0.15))
    mixed = oracle_vadv(oracle, f, dtr, 3, ni, nj, nk, "mixed")
    wide = oracle_vadv(oracle, f, dtr, 3, ni, nj, nk, "wide")
    got = run_vadv(f, grid, 0.15, 1)[:nk, 3:3 + nj, 3:3 + ni]
    assert np.array_equal(got, mixed)
    err = rel_err(got, wide)
    print(f"[precision] vadv FP32-storage/FP64-compute max rel err vs FP64 oracle: {err:.3e}", flush=True)
    assert err <= MIXED_TOL


@pytest.mark.parametrize("ni,nj,nk", [(12, 33, 61), (256, 256, 80)])
def test_laplacian_parity(gpu, oracle, ni, nj, nk):
    from gridtools_amd import inputs, stencils

    print(f"[case] laplacian {ni}x{nj}x{nk}", flush=True)
    h = 1
    b = gpu.builder().type(np.float64).dimensions(ni + 2, nj + 2, nk).halos(h, h, 0)
    inp = b.uniform(SEED, inputs.FIELD_IDS["in"], -1.0, 1.0).build()
    out = b.value(0).build()
    grid = gpu.make_grid(gpu.halo_descriptor.symmetric(h, ni + 2), gpu.halo_descriptor.symmetric(h, nj + 2), nk)
    stencils.laplacian(out, inp, grid)
    a = inp.padded_numpy()
    res = np.zeros_like(a)
    oracle.laplacian(oracle.View(res, h, h), oracle.View(a, h, h), ni, nj, nk)
    assert np.array_equal(out.padded_numpy()[:, 1:1 + nj, 1:1 + ni], res[:, 1:1 + nj, 1:1 + ni])


def test_fill_uniform_matches_host_generator(gpu):
    from gridtools_amd import inputs

    for dtype in (np.float64, np.float32):
        ds = gpu.builder().type(dtype).dimensions(37, 21, 5).halos(2, 2, 0).uniform(7, 3, -0.5, 0.5).build()
        assert np.array_equal(ds.to_numpy(), inputs.uniform((37, 21, 5), 7, 3, -0.5, 0.5, dtype=dtype))


def test_repeated_launches_are_deterministic(gpu):
    from gridtools_amd import stencils

    inp, coeff, out, grid = hdiff_case(gpu, 100, 37, 11, np.float64)
    stencils.hdiff(out, inp, coeff, grid)
    first = out.padded_numpy()
    for _ in range(3):
        stencils.hdiff(out, inp, coeff, grid)
    assert np.array_equal(out.padded_numpy(), first)


def test_autotuner_picks_a_config_and_dispatch_stays_exact(gpu, oracle, tmp_path):
    from gridtools_amd import stencils, tune

    tune.clear()
    ni, nj, nk = 200, 120, 12
    inp, coeff, out, grid = hdiff_case(gpu, ni, nj, nk, np.float64)
    cfg, us = tune.hdiff(out, inp, coeff, grid, reps=3)
    print(f"[tune] {ni}x{nj}x{nk} fp64 -> {cfg} {us:.2f} us", flush=True)
    assert cfg["impl"] in ("lds-tile", "register-sweep", "register-sweep-pairs") and us > 0
    stencils.hdiff(out, inp, coeff, grid)  # now dispatched with the tuned configuration
    got = out.padded_numpy()[:nk, 2:2 + nj, 2:2 + ni]
    assert np.array_equal(got, oracle_hdiff(oracle, inp, coeff, 2, ni, nj, nk, "native"))
    path = tmp_path / "tune.txt"
    tune.save(path)
    assert f"hdiff 8 8 {ni} {nj} {nk}" in path.read_text()
    tune.clear()
