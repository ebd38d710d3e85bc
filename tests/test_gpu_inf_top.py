"""Regression: infinite (and NaN) u_stage at the top level.

The hand-written vertical sweeps evaluate the top level with the interior formula and zero k+1 winds.  The k+1
correction term is selected away at the top (not computed from a substituted u_stage), so the top level has the bits
of the reference's apply(last_level) (vertical_advection_dycore.cpp:69-81) even when u_stage there is infinite,
where 0 * (inf - inf) would be a NaN.  Expected values: the fixture tests/golden/vadv_inf_top.npz (an independent
pure-Python transcription of the reference functors) and, at larger sizes, the oracle, which
tests/test_oracle_vadv.py pins to that fixture.  NaN and infinity positions must agree exactly."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 3


def _store(gt, arr_kji, dtype):
    """A storage of halo H whose [k, j, i] contents (storage index space) are arr_kji."""
    nk, nj, ni = arr_kji.shape
    ds = gt.builder().type(dtype).dimensions(ni, nj, nk).halos(H, H, 0).build()
    ds.from_numpy(np.ascontiguousarray(arr_kji.transpose(2, 1, 0)))
    return ds


def _inject(rng, us, ni, nj, nk, count):
    top = nk - 1
    for n in range(count):
        j, i = H + rng.integers(nj), H + rng.integers(ni)
        us[top, j, i] = (np.inf, -np.inf, np.nan)[n % 3]
    us[top - 1, H + rng.integers(nj), H + rng.integers(ni)] = np.inf


def _oracle(oracle, fl, dtr, ni, nj, nk, mode):
    pad = [np.ascontiguousarray(a) for a in fl]
    res = pad[0].copy()
    V = oracle.View
    with np.errstate(invalid="ignore", over="ignore"):
        oracle.vadv(V(res, H, H), V(pad[1], H, H), V(pad[2], H, H), V(pad[3], H, H), V(pad[4], H, H), dtr, ni, nj,
                    nk, mode=mode, threads=8)
    return res[:, H:H + nj, H:H + ni]


def _run(gt, fl, dtr, ni, nj, nk, dtype, precision):
    from gridtools_amd import stencils

    f = [_store(gt, a, dtype) for a in fl]
    grid = gt.make_grid(gt.halo_descriptor.symmetric(H, ni + 2 * H), gt.halo_descriptor.symmetric(H, nj + 2 * H), nk)
    stencils.vadv(f[0], f[1], f[2], f[3], f[4], dtr, grid, precision=precision)
    return f[0].to_numpy().transpose(2, 1, 0)[:, H:H + nj, H:H + ni]


def test_vadv_inf_top_fixture(gpu):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "vadv_inf_top.npz"))
    ni, nj, nk, h = (int(x) for x in z["dims"])
    assert h == H
    fl = [z[n] for n in ("utens_stage", "u_stage", "wcon", "u_pos", "utens")]
    print(f"[case] vadv inf-top fixture {ni}x{nj}x{nk}", flush=True)
    got = _run(gpu, fl, float(z["dtr"]), ni, nj, nk, np.float64, 0)
    np.testing.assert_array_equal(got, z["expected"][:, H:H + nj, H:H + ni])


# (ni, nj, nk, env): onchip buckets, the persistent strips (>= 2 rounds of strips per SIMD), the deep scratch path
# and, for FP32 storage with ni % 4 == 0 and one round of strips, the 16-byte-request kernels (LDS slots and the
# register ring; 128 columns: both strips whole, every row read inside the halo)
CASES = [(67, 5, 17, {}), (67, 5, 80, {}), (130, 700, 67, {}), (70, 9, 81, {}),
         (256, 512, 80, {"GT355_TUNE_vadv_persist_d": "16"}), (128, 40, 77, {}),
         (128, 40, 77, {"GT355_TUNE_vadv_wide_lds": "0"})]


@pytest.mark.parametrize("ni,nj,nk,env", CASES)
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1)])
def test_vadv_inf_top_all_paths(gpu, oracle, monkeypatch, ni, nj, nk, env, dtype, precision):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    print(f"[case] vadv inf-top {np.dtype(dtype).name} prec={precision} {ni}x{nj}x{nk} {env}", flush=True)
    rng = np.random.default_rng(ni * 7 + nk)
    shp = (nk, nj + 2 * H, ni + 2 * H)
    fl = [rng.uniform(-1, 1, shp) for _ in range(2)] + [rng.uniform(-0.5, 0.5, (nk + 1,) + shp[1:])] + \
         [rng.uniform(-1, 1, shp) for _ in range(2)]
    fl = [a.astype(dtype) for a in fl]
    _inject(rng, fl[1], ni, nj, nk, 12)
    fl[2][nk, H:H + nj, H:H + ni + 1] = np.inf  # never read
    dtr = 0.15 if dtype == np.float64 else float(np.float32(0.15))
    ref = _oracle(oracle, [a.astype(np.float64) for a in fl] if precision else fl, dtr, ni, nj, nk, "native")
    if precision:
        ref = ref.astype(np.float32)
    got = _run(gpu, fl, 0.15, ni, nj, nk, dtype, precision)
    assert np.isnan(ref).any() and np.isinf(ref).any()
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("fused", [0, 1])
def test_compose_inf_top(gpu, oracle, monkeypatch, fused):
    """The composition's vertical sweep (two launches, and the fused strip sweep) with infinite `in` at the top
    level, so that hdiff hands the sweep infinite / NaN u_stage there."""
    from gridtools_amd import stencils

    monkeypatch.setenv("GT355_TUNE_hv_fused", str(fused))
    ni, nj, nk = 70, 9, 33
    print(f"[case] compose inf-top fused={fused}", flush=True)
    rng = np.random.default_rng(5)
    shp = (nk, nj + 2 * H, ni + 2 * H)
    inp = rng.uniform(-1, 1, shp)
    inp[nk - 1, H + 2, H + 10] = np.inf
    inp[nk - 1, H + 6, H + 40] = -np.inf
    coeff = rng.uniform(0, 0.25, shp)
    uts, up, ut = (rng.uniform(-1, 1, shp) for _ in range(3))
    wc = rng.uniform(-0.5, 0.5, (nk + 1,) + shp[1:])
    V = oracle.View
    ust = np.zeros(shp)
    with np.errstate(invalid="ignore", over="ignore"):
        oracle.hdiff(V(ust, H, H), V(np.ascontiguousarray(inp), H, H), V(coeff, H, H), ni, nj, nk, threads=8)
    ref = _oracle(oracle, [uts, ust, wc, up, ut], 0.15, ni, nj, nk, "native")
    f = {n: _store(gpu, a, np.float64) for n, a in
         dict(in_=inp, coeff=coeff, u_stage=np.zeros(shp), utens_stage=uts, wcon=wc, u_pos=up, utens=ut).items()}
    grid = gpu.make_grid(gpu.halo_descriptor.symmetric(H, ni + 2 * H), gpu.halo_descriptor.symmetric(H, nj + 2 * H),
                         nk)
    stencils.hdiff_vadv(f["in_"], f["coeff"], f["u_stage"], f["utens_stage"], f["wcon"], f["u_pos"], f["utens"],
                        0.15, grid)
    got = f["utens_stage"].to_numpy().transpose(2, 1, 0)[:, H:H + nj, H:H + ni]
    assert np.isnan(ref).any() or np.isinf(ref).any()
    np.testing.assert_array_equal(got, ref)
