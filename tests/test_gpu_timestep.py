"""Timestep sequences (hdiff -> vadv -> periodic-i / zero-gradient-j halo fill -> swap) on the GPU.

* several steps, bit-identical to the same sequence replayed with the oracle (FP64 and the precision form; with
  the two-launch composition, which stores u_stage in FP32 in the precision form, and with the fused one, which
  keeps it in FP64 between the stages);
* a hipGraph of an even number of steps, replayed several times with eager work in between, equal after every
  replay to eager stepping of an identical second instance;
* the 1536x1536x80 FP32 domain: its first two steps bit-identical to the oracle replay, and after 20 steps every
  value finite and the halo satisfying the boundary conditions exactly.
The timestepper's winds keep the implicit solve diagonally dominant, so the tendency that feeds back into each step
stays bounded; every compared field is asserted finite."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_steps(oracle, ts, nsteps, fused):
    """Replays nsteps of the timestepper's operator on host copies of its current fields.  In the precision form
    the two-launch composition stores u_stage in FP32 before the vertical sweep reads it; the fused one does not."""
    H = ts.halo
    ni, nj, nk = ts.ni, ts.nj, ts.nk
    f = {n: getattr(ts, n).padded_numpy() for n in ("u", "u_stage", "coeff", "utens_stage", "u_pos", "utens", "wcon")}
    wide = ts.precision == 1
    dtr = float(np.float32(ts.dtr)) if f["u"].dtype == np.float32 else ts.dtr
    hal = [h.as_tuple() for h in ts.halos]
    V = oracle.View
    for _ in range(nsteps):
        cv = (lambda a: a.astype(np.float64)) if wide else (lambda a: a)
        ust = np.zeros_like(cv(f["u"]))
        oracle.hdiff(V(ust, H, H), V(cv(f["u"]), H, H), V(cv(f["coeff"]), H, H), ni, nj, nk, threads=8)
        uts = cv(f["utens_stage"]).copy()
        ust_v = ust.astype(np.float32).astype(np.float64) if wide and not fused else ust
        oracle.vadv(V(uts, H, H), V(ust_v, H, H), V(cv(f["wcon"]), H, H), V(cv(f["u_pos"]), H, H),
                    V(cv(f["utens"]), H, H), dtr, ni, nj, nk, threads=8)
        # write back the interiors, keep the halos, then apply the boundary conditions
        new_ust = f["u_stage"].copy()
        new_ust[:nk, H:H + nj, H:H + ni] = ust[:nk, H:H + nj, H:H + ni]
        f["utens_stage"][:nk, H:H + nj, H:H + ni] = uts[:nk, H:H + nj, H:H + ni]
        for a in (new_ust, f["utens_stage"]):
            oracle.halo_fill(a, hal, (2, 1, 0))
        f["u"], f["u_stage"] = new_ust, f["u"]
    return f


@pytest.mark.parametrize("bc", ["fill", "exchange"])
@pytest.mark.parametrize("fused", [0, 1, 2])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1)])
def test_steps_match_oracle(gpu, oracle, monkeypatch, dtype, precision, fused, bc):
    """bc="exchange" applies the boundary conditions through the halo-pack kernels (pack / single-rank periodic
    exchange / unpack) and must give the bits of the oracle's periodic-i / zero-gradient-j halo fill."""
    from gridtools_amd.timestep import HdiffVadvTimestepper

    monkeypatch.setenv("GT355_TUNE_hv_fused", str(fused))
    ts = HdiffVadvTimestepper(70, 45, 24, dtype=dtype, precision=precision, seed=3, bc=bc)
    ref = oracle_steps(oracle, ts, 4, fused)
    for _ in range(4):
        ts.step()
    for name in ("u", "utens_stage"):
        got = getattr(ts, name).to_numpy()
        exp = ref[name][:, :, :got.shape[0]].transpose(2, 1, 0)[:, :, :]
        exp = exp[:got.shape[0], :got.shape[1], :got.shape[2]]
        assert np.isfinite(got).all(), name
        assert np.array_equal(got, exp), name


@pytest.mark.parametrize("ni,nj,nk", [(130, 67, 40), (128, 67, 80)])
def test_graph_replays_match_eager(gpu, ni, nj, nk):
    """Four captured timesteps (FP32 storage, the precision form) replayed three times equal eager steps bit for bit:
    at 130x67x40 the vadv runs in a register bucket of the one-row kernel; at 128x67x80 it is the 16-byte-request
    kernel with its rows loaded straight into LDS."""
    from gridtools_amd.timestep import HdiffVadvTimestepper
    from gridtools_amd.timing import Stream, synchronize

    print(f"[case] graph replays {ni}x{nj}x{nk}", flush=True)
    st = Stream()
    a = HdiffVadvTimestepper(ni, nj, nk, seed=11)
    b = HdiffVadvTimestepper(ni, nj, nk, seed=11)
    g = a.capture(4, st)
    for rep in range(3):
        g.launch(st)
        st.synchronize()
        for _ in range(4):
            b.step()
        synchronize()
        for name in ("u", "u_stage", "utens_stage"):
            assert np.array_equal(getattr(a, name).to_numpy(), getattr(b, name).to_numpy(), equal_nan=True), \
                (rep, name)


def test_large_domain_first_steps_match_oracle(gpu, oracle):
    """The full-size large-domain workload (1536x1536x80, FP32 storage, FP64 compute, the default two-launch
    composition and the halo fill): its first two steps equal the oracle replay bit for bit."""
    from gridtools_amd.timestep import HdiffVadvTimestepper

    print("[case] large domain 1536x1536x80: 2 steps vs the oracle", flush=True)
    ts = HdiffVadvTimestepper(1536, 1536, 80, seed=5)
    ref = oracle_steps(oracle, ts, 2, 0)
    for _ in range(2):
        ts.step()
    for name in ("u", "utens_stage"):
        got = getattr(ts, name).padded_numpy()
        assert np.isfinite(got).all(), name
        assert np.array_equal(got, ref[name].astype(got.dtype)), name


def test_large_domain_boundaries_hold(gpu):
    from gridtools_amd.timestep import HdiffVadvTimestepper

    ts = HdiffVadvTimestepper(1536, 1536, 80, seed=5)
    for _ in range(20):
        ts.step()
    H, ni, nj = 3, 1536, 1536
    for name in ("u", "utens_stage"):
        a = getattr(ts, name).to_numpy()  # [i, j, k] with halo
        assert np.isfinite(a).all(), name
        # periodic in i for the interior rows, zero-gradient in j for every column (corners included)
        assert np.array_equal(a[:H, H:H + nj], a[ni:ni + H, H:H + nj], equal_nan=True)
        assert np.array_equal(a[H + ni:, H:H + nj], a[H:2 * H, H:H + nj], equal_nan=True)
        for jj in range(H):
            assert np.array_equal(a[:, jj], a[:, H], equal_nan=True)
            assert np.array_equal(a[:, H + nj + jj], a[:, H + nj - 1], equal_nan=True)
