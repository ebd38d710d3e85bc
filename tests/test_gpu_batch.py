"""GPU parity of the batching form: uneven domains in one persistent launch (laplacian, hdiff and vadv per domain)
against the oracle chain on each domain's seeded inputs, bit for bit, and its timestep loop eager vs replayed graph.

The work table has more items than workgroups in the wide-domain cases (wg_per_cu = 1), so workgroups walk
several items and reuse their LDS tiles and scratch slot; k depths cover the first/last-level special cases
(nk = 2), partial ring blocks (nk not a multiple of the ring depth) and the deepest columns (137, 144).
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

H = 3
SIZES = [(23, 11, 43), (70, 9, 81), (130, 37, 2), (64, 4, 137), (65, 5, 96), (48, 48, 144), (1, 1, 3)]
WIDE = [(512, 256, 40), (97, 300, 61)]
# strip-form edge cases: ni of exactly one strip, one column past a strip, depths 78..81 around the register part
NARROW = [(64, 3, 80), (65, 2, 78), (128, 2, 81), (63, 4, 79), (2, 2, 137)]
# depths around a 32-level register part
NARROW32 = [(64, 3, 32), (65, 2, 31), (128, 2, 33), (63, 4, 36), (30, 5, 2)]
# depths around the vertical phase's 96-level register part (taken from 89 levels on by default)
DEEP96 = [(64, 3, 88), (65, 2, 89), (128, 2, 95), (63, 4, 97), (2, 2, 100), (70, 3, 144)]


def oracle_domain(oracle, host, ni, nj, nk, dtr, mode):
    """laplacian, hdiff and vadv of one domain (host arrays [k, j, i_padded]); returns interiors."""
    V = oracle.View
    lap = np.zeros_like(host["in_"])
    ust = np.zeros_like(host["in_"])
    oracle.laplacian(V(lap, H, H), V(host["in_"], H, H), ni, nj, nk)
    oracle.hdiff(V(ust, H, H), V(host["in_"], H, H), V(host["coeff"], H, H), ni, nj, nk, mode=mode, threads=8)
    uts = host["utens_stage"].copy()
    oracle.vadv(V(uts, H, H), V(ust, H, H), V(host["wcon"], H, H), V(host["u_pos"], H, H), V(host["utens"], H, H),
                dtr, ni, nj, nk, mode=mode, threads=8)
    box = np.s_[:nk, H:H + nj, H:H + ni]
    return lap[box], ust[box], uts[box]


def oracle_domain_rounded_ustage(oracle, host, ni, nj, nk, dtr):
    """The precision form of the two-phase batch: lap and u_stage are the FP64 chain on the FP32-rounded inputs,
    each rounded once; the vertical phase reads u_stage back rounded to FP32, so vadv runs in FP64 on that."""
    V = oracle.View
    lap = np.zeros_like(host["in_"])
    ust = np.zeros_like(host["in_"])
    oracle.laplacian(V(lap, H, H), V(host["in_"], H, H), ni, nj, nk)
    oracle.hdiff(V(ust, H, H), V(host["in_"], H, H), V(host["coeff"], H, H), ni, nj, nk, mode="native", threads=8)
    ust_r = ust.astype(np.float32).astype(np.float64)
    uts = host["utens_stage"].copy()
    oracle.vadv(V(uts, H, H), V(ust_r, H, H), V(host["wcon"], H, H), V(host["u_pos"], H, H), V(host["utens"], H, H),
                dtr, ni, nj, nk, mode="native", threads=8)
    box = np.s_[:nk, H:H + nj, H:H + ni]
    return tuple(r[box].astype(np.float32) for r in (lap, ust, uts))


def gpu_domain(m):
    g = m["grid"]
    box = np.s_[:g.k_size, H:H + g.j_size, H:H + g.i_size]
    return tuple(m[n].padded_numpy()[box] for n in ("lap", "u_stage", "utens_stage"))


def make_members(gt, sizes, dtype, seed=11):
    from gridtools_amd import inputs

    members = []
    for d, (ni, nj, nk) in enumerate(sizes):
        s = seed + d
        b = gt.builder().type(dtype).dimensions(ni + 2 * H, nj + 2 * H, nk).halos(H, H, 0)
        members.append({
            "in_": b.uniform(s, inputs.FIELD_IDS["in"], -1.0, 1.0).build(),
            "coeff": b.uniform(s, inputs.FIELD_IDS["coeff"], 0.0, 0.25).build(),
            "lap": b.value(0).build(),
            "u_stage": b.value(0).build(),
            "utens_stage": b.uniform(s, inputs.FIELD_IDS["utens_stage"], -1.0, 1.0).build(),
            "u_pos": b.uniform(s, inputs.FIELD_IDS["u_pos"], -1.0, 1.0).build(),
            "utens": b.uniform(s, inputs.FIELD_IDS["utens"], -1.0, 1.0).build(),
            "wcon": gt.builder().type(dtype).dimensions(ni + 2 * H, nj + 2 * H, nk + 1).halos(H, H, 0)
            .uniform(s, inputs.FIELD_IDS["wcon"], -0.5, 0.5).build(),
            "grid": gt.make_grid(gt.halo_descriptor.symmetric(H, ni + 2 * H),
                                 gt.halo_descriptor.symmetric(H, nj + 2 * H), nk),
        })
    return members


def check_batch(gt, oracle, sizes, dtype, precision, wg_per_cu, rounded_ustage=False):
    from gridtools_amd.batch import Batch

    print(f"[case] batch {np.dtype(dtype).name} precision={precision} wg_per_cu={wg_per_cu} {sizes}", flush=True)
    members = make_members(gt, sizes, dtype)
    hosts = [{n: m[n].padded_numpy() for n in ("in_", "coeff", "utens_stage", "u_pos", "utens", "wcon")}
             for m in members]
    b = Batch(members, precision=precision, wg_per_cu=wg_per_cu)
    nitems, grid = b.info()
    print(f"[case] items {nitems} workgroups {grid}", flush=True)
    b.run(0.15)
    gt.synchronize()
    dtr = 0.15 if dtype == np.float64 else float(np.float32(0.15))
    for (ni, nj, nk), m, host in zip(sizes, members, hosts):
        got = gpu_domain(m)
        if precision == 1 and rounded_ustage:
            ref = oracle_domain_rounded_ustage(oracle, {k: v.astype(np.float64) for k, v in host.items()}, ni, nj,
                                               nk, dtr)
        elif precision == 1:
            # FP64 chain on the FP32-rounded inputs, each output rounded once
            ref = oracle_domain(oracle, {k: v.astype(np.float64) for k, v in host.items()}, ni, nj, nk, dtr,
                                "native")
            ref = tuple(r.astype(np.float32) for r in ref)
        else:
            ref = oracle_domain(oracle, host, ni, nj, nk, dtr, "native")
        for name, g, r in zip(("lap", "u_stage", "utens_stage"), got, ref):
            assert np.array_equal(g, r), (name, (ni, nj, nk), rel_err(g, r))
    b.close()


@pytest.mark.parametrize("impl", [2, 1])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1), (np.float32, 0)])
def test_batch_uneven_domains_bit_exact(gpu, oracle, monkeypatch, dtype, precision, impl):
    """impl 2: 64-column strips with ccol/dcol in registers (tail scratch past 80 levels); impl 1: 64x4 tiles
    with ccol/dcol in scratch."""
    monkeypatch.setenv("GT355_TUNE_batch_impl", str(impl))
    check_batch(gpu, oracle, SIZES + NARROW, dtype, precision, wg_per_cu=2)


@pytest.mark.parametrize("nk_reg", [80, 32])
@pytest.mark.parametrize("hjb", [16, 8, 32])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1), (np.float32, 0)])
def test_batch_two_phase_bit_exact(gpu, oracle, monkeypatch, dtype, precision, hjb, nk_reg):
    """impl 3: the stencil phase (one-column register sweep over every domain's items) then the vertical phase
    (strip sweep reading u_stage back), with both register depths of the vertical phase and both item heights."""
    monkeypatch.setenv("GT355_TUNE_batch_impl", "3")
    monkeypatch.setenv("GT355_TUNE_batch_hjb", str(hjb))
    monkeypatch.setenv("GT355_TUNE_batch_nk", str(nk_reg))
    print(f"[case] batch two-phase hjb={hjb} nk_reg={nk_reg} {np.dtype(dtype).name} prec={precision}", flush=True)
    check_batch(gpu, oracle, SIZES + NARROW + NARROW32, dtype, precision, wg_per_cu=2, rounded_ustage=True)
    check_batch(gpu, oracle, WIDE + SIZES[:3], dtype, precision, wg_per_cu=1, rounded_ustage=True)


@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1), (np.float32, 0)])
def test_batch_vertical_phase_column_list_items_bit_exact(gpu, oracle, monkeypatch, dtype, precision):
    """GT355_TUNE_batch_flat=1: the vertical phase's items are 64 consecutive entries of a domain's column list
    (spanning row ends; the last item of a domain repeats its last column), every register depth."""
    monkeypatch.setenv("GT355_TUNE_batch_impl", "3")
    monkeypatch.setenv("GT355_TUNE_batch_flat", "1")
    print(f"[case] batch column-list items {np.dtype(dtype).name} prec={precision}", flush=True)
    check_batch(gpu, oracle, SIZES + NARROW + DEEP96, dtype, precision, wg_per_cu=2, rounded_ustage=True)
    check_batch(gpu, oracle, WIDE + SIZES[:3], dtype, precision, wg_per_cu=1, rounded_ustage=True)


@pytest.mark.parametrize("deep_nk", [89, 65, 145])
@pytest.mark.parametrize("dtype,precision", [(np.float32, 1), (np.float32, 0)])
def test_batch_vertical_96_register_levels_bit_exact(gpu, oracle, monkeypatch, dtype, precision, deep_nk):
    """The two-phase form's vertical phase with FP32 storage keeps 96 levels of ccol/dcol in registers for columns
    of at least deep_nk levels (89 by default; 65 sends every column past 64 levels there; 145 none): the depths
    cover 78..100 and 137..144, with and without tail levels."""
    monkeypatch.setenv("GT355_TUNE_batch_impl", "3")
    monkeypatch.setenv("GT355_TUNE_batch_deep_nk", str(deep_nk))
    print(f"[case] batch vertical deep_nk={deep_nk} {np.dtype(dtype).name} prec={precision}", flush=True)
    check_batch(gpu, oracle, SIZES + NARROW + DEEP96, dtype, precision, wg_per_cu=2, rounded_ustage=True)


@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1), (np.float32, 0)])
def test_batch_strip_32_register_levels_bit_exact(gpu, oracle, monkeypatch, dtype, precision):
    """The strip form with 32 levels of ccol/dcol in registers (two wavefronts per SIMD): every level past 32 goes
    through the tail scratch; the depths cover 2..144, around 32 and around 80."""
    monkeypatch.setenv("GT355_TUNE_batch_impl", "2")
    monkeypatch.setenv("GT355_TUNE_batch_nk", "32")
    print(f"[case] batch nk_reg=32 {np.dtype(dtype).name} prec={precision}", flush=True)
    check_batch(gpu, oracle, SIZES + NARROW + NARROW32, dtype, precision, wg_per_cu=2)
    check_batch(gpu, oracle, WIDE + SIZES[:3], dtype, precision, wg_per_cu=1)


@pytest.mark.parametrize("impl", [2, 1])
@pytest.mark.parametrize("dtype,precision", [(np.float64, 0), (np.float32, 1)])
def test_batch_more_items_than_workgroups(gpu, oracle, monkeypatch, dtype, precision, impl):
    monkeypatch.setenv("GT355_TUNE_batch_impl", str(impl))
    check_batch(gpu, oracle, WIDE + SIZES[:3], dtype, precision, wg_per_cu=1)


@pytest.mark.parametrize("impl", [3, 2])
def test_batch_matches_single_domain_entry_points(gpu, monkeypatch, impl):
    """On the GPU the batch equals gt355_laplacian + gt355_hdiff_vadv run per domain: the two-phase form (impl 3)
    the two-launch composition (u_stage stored in FP32 between the launches in the precision form), the
    single-launch strip form (impl 2) the fused composition (nk <= 80), which keeps u_stage in FP64."""
    monkeypatch.setenv("GT355_TUNE_batch_impl", str(impl))
    if impl == 2:
        monkeypatch.setenv("GT355_TUNE_hv_fused", "1")
    from gridtools_amd import stencils
    from gridtools_amd.batch import Batch

    sizes = [(100, 60, 80), (256, 256, 80), (77, 130, 33)]
    a = make_members(gpu, sizes, np.float32)
    b = make_members(gpu, sizes, np.float32)
    Batch(a, precision=1).run(0.15)
    for m in b:
        g = m["grid"]
        stencils.laplacian(m["lap"], m["in_"], g, precision=1)
        stencils.hdiff_vadv(m["in_"], m["coeff"], m["u_stage"], m["utens_stage"], m["wcon"], m["u_pos"], m["utens"],
                            0.15, g, precision=1)
    gpu.synchronize()
    for ma, mb in zip(a, b):
        for x, y in zip(gpu_domain(ma), gpu_domain(mb)):
            assert np.array_equal(x, y)


def test_batch_timesteps_graph_replays_match_eager(gpu, oracle):
    """BatchedTimesteps: the first eager step equals the oracle chain; a captured 2-step graph replayed three
    times matches an identical twin stepped eagerly, compared after every replay."""
    from gridtools_amd.batch import BatchedTimesteps
    from gridtools_amd.timing import Stream

    sizes = [(40, 30, 50), (90, 17, 96), (64, 64, 41)]
    eager = BatchedTimesteps(sizes, seed=5, halo=H)
    # step 1 against the oracle: the FP64 chain on the FP32 inputs with every output rounded once, the vertical
    # phase reading u_stage back in FP32 (the default two-phase form)
    host = []
    for f in eager.domains:
        h = {n: f[n].padded_numpy().astype(np.float64) for n in ("coeff", "utens_stage", "u_pos", "utens", "wcon")}
        h["in_"] = f["u"].padded_numpy().astype(np.float64)
        host.append(h)
    eager.step()
    gpu.synchronize()
    for d, ((ni, nj, nk), h) in enumerate(zip(sizes, host)):
        _, ref_u, ref_t = oracle_domain_rounded_ustage(oracle, h, ni, nj, nk, float(np.float32(0.15)))
        box = np.s_[:nk, H:H + nj, H:H + ni]
        assert np.array_equal(eager.current(d).padded_numpy()[box], ref_u)
        assert np.array_equal(eager.domains[d]["utens_stage"].padded_numpy()[box], ref_t)
    twin = BatchedTimesteps(sizes, seed=5, halo=H)
    graphed = BatchedTimesteps(sizes, seed=5, halo=H)
    st = Stream()
    g = graphed.capture(2, st)
    for r in range(3):
        g.launch(st)
        twin.step()
        twin.step()
        st.synchronize()
        gpu.synchronize()
        for d in range(len(sizes)):
            assert np.array_equal(graphed.current(d).padded_numpy(), twin.current(d).padded_numpy()), (r, d)
            assert np.array_equal(graphed.domains[d]["utens_stage"].padded_numpy(),
                                  twin.domains[d]["utens_stage"].padded_numpy()), (r, d)


def test_batch_tune_picks_a_grid_and_keeps_results(gpu):
    """The grid-size tuner returns a valid choice, and the tuned batch computes the same bits (every item is
    independent of which workgroup runs it)."""
    from gridtools_amd.batch import Batch

    sizes = [(300, 200, 60), (64, 300, 120), (100, 100, 90)]
    a = make_members(gpu, sizes, np.float32)
    b = make_members(gpu, sizes, np.float32)
    tuner = Batch(make_members(gpu, sizes, np.float32), precision=1, wg_per_cu=1)
    w, us = tuner.tune(reps=3)
    print(f"[tune] batch wg_per_cu {w} ({us:.1f} us)", flush=True)
    assert 1 <= w <= 4 and us > 0
    Batch(a, precision=1, wg_per_cu=1).run(0.15)
    Batch(b, precision=1, wg_per_cu=w).run(0.15)
    gpu.synchronize()
    for ma, mb in zip(a, b):
        for x, y in zip(gpu_domain(ma), gpu_domain(mb)):
            assert np.array_equal(x, y)


@pytest.mark.timeout(600)
def test_batch_c4_workload_step_and_graph(gpu, oracle):
    """The batching-form workload as stated: 64 domains with ni, nj ~ U{48..384} and nk ~ U{40..137} drawn from seed
    2024, FP32 storage / FP64 compute, the default two-phase form.  One eager timestep (on the single-stencil
    distributions, wcon ~ U(-0.5, 0.5)) equals the oracle chain on every domain bit for bit (lap, u_stage and
    utens_stage); then a captured 10-timestep graph equals 10 eager timesteps of an identical twin on every domain,
    bit for bit, with the timestep loop's bounded wind (the state stays finite)."""
    from gridtools_amd.batch import BatchedTimesteps, draw_sizes
    from gridtools_amd.timing import Stream

    sizes = draw_sizes(2024)
    assert len(sizes) == 64
    print(f"[case] batch C4: 64 domains, {sum(a * b * c for a, b, c in sizes) / 1e6:.1f} M points", flush=True)
    eager = BatchedTimesteps(sizes, seed=2024, halo=H, wcon_max=0.5)
    host = []
    for f in eager.domains:
        h = {n: f[n].padded_numpy() for n in ("coeff", "utens_stage", "u_pos", "utens", "wcon")}
        h["in_"] = f["u"].padded_numpy()
        host.append(h)
    eager.step()
    gpu.synchronize()
    dtr = float(np.float32(0.15))
    for d, ((ni, nj, nk), h) in enumerate(zip(sizes, host)):
        h64 = {k: v.astype(np.float64) for k, v in h.items()}
        ref_l, ref_u, ref_t = oracle_domain_rounded_ustage(oracle, h64, ni, nj, nk, dtr)
        box = np.s_[:nk, H:H + nj, H:H + ni]
        f = eager.domains[d]
        assert np.array_equal(f["lap"].padded_numpy()[box], ref_l), ("lap", d)
        assert np.array_equal(eager.current(d).padded_numpy()[box], ref_u), ("u_stage", d)
        assert np.array_equal(f["utens_stage"].padded_numpy()[box], ref_t), ("utens_stage", d)
        host[d] = None
    del eager
    print("[case] batch C4: one step equals the oracle; 10-step graph against 10 eager steps", flush=True)
    twin = BatchedTimesteps(sizes, seed=2024, halo=H)
    graphed = BatchedTimesteps(sizes, seed=2024, halo=H)
    st = Stream()
    g = graphed.capture(10, st)
    g.launch(st)
    for _ in range(10):
        twin.step()
    st.synchronize()
    gpu.synchronize()
    for d in range(len(sizes)):
        for a, b in ((graphed.current(d), twin.current(d)),
                     (graphed.domains[d]["utens_stage"], twin.domains[d]["utens_stage"])):
            x, y = a.padded_numpy(), b.padded_numpy()
            assert np.all(np.isfinite(x)), d
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), d


def test_batch_tuner_configuration_persists(gpu, tmp_path):
    """The batch tuner's choice (grid, the vertical phase's 96-level depth, the stencil phase's rows per item and
    workgroups per CU) is valid, recorded under the domain list's signature (a new batch of the same sizes starts
    from it, also after gt355_tune_save / clear / load), and every configuration computes the same bits."""
    from gridtools_amd import _abi
    from gridtools_amd.batch import Batch

    sizes = [(300, 200, 60), (64, 300, 120), (100, 100, 90), (70, 50, 137)]
    _abi.call("gt355_tune_clear")
    tuner = Batch(make_members(gpu, sizes, np.float32), precision=1, wg_per_cu=1)
    default = tuner.config()
    assert default["form"] == 3 and default["hjb"] == 16 and default["hw"] == 8
    w, us = tuner.tune(reps=2)
    cfg = tuner.config()
    print(f"[tune] batch {cfg} ({us:.1f} us)", flush=True)
    assert cfg["w"] == w and 1 <= w <= 4 and us > 0
    assert cfg["hjb"] in (8, 16, 32) and cfg["hw"] in (2, 4, 8) and cfg["deep_nk"] in (81, 89, 97, 1 << 20)
    again = Batch(make_members(gpu, sizes, np.float32), precision=1, wg_per_cu=1)
    assert again.config() == cfg
    other = Batch(make_members(gpu, sizes[:3], np.float32), precision=1, wg_per_cu=1)
    assert other.config() == default  # another domain list: no entry
    path = str(tmp_path / "tune.txt")
    _abi.call("gt355_tune_save", path.encode())
    assert any(ln.startswith("batch ") for ln in open(path))
    _abi.call("gt355_tune_clear")
    assert Batch(make_members(gpu, sizes, np.float32), precision=1, wg_per_cu=1).config() == default
    n = ctypes.c_int()
    _abi.call("gt355_tune_load", path.encode(), ctypes.byref(n))
    assert Batch(make_members(gpu, sizes, np.float32), precision=1, wg_per_cu=1).config() == cfg
    _abi.call("gt355_tune_clear")
    # the same bits in every configuration
    ref = make_members(gpu, sizes, np.float32)
    Batch(ref, precision=1, wg_per_cu=1).run(0.15)
    gpu.synchronize()
    want = [gpu_domain(m) for m in ref]
    for c in ({"hjb": 8, "hw": 2, "w": 4, "deep_nk": 81}, {"hjb": 32, "hw": 4, "w": 2, "deep_nk": 1 << 20},
              {"hjb": 16, "hw": 8, "w": 3, "deep_nk": 97}):
        m = make_members(gpu, sizes, np.float32)
        b = Batch(m, precision=1, wg_per_cu=1)
        b.set_config(**c)
        assert b.config() == dict(c, form=3)
        b.run(0.15)
        gpu.synchronize()
        for x, y in zip([gpu_domain(mm) for mm in m], want):
            for p, q in zip(x, y):
                assert np.array_equal(p, q), c
    with pytest.raises(RuntimeError, match="rows per item"):
        tuner.set_config(hjb=12)
