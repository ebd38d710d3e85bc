#!/usr/bin/env python3
"""Benchmark: horizontal diffusion + vertical advection, FP64, 256x256x80 on one MI355X.

One step = one hdiff launch and one vadv launch over the 256x256x80 interior (hdiff halo 2, vadv halo 3),
the plain FP64 drop-in path.  Inputs are synthetic (seeded SplitMix64, the distributions of the workload
definition) and resident in HBM before timing.  Steps cycle over NSETS independent field sets so that no
step finds its inputs in the 256 MB Infinity Cache left there by the previous one.

Printed JSON (one line): value = interior points / step time (one combined hdiff+vadv update per point), the
per-kernel figures for FP64 and the FP32 precision form (FP32 storage, FP64 compute), the roofline of the
dominant kernel (vadv FP64, algorithmic bytes / mean launch time from hipEvents on the launch stream) and the
CPU baseline (the oracle port, OpenMP, on this host) on a bounded sample of the same workload.
"""
import argparse
import contextlib
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NI, NJ, NK = 256, 256, 80
P = NI * NJ * NK
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
# algorithmic (compulsory) bytes per interior point: hdiff reads in, coeff and writes out; vadv reads
# utens_stage, u_stage, wcon, u_pos, utens and writes utens_stage
BYTES_PER_PT = {"hdiff": 3, "vadv": 6}
SEED = 42
# launches per per-kernel timing window (hdiff FP64 256x256x80: 250 x 22 us = 5.5 ms)
KERNEL_WINDOW = 250
# steps of the long step window: the comparison of the two forms uses it whatever --steps is (a window of 200
# steps, 8 ms of FP32-form steps, measured 39-42 us per step where 2000 steps give 36.8 us, the sum of its kernels)
STEP_WINDOW = 2000


def make_set(gt, dtype, s):
    from gridtools_amd import inputs

    seed = SEED + s
    hb = gt.builder().type(dtype).dimensions(NI + 4, NJ + 4, NK).halos(2, 2, 0)
    hd = {
        "in": hb.uniform(seed, inputs.FIELD_IDS["in"], -1.0, 1.0).build(),
        "coeff": hb.uniform(seed, inputs.FIELD_IDS["coeff"], 0.0, 0.25).build(),
        "out": hb.value(0).build(),
    }
    vb = gt.builder().type(dtype).dimensions(NI + 6, NJ + 6, NK).halos(3, 3, 0)
    va = {n: vb.uniform(seed, inputs.FIELD_IDS[n], -1.0, 1.0).build()
          for n in ("utens_stage", "u_stage", "u_pos", "utens")}
    va["wcon"] = gt.builder().type(dtype).dimensions(NI + 6, NJ + 6, NK + 1).halos(3, 3, 0) \
        .uniform(seed, inputs.FIELD_IDS["wcon"], -0.5, 0.5).build()
    return hd, va


def grids(gt):
    gh = gt.make_grid(gt.halo_descriptor.symmetric(2, NI + 4), gt.halo_descriptor.symmetric(2, NJ + 4), NK)
    gv = gt.make_grid(gt.halo_descriptor.symmetric(3, NI + 6), gt.halo_descriptor.symmetric(3, NJ + 6), NK)
    return gh, gv


_LAUNCH_STREAM = []


def launch_stream():
    """The stream every timed leg launches on and records its events on: one created (non-blocking) HIP stream of
    the default priority.  Launches on the legacy null stream carry its implicit synchronisation and measured
    slower (FP64 step 65.2 vs 64.3 us, FP32-form step 40.5 vs 38.7 us)."""
    from gridtools_amd.timing import Stream

    if not _LAUNCH_STREAM:
        _LAUNCH_STREAM.append(Stream())
    return _LAUNCH_STREAM[0]


def timed_steps(gt, sets, precision, steps, warmup):
    """Runs warmup + steps steps; returns (total ms of the timed steps, mean ms per hdiff launch, mean ms per vadv
    launch, mean ms per step over a window of at least STEP_WINDOW steps).  The step loop is bracketed by one event
    pair (synchronised on both sides).  Per-kernel times come
    from separate runs of `steps` back-to-back launches of one kernel, also bracketed by one event pair on the
    launch stream, so that they measure launch-to-launch kernel time without the dispatch gap an event pair
    around every single launch would add."""
    from gridtools_amd import stencils
    from gridtools_amd.timing import Event, synchronize

    gh, gv = grids(gt)
    st = launch_stream()

    def hdiff(step):
        hd = sets[step % len(sets)][0]
        stencils.hdiff(hd["out"], hd["in"], hd["coeff"], gh, precision=precision, stream=st.handle)

    def vadv(step):
        va = sets[step % len(sets)][1]
        stencils.vadv(va["utens_stage"], va["u_stage"], va["wcon"], va["u_pos"], va["utens"], 0.15, gv,
                      precision=precision, stream=st.handle)

    def timed(fn_list, n, offset):
        start, stop = Event(), Event()
        synchronize()
        st.synchronize()
        start.record(st)
        for s in range(n):
            for fn in fn_list:
                fn(offset + s)
        stop.record(st)
        stop.synchronize()
        st.synchronize()
        return start.elapsed_ms(stop)

    gt_sync()  # fields were filled on the null stream

    for s in range(warmup):
        hdiff(s)
        vadv(s)
    total = timed([hdiff, vadv], steps, warmup)
    # per-kernel windows of at least KERNEL_WINDOW launches (several ms), whatever --steps is, so that the start-up
    # of a window does not weigh on the mean
    n = max(steps, KERNEL_WINDOW)
    hd_ms = timed([hdiff], n, warmup + steps) / n
    va_ms = timed([vadv], n, warmup + steps + n) / n
    nw = max(steps, STEP_WINDOW)
    win_ms = timed([hdiff, vadv], nw, warmup + steps + 2 * n) / nw
    return total, hd_ms, va_ms, win_ms


def graphed_steps(gt, sets, precision, nsteps=40, reps=10):
    """Milliseconds per step of a graph of nsteps steps (cycling over the field sets, the two launches of a step one
    after the other on one stream), replayed reps times after one warm-up replay: the step without host launch
    gaps."""
    from gridtools_amd import stencils
    from gridtools_amd.timestep import capture
    from gridtools_amd.timing import Event, Stream

    gh, gv = grids(gt)

    def steps(st):
        for s in range(nsteps):
            hd, va = sets[s % len(sets)]
            stencils.hdiff(hd["out"], hd["in"], hd["coeff"], gh, precision=precision, stream=st.handle)
            stencils.vadv(va["utens_stage"], va["u_stage"], va["wcon"], va["u_pos"], va["utens"], 0.15, gv,
                          precision=precision, stream=st.handle)

    st = Stream()
    gt_sync()
    steps(st)
    g = capture(st, steps)
    g.launch(st)
    st.synchronize()
    start, stop = Event(), Event()
    start.record(st)
    for _ in range(reps):
        g.launch(st)
    stop.record(st)
    stop.synchronize()
    return start.elapsed_ms(stop) / (reps * nsteps)


def autotune(gt, dtype, precision, shapes):
    """Runs the project's autotuner for hdiff on each (n, nk) shape (outside every timed region) on scratch fields
    of that shape; later launches of that shape use the recorded winner.  Returns {shape: (config, us)}."""
    from gridtools_amd import tune

    out = {}
    for n, nk in shapes:
        hb = gt.builder().type(dtype).dimensions(n + 4, n + 4, nk).halos(2, 2, 0)
        i_ = hb.uniform(SEED, 1, -1.0, 1.0).build()
        c_ = hb.uniform(SEED, 2, 0.0, 0.25).build()
        o_ = hb.value(0).build()
        g = gt.make_grid(gt.halo_descriptor.symmetric(2, n + 4), gt.halo_descriptor.symmetric(2, n + 4), nk)
        cfg, us = tune.hdiff(o_, i_, c_, g, precision=precision, reps=60)
        out[f"{n}x{n}x{nk}"] = {"config": cfg, "us": round(us, 2)}
    return out


def kernel_stats(name, ms, bytes_per_elem):
    algo = BYTES_PER_PT[name] * bytes_per_elem * P
    return {
        "mean_us": round(ms * 1e3, 2),
        "gpts": round(P / (ms * 1e-3) / 1e9, 2),
        "algo_bytes": algo,
        "achieved_gbs": round(algo / (ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
    }


def measured_traffic(kernel_prefix):
    """HBM bytes per launch of a kernel as measured by the calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE passes
    committed under profiles/ (line 'TRAFFIC <kernel> <workload> = <bytes>'); None if no such line exists."""
    import glob
    import re

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            m = re.match(r"TRAFFIC (\S+) (\S+) = (\d+)", line.strip())
            if m and m.group(1).startswith(kernel_prefix) and m.group(2) == f"{NI}x{NJ}x{NK}":
                return {"bytes": int(m.group(3)), "source": os.path.relpath(path, ROOT), "kernel": m.group(1)}
    return None


def copy_leg(gt, steps, warmup, nsets):
    """copy_stencil FP64 over the same 256x256x80 interior (16 B/point): the bandwidth a pure stream reaches."""
    from gridtools_amd import stencils
    from gridtools_amd.timing import Event, synchronize

    b = gt.builder().type(np.float64).dimensions(NI, NJ, NK).halos(0, 0, 0)
    pairs = [(b.uniform(SEED + s, 1, -1.0, 1.0).build(), b.value(0).build()) for s in range(max(nsets, 4) * 2)]
    g = gt.make_grid(NI, NJ, NK)
    st = launch_stream()
    gt_sync()  # fields were filled on the null stream
    for s in range(warmup):
        src, dst = pairs[s % len(pairs)]
        stencils.copy(dst, src, g, stream=st.handle)
    synchronize()
    st.synchronize()
    start, stop = Event(), Event()
    start.record(st)
    for s in range(steps):
        src, dst = pairs[(warmup + s) % len(pairs)]
        stencils.copy(dst, src, g, stream=st.handle)
    stop.record(st)
    stop.synchronize()
    ms = start.elapsed_ms(stop) / steps
    gbs = 16 * P / (ms * 1e-3) / 1e9
    return {"mean_us": round(ms * 1e3, 2), "gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def cartesian_leg(gt, steps, warmup, nsets):
    """The further cartesian regression stencils at 256x256x80 FP64, each timed over cycling field sets (MALL-cold):
    simple_hori_diff (in, coeff read, out written: 24 B/point; the j-only metric is negligible), the expandable_run of
    prepare_tracers over 11 members (rho once + 11 reads + 11 writes: 184 B/point), the five-field copy of
    copy_stencil_tuple (80 B/point), the extended_4D projection (144 B/point) and transform_layout of a storage
    interior into a k-fastest dense buffer (16 B/point)."""
    from gridtools_amd import stencils

    h = 2
    b = gt.builder().type(np.float64).dimensions(NI + 2 * h, NJ + 2 * h, NK).halos(h, h, 0)
    g = gt.make_grid(gt.halo_descriptor.symmetric(h, NI + 2 * h), gt.halo_descriptor.symmetric(h, NJ + 2 * h), NK)
    n = max(nsets, 4)
    shd = [(b.uniform(SEED + s, 21, -1.0, 1.0).build(), b.selector(0, 1, 0).uniform(SEED + s, 23, 0.5, 1.5).build(),
            b.selector(0, 1, 0).uniform(SEED + s, 24, 0.5, 1.5).build(), b.uniform(SEED + s, 22, 0.0, 0.25).build(),
            b.value(0).build()) for s in range(n)]

    def run_shd(s):
        inp, cro, cru, coeff, out = shd[s % n]
        stencils.simple_hori_diff(out, inp, cro, cru, coeff, g, stream=launch_stream().handle)

    def stats(ms, bytes_per_pt):
        gbs = bytes_per_pt * P / (ms * 1e-3) / 1e9
        return {"mean_us": round(ms * 1e3, 2), "gpts_per_s": round(P / (ms * 1e-3) / 1e9, 3),
                "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                "bytes_per_point": bytes_per_pt}

    gt_sync()  # fields were filled on the null stream

    for s in range(warmup):
        run_shd(s)
    res = {"simple_hori_diff_fp64": stats(_event_time(run_shd, steps), 24)}
    del shd
    bt = gt.builder().type(np.float64).dimensions(NI, NJ, NK)
    gt_ = gt.make_grid(NI, NJ, NK)
    tracers = [[bt.uniform(SEED + 11 * c + m, 31, -1.0, 1.0).build() for m in range(11)] for c in range(2)]
    outs = [[bt.value(0).build() for _ in range(11)] for _ in range(2)]
    rho = [bt.uniform(SEED + c, 32, 0.5, 1.5).build() for c in range(2)]

    def run_tracers(s):
        stencils.prepare_tracers(outs[s % 2], tracers[s % 2], rho[s % 2], gt_, stream=launch_stream().handle)

    gt_sync()  # fields were filled on the null stream

    for s in range(min(warmup, 4)):
        run_tracers(s)
    res["prepare_tracers_fp64_x11"] = stats(_event_time(run_tracers, max(10, steps // 4)), (2 * 11 + 1) * 8)
    del tracers, outs, rho
    ins = [[bt.uniform(SEED + 5 * c + t, 41, -1.0, 1.0).build() for t in range(5)] for c in range(2)]
    outs = [[bt.value(0).build() for _ in range(5)] for _ in range(2)]

    def run_copy(s):
        stencils.copy_fields(outs[s % 2], ins[s % 2], gt_, stream=launch_stream().handle)

    gt_sync()  # fields were filled on the null stream

    for s in range(min(warmup, 4)):
        run_copy(s)
    res["copy_tuple_fp64_x5"] = stats(_event_time(run_copy, max(10, steps // 2)), 5 * 16)
    del ins, outs
    # extended_4D: 8 f values and 2 jac values read, 8 results written per point (144 B/point FP64)
    b6 = gt.builder().type(np.float64).dimensions(NI, NJ, NK, 2, 2, 2)
    rng = np.random.default_rng(SEED)
    phi, psi = rng.uniform(-1, 1, size=(2, 2, 2, 2)), rng.uniform(-1, 1, size=(2, 2, 2, 2))
    ext = [(b6.build(), b6.uniform(SEED + c, 51, -1.0, 1.0).build(),
            gt.builder().type(np.float64).dimensions(NI, NJ, NK, 2).uniform(SEED + c, 52, 0.5, 1.5).build())
           for c in range(2)]

    def run_ext(s):
        r, f, jac = ext[s % 2]
        stencils.extended_4d(r, f, jac, phi, psi, gt_, stream=launch_stream().handle)

    gt_sync()  # fields were filled on the null stream

    for s in range(min(warmup, 4)):
        run_ext(s)
    res["extended_4d_fp64"] = stats(_event_time(run_ext, max(10, steps // 2)), 18 * 8)
    del ext
    # transform_layout: a 256x256x80 FP64 storage interior (i fastest, halo 3, padded rows) into a k-fastest dense
    # buffer -- the transpose path (16 B/point)
    from gridtools_amd import layout

    h3 = 3
    bs = gt.builder().type(np.float64).dimensions(NI + 2 * h3, NJ + 2 * h3, NK).halos(h3, h3, 0)
    stores = [bs.uniform(SEED + c, 61, -1.0, 1.0).build() for c in range(4)]
    dense = [gt.builder().type(np.float64).dimensions(NK, NJ, NI).build() for _ in range(4)]

    def run_lt(s):
        src, dst = stores[s % 4], dense[s % 4]
        layout.transform_layout(dst.address(0, 0, 0), src.address(h3, h3, 0), [NI, NJ, NK],
                                [dst.strides[2], dst.strides[1], 1], list(src.strides),
                                stream=launch_stream().handle)

    gt_sync()  # fields were filled on the null stream

    for s in range(min(warmup, 4)):
        run_lt(s)
    res["transform_layout_fp64_transpose"] = stats(_event_time(run_lt, max(10, steps // 2)), 16)
    return res


def generic_engine_leg(reps=200):
    """The generic stencil engine (include/gt355/stencil.hpp) on the reference's own specs at 256x256x80 FP64 --
    horizontal_diffusion.cpp's hdiff, vertical_advection_dycore.cpp's vadv and tridiagonal.cpp's Thomas solve -- timed
    by the C++ program tests/cpp/gt355_generic --bench (hipEvents over back-to-back runs on 4 rotating field sets).
    Twice: as a user gets it (the hdiff and vadv specs are recognised and run on the hand-written kernels; the
    tridiagonal spec, which matches no hand-written kernel, runs on the engine's kernels), and with
    GT355_ENGINE_FAST=0 (every spec on the engine's own kernels)."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "cpp", "gt355_generic")
    if not os.path.exists(exe):
        return None
    out = {}
    for key, fast in (("drop_in", "1"), ("engine_kernels", "0")):
        env = dict(os.environ, GT355_ENGINE_FAST=fast)
        res = subprocess.run([exe, "--bench", str(reps)], capture_output=True, text=True, timeout=240, env=env)
        if res.returncode != 0:
            raise RuntimeError(f"gt355_generic --bench failed: {res.stdout[-500:]} {res.stderr[-500:]}")
        r = json.loads(res.stdout.strip().splitlines()[-1])
        for k in ("hdiff_gbs", "vadv_gbs", "tridiagonal_gbs"):
            r[k.replace("gbs", "hbm_frac")] = round(r[k] / HBM_PEAK_GBS, 4)
        out[key] = r
    return out


def _event_time(fn, n, min_ms=5.0):
    """Mean ms of n calls of fn() (which launch on launch_stream()) bracketed by one event pair on that stream
    (synchronised both sides).  A window shorter than min_ms is measured again with enough calls to fill it."""
    from gridtools_amd.timing import Event, synchronize

    st = launch_stream()

    def window(count):
        synchronize()
        st.synchronize()
        start, stop = Event(), Event()
        start.record(st)
        for s in range(count):
            fn(s)
        stop.record(st)
        stop.synchronize()
        st.synchronize()
        return start.elapsed_ms(stop)

    total = window(n)
    if total < min_ms:
        n = int(np.ceil(min_ms / max(total / n, 1e-4)))
        total = window(n)
    return total / n


def compose_leg(gt, steps, warmup, n=512, nk=80, nsets=2):
    """The precision form on the hdiff -> vadv composition at n x n x nk (gt355_hdiff_vadv: an hdiff launch then
    a vadv launch) with FP32 storage and FP64 compute, against the same composition in FP64; the fused single-launch
    kernel is timed beside it.  Errors: each
    output of the FP32 form against the FP64 path run on the same FP32-rounded inputs (the FP64 path equals the
    FP64 oracle bit for bit), max and RMS, relative to the field's max |value| and RMS."""
    from gridtools_amd import _abi, inputs, stencils

    h = 3
    grid = gt.make_grid(gt.halo_descriptor.symmetric(h, n + 2 * h), gt.halo_descriptor.symmetric(h, n + 2 * h), nk)

    def make(dtype, seed):
        b = gt.builder().type(dtype).dimensions(n + 2 * h, n + 2 * h, nk).halos(h, h, 0)
        f = {k: b.uniform(seed, inputs.FIELD_IDS[k], -1.0, 1.0).build()
             for k in ("in", "utens_stage", "u_pos", "utens")}
        f["coeff"] = b.uniform(seed, inputs.FIELD_IDS["coeff"], 0.0, 0.25).build()
        f["u_stage"] = b.value(0).build()
        f["wcon"] = gt.builder().type(dtype).dimensions(n + 2 * h, n + 2 * h, nk + 1).halos(h, h, 0) \
            .uniform(seed, inputs.FIELD_IDS["wcon"], -0.5, 0.5).build()
        return f

    # dtr_stage as the FP32 form holds it, so that the FP64 reference solves the same system (the solve is
    # ill-conditioned enough that the 4e-8 relative change of dtr alone moves utens_stage by ~1e-2)
    dtr = float(np.float32(0.15))

    def run(f, precision):
        stencils.hdiff_vadv(f["in"], f["coeff"], f["u_stage"], f["utens_stage"], f["wcon"], f["u_pos"], f["utens"],
                            dtr, grid, precision=precision, stream=launch_stream().handle)

    res, two = {}, {}
    for name, dtype, prec in (("fp64", np.float64, gt.NATIVE), ("fp32_form", np.float32, gt.WIDEN)):
        sets = [make(dtype, SEED + s) for s in range(nsets)]
        gt_sync()  # fields were filled on the null stream
        for s in range(warmup):
            run(sets[s % nsets], prec)
        # default dispatch: hdiff launch, then vadv launch (u_stage through HBM)
        res[name] = _event_time(lambda s: run(sets[s % nsets], prec), steps)
        # the fused single-launch kernels, for comparison: LDS tile (1) and DPP register sweep (2)
        two[name] = {}
        for variant, key in (("1", "lds_tile"), ("2", "dpp_sweep")):
            os.environ["GT355_TUNE_hv_fused"] = variant
            _abi.call("gt355_tune_env_refresh")
            try:
                for s in range(max(warmup, 2)):  # first launches load the variant's code object
                    run(sets[s % nsets], prec)
                two[name][key] = round(_event_time(lambda s: run(sets[s % nsets], prec), steps) * 1e3, 2)
            finally:
                del os.environ["GT355_TUNE_hv_fused"]
                _abi.call("gt355_tune_env_refresh")
        del sets
    pts = n * n * nk
    # errors on one fresh input set: FP32 form vs the FP64 path on the widened FP32 inputs
    f32 = make(np.float32, SEED)
    f64 = {}
    for k, v in f32.items():
        f64[k] = gt.builder().type(np.float64).dimensions(*v.lengths).halos(*v.halos).build()
        stencils.convert(f64[k], v)
    gt_sync()  # fields were filled and converted on the null stream
    run(f32, gt.WIDEN)
    run(f64, gt.NATIVE)
    gt_sync()
    box = np.s_[:nk, h:h + n, h:h + n]
    errors = {}
    for k in ("u_stage", "utens_stage"):
        a = f32[k].padded_numpy()[box].astype(np.float64)
        r = f64[k].padded_numpy()[box]
        d = a - r
        errors[k] = {"max_rel": float(np.max(np.abs(d)) / np.max(np.abs(r))),
                     "rms_rel": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(r * r)))}
    # algorithmic bytes per point of the two launches: hdiff reads in, coeff and writes u_stage; vadv reads
    # utens_stage, u_stage, wcon, u_pos, utens and writes utens_stage (9 accesses)
    out = {"workload": f"hdiff->vadv composition {n}x{n}x{nk} (hdiff launch + vadv launch per step)"}
    for name, esize in (("fp64", 8), ("fp32_form", 4)):
        ms = res[name]
        gbs = 9 * esize * pts / (ms * 1e-3) / 1e9
        out[name] = {"mean_us": round(ms * 1e3, 2), "gpts_per_s": round(pts / (ms * 1e-3) / 1e9, 3),
                     "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["fused_single_launch_us"] = two
    out["speedup_fp32_form_vs_fp64"] = round(res["fp64"] / res["fp32_form"], 3)
    out["errors_vs_fp64"] = errors
    return out


def gt_sync():
    from gridtools_amd.timing import synchronize

    synchronize()


def batch_leg(gt, nsteps=10, reps=3, seed=2024, per_domain_calls=True):
    """The batching form: 64 domains, ni, nj ~ U{48..384}, nk ~ U{40..137} (seeded), laplacian + hdiff + vadv in
    FP32 storage / FP64 compute, in the batch's default two-phase form (per timestep one persistent launch of the
    stencil phase and one of the vertical phase, each over every domain); the grid size from the batch tuner; a
    graph of nsteps timesteps (u_stage -> in swap) replayed `reps` times."""
    from gridtools_amd.batch import BatchedTimesteps, draw_sizes
    from gridtools_amd.timing import Event, Stream

    sizes = draw_sizes(seed)
    bt = BatchedTimesteps(sizes, seed=seed)
    w0, us0 = bt.batches[0].tune(reps=3)
    w1, us1 = bt.batches[1].tune(reps=3)
    st = Stream()
    g = bt.capture(nsteps, st)
    g.launch(st)  # warm-up replay
    st.synchronize()
    start, stop = Event(), Event()
    start.record(st)
    for _ in range(reps):
        g.launch(st)
    stop.record(st)
    stop.synchronize()
    ms = start.elapsed_ms(stop) / reps
    pts = bt.points()
    items, grid = bt.batches[0].info()
    # algorithmic bytes per point and timestep (FP32 storage): in, coeff, utens_stage, wcon, u_pos, utens read;
    # lap, u_stage, utens_stage written
    gbs = 9 * 4 * pts * nsteps / (ms * 1e-3) / 1e9
    calls = None
    if per_domain_calls:
        scoped = deferred_calls(bt, nsteps, reps, True)
        direct = deferred_calls(bt, nsteps, reps, False)
        calls = {
            "workload": "the same domains as one gt355_hdiff_vadv call per domain and timestep (no laplacian), "
                        f"{nsteps} timesteps in a graph",
            "ms_per_timestep_in_deferred_scope": round(scoped / nsteps, 4),
            "ms_per_timestep_direct": round(direct / nsteps, 4),
        }
    return {
        "workload": f"64 domains (seed {seed}), {pts / 1e6:.1f} M points, {nsteps}-timestep graph, two-phase form "
                    "(stencil-phase + vertical-phase launch per timestep, each over every domain)",
        "ms_per_graph": round(ms, 4),
        "ms_per_timestep": round(ms / nsteps, 4),
        "gpts_per_s": round(pts * nsteps / (ms * 1e-3) / 1e9, 3),
        "achieved_gbs": round(gbs, 1),
        "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
        "items": items,
        "tuned_wg_per_cu": [w0, w1],
        "tuned_config": [bt.batches[0].config(), bt.batches[1].config()],
        "tune_us": [round(us0, 1), round(us1, 1)],
        "workgroups": grid,
        "per_domain_calls": calls,
    }


def deferred_calls(bt, nsteps, reps, scoped):
    """Milliseconds per graph of nsteps timesteps that call gt355_hdiff_vadv once per domain (u and u_stage swap
    roles every step), inside a deferred-run scope per timestep (two persistent launches) or directly (two launches
    per domain).  The graph is captured after one eager run, which builds the scope's batches."""
    from gridtools_amd import stencils
    from gridtools_amd.batch import deferred
    from gridtools_amd.timestep import capture
    from gridtools_amd.timing import Event, Stream

    def steps(s):
        for n in range(nsteps):
            with deferred(s) if scoped else contextlib.nullcontext():
                for f in bt.domains:
                    a, b = (f["u"], f["v"]) if n % 2 == 0 else (f["v"], f["u"])
                    stencils.hdiff_vadv(a, f["coeff"], b, f["utens_stage"], f["wcon"], f["u_pos"], f["utens"], bt.dtr,
                                        f["grid"], precision=bt.batches[0].precision, stream=s.handle)

    st = Stream()
    steps(st)
    g = capture(st, steps)
    g.launch(st)
    st.synchronize()
    start, stop = Event(), Event()
    start.record(st)
    for _ in range(reps):
        g.launch(st)
    stop.record(st)
    stop.synchronize()
    return start.elapsed_ms(stop) / reps


def large_leg(gt, nsteps=20, n=1536, nk=80):
    """Large single domain, FP32 storage / FP64 compute, n x n x nk, nsteps timesteps of composition + boundary
    conditions (periodic i, zero-gradient j, 3-point halo) captured as one graph; the boundary conditions by one
    halo-fill launch, and beside it by the halo-pack kernels (pack / single-rank periodic exchange / unpack)."""
    from gridtools_amd.timestep import HdiffVadvTimestepper
    from gridtools_amd.timing import Event, Stream

    def run_graph(bc):
        ts = HdiffVadvTimestepper(n, n, nk, seed=SEED, bc=bc)
        st = Stream()
        g = ts.capture(nsteps, st)
        g.launch(st)
        st.synchronize()
        start, stop = Event(), Event()
        start.record(st)
        g.launch(st)
        stop.record(st)
        stop.synchronize()
        return start.elapsed_ms(stop), ts.points()

    ms, pts = run_graph("fill")
    ms_x, _ = run_graph("exchange")
    gbs = 8 * 4 * pts * nsteps / (ms * 1e-3) / 1e9  # fused launch traffic only (halo fill excluded)
    bc = boundary_conditions(n, nk)
    return {
        "workload": f"{n}x{n}x{nk} FP32 storage / FP64 compute, {nsteps} timesteps (compose + halo BCs) in a graph",
        "ms_total": round(ms, 3),
        "ms_per_timestep": round(ms / nsteps, 4),
        "gpts_per_s": round(pts * nsteps / (ms * 1e-3) / 1e9, 3),
        "achieved_gbs_compose": round(gbs, 1),
        "hbm_frac_compose": round(gbs / HBM_PEAK_GBS, 4),
        "ms_per_timestep_bc_by_halo_exchange": round(ms_x / nsteps, 4),
        "boundary_conditions": bc,
    }


def boundary_conditions(n, nk, h=3, reps=KERNEL_WINDOW):
    """The timestep's boundary conditions alone on its two n x n x nk FP32 fields with an h-point halo (periodic i,
    zero-gradient j): one halo-fill launch, and the halo-pack route (zero-gradient j by halo fill, then the periodic
    i faces by the single-rank pack / exchange / unpack).  Mean microseconds over `reps` back-to-back calls on the
    launch stream; algorithmic bytes = every halo point of both fields written once and its source read once."""
    from gridtools_amd import boundaries
    from gridtools_amd.grid import halo_descriptor
    from gridtools_amd.timestep import HdiffVadvTimestepper
    from gridtools_amd.timing import Event

    ts = HdiffVadvTimestepper(n, n, nk, seed=SEED, bc="exchange", halo=h)
    st = launch_stream()
    fields = [ts.u_stage, ts.utens_stage]
    P, Z, N = boundaries.PERIODIC, boundaries.ZERO_GRADIENT, boundaries.NONE

    def timed(fn):
        fn()
        st.synchronize()
        start, stop = Event(), Event()
        start.record(st)
        for _ in range(reps):
            fn()
        stop.record(st)
        stop.synchronize()
        return 1e3 * start.elapsed_ms(stop) / reps

    fill_us = timed(lambda: boundaries.halo_fill(fields, ts.halos, (P, Z, N), stream=st.handle))
    xchg_us = timed(lambda: (boundaries.halo_fill(fields, ts.halos, (N, Z, N), stream=st.handle),
                             boundaries.halo_exchange_local(fields, ts._xhalos, ts._xbuf.value, ts._xmask,
                                                            stream=st.handle)))
    halo_pts = ((n + 2 * h) ** 2 - n * n) * nk
    algo = 2 * 2 * 4 * halo_pts
    return {
        "fields": f"2 x {n}x{n}x{nk} FP32, halo {h}, periodic i / zero-gradient j",
        "halo_points_per_field": halo_pts,
        "algo_bytes": algo,
        "halo_fill_us": round(fill_us, 2),
        "halo_fill_gbs": round(algo / (fill_us * 1e-6) / 1e9, 1),
        "pack_exchange_route_us": round(xchg_us, 2),
    }


def host_cpu():
    """CPU model, physical cores of the host, and the CPUs this process may run on."""
    model, cores = "unknown", set()
    try:
        phys = None
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name") and model == "unknown":
                    model = ln.split(":", 1)[1].strip()
                elif ln.startswith("physical id"):
                    phys = ln.split(":", 1)[1].strip()
                elif ln.startswith("core id"):
                    cores.add((phys, ln.split(":", 1)[1].strip()))
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, len(cores) or None, allowed


def cpu_baseline(budget_s=12.0):
    """The oracle ports (OpenMP, same arithmetic) on this host, hdiff+vadv FP64 256x256x80, bounded samples: the
    cpu_ifirst traversal and the cpu_kfirst traversal (8x8 blocks, k innermost); the faster one is the value.

    Threads: OMP_NUM_THREADS, the host-CPU share of one GPU on the measuring machine (16 there), else the CPUs this
    process may use; bound one per core (OMP_PROC_BIND=true, OMP_PLACES=cores, set in main() before the OpenMP
    runtime loads), as the reference's perf runs bind them (jenkins/envs/base.sh:18-19)."""
    from gridtools_amd import inputs
    from oracle import binding as ob

    model, phys, allowed = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or allowed
    threads = max(1, min(threads, allowed))
    h = 3
    shp = (NK, NJ + 2 * h, NI + 2 * h)
    rnd = lambda fid, a, b, nk=NK: np.ascontiguousarray(  # noqa: E731
        inputs.uniform((NI + 2 * h, NJ + 2 * h, nk), SEED, fid, a, b).transpose(2, 1, 0))
    in_, co = rnd(1, -1, 1), rnd(2, 0, 0.25)
    out = np.zeros(shp)
    uts0, us, up, ut = (rnd(i, -1, 1) for i in (6, 3, 4, 5))
    wc = rnd(7, -0.5, 0.5, NK + 1)
    V = ob.View
    res = {}
    for mode, name in (("native", "cpu_ifirst"), ("kfirst", "cpu_kfirst")):
        times = []
        t_end = time.perf_counter() + budget_s / 2
        reps = 0
        while (reps < 3 or time.perf_counter() < t_end) and reps < 5000:
            uts = uts0.copy()
            t0 = time.perf_counter()
            ob.hdiff(V(out, h, h), V(in_, h, h), V(co, h, h), NI, NJ, NK, mode=mode, threads=threads)
            ob.vadv(V(uts, h, h), V(us, h, h), V(wc, h, h), V(up, h, h), V(ut, h, h), 0.15, NI, NJ, NK, mode=mode,
                    threads=threads)
            times.append(time.perf_counter() - t0)
            reps += 1
        med = statistics.median(times)
        res[name] = {"value": round(P / med / 1e9, 4), "reps": reps, "median_ms": round(med * 1e3, 2)}
    best = max(res, key=lambda n: res[n]["value"])
    return {
        "value": res[best]["value"],
        "unit": "Ggridpoint-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"hdiff+vadv FP64 256x256x80, oracle/gt_oracle.c OpenMP ports (the reference's cpu_ifirst "
                  f"cannot be built here: its headers need Boost.Preprocessor, absent from the image) on {threads} "
                  f"threads bound to cores ({model}; {phys} physical cores in the host, {allowed} usable by this "
                  f"process): cpu_ifirst traversal {res['cpu_ifirst']['reps']} reps (median "
                  f"{res['cpu_ifirst']['median_ms']} ms), cpu_kfirst traversal {res['cpu_kfirst']['reps']} reps "
                  f"(median {res['cpu_kfirst']['median_ms']} ms); value = the faster ({best})",
        "cpu_model": model,
        "host_physical_cores": phys,
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "variants": res,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # 1000 untimed steps (63 ms FP64, 37 ms FP32 form): after 100 the 200 timed FP32-form steps still read 40.3-40.4
    # us per step where 2000 steps give 36.7 us, and the FP64 step 63.4-63.7 against 62.6-62.8 us; after 1000 the
    # timed steps read the steady state
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--sets", type=int, default=4)
    ap.add_argument("--legs", default="tune,fp64,fp32,graph,copy,cartesian,generic,compose,batch,deferred,large,cpu",
                    help="comma list of legs: tune (hdiff autotuner first, untimed), fp64 (always run), fp32 "
                         "(precision form per kernel), graph (the steps captured in a graph), copy (bandwidth "
                         "yardstick), cartesian (simple_hori_diff, prepare_tracers x11, copy_stencil_tuple x5, extended_4D, "
                         "transform_layout), generic (the generic engine on user hdiff / vadv specs), "
                         "compose (precision form on the fused composition, 512x512x80), batch (64 "
                         "uneven domains, persistent launch, graph of 10 timesteps), deferred (with batch: the same "
                         "domains as per-domain calls in a deferred-run scope and directly), large (1536x1536x80, 20 "
                         "timesteps with boundary conditions), cpu (CPU baseline)")
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    args.no_fp32 = "fp32" not in legs
    args.no_cpu = "cpu" not in legs
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    # the CPU baseline's OpenMP threads are bound one per core; libgomp reads these when the oracle loads
    os.environ.setdefault("OMP_PROC_BIND", "true")
    os.environ.setdefault("OMP_PLACES", "cores")

    import gridtools_amd as gt
    from gridtools_amd.timing import Range

    gt.init(0)
    info = gt.device_info()
    print(f"[bench] device {info['name']} {info['arch']} CUs={info['cu_count']}", file=sys.stderr, flush=True)

    tuned = None
    if "tune" in legs:
        # hdiff tiles chosen by the autotuner for every hdiff shape this run times (FP64 and the FP32 form)
        print("[bench] autotune hdiff", file=sys.stderr, flush=True)
        shapes = [(NI, NK), (512, NK)]
        tuned = {"fp64": autotune(gt, np.float64, gt.NATIVE, shapes),
                 "fp32_form": autotune(gt, np.float32, gt.WIDEN, shapes)}
    sets64 = [make_set(gt, np.float64, s) for s in range(args.sets)]
    with Range("bench: hdiff + vadv FP64 256x256x80"):
        total, hd_ms, va_ms, win64 = timed_steps(gt, sets64, gt.NATIVE, args.steps, args.warmup)
    print(f"[bench] fp64 {args.steps} steps: {total:.3f} ms", file=sys.stderr, flush=True)
    ms_per_step = total / args.steps
    hd64 = kernel_stats("hdiff", hd_ms, 8)
    va64 = kernel_stats("vadv", va_ms, 8)
    graphed = None
    if "graph" in legs:
        graphed = {"fp64": graphed_steps(gt, sets64, gt.NATIVE)}
    del sets64

    fp32 = None
    if not args.no_fp32:
        sets32 = [make_set(gt, np.float32, s) for s in range(args.sets)]
        total32, hd_ms32, va_ms32, win32 = timed_steps(gt, sets32, gt.WIDEN, args.steps, args.warmup)
        print(f"[bench] fp32 form {args.steps} steps: {total32:.3f} ms", file=sys.stderr, flush=True)
        fp32 = {
            "form": "FP32 storage, FP64 compute (pivots and divisions in FP64)",
            "ms_per_step": round(total32 / args.steps, 5),
            "gpts_per_s": round(P / (total32 / args.steps * 1e-3) / 1e9, 3),
            "speedup_vs_fp64_k_steps": round(total / total32, 3),
            # both forms over the same long window of steps (step_window below)
            "ms_per_step_long_window": round(win32, 5),
            "speedup_vs_fp64": round(win64 / win32, 3),
            "hdiff": kernel_stats("hdiff", hd_ms32, 4),
            "vadv": kernel_stats("vadv", va_ms32, 4),
        }
        if graphed is not None:
            graphed["fp32_form"] = graphed_steps(gt, sets32, gt.WIDEN)
        del sets32
    in_graph = None
    if graphed is not None:
        in_graph = {"workload": "the same steps captured in a graph (40 steps, replayed 10 times)"}
        for form, v in graphed.items():
            in_graph[form] = {"ms_per_step": round(v, 5), "gpts_per_s": round(P / (v * 1e-3) / 1e9, 3)}
        if "fp32_form" in graphed:
            in_graph["fp32_form_speedup_vs_fp64"] = round(graphed["fp64"] / graphed["fp32_form"], 3)

    copy = copy_leg(gt, args.steps, args.warmup, args.sets) if "copy" in legs else None
    cartesian = None
    if "cartesian" in legs:
        print("[bench] cartesian regression stencils leg", file=sys.stderr, flush=True)
        with Range("bench: cartesian regression stencils 256x256x80"):
            cartesian = cartesian_leg(gt, args.steps, args.warmup, args.sets)
    generic = None
    if "generic" in legs:
        print("[bench] generic engine leg", file=sys.stderr, flush=True)
        with Range("bench: generic stencil engine 256x256x80"):
            generic = generic_engine_leg()
    compose = None
    if "compose" in legs:
        print("[bench] compose leg", file=sys.stderr, flush=True)
        with Range("bench: precision-form composition 512x512x80"):
            compose = compose_leg(gt, max(200, args.steps // 4), min(args.warmup, 20))
    batch = None
    if "batch" in legs:
        print("[bench] batch leg", file=sys.stderr, flush=True)
        with Range("bench: batching form, 64 domains"):
            batch = batch_leg(gt, per_domain_calls="deferred" in legs)
    large = None
    if "large" in legs:
        print("[bench] large-domain leg", file=sys.stderr, flush=True)
        with Range("bench: large domain 1536x1536x80"):
            large = large_leg(gt)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline()

    value = P / (ms_per_step * 1e-3)
    dom = va64
    traffic = measured_traffic("vadv_onchip_kernel<double,double,80")
    line = {
        "metric": "gridpoint-updates/s + % HBM roofline, hdiff & vadv FP64 vs FP32 @256x256x80",
        "value": round(value, 1),
        "unit": "gridpoint-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64: in,u_* ~ U(-1,1), coeff ~ U(0,0.25), wcon ~ U(-0.5,0.5), dtr 0.15)",
        "config": {"workload": "hdiff + vadv FP64, 256x256x80 interior (halo 2 / 3), one launch each per step",
                   "field_sets": args.sets, "parallelism": "single GPU"},
        "kernels": {"hdiff_fp64": hd64, "vadv_fp64": va64},
        # the whole step: hdiff's 24 B/point and vadv's 48 B/point of algorithmic traffic over the step time
        "step_fp64": {"algo_bytes": 72 * P, "achieved_gbs": round(72 * P / (ms_per_step * 1e-3) / 1e9, 1),
                      "hbm_frac": round(72 * P / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        # the FP64 step again over a window of max(steps, STEP_WINDOW) steps, the FP32 form's comparison basis
        "step_window": {"steps": max(args.steps, STEP_WINDOW), "fp64_ms_per_step": round(win64, 5)},
        "autotuned_hdiff": tuned,
        "fp32_form": fp32,
        "graphed_steps": in_graph,
        "copy_fp64": copy,
        "cartesian_stencils": cartesian,
        "generic_engine": generic,
        "precision_form_composition": compose,
        "batching_form": batch,
        "large_domain": large,
        "roofline": {
            "kernel": "vadv FP64 (vadv_*_kernel<double,double>)",
            "bound": "hbm",
            "achieved": dom["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic["bytes"] if traffic else None,
            "traffic_source": traffic,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
