#!/usr/bin/env python3
"""Benchmark of the MI355X D2Q9 step (one JSON line on stdout).

Default workload (--workload batched): the batched configuration -- 256
seeded disk packings in one batch, side drawn from {128, 256, 512, 1024},
porosity U[0.4, 0.8], disk radius U[3, 16], Kn log-uniform in [0.01, 1],
regularised MRT collision with specular bounce-back, periodic in x and y,
FP32 deviation storage.  A "step" is one launch of the batched transport
kernel over every lattice of the batch (about 53 M fluid cells).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--repeats R]
                    [--workload batched|quadtree|shanchen|transport|fp64|warmstart|shape]
                    [--no-cpu] [--no-steady] [--no-fp64] [--eager]

value = fluid-cell updates per second (MLUPS) over a window of exactly K
steps, timed with HIP events on the batch's stream with the populations
resident in HBM; R windows are timed and the median is reported.
roofline: 72 B per cell update (nine FP32 populations read and written;
metadata not counted) over the median window's time per launch, beside the
measured copy ceiling of the same GPU; traffic: HBM bytes per launch from
the committed rocprofv3 counter summary of the same workload.  Before the
windows the batch runs every lattice from rest to steady state (the wall
time to steady permeability), and sampled lattices are compared with the
FP64 single-block path (permeability error).  The CPU baseline is the oracle
(the C++ restatement of the reference, -Ofast -fopenmp) on one lattice of the
batch, with the host's threads and with one.

Other workloads: quadtree (multi-block transport of a 1024^2 packing, a
quadtree of at most --levels levels, 3 by default; a step is one coarse
step), shanchen (Shan-Chen adsorption, 2048^2, 88 B per
update), transport (single-block FP32 transport, --n), fp64 (single-block
FP64 transport, 256^2, 144 B per update), warmstart (the learned warm start
against plain runs), shape (Shape::calculateProperties, 1024^2).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def shanchen_inputs(n=2048, seed=2027):
    from mantiis_amd import geometry as G
    rng = np.random.default_rng(seed)
    pore = G.disk_packing(n, 0.5, 4, 12, seed=seed)
    G_ads = float(rng.uniform(-1.5, -0.5))
    rho0 = float(rng.uniform(0.1, 0.5))
    kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
    rho = rho0 * (1.0 + 0.01 * rng.standard_normal(pore.shape))
    return pore, kn, lp, G_ads, rho0, rho


def transport_inputs(n=1024, seed=2026):
    """Quadtree-workload geometry (porosity U[0.4, 0.8], Kn = 0.5) run single-block."""
    from mantiis_amd import geometry as G
    rng = np.random.default_rng(seed)
    phi = float(rng.uniform(0.4, 0.8))
    pore = G.disk_packing(n, phi, 4, 12, seed=seed)
    kn, lp = G.uniform_kn(pore, 0.5, 0.7730846760658538)
    return pore, kn, lp


def quadtree_inputs(n=1024, seed=2025):
    """Quadtree configuration: seeded disk packing, porosity U[0.4, 0.8], Kn = 0.5."""
    from mantiis_amd import geometry as G
    rng = np.random.default_rng(seed)
    phi = float(rng.uniform(0.4, 0.8))
    pore = G.disk_packing(n, phi, 4, 12, seed=seed)
    kn, lp = G.uniform_kn(pore, 0.5, 0.7730846760658538)
    return pore, kn, lp, phi


def copy_peak_gbs(launches=10):
    """Measured copy bandwidth of this GPU: the library's 16-byte-per-lane
    nontemporal copy of 1 GiB (read + write bytes per second), timed with HIP
    events on the stream it runs on.  The roofline's practical ceiling, beside
    the nominal 8 TB/s."""
    from mantiis_amd.hip import DeviceArray, load_library
    from mantiis_amd.hiptime import StreamTimer
    n = 1 << 28
    a, b = DeviceArray((n,)), DeviceArray((n,))
    lib = load_library()
    if lib.mts_probe_copy16(b.ptr, a.ptr, n, 2) != 0:
        raise RuntimeError(lib.mts_last_error().decode())
    t = StreamTimer(0)
    t.start()
    rc = lib.mts_probe_copy16(b.ptr, a.ptr, n, launches)
    ms = t.stop()
    t.close()
    a.close()
    b.close()
    if rc != 0:
        raise RuntimeError(lib.mts_last_error().decode())
    return 2 * 4 * n * launches / (ms / 1e3) / 1e9


def mean_ux(lat):
    d = lat.diagnostics()  # sum rho, sum j_x, sum j_y, cells
    return d[1] / d[0]


def batched_one(args):
    """One lattice of the batched configuration: side L in {128, 256, 512,
    1024}, porosity U[0.4, 0.8], disk radius U[3, 16], Kn log-uniform in
    [0.01, 1] (LocPore = (mfp/dx)/Kn)."""
    from mantiis_amd import geometry as G
    seed, = args
    rng = np.random.default_rng(seed)
    L = int(rng.choice([128, 256, 512, 1024]))
    phi = float(rng.uniform(0.4, 0.8))
    kn = float(np.exp(rng.uniform(np.log(0.01), np.log(1.0))))
    pore = G.disk_packing(L, phi, 3, 16, seed=seed)
    knf, lp = G.uniform_kn(pore, kn, 0.7730846760658538)
    return pore, knf, lp, kn


def batched_inputs(nb=256, seed=4000, workers=8):
    if workers <= 1:  # no forked workers (e.g. under a profiler that follows forks)
        return [batched_one((seed + i,)) for i in range(nb)]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(batched_one, [(seed + i,) for i in range(nb)])


def run_to_steady(lat, tol, chunk, max_steps, max_s):
    """Steps until the mean velocity changes by less than tol per step,
    |<u_x>(t) - <u_x>(t - chunk)| / (chunk |<u_x>(t)|) < tol, checked every
    chunk steps.  Returns (steps, seconds, criterion, reference residual)."""
    lat.synchronize()
    t0 = time.perf_counter()
    n = 0
    u_prev = mean_ux(lat)
    crit = 1.0
    while n < max_steps and time.perf_counter() - t0 < max_s:
        lat.step("transport", chunk)
        n += chunk
        u = mean_ux(lat)
        crit = abs(u - u_prev) / (chunk * abs(u)) if u != 0 else 1.0
        u_prev = u
        if crit < tol:
            break
    lat.synchronize()
    return n, time.perf_counter() - t0, crit, lat.residual()


PROFILE_FOR = {"batched": "k_cbat_batched_256.md", "quadtree": "k_cmbx_quadtree_transport_1024.md",
               "quadtree3": "k_cmb_quadtree3_transport_1024.md", "shanchen": "k_csc_shanchen_adsorption_2048.md",
               "transport": "k_ctr_transport_4096.md"}


def profiled_traffic(workload, n):
    """HBM bytes per launch of the dominant kernel (per coarse step, for the
    three-level quadtree) from the committed rocprofv3 summary of the same
    workload (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration);
    None when there is none."""
    name = PROFILE_FOR.get(workload)
    if name is None or (workload == "transport" and n != 4096):
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        return None
    return None


def w9():
    return np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def make_shanchen(precision="fp32dev"):
    from mantiis_amd.hip import Lattice
    pore, kn, lp, G_ads, rho0, rho = shanchen_inputs()
    lat = Lattice(pore, kn, lp, precision=precision, collision=1, wall=0, dx=1e-9, Fbody=0.0, Gfs=G_ads, periodic_y=1)
    fl = pore.reshape(-1) == 1
    r = rho.reshape(-1)[fl]
    lat.set("f", (r[:, None] * w9()[None, :]).reshape(-1))
    lat.set("rho", r)
    return lat, dict(pore=pore, kn=kn, lp=lp, G_ads=G_ads, rho0=rho0, rho=rho)


def time_steps(lat, kind, steps, warmup):
    """Event time of K steps in one call (the value), then the same K steps
    again with the kernel timer on: the summed start-to-end event time and
    the count of the dominant kernel's launches (the roofline).  The timer's
    per-launch events are kept out of the first region."""
    from mantiis_amd.hiptime import StreamTimer
    t = StreamTimer(lat.stream_handle())
    if warmup > 0:
        lat.step(kind, warmup)
    lat.synchronize()
    t.start()
    lat.step(kind, steps)
    ms = t.stop()
    lat.kernel_timer(True)
    lat.kernel_timer_read()  # clear
    lat.step(kind, steps)
    kms, klaunch = lat.kernel_timer_read()
    lat.kernel_timer(False)
    t.close()
    return ms, kms, klaunch


def quadtree_updates(level):
    """Cell updates of one coarse step: a cell of level l of L updates
    2^(L - l) times (coreLB.hpp:1310-1331)."""
    L = int(level.max())
    return int(sum((1 << (L - l)) * int((level == l).sum()) for l in range(1, L + 1)))


def cpu_baseline_quadtree(sample_n, sample_steps, levels):
    """Oracle multi-block transport on a sample_n^2 lattice of the same
    distribution (the reference's quadtree build is quadratic in the
    lattice size, so the 1024^2 lattice itself takes minutes to set up)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    pore, kn, lp, phi = quadtree_inputs(sample_n)
    o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-6, fast=True, periodic_y=True, multiblock=levels)
    upd = quadtree_updates(o.grid()["level"])
    o.run("transport", 1, 0.0)  # warm-up
    t0 = time.perf_counter()
    o.run("transport", sample_steps, 0.0)
    dt = time.perf_counter() - t0
    threads = int(O.load(fast=True).orc_threads())
    o.close()
    return dict(value=round(upd * sample_steps / dt / 1e6, 4), unit="MLUPS", cores=threads, kind="port",
                sample=f"{sample_steps} coarse steps (after 1 warm-up) of a {sample_n}x{sample_n} lattice of the same "
                       f"distribution (porosity {phi:.3f}, {upd} cell updates per step); oracle built -Ofast "
                       f"-fopenmp, OMP threads {threads}")


def cpu_baseline(inputs, sample_steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    o = O.Oracle(inputs["pore"], inputs["kn"], inputs["lp"], dx=1e-9, Fbody=0.0, fast=True, periodic_y=True)
    o.set_interaction(-1.0, inputs["G_ads"])
    fl = inputs["pore"].reshape(-1) == 1
    r = inputs["rho"].reshape(-1)[fl]
    o.set("f", (r[:, None] * w9()[None, :]).reshape(-1))
    o.set("rho", r)
    o.run("adsorption", 1, 0.0)  # warm-up
    t0 = time.perf_counter()
    o.run("adsorption", sample_steps, 0.0)
    dt = time.perf_counter() - t0
    n = o.ncell
    threads = int(O.load(fast=True).orc_threads())
    o.close()
    return dict(value=round(n * sample_steps / dt / 1e6, 4), unit="MLUPS", cores=threads, kind="port",
                sample=f"{sample_steps} adsorption steps of the same 2048^2 lattice ({n} fluid cells) after 1 warm-up "
                       f"step; oracle built -Ofast -fopenmp, OMP threads {threads} (forces and Guo source loops are "
                       f"serial as in the reference)")


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# The oracle's -Ofast build against the reference itself, measured in the
# build container (Intel Xeon @ 2.70 GHz, 8 cores) on a 256^2 transport
# lattice (RMRT + SRBB): the reference's 1.40 / 7.66 MLUPS at 1 / 8 threads
# (stubbed-header build of the survey) against the oracle's 2.53 / 11.8.
CPU_CALIBRATION = {"reference_mlups_1t": 1.40, "reference_mlups_8t": 7.66, "oracle_mlups_1t": 2.53,
                   "oracle_mlups_8t": 11.83, "ratio_1t": 1.81, "ratio_8t": 1.54,
                   "host": "build container, Intel(R) Xeon(R) Processor @ 2.70GHz, 8 cores",
                   "note": "the oracle runs 1.5-1.8x faster than the reference on the same host, so this baseline "
                           "overstates the reference's CPU speed"}


def cpu_baseline_batched(geoms, g, steps_n, steps_1):
    """Oracle (-Ofast -fopenmp) transport on lattice g of the batch: steps_n
    steps with the host's OpenMP threads, steps_1 steps with one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    pore, kn, lp, knv = geoms[g]
    lib = O.load(fast=True)
    threads = int(lib.orc_threads())
    out = {}
    for th, st in ((threads, steps_n), (1, steps_1)):
        lib.orc_set_threads(th)
        o = O.Oracle(pore, kn, lp, dx=1e-9, Fbody=1e-6, fast=True, periodic_y=True)
        o.run("transport", 2, 0.0)  # warm-up
        t0 = time.perf_counter()
        o.run("transport", st, 0.0)
        dt = time.perf_counter() - t0
        out[th] = (o.ncell * st / dt / 1e6, o.ncell, dt)
        o.close()
    lib.orc_set_threads(threads)
    n = pore.shape[0]
    return dict(value=round(out[threads][0], 4), unit="MLUPS", cores=threads, kind="port",
                one_thread=round(out[1][0], 4),
                sample=f"lattice {g} of the batch ({n}x{n}, Kn {knv:.4g}, {out[threads][1]} fluid cells): "
                       f"{steps_n} transport steps with {threads} OpenMP threads ({out[threads][2]:.1f} s), "
                       f"{steps_1} with 1 thread ({out[1][2]:.1f} s); oracle built -Ofast -fopenmp",
                host_cpu=host_cpu_model(), calibration=CPU_CALIBRATION)


def fp64_permeability_check(geoms, b, steps, sample):
    """Lattices `sample` of the batch against the FP64 single-block path of
    the library run from rest for the same number of steps: relative error of
    the permeability k = <u_x> nu / F (the FP32 batch's <u_x> against FP64's;
    nu and F are the same) and of the velocity field."""
    from mantiis_amd.hip import Lattice
    out = []
    for g in sample:
        pore, kn, lp, knv = geoms[g]
        st = int(abs(steps[g]))
        lat = Lattice(pore, kn, lp, precision="fp64", collision=2, wall=2, dx=1e-9, Fbody=1e-6, periodic_y=1)
        lat.step("transport", st)
        ux, uy, rho = lat.complete_domain()
        lat.close()
        brho, bux, buy = b.fields(g)
        fl = pore.reshape(-1) == 1
        k64 = float((rho[fl] * ux[fl]).sum() / rho[fl].sum())
        k32 = float((brho.reshape(-1)[fl] * bux.reshape(-1)[fl]).sum() / brho.reshape(-1)[fl].sum())
        du = float(np.linalg.norm(bux.reshape(-1)[fl] - ux[fl]) / np.linalg.norm(ux[fl]))
        out.append({"lattice": int(g), "side": int(pore.shape[0]), "kn": round(knv, 5), "steps": st,
                    "k_rel_err": abs(k32 - k64) / abs(k64), "ux_rel_l2": du})
    return out


def bench_batched(args):
    """The batched configuration (the default workload): 256 lattices of the
    batched distribution in one batch.  (1) every lattice from rest to steady
    state (mean-velocity change below 1e-6 per step, checked every 100 steps)
    -- the wall time to steady permeability; (2) all lattices re-activated,
    W warm-up steps, then `repeats` windows of exactly K steps, each one
    k_cbat launch per step over the whole batch, timed with HIP events on the
    batch's stream: value, ms_per_step and the roofline come from the median
    window; (3) the FP32 permeability of sampled lattices against the FP64
    single-block path; (4) the oracle on one lattice as the CPU baseline."""
    from mantiis_amd.hip import Batch
    from mantiis_amd.hiptime import StreamTimer
    t0 = time.perf_counter()
    geoms = batched_inputs(args.batch, workers=args.workers)
    t_gen = time.perf_counter() - t0
    print(f"[bench] {len(geoms)} lattices generated in {t_gen:.1f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = Batch([(g[0], g[1], g[2]) for g in geoms], collision=2, wall=2, dx=1e-9, Fbody=1e-6,
              batch_steps=args.batch_steps)
    t_setup = time.perf_counter() - t0
    cells = b.cells()
    ncell = int(cells.sum())
    print(f"[bench] batch of {ncell} cells set up in {t_setup:.1f} s", file=sys.stderr, flush=True)
    if args.launch:  # a fixed launch shape (profiler runs: the tuner's timings are distorted under counters)
        thr, xrun = (int(v) for v in args.launch.split(","))
        b.set_config(thr, xrun)
        tab = None
    else:
        thr, xrun, tab = b.tune(8)
    spl, units = b.kernel()  # steps per launch of the step kernel (2: the tile kernels)
    timer = StreamTimer(b.stream_handle())
    steady = None
    r = None
    if not args.no_steady:
        timer.start()
        r = b.run(400000, check_every=100, tol=1e-6, graph=not args.eager)
        run_ms = timer.stop()
        conv = r["steps"] > 0
        sizes = np.array([g[0].shape[0] for g in geoms])
        by_size = {int(L): {"lattices": int((sizes == L).sum()),
                            "median_steps": int(np.median(np.abs(r["steps"][sizes == L]))) if (sizes == L).any() else 0,
                            "max_steps": int(np.abs(r["steps"][sizes == L]).max()) if (sizes == L).any() else 0}
                   for L in (128, 256, 512, 1024)}
        steady = {"criterion": "|d<u_x>/dt| / |<u_x>| < 1e-6 per step, checked every 100 steps, per lattice",
                  "wall_s": round(run_ms / 1e3, 3), "converged": int(conv.sum()),
                  "max_steps": int(np.abs(r["steps"]).max()), "cell_updates": int(r["updates"]),
                  "mlups_over_run": round(r["updates"] / (run_ms / 1e3) / 1e6, 2), "graph": not args.eager,
                  "by_side": by_size}
        print(f"[bench] steady state of all lattices in {run_ms / 1e3:.2f} s", file=sys.stderr, flush=True)
    # the timed windows: every lattice active
    b.activate(np.ones(len(geoms), np.int32))
    if args.warmup > 0:
        b.step(args.warmup)
    reps = []
    for _ in range(max(args.repeats, 1)):
        timer.start()
        b.step(args.steps)
        reps.append(timer.stop())
    timer.close()
    ms = float(np.median(reps))
    ms_step = ms / args.steps
    mlups = ncell * args.steps / (ms / 1e3) / 1e6
    achieved = 72 * ncell / (ms_step / 1e3) / 1e9
    print(f"[bench] {args.repeats} windows of {args.steps} steps: median {ms_step:.4f} ms per step",
          file=sys.stderr, flush=True)
    fp64 = None
    if r is not None and not args.no_fp64:
        sizes = [g[0].shape[0] for g in geoms]
        sample = [sizes.index(L) for L in (128, 256, 512, 1024) if L in sizes]
        fp64 = fp64_permeability_check(geoms, b, r["steps"], sample)
        print("[bench] FP32 vs FP64 permeability checked", file=sys.stderr, flush=True)
    b.close()
    copy = copy_peak_gbs()
    cpu = None
    if not args.no_cpu:
        sizes = [g[0].shape[0] for g in geoms]
        g512 = sizes.index(512) if 512 in sizes else 0
        cpu = cpu_baseline_batched(geoms, g512, args.cpu_steps or 600, max((args.cpu_steps or 600) // 10, 10))
    traffic = profiled_traffic("batched", 0)
    out = {
        "metric": "MLUPS & %HBM roofline",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (populations)",
        "data": "synthetic (seeded disk packings)",
        "config": {"workload": f"batched: {args.batch} lattices, side U{{128,256,512,1024}}, porosity U[0.4,0.8], "
                               "disk radius U[3,16], Kn log-uniform [0.01,1], RMRT+SRBB, torus, FP32 deviations; "
                               "a step is one k_cbat launch over every lattice of the batch",
                   "lattices": args.batch, "fluid_cells": ncell, "parallelism": "single GPU",
                   "repeats": len(reps), "window_ms": [round(v, 4) for v in reps]},
        "roofline": {"bound": "hbm",
                     "kernel": ("k_ctile<RMRT,SRBB,2> (batched FP32 transport: a 32x16-site tile and its two-site "
                                "ring staged in LDS, two steps per launch over the tile queue)") if spl == 2 else
                               ("k_cbat<RMRT,SRBB> (batched FP32 transport: gather-stream + SRBB + RMRT, one step "
                                "per launch over the queue of 256-cell blocks)"),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_cell_update": 72,
                     "steps_per_launch": spl, "cells_per_launch": ncell,
                     "algorithmic_bytes_per_launch": 72 * ncell * spl,
                     "kernel_ms": round(ms_step * spl, 6), "launches_timed": args.steps // spl * len(reps),
                     "copy_peak": round(copy, 1), "frac_of_copy": round(achieved / copy, 4)},
        "cpu_baseline": cpu,
        "steady": steady,
        "fp32_vs_fp64": fp64,
        "autotune": {"threads_per_block": thr, "xcd_run": xrun,
                     "ms_per_step": [[round(v, 4) for v in row] for row in tab.tolist()],
                     "candidates": "rows: threads per 256-cell block 64, 128, 256; columns: XCD run 0, 8, 32; "
                                   "8 steps each from the first state, state restored"}
                    if tab is not None else {"threads_per_block": thr, "xcd_run": xrun, "fixed": True},
        "batch_setup": {"generation_s": round(t_gen, 1), "setup_s": round(t_setup, 1)},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_warmstart(args):
    """Learned warm start: train the encoder-decoder on steady states of
    --train 256^2 lattices of the batched distribution (from the batched
    solver), then solve --eval unseen 512^2 lattices from rest and from the
    network's prediction to the same tolerance; value = plain wall time /
    warm-started wall time of the batched solve."""
    from mantiis_amd import warmstart as WS
    from mantiis_amd.hip import globals_for
    F = 1e-6
    rho_in = globals_for(dx=1e-9)[1]["rhoin"]
    t0 = time.perf_counter()
    train_s = WS.samples(range(10000, 10000 + args.train), 256)
    print(f"[bench] {len(train_s)} training lattices generated", file=sys.stderr, flush=True)
    r, fields = WS.solve(train_s, F, graph=not args.eager)
    t_data = time.perf_counter() - t0
    print(f"[bench] training lattices solved to steady state ({t_data:.1f} s)", file=sys.stderr, flush=True)
    ds = WS.Dataset([WS.features(s, F) for s in train_s], [WS.targets(s, f, F) for s, f in zip(train_s, fields)])
    del fields
    c0, B = 24, 8
    t0 = time.perf_counter()
    net, hist = WS.train(ds, c0=c0, batch=B, epochs=args.epochs, lr=1e-3,
                         log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
    t_train = time.perf_counter() - t0
    steps_train = args.epochs * (ds.n // B)
    flops = net.flops(B, 256, 256)
    tflops = 3 * flops * steps_train / t_train / 1e12  # forward + two backward products per layer
    params = net.params()
    net.close()
    ds = None
    eval_s = WS.samples(range(20000, 20000 + args.eval), 512)
    t0 = time.perf_counter()
    print(f"[bench] {len(eval_s)} evaluation lattices generated", file=sys.stderr, flush=True)
    rc, fc = WS.solve(eval_s, F, graph=not args.eager)
    print("[bench] evaluation lattices solved from rest", file=sys.stderr, flush=True)
    t_cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    pred = WS.predict(params, c0, eval_s, F)
    t_pred = time.perf_counter() - t0
    t0 = time.perf_counter()
    t_seed = 0.0  # the populations are formed on the device inside the warm-started solve (mts_batch_seed)
    print("[bench] predicted", file=sys.stderr, flush=True)
    rw, fw = WS.solve(eval_s, F, warm_u=pred, rho_in=rho_in, graph=not args.eager)
    t_warm = time.perf_counter() - t0
    kc = np.array([WS.permeability(s, f[1], F) for s, f in zip(eval_s, fc)])
    kw = np.array([WS.permeability(s, f[1], F) for s, f in zip(eval_s, fw)])
    kp = np.array([WS.permeability(s, p[0], F) for s, p in zip(eval_s, pred)])
    warm_total = t_pred + t_seed + t_warm
    out = {
        "metric": "MLUPS & %HBM roofline; wall-time to steady permeability (warm-start vs plain)",
        "value": round(t_cold / warm_total, 3),
        "unit": "x (plain / warm-started wall time to steady state, batched)",
        "n_gpus": 1, "steps": int(rw["steps"].sum()), "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network, fp32 accumulation) / fp32 (LBM)",
        "data": "synthetic (seeded disk packings); training targets from the batched solver",
        "config": {"workload": f"learned warm start: train {args.train} x 256^2 lattices of the batched distribution "
                               f"({args.epochs} epochs, U-Net c0={c0}, {len(params)} parameters), infer {args.eval} "
                               "unseen 512^2 lattices, then LBM to |d<u_x>/dt|/|<u_x>| < 1e-6 per step",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": "k_gemm (BF16 MFMA implicit GEMM, training step)",
                     "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 5),
                     "traffic": None},
        "cpu_baseline": None,
        "warm_start": {"plain_s": round(t_cold, 3), "warm_s": round(warm_total, 3), "predict_s": round(t_pred, 3),
                       "seed_s": round(t_seed, 3), "warm_lbm_s": round(t_warm, 3),
                       "plain_steps": int(rc["steps"].sum()), "warm_steps": int(rw["steps"].sum()),
                       "k_error_network": {"median": float(np.median(np.abs(kp / kc - 1))),
                                           "max": float(np.abs(kp / kc - 1).max())},
                       "k_error_warm_vs_plain": {"median": float(np.median(np.abs(kw / kc - 1))),
                                                 "max": float(np.abs(kw / kc - 1).max())},
                       "training": {"data_s": round(t_data, 1), "train_s": round(t_train, 1),
                                    "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1]}},
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_shape(args):
    """Shape::calculateProperties (shape.hpp:127-251) on the device: LocPore
    and Kn of an n x n disk packing, domain and outputs resident in HBM.  A
    step is one full run (canvas, thinning to convergence, two distance
    transforms, nearest-skeleton search)."""
    from mantiis_amd import geometry as G
    from mantiis_amd.hip import DeviceArray, ShapeProperties
    from mantiis_amd.hiptime import StreamTimer
    n = args.n or 1024
    pore = G.disk_packing(n, 0.6, 3, 16, seed=31)
    mfp = 0.7730846760658538e-9
    h = ShapeProperties(n, n)
    d = DeviceArray.from_numpy(pore.astype(np.uint8))
    lp, kn = DeviceArray((n, n), np.float64), DeviceArray((n, n), np.float64)
    for _ in range(args.warmup):
        it = h.run(d, 1e-9, mfp, lp, kn)
    timer = StreamTimer(h.stream())
    t0 = time.perf_counter()
    timer.start()
    for _ in range(args.steps):
        it = h.run(d, 1e-9, mfp, lp, kn)
    ms = timer.stop()
    wall = time.perf_counter() - t0
    timer.close()
    pix = n * n
    value = pix * args.steps / (ms / 1e3) / 1e6
    bytes_per_pixel = 17  # domain byte in, LocPore and Kn (FP64) out
    achieved = bytes_per_pixel * pix * args.steps / (ms / 1e3) / 1e9
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        cn = args.cpu_n if args.cpu_n != 256 else n
        cpore = pore if cn == n else G.disk_packing(cn, 0.6, 3, 16, seed=31)
        t1 = time.perf_counter()
        O.shape_properties(cpore, 1e-9, mfp, fast=True)
        cs = time.perf_counter() - t1
        cpu = {"value": round(cn * cn / cs / 1e6, 4), "unit": "Mpix/s", "cores": 1, "kind": "port",
               "sample": f"one calculateProperties of a {cn}x{cn} disk packing by the oracle (-Ofast, 1 thread), "
                         f"{cs:.2f} s"}
    out = {
        "metric": "Mpix/s (calculateProperties)",
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 bit planes (thinning) / int32 squares (distances) / fp64 (LocPore, Kn)",
        "data": "synthetic (seeded disk packing)",
        "config": {"workload": f"Shape::calculateProperties of a {n}x{n} disk packing (porosity 0.6, radius "
                               f"U[3,16]); canvas {3 * n}x{3 * n}, {it} thinning iterations on the device",
                   "lattice": [n, n], "parallelism": "single GPU", "wall_ms_per_step": round(wall * 1e3 / args.steps, 4)},
        "roofline": {"bound": "hbm", "kernel": "whole run (k_thin is ~80% of it: a chain of 2 x iterations dependent "
                                               "subiterations, latency- not bandwidth-bound)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                     "bytes_per_pixel": bytes_per_pixel, "pixels_per_run": pix},
        "cpu_baseline": cpu,
    }
    h.close()
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--repeats", type=int, default=5, help="batched: timed windows of --steps steps (median reported)")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "quadtree", "shanchen", "warmstart", "transport", "fp64", "shape"])
    ap.add_argument("--no-fp64", action="store_true", help="batched: skip the FP64 permeability check")
    ap.add_argument("--batch-steps", type=int, default=1, choices=[1, 2],
                    help="batched: steps per launch (1: one-step cell-order kernel, 2: tile kernels)")
    ap.add_argument("--train", type=int, default=2000, help="warmstart: training lattices (256^2)")
    ap.add_argument("--eval", type=int, default=64, help="warmstart: unseen evaluation lattices (512^2)")
    ap.add_argument("--epochs", type=int, default=20, help="warmstart: training epochs")
    ap.add_argument("--batch", type=int, default=256, help="batched: number of lattices")
    ap.add_argument("--workers", type=int, default=8, help="batched/warmstart: processes generating the geometries")
    ap.add_argument("--eager", action="store_true",
                    help="batched / warmstart: launch the steps one by one instead of replaying a captured graph "
                         "per chunk (a rocprofv3 kernel trace of the graph path ends in a segmentation fault inside "
                         "the HIP runtime, below mts_batch_s::chunk_graph)")
    ap.add_argument("--no-steady", action="store_true", help="batched / quadtree: skip the run to steady state")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--launch", default="",
                    help="batched: THREADS,XCD_RUN -- a fixed launch shape instead of the autotuner's choice")
    ap.add_argument("--cpu-steps", type=int, default=0, help="CPU sample steps (default: quadtree 1000, shanchen 16)")
    ap.add_argument("--cpu-n", type=int, default=256, help="quadtree: lattice side of the CPU sample")
    ap.add_argument("--n", type=int, default=0, help="lattice side of the transport workload (default 1024)")
    ap.add_argument("--levels", type=int, default=3, choices=[2, 3],
                    help="quadtree: most levels of the quadtree (3: the benchmark configuration; 2: the reference's)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a plain 1 GiB device copy (torch) to calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py is a single-process, single-GPU benchmark.", file=sys.stderr)
        return 2
    if args.steps < 2:
        print("--steps must be at least 2", file=sys.stderr)
        return 2

    if args.copy_probe:
        # the library's 4-byte-per-lane copy: reads 1 GiB, writes 1 GiB per launch
        from mantiis_amd.hip import DeviceArray, load_library
        n = 1 << 28
        a = DeviceArray((n,))
        b = DeviceArray((n,))
        a.upload(np.ones(n, np.float32))
        lib = load_library()
        rc = lib.mts_probe_copy(b.ptr, a.ptr, n, 5)
        if rc != 0:
            print("probe copy failed:", lib.mts_last_error().decode(), file=sys.stderr)
            return 1
        print(json.dumps({"copy_probe_bytes": n * 4, "launches": 5}))
        return 0
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "warmstart":
        return bench_warmstart(args)
    if args.workload == "shape":
        return bench_shape(args)
    if args.workload == "shanchen":
        lat, inputs = make_shanchen()
        kind = "adsorption"
        bytes_per_cell = 88
        workload = ("Shan-Chen adsorption, 2048x2048 seeded disk packing porosity 0.5, "
                    f"G_ads={inputs['G_ads']:.4f} (U[-1.5,-0.5]), rho0={inputs['rho0']:.4f} (U[0.1,0.5]) +1% noise, "
                    "MRT + Guo source + bounce-back, FP32 deviation storage, FP64 psi")
        kernel = "k_csc (tiled Shan-Chen: forces, MRT + Guo source, bounce-back, generalized periodic)"
        dtype = "fp32 (populations) / fp64 (psi, moments)"
    elif args.workload == "quadtree":
        from mantiis_amd.hip import Lattice
        pore, kn, lp, phi = quadtree_inputs(args.n or 1024)
        lat = Lattice(pore, kn, lp, precision="fp32dev", collision=2, wall=2, dx=1e-9, Fbody=1e-6, periodic_y=1,
                      multiblock=args.levels)
        lvl = lat.grid()["level"]
        L = int(lvl.max())
        nl = [int((lvl == l).sum()) for l in range(1, L + 1)]
        kind, inputs = "transport", None
        # cell updates per coarse step: 2^(L - l) substeps of level l
        updates = quadtree_updates(lvl)
        bytes_per_cell = 72
        kernel_cells = nl[0] + nl[1] if L == 2 else updates  # two levels: one launch, first substep of both
        counts = ", ".join(f"{c} level-{l + 1}" for l, c in enumerate(nl))
        workload = (f"quadtree multi-block transport RMRT+SRBB, {lat.nx}x{lat.ny} disk packing porosity {phi:.3f} "
                    f"(U[0.4,0.8]), Kn=0.5, torus, FP32 deviation; at most {args.levels} levels: {counts} cells "
                    f"({1 << (L - 1)} level-1 substeps per coarse step)")
        if L == 2:
            kernel = ("k_cmbx (interface of the previous coarse step + first substep of both levels, RMRT+SRBB, "
                      "stored-field write)")
        else:
            kernel = ("coarse step: k_cmb per level substep (coarsest first) + k_ifx per pair of levels "
                      "(interface overrides)")
        dtype = "fp32 (populations)"
    elif args.workload == "transport":
        from mantiis_amd.hip import Lattice
        pore, kn, lp = transport_inputs(args.n or 1024)
        lat = Lattice(pore, kn, lp, precision="fp32dev", collision=2, wall=2, dx=1e-9, Fbody=1e-6, periodic_y=1)
        kind, bytes_per_cell, inputs = "transport", 72, None
        workload = (f"transport RMRT+SRBB, {lat.nx}x{lat.ny} disk packing (porosity U[0.4,0.8]), Kn=0.5, torus, "
                    "single-block, FP32 deviation")
        kernel = "k_ctr (fused gather + RMRT, SRBB walls)"
        dtype = "fp32 (populations)"
    else:
        from mantiis_amd import geometry as G
        from mantiis_amd.hip import Lattice
        pore = G.disk_packing(256, 0.6, 4, 12, seed=1)
        kn, lp = G.uniform_kn(pore, 0.1, 0.7730846760658538)
        lat = Lattice(pore, kn, lp, precision="fp64", collision=2, wall=2, dx=1e-9, Fbody=1e-6, periodic_y=1)
        kind, bytes_per_cell, inputs = "transport", 144, None
        workload = "transport RMRT+SRBB, 256x256 disk packing r~U[4,12] porosity 0.6, Kn=0.1, FP64"
        kernel = "k_transport<double,1,RMRT,SRBB>"
        dtype = "fp64"

    n = lat.ncell
    if args.workload != "quadtree":
        updates = kernel_cells = n
    ms, kms, klaunch = time_steps(lat, kind, args.steps, args.warmup)
    mlups = updates * args.steps / (ms / 1e3) / 1e6
    t_kernel_ms = kms / max(klaunch, 1)
    # (the three-level quadtree step times no single launch: klaunch 0)
    achieved = bytes_per_cell * kernel_cells / (t_kernel_ms / 1e3) / 1e9 if t_kernel_ms > 0 else 0.0
    step_achieved = bytes_per_cell * updates * args.steps / (ms / 1e3) / 1e9
    dominant = None
    if args.workload == "quadtree" and L > 2:
        # the step's launches are timed together: the roofline is taken over
        # the whole coarse step from the events around the K-step call
        t_kernel_ms = ms / args.steps
        achieved = step_achieved
    elif args.workload == "quadtree":
        # The quadtree step is two short launches (~16 and ~32 us), for which
        # per-launch events add several us; the roofline is taken over the
        # whole coarse step from the events around the K-step call, and the
        # dominant launch is reported beside it.
        dominant = {"kernel": kernel, "event_ms": round(t_kernel_ms, 6), "cells_per_launch": kernel_cells,
                    "achieved": round(achieved, 1)}
        kernel = "coarse step: k_cmbx (interface + first substeps) + k_cmb (second level-1 substep)"
        t_kernel_ms = ms / args.steps
        kernel_cells = updates
        achieved = step_achieved
    steady = None
    if args.workload == "quadtree" and not args.no_steady:
        from mantiis_amd.hip import Lattice
        lat2 = Lattice(pore, kn, lp, precision="fp32dev", collision=2, wall=2, dx=1e-9, Fbody=1e-6, periodic_y=1,
                       multiblock=args.levels)
        st, sec, crit, r = run_to_steady(lat2, 1e-6, 100, 400000, 120.0)
        steady = {"criterion": "|d<u_x>/dt| / |<u_x>| < 1e-6 per step, checked every 100 steps", "steps": st,
                  "wall_s": round(sec, 3), "final": crit, "reference_residual": r, "plain": True}
        # the FP64 library path for the same coarse steps: the steady
        # permeability's FP32 error
        lat3 = Lattice(pore, kn, lp, precision="fp64", collision=2, wall=2, dx=1e-9, Fbody=1e-6, periodic_y=1,
                       multiblock=args.levels)
        lat3.step("transport", st)
        k32, k64 = lat2.permeability()["k_lattice"], lat3.permeability()["k_lattice"]
        steady["fp32_vs_fp64"] = {"k_fp32": k32, "k_fp64": k64, "k_rel_err": abs(k32 - k64) / abs(k64),
                                  "ux_rel_l2": float(np.linalg.norm(lat2.get("ux") - lat3.get("ux")) /
                                                     np.linalg.norm(lat3.get("ux")))}
        lat3.close()
        lat2.close()
    diag = lat.diagnostics()

    cpu = None
    if not args.no_cpu and args.workload == "quadtree":
        cpu = cpu_baseline_quadtree(args.cpu_n, args.cpu_steps or (1000 if args.levels == 2 else 500), args.levels)
    elif not args.no_cpu and inputs is not None:
        cpu = cpu_baseline(inputs, args.cpu_steps or 16)

    out = {
        "metric": "MLUPS & %HBM roofline",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded disk packing)",
        "config": {"workload": workload, "lattice": [lat.nx, lat.ny], "fluid_cells": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(args.workload + ("3" if args.workload == "quadtree" and L > 2 else ""),
                                                 args.n),
                     "bytes_per_cell_update": bytes_per_cell, "cells_per_launch": kernel_cells,
                     "kernel_ms": round(t_kernel_ms, 6), "launches_timed": klaunch,
                     "step_achieved": round(step_achieved, 1)},
        "cpu_baseline": cpu,
        "mass": float(diag[0]),
    }
    copy = copy_peak_gbs()
    out["roofline"]["copy_peak"] = round(copy, 1)
    out["roofline"]["frac_of_copy"] = round(achieved / copy, 4)
    if args.workload == "quadtree":
        out["config"]["cell_updates_per_step"] = updates
        out["roofline"]["dominant_launch"] = dominant
    if steady is not None:
        out["steady"] = steady
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
