#!/usr/bin/env python3
"""Benchmark: STEM probe positions per second at 512^2 x 200 slices.

Workload (stem_scan_512): 512x512 complex64 grid over 8x8 perovskite cells
(31.24 A), 200 slices of two alternating types with 8 frozen-phonon variants
each, 64x64 probe positions over 2x2 cells, 8 passes per position (weight
1/8, variant sequences from one generator seeded 42), three annular
detectors (0-11, 11-22, 70-200 mrad).  One step = 512 positions (4096 probe
passes); the default 8 timed steps cover the full 4096-position scan.
All inputs (gratings, propagator, detectors, probe, positions, variant
sequences) are resident in HBM before the timed region.

Prints ONE JSON line.  The timed region replays each batch's captured HIP
graph (the batch's whole launch sequence; --no-graph: eager launches), with
HIP events around every replay.  A second pass over the same steps with eager
launches and HIP events around every kernel launch (on the engine's stream)
gives the per-kernel durations behind `roofline`: the dominant kernel's
algorithmic bytes per launch (16 n per wave: one read and one write of the
complex64 wave; 8 n for the initial column pass, which reads the shared
probe) over its average launch duration, against MI355X's 8 TB/s HBM peak.
Its `traffic` is the HBM bytes per launch measured with rocprofv3
FETCH_SIZE / WRITE_SIZE (scaled by a plain copy of known size) and committed
in profiles/hbm_traffic_stem_scan_512.md, per wave times the waves per
launch.  `fft_rate` is 10 n log2 n flops per probe pass and slice over the
timed region against the FP32 vector peak.  Outside the timed region,
`accuracy` compares a subset of the scan (8 positions x 8 passes) with the
CPU oracle's float32 run and its FP64 twin, detector by detector, and
`cpu_baseline` is that float32 oracle run (this project's restatement of the
reference CPUMultislice, one std::thread per core), with the CPU model and a
one-thread figure.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "STEM probe-positions/s at 512^2 x 200 slices; surrogate vs plain at <=1% error"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_stem_scan_512.md")


def measured_traffic_per_wave(kernel):
    """HBM bytes per wave per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement (copy-calibrated), or None."""
    try:
        txt = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_wave\[" + re.escape(kernel) + r"\]\s*=\s*([0-9.]+)", txt)
    return float(m.group(1)) if m else None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# raster indices of the accuracy subset of the 64x64 scan: the A column at
# the origin, points between columns, a B column, the far corner
ACC_IDX = np.array([0, 37, 500, 1056, 1500, 2080, 3000, 4095])


def accuracy_and_cpu_baseline(eng, sp, dx_all, dy_all, npass, cores):
    """Outside the timed region.  On the accuracy subset (8 positions x
    npass passes of the headline scan, own variant sequences from seed 42):
    the GPU result, the oracle's float32 run (timed: the CPU baseline, on
    `cores` threads) and its FP64 twin; per-detector relative L2 errors of
    GPU vs oracle f32, GPU vs FP64 and oracle f32 vs FP64.  Then one probe
    pass of the oracle on one thread (the single-core figure)."""
    from jmultislicelib_amd import configs
    from oracle import oracle as O
    dx, dy = dx_all[ACC_IDX], dy_all[ACC_IDX]
    nacc = len(ACC_IDX)
    w = np.full(npass, 1.0 / npass, np.float32)
    var = configs.variant_sequences(sp, nacc * npass, seed=42)
    eng.set_graph(False)
    eng.set_timing(False)
    eng.stage_scan(dx, dy, None, w, var)
    eng.run_staged(0, nacc)
    eng.sync()
    g = eng.scan_results()[:, 0, :].astype(np.float64)
    t0 = time.perf_counter()
    r32 = O.stem_scan(sp, dx, dy, w, var, prec=32, nthreads=cores)[:, 0, :].astype(np.float64)
    dt = time.perf_counter() - t0
    r64 = O.stem_scan(sp, dx, dy, w, var, prec=64, nthreads=cores)[:, 0, :].astype(np.float64)
    t1 = time.perf_counter()
    O.stem_scan(sp, dx[:1], dy[:1], [1.0], var[:1], prec=32, nthreads=1)
    dt1 = time.perf_counter() - t1

    def rel(a, b):
        return [float(np.linalg.norm(a[:, d] - b[:, d]) / np.linalg.norm(b[:, d])) for d in range(a.shape[1])]

    names = ["BF 0-11 mrad", "ABF 11-22 mrad", "HAADF 70-200 mrad"]
    acc = {"subset": f"{nacc} positions (raster indices {ACC_IDX.tolist()}) x {npass} passes, exit plane",
           "detectors": names,
           "gpu_vs_oracle_f32": [float(f"{e:.3e}") for e in rel(g, r32)],
           "gpu_vs_fp64": [float(f"{e:.3e}") for e in rel(g, r64)],
           "oracle_f32_vs_fp64": [float(f"{e:.3e}") for e in rel(r32, r64)],
           "tolerance": 1e-4}
    cpu = {"value": round(nacc / dt, 4), "unit": "positions/s", "cores": cores, "kind": "port",
           "sample": f"oracle float32 scan of the accuracy subset: {nacc} positions x {npass} passes "
                     f"(512^2, 200 slices, 3 detectors) in {dt:.1f} s on {cores} threads",
           "cpu_model": cpu_model(),
           "one_thread": {"value": round(1.0 / (dt1 * npass), 5), "unit": "positions/s",
                          "sample": f"1 probe pass in {dt1:.2f} s on 1 thread; positions/s = passes/s / {npass}"}}
    return acc, cpu


def run_thick2048(args):
    """Large thick specimen: 2048^2 over 125 A, 1000 distinct 2 A amorphous
    slices (0.06 atoms/A^3, Z 6/8/14), gratings generated on the device and
    resident in HBM (~33.5 GB), batches of 32 probe positions (25 mrad) plus
    one plane wave, detectors 0-11, 11-22, 70-200 mrad."""
    from jmultislicelib_amd import Engine, configs
    t0 = time.perf_counter()
    sp = configs.thick_amorphous()
    t1 = time.perf_counter()
    eng = Engine().setup(sp)
    eng.sync()
    t2 = time.perf_counter()
    log(f"setup: host {t1 - t0:.1f} s, device gratings + upload {t2 - t1:.1f} s "
        f"({len(sp.objslc)} slices of {sp.nx}x{sp.ny})")
    npos = 32
    g = np.arange(npos)
    dx = ((g % 8) * 0.15).astype(np.float32)  # nm
    dy = ((g // 8) * 0.15).astype(np.float32)
    var = np.zeros((npos * args.steps, len(sp.objslc)), np.int32)
    eng.set_batch(npos)
    eng.stage_scan(np.tile(dx, args.steps), np.tile(dy, args.steps), None, np.ones(1, np.float32), var)
    eng.sync()
    # warm-up on the last 20 slices only (same kernels, short)
    eng.run_staged(0, npos, len(sp.objslc) - 20)
    eng.sync()
    eng.set_timing(True)
    t3 = time.perf_counter()
    for s_ in range(args.steps):
        eng.run_staged(s_ * npos, npos)
    eng.sync()
    wall = time.perf_counter() - t3
    kinds = {0: "row_pass", 1: "col_pass", 2: "combine"}
    tm = {k: eng.timing(k) for k in kinds}
    for k, (ms, cnt) in tm.items():
        log(f"kernel {kinds[k]}: {cnt} launches, {ms:.1f} ms total, {ms / max(cnt, 1):.3f} ms avg")
    n = sp.nx * sp.ny
    positions = args.steps * npos
    slice_bytes = 32.0 * n * positions * len(sp.objslc)
    # the plane wave of each batch (single-wave path)
    pw = np.zeros((sp.ny, sp.nx), np.complex64)
    pw[0, 0] = 1.0
    eng.set_timing(False)
    t4 = time.perf_counter()
    eng.wave_set(pw)
    eng.multislice(np.zeros(len(sp.objslc), np.int32))
    pw_s = time.perf_counter() - t4
    res = eng.scan_results()
    log(f"scan: {positions} positions in {wall:.2f} s -> {positions / wall:.2f} positions/s; "
        f"slice-loop HBM rate {slice_bytes / wall / 1e9:.0f} GB/s ({slice_bytes / wall / 1e9 / HBM_PEAK_GBS:.1%}); "
        f"plane wave {pw_s:.2f} s; finite {bool(np.isfinite(res).all())}")
    dom = max((0, 1), key=lambda k: tm[k][0])
    ms, cnt = tm[dom]
    waves_per_launch = npos
    achieved = 16.0 * n * waves_per_launch / (ms / cnt * 1e-3) / 1e9
    out = {"metric": "STEM probe-positions/s at 2048^2 x 1000 distinct slices (large thick specimen)",
           "value": round(positions / wall, 3), "unit": "positions/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 0, "ms_per_step": round(wall / args.steps * 1e3, 1), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "thick2048: 2048x2048 complex64 over 125 A, 1000 distinct 2 A amorphous "
                                  "slices resident in HBM, 32 probes per batch + 1 plane wave",
                      "grid": [sp.nx, sp.ny], "slices": len(sp.objslc), "positions_per_step": npos,
                      "plane_wave_s": round(pw_s, 3)},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kinds[dom],
                        "avg_launch_ms": round(ms / cnt, 4)},
           "cpu_baseline": None}
    eng.close()
    print(json.dumps(out), flush=True)


def run_jobs64(args):
    """Uneven batch: 64 random orthorhombic crystal jobs (seed 2), grid edge
    from {360, 384, 480, 512, 640, 720} at 0.06 A, 50-400 slices of c/2,
    16-256 probe positions each (one pass), 20-30 mrad apertures, detectors
    0-11, 11-22, 70-200 mrad.  Every job has its own engine.  By default the
    jobs of each grid edge form a group (jms_hip_group): one row and one
    column launch per slice step over the waves of all its jobs, the six
    groups on six streams.  --streams: each job's own launch sequence on its
    own stream, all enqueued before the first wait.  One step = the whole
    batch."""
    from jmultislicelib_amd import Engine, Group, configs
    t0 = time.perf_counter()
    jobs = configs.uneven_jobs(64, seed=2)
    engines = []
    for j in jobs:
        eng = Engine().setup(j.setup)
        eng.set_batch(len(j.dx))
        eng.set_graph(args.graph)
        var = np.zeros((len(j.dx), len(j.setup.objslc)), np.int32)
        eng.stage_scan(j.dx, j.dy, None, np.ones(1, np.float32), var)
        engines.append(eng)
    for eng in engines:
        eng.sync()
    grouped = not args.streams
    groups = []
    if grouped:  # one group per grid edge: one launch per pass and step over all its jobs
        for e in sorted(set(j.setup.nx for j in jobs)):
            groups.append(Group([eng for eng, j in zip(engines, jobs) if j.setup.nx == e]))
    positions = sum(len(j.dx) for j in jobs)
    work = sum(len(j.dx) * len(j.setup.objslc) * 32.0 * j.setup.nx * j.setup.ny for j in jobs)
    log(f"setup {time.perf_counter() - t0:.1f} s: {len(jobs)} jobs, {positions} positions, "
        f"edges {sorted(set(j.setup.nx for j in jobs))}, slices {min(len(j.setup.objslc) for j in jobs)}-"
        f"{max(len(j.setup.objslc) for j in jobs)}")
    def step():
        if grouped:
            for g in groups:  # each group on its first member's stream: the edges run concurrently
                g.run()
        else:
            for eng, j in zip(engines, jobs):
                eng.run_staged(0, len(j.dx))
        for eng in engines:
            eng.sync()

    for _ in range(args.warmup):
        step()
    times = []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t1)
    wall = sum(times)
    finite = all(np.isfinite(eng.scan_results()).all() for eng in engines)
    rate = work * args.steps / wall / 1e9
    log(f"jobs64: {positions * args.steps} positions in {wall:.2f} s -> {positions * args.steps / wall:.1f} "
        f"positions/s; slice-loop HBM rate {rate:.0f} GB/s ({rate / HBM_PEAK_GBS:.1%}); finite {finite}")
    out = {"metric": "STEM probe-positions/s, uneven batch of 64 random-crystal jobs (360-720 point grids, "
                     "50-400 slices)",
           "value": round(positions * args.steps / wall, 2), "unit": "positions/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 1),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "jobs64: 64 random orthorhombic crystal jobs, seed 2, " +
                                  ("grouped launches per grid edge (one row and one column launch per slice "
                                   "step over all jobs of the edge), the six edges on six streams" if grouped else
                                   "one engine and stream per job, all enqueued before the first wait"),
                      "jobs": len(jobs), "positions_per_step": positions, "grouped": grouped},
           "roofline": {"bound": "hbm", "achieved": round(rate, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(rate / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "whole slice loop over all jobs (32 n bytes per position and slice)"},
           "cpu_baseline": None}
    for g in groups:
        g.close()
    for eng in engines:
        eng.close()
    print(json.dumps(out), flush=True)


def run_surrogate(args):
    """Learned surrogate with multislice fallback.  Training data: static
    random perovskite-like lattices (a in [3.6, 4.2] A, A-site Z in [20, 82],
    B-site Z in [22, 74], O), 512^2 over 8x8 cells, 200 slices, 25 mrad probe,
    `--train-positions` random positions per lattice, one multislice pass per
    position (no frozen phonons, so the plain reference is deterministic).
    An ensemble of 4 networks is trained by the library's kernels, by
    default on the 8 images of every training position under the lattices'
    point group 4mm (the intensities are equal there), and by default each
    prediction is averaged over the 8 images of its position.  The
    evaluation is a scan x scan raster over one cell of an unseen lattice:
    the plain batched multislice of every position against the surrogate
    scan (device windows, ensemble inference, 5% random audit plus every
    position whose ensemble spread exceeds 2% on ABF or HAADF re-simulated by
    the plain multislice; with --calib K the re-simulated positions also
    calibrate the other predictions, inside the timed scan)."""
    from jmultislicelib_amd import Engine
    from jmultislicelib_amd import surrogate as S
    rng = np.random.default_rng(2024)
    train = [S.random_lattice(rng) for _ in range(args.lattices)]
    test_lat = S.random_lattice(np.random.default_rng(777))
    eng = Engine()
    t0 = time.perf_counter()
    smp = S.simulate(train, args.train_positions, seed=5, eng=eng)
    t1 = time.perf_counter()
    log(f"training data: {len(smp.t)} positions on {len(train)} lattices in {t1 - t0:.1f} s")
    ens = S.Ensemble(nmembers=4, max_batch=2048, seed=1, step=args.window_step)  # batches of 2048 at inference
    for net in ens.nets:
        net.set_kernels(not args.fma)
    hist = ens.train(smp, epochs=args.epochs, batch=256, lr=args.lr, log=log, augment=args.augment)
    t2 = time.perf_counter()
    log(f"training: {args.epochs} epochs x 4 members in {t2 - t1:.1f} s; final losses {hist[-1]}")
    # evaluation raster
    ns = args.scan
    g = (np.arange(ns) + 0.5) * test_lat.a / ns
    py, px = [a.ravel().astype(np.float32) for a in np.meshgrid(g, g, indexing="ij")]
    eng.setup(S.lattice_setup(test_lat))
    eng.set_batch(2048)
    S.plain_scan(eng, px[:64], py[:64])  # warm-up
    eng.set_batch(2048)
    t3 = time.perf_counter()
    plain = S.plain_scan(eng, px, py)
    t4 = time.perf_counter()
    if args.sym:
        x8, y8 = S.symmetric_images(px, py, test_lat.a)
        ws = S.Samples([test_lat], np.zeros(8 * len(px), np.int32), x8.ravel(), y8.ravel(),
                       np.tile(plain, (8, 1))).windows(step=ens.step)
        net_only, spread = ens.predict(ws, 8)
    else:
        ws = S.Samples([test_lat], np.zeros(len(px), np.int32), px, py, plain).windows(step=ens.step)
        net_only, spread = ens.predict(ws)
    ws.close()
    eng.set_batch(2048)
    out, idx, tm = S.surrogate_scan(ens, eng, test_lat, px, py, audit=0.05, spread_max=0.02, seed=9, sym=args.sym,
                                    calib=args.calib)
    t_plain, t_surr = t4 - t3, tm["total_s"]
    raw = tm.pop("uncalibrated")
    rel = lambda p, d: float(np.mean(np.abs(p[:, d] - plain[:, d]) / np.abs(plain[:, d])))
    pred = np.setdiff1d(np.arange(len(px)), idx)  # positions the scan did not re-simulate
    rel_p = lambda p, d: float(np.mean(np.abs(p[pred, d] - plain[pred, d]) / np.abs(plain[pred, d])))
    errs = {S.DET_NAMES[d]: {"surrogate_scan": round(rel(out, d), 5), "network_only": round(rel(net_only, d), 5),
                             "fallback_without_calibration": round(rel(raw, d), 5),
                             "predicted_positions_only": round(rel_p(out, d), 5)}
            for d in range(3)}
    log(f"plain {t_plain:.2f} s, surrogate {t_surr:.2f} s ({tm}); re-simulated {len(idx)} of {len(px)}; "
        f"errors {errs}")
    out_json = {"metric": "surrogate+fallback vs plain batched multislice: scan time ratio (512^2 x 200 slices, "
                          "mean relative HAADF/ABF error stated)",
                "value": round(t_plain / t_surr, 2), "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 1,
                "ms_per_step": round(t_surr * 1e3, 1), "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "bf16 (network, FP32 accumulation); f32 (multislice)",
                "data": "synthetic",
                "config": {"workload": f"surrogate: {len(train)} training lattices x {args.train_positions} "
                                       f"positions, unseen lattice {ns}x{ns} raster over one cell",
                           "positions": len(px), "plain_s": round(t_plain, 3), "surrogate_s": round(t_surr, 3),
                           "resimulated": int(len(idx)), "audit": 0.05, "spread_max": 0.02,
                           "window": {"pixels": [S.WIN, S.WIN], "step_ang": ens.step,
                                      "configured": abs(ens.step - 0.1) < 1e-9},
                           "errors": errs, "train_data_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
                           "epochs": args.epochs, "augment_4mm": bool(args.augment), "predict_4mm": bool(args.sym),
                           "calibration_kmax": args.calib,
                           "conv_kernels": "fma" if args.fma else "mfma",
                           "timings": {k: round(v, 3) for k, v in tm.items()},
                           "test_lattice": {"a": test_lat.a, "za": test_lat.za, "zb": test_lat.zb}},
                "roofline": None, "cpu_baseline": None}
    ens.close()
    eng.close()
    print(json.dumps(out_json), flush=True)


def run_stem512(args):
    """The headline scan (config: 512^2 x 200 slices, 8 frozen-phonon passes
    per position).  Timed region: `steps` steps of positions_per_step
    positions, each batch replayed from its captured HIP graph (--no-graph:
    eager launches), HIP events around every replay.  Then, untimed for
    `value`, the same steps with eager launches and HIP events around every
    kernel launch: the per-kernel durations behind `roofline`."""
    from jmultislicelib_amd import Engine, configs

    npass = 8
    graph = not args.no_graph
    t0 = time.perf_counter()
    sp = configs.stem_scan_512()
    dx_all, dy_all = configs.scan_positions(64, 2)
    npos_total = len(dx_all)
    pps = args.positions_per_step
    nsteps_needed = args.warmup + args.steps
    # position p of step s: the scan is walked cyclically
    order = np.arange(nsteps_needed * pps) % npos_total
    dx, dy = dx_all[order], dy_all[order]
    var = configs.variant_sequences(sp, len(order) * npass, seed=42)
    log(f"setup {time.perf_counter() - t0:.1f} s: {sp.nx}x{sp.ny}, {len(sp.objslc)} slices, "
        f"{len(order)} positions x {npass} passes")

    eng = Engine().setup(sp)
    eng.set_batch(args.batch)
    weights = np.full(npass, 1.0 / npass, np.float32)
    eng.stage_scan(dx, dy, None, weights, var)
    eng.sync()

    def timed_steps(use_graph):
        eng.set_graph(use_graph)
        eng.set_timing(False)
        for s_ in range(args.warmup):  # captures the batch graphs when use_graph
            eng.run_staged(s_ * pps, pps)
        eng.sync()
        eng.set_timing(True)
        t1 = time.perf_counter()
        for s_ in range(args.warmup, nsteps_needed):
            eng.run_staged(s_ * pps, pps)
        eng.sync()
        return time.perf_counter() - t1

    wall = timed_steps(graph)
    res = eng.scan_results()
    finite = bool(np.isfinite(res).all())
    n = sp.nx * sp.ny
    S = len(sp.objslc)
    positions = args.steps * pps
    value = positions / wall
    replay = None
    if graph:
        ms, cnt = eng.timing(3)
        replay = {"replays": cnt, "avg_replay_ms": round(ms / max(cnt, 1), 3),
                  "waves_per_replay": args.batch, "gpu_ms_total": round(ms, 1)}
        log(f"graph replays: {cnt}, {ms:.1f} ms total, {ms / max(cnt, 1):.2f} ms avg")
        wall_eager = timed_steps(False)
        log(f"eager launches (kernel timing pass): {positions / wall_eager:.1f} positions/s")
    else:
        wall_eager = wall
    kinds = {0: "row_pass", 1: "col_pass", 2: "combine"}
    tm = {k: eng.timing(k) for k in kinds}
    for k, (ms, cnt) in tm.items():
        log(f"kernel {kinds[k]}: {cnt} launches, {ms:.1f} ms total, {ms / max(cnt, 1):.3f} ms avg")
    dom = max((0, 1), key=lambda k: tm[k][0])
    ms, cnt = tm[dom]
    passes = positions * npass
    if dom == 0:  # S row passes per wave, 16 n bytes each
        total_bytes = 16.0 * n * passes * S
    else:  # S column passes of 16 n bytes plus the initial one: probe read shared, 8 n written
        total_bytes = (16.0 * n * S + 8.0 * n) * passes
    bytes_per_launch = total_bytes / max(cnt, 1)
    achieved = bytes_per_launch / (ms / cnt * 1e-3) / 1e9
    nwave_avg = passes * (S + (1 if dom == 1 else 0)) / max(cnt, 1)
    tpw = measured_traffic_per_wave(kinds[dom])
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if tpw is None else round(tpw * nwave_avg),
            "traffic_source": None if tpw is None else os.path.relpath(TRAFFIC_PROFILE, ROOT),
            "kernel": kinds[dom], "bytes_per_launch": round(bytes_per_launch),
            "avg_launch_ms": round(ms / cnt, 4),
            "timing": "HIP events on the engine stream around every eager launch"}
    total_kernel_ms = sum(v[0] for v in tm.values())
    log(f"kernel time {total_kernel_ms / 1e3:.2f} s of {wall_eager:.2f} s wall (eager pass)")
    # whole-slice view: 32 n bytes per probe pass per slice
    slice_bytes = 32.0 * n * passes * S
    flops = 10.0 * n * np.log2(n) * passes * S  # two 2-D FFTs of 5 n log2 n per probe pass and slice
    log(f"scan: {positions} positions in {wall:.2f} s -> {value:.1f} positions/s; "
        f"slice-loop HBM rate {slice_bytes / wall / 1e9:.0f} GB/s ({slice_bytes / wall / 1e9 / HBM_PEAK_GBS:.1%} of peak); "
        f"FFT rate {flops / wall / 1e12:.1f} TFLOP/s")
    log(f"results finite: {finite}; first position detectors {res[0, 0]}")

    acc = cpu = None
    if not args.no_cpu_baseline:
        try:
            cores = min(16, len(os.sched_getaffinity(0)))
            acc, cpu = accuracy_and_cpu_baseline(eng, sp, dx_all, dy_all, npass, cores)
            log(f"accuracy: {acc}")
            log(f"cpu baseline: {cpu['value']:.3f} positions/s ({cpu['sample']}); 1 thread "
                f"{cpu['one_thread']['value']:.4f} positions/s; {cpu['cpu_model']}")
        except Exception as ex:  # the baseline must not hide the GPU result
            log(f"accuracy / cpu baseline failed: {ex}")
    eng.close()

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "positions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "stem_scan_512: 512x512 complex64, 200 slices (2 types x 8 frozen-phonon "
                               "variants), 64x64 positions x 8 passes, 3 annular detectors",
                   "grid": [sp.nx, sp.ny], "slices": int(S), "positions_per_step": pps,
                   "passes": npass, "batch_waves": args.batch, "graph": graph,
                   "graph_replays": replay,
                   "eager_positions_per_s": round(positions / wall_eager, 2),
                   "slice_loop_hbm_frac": round(slice_bytes / wall / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roof,
        "fft_rate": {"achieved": round(flops / wall / 1e12, 2), "peak": 157.3, "unit": "TFLOP/s",
                     "frac": round(flops / wall / 1e12 / 157.3, 4),
                     "flops_per_probe_pass_slice": int(10 * n * np.log2(n)),
                     "note": "10 n log2 n per probe pass and slice over the timed region; FP32 vector peak"},
        "accuracy": acc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--positions-per-step", type=int, default=512)
    ap.add_argument("--batch", type=int, default=2048, help="probe passes resident per batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", action="store_true", help="jobs64 with --streams: replay batches from HIP graphs")
    ap.add_argument("--streams", action="store_true",
                    help="jobs64: one engine launch sequence per job on its own stream instead of grouped launches")
    ap.add_argument("--no-graph", action="store_true",
                    help="stem512: time eager launches instead of HIP graph replays")
    ap.add_argument("--workload", choices=["stem512", "thick2048", "jobs64", "surrogate"], default="stem512",
                    help="stem512: the headline scan (default); thick2048: 2048^2 x 1000 distinct slices; "
                         "jobs64: uneven batch of 64 random-crystal jobs (edges 360-720); surrogate: "
                         "train the detector-intensity surrogate and time surrogate+fallback vs plain")
    ap.add_argument("--lattices", type=int, default=500, help="surrogate: training lattices")
    ap.add_argument("--train-positions", type=int, default=400, help="surrogate: positions per lattice")
    ap.add_argument("--epochs", type=int, default=40, help="surrogate: training epochs")
    ap.add_argument("--scan", type=int, default=256, help="surrogate: evaluation scan is scan x scan")
    ap.add_argument("--lr", type=float, default=1e-3, help="surrogate: initial learning rate")
    ap.add_argument("--no-augment", dest="augment", action="store_false",
                    help="surrogate: train on the positions as drawn (default: also on their 8 symmetry (4mm) images)")
    ap.add_argument("--no-sym", dest="sym", action="store_false",
                    help="surrogate: one prediction per position (default: averaged over its 8 symmetry images)")
    ap.add_argument("--fma", action="store_true", help="surrogate: FMA convolution kernels instead of MFMA")
    ap.add_argument("--calib", type=int, default=6, metavar="K",
                    help="surrogate: calibrate the predictions with the re-simulated positions (periodic 4mm "
                         "basis up to order K; 0 = off)")
    ap.add_argument("--window-step", type=float, default=0.1,
                    help="surrogate: window pixel spacing [A] (0.1 A: 32 x 32 pixels over 3.2 A, the configured input)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference "
                         "kernel for scaling HBM counter readings) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        nel = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 copies of {nel * 4} bytes (read {nel * 4} + write {nel * 4} per copy)")
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(positions shard independently across GPUs)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "thick2048":
        run_thick2048(args)
        return
    if args.workload == "jobs64":
        run_jobs64(args)
        return
    if args.workload == "surrogate":
        run_surrogate(args)
        return

    run_stem512(args)


if __name__ == "__main__":
    main()
