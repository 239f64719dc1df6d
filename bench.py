#!/usr/bin/env python3
"""Benchmark: STEM probe.slices/s at 512^2, batched vs plain, on one MI355X.

Workload (the batched HAADF STEM configuration): SrTiO3 [001] tiled 16x16x40
(51,200 atoms), 512^2 complex64 waves, 200 kV, 20 mrad aberration-free probe,
64x64 probe raster over one unit cell, 60-200 mrad annular detector, frozen
phonons with u_rms 0.08 A (seeded Philox displacements), projected Kirkland
potential, 83 slices of 1.9525 A.

A step is one frozen-phonon configuration of the whole job, with the displaced
coordinates already generated: GPU atom binning, the 83 band-limited transmission
functions, and all 4096 probes through 83 slices in batches of B waves with the
detector reduced in FP64 and accumulated across configurations.  value is the
whole-job probe.slices/s over the timed steps.

Also reported: the slice-loop rate for B in --sweep, the plain form (one probe per
launch sequence with a host sync after every slice, as clTEM's doMultiSliceStep
does), the dominant kernel's roofline from hipEvent timings, and the CPU baseline
(the FP32 C++ restatement timed on the host cores) with the GPU-vs-FP64 detector
error on sampled probes.

--workload sio2 measures the full-3D frozen-phonon CTEM case instead (amorphous SiO2,
20,000 atoms, 1024^2, 106 slices, 300 kV, Lobato with 20 integrals, 30 configurations
batched through one graph-captured slice loop): value is configuration.slices/s.
--workload mixed measures the mixed-size batched CBED case (48 SrTiO3 instances on
256^2..768^2 grids, 5 configurations each, the slice loops of all sizes in one graph
launch): value is wave.slices/s of the whole job (binning, potentials, probes, slice loops,
FP64 sums).
--workload potentials measures the potential builders alone (projected Kirkland at 512^2,
full-3D Lobato at 1024^2): value is atom.pixel evaluations/s of the projected kernels.
--workload adjoint measures position gradients of the Au/C adjoint case (2000 atoms,
256^2, 46 slices): one step rebuilds the transmission functions from the positions and
runs the checkpointed slice loop and the adjoint slice loop; value is gradients/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import cltem_amd  # noqa: E402
from cltem_amd import workloads as wl  # noqa: E402

METRIC = "STEM probe·slices/s at 512², batched vs plain, rel-L2≤1e-4 vs FP64"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def setup(engine, w):
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    engine.set_geometry(g)
    p = cltem_amd.load_params(w["params"])
    engine.set_parameterisation(p)
    return g, p


def build_config(engine, c, spec):
    engine.sort_atoms(c, spec.x, spec.y, spec.z, spec.Z, want_ids=False)
    engine.build_transmission(c)


def plain_rate(engine, g, pos, probe, dets):
    """One probe per launch sequence, host sync after every slice (clTEM's structure)."""
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    ns = g.n_slices
    out = []
    engine.synchronize()
    t0 = time.perf_counter()
    for p in pos:
        engine.init_probes(p[None, :], probe)
        for s in range(ns):
            engine.multislice(s, 1)
            engine.synchronize()
        out.append(engine.stem_detect(dets)[0])
    engine.synchronize()
    dt = time.perf_counter() - t0
    return len(pos) * ns / dt, np.array(out)


PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "slice_sweeps_512_B64.md")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (calibrated against a copy kernel of known size), or None."""
    try:
        text = open(PROFILE_SUMMARY).read()
    except OSError:
        return None
    for line in text.splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            for item in line.split(":", 1)[1].split():
                k, v = item.split("=")
                if k == kernel:
                    return int(v)
    return None


def roofline(engine, g, pos, batch, probe, dets):
    """Dominant sweep's achieved algorithmic bandwidth: average launch durations of the row
    and column passes timed back to back with hipEvents on the engine stream (a batch of B
    probes of one configuration, current kernel variants)."""
    N = g.resolution * g.resolution
    engine.stem_batch(0, pos[:batch], batch, probe, dets)  # B waves of configuration 0
    row_us, col_us = engine.time_sweeps(g.n_slices // 2, reps=20)
    # algorithmic bytes per launch (8 bytes per complex64):
    #   row pass: read + write every wave (16N each) and the shared t slice once (8N)
    #   column pass: read + write every wave
    row_bytes = batch * 16 * N + 8 * N
    col_bytes = batch * 16 * N
    kern = {"row_pass": (row_us, row_bytes), "column_pass": (col_us, col_bytes)}
    name = max(kern, key=lambda k: kern[k][0])
    us, b = kern[name]
    achieved = b / (us * 1e-6) / 1e9
    # the same access pattern with no arithmetic: what this part's memory system delivers to it
    copy_us = engine.bandwidth_probe(1 if name == "column_pass" else 0, batch, reps=10)
    return {
        "kernel": f"{name} (512^2, B={batch})",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": profiled_traffic(name) if (batch == 64 and g.resolution == 512) else None,
        "traffic_source": "profiles/slice_sweeps_512_B64.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                          "calibrated with a copy kernel)",
        "avg_launch_us": round(us, 2),
        "bytes_per_launch": int(b),
        "row_pass_avg_us": round(row_us, 2),
        "column_pass_avg_us": round(col_us, 2),
        "row_pass_GBs": round(row_bytes / (row_us * 1e-6) / 1e9, 1),
        "column_pass_GBs": round(col_bytes / (col_us * 1e-6) / 1e9, 1),
        "copy_pattern_avg_us": round(copy_us, 2),
        "copy_pattern_GBs": round(batch * 16 * N / (copy_us * 1e-6) / 1e9, 1),
        "frac_of_copy_pattern": round(copy_us / us, 4),
        "timing": "hipEvents around 20 back-to-back launches of each sweep on the engine stream; the copy "
                  "pattern is the same kernel's tile loads and stores without arithmetic (10 launches)",
    }


def host_cores():
    """(threads used, CPUs visible).  The CPU baseline runs on every CPU the process may use,
    up to the 16-core share that one GPU's job is given on the GPU hosts (more threads than
    that share contend for the same cores)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, visible)), visible


def cpu_baseline(w, g, spec0, sample, check_pos, gpu_det):
    """FP32 C++ restatement timed on the host cores over the probes `sample`, and the FP64
    oracle as checker of the GPU detector values `gpu_det` at `check_pos`."""
    from oracle import oracle_py as orc
    cores, visible = host_cores()
    og = orc.geometry(orc.scenario(**w["scenario"]))
    p = cltem_amd.load_params(w["params"])
    t0 = time.perf_counter()
    srt = orc.sort_atoms(og, spec0.x, spec0.y, spec0.z, spec0.Z)
    trans = orc.transmission(og, srt, p.values, p.form, p.i_per_atom, nthreads=cores)
    t_pot = time.perf_counter() - t0
    P = orc.propagator(og)
    det4 = np.array([list(d) + [0.0] * (4 - len(d)) for d in w["detectors"]])
    t0 = time.perf_counter()
    orc.run_stem(og, trans, P, sample, det4, cond_ap_mrad=w["cond_ap_mrad"], nthreads=cores, fp32=True)
    t_f32 = time.perf_counter() - t0
    ref = orc.run_stem(og, trans, P, check_pos, det4, cond_ap_mrad=w["cond_ap_mrad"], nthreads=cores)
    err = float(np.max(np.abs(gpu_det - ref) / np.abs(ref)))
    return {
        "value": round(len(sample) * og.n_slices / t_f32, 2),
        "unit": "probe·slices/s",
        "cores": cores,
        "host_cpus_visible": visible,
        "kind": "port",
        "sample": f"{len(sample)} probes x {og.n_slices} slices of phonon configuration 0 (FP32 C++ "
                  f"restatement with an iterative radix-4 Stockham FFT, std::thread over probes on "
                  f"{cores} threads); potential build of {og.n_slices} slices took {t_pot:.1f} s in FP64",
    }, err, len(check_pos)


def kernel_roofline(engine, g, n_waves, label):
    """Row/column pass launch durations over all waves (each wave reading its own
    transmission function), timed back to back with hipEvents, as achieved algorithmic GB/s."""
    N = g.resolution * g.resolution
    row_us, col_us = engine.time_sweeps(g.n_slices // 2, reps=10)
    row_bytes = n_waves * 24 * N   # read + write the wave, read its own t
    col_bytes = n_waves * 16 * N
    kern = {"row_pass": (row_us, row_bytes), "column_pass": (col_us, col_bytes)}
    name = max(kern, key=lambda k: kern[k][0])
    us, b = kern[name]
    achieved = b / (us * 1e-6) / 1e9
    return {
        "kernel": f"{name} ({label})", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "avg_launch_us": round(us, 2), "bytes_per_launch": int(b),
        "row_pass_avg_us": round(row_us, 2), "column_pass_avg_us": round(col_us, 2),
        "row_pass_GBs": round(row_bytes / (row_us * 1e-6) / 1e9, 1),
        "column_pass_GBs": round(col_bytes / (col_us * 1e-6) / 1e9, 1),
        "timing": "hipEvents around 10 back-to-back launches of each sweep on the engine stream",
    }


def bench_sio2(args):
    """A step is the whole job: GPU binning and full-3D potentials of every configuration
    (displaced coordinates generated beforehand, as for the STEM workload), the batched
    graph-captured slice loop of all configurations and the FP64 phonon-averaged intensity."""
    w = wl.ctem_amorphous_sio2(n_configs=args.configs)
    spec = w["specimen"]
    n_cfg = w["n_configs"]
    engine = cltem_amd.Engine(0)
    dev, cus, _ = engine.device_info()
    g, p = setup(engine, w)
    N = g.resolution * g.resolution
    ns = g.n_slices
    log(f"{dev}: {len(spec)} atoms, {g.resolution}^2, {ns} slices, {n_cfg} configurations, full 3D")
    displaced = [spec.displaced(w["u_rms"], w["seed"], c) for c in range(n_cfg)]

    def build_all():
        for c in range(n_cfg):
            build_config(engine, c, displaced[c])

    def loop():
        engine.init_plane_wave(1.0)
        engine.multislice(graph=True)
        return engine.exit_intensity_sum()

    build_all()
    tune = engine.autotune(0, n_cfg, reps=2)
    engine.set_wave_count(n_cfg)
    engine.set_wave_configs(list(range(n_cfg)))
    mean = None
    for _ in range(args.warmup):
        build_all()
        mean = loop()
    engine.synchronize()
    t0 = time.perf_counter()
    t_build = 0.0
    for _ in range(args.steps):
        tb = time.perf_counter()
        build_all()
        engine.synchronize()
        t_build += time.perf_counter() - tb
        mean = loop()
    engine.synchronize()
    t_total = time.perf_counter() - t0
    value = args.steps * n_cfg * ns / t_total
    log(f"timed {args.steps} whole jobs in {t_total:.3f} s -> {value:.1f} configuration.slices/s "
        f"(potentials {t_build / args.steps:.3f} s per job)")
    engine.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loop()
    engine.synchronize()
    loop_rate = args.steps * n_cfg * ns / (time.perf_counter() - t0)
    log(f"slice loop alone: {loop_rate:.1f} configuration.slices/s")
    roof = kernel_roofline(engine, g, n_cfg, f"1024^2, {n_cfg} configurations")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle_py as orc
        og = orc.geometry(orc.scenario(**w["scenario"]))
        cores, visible = host_cores()
        k = max(1, min(args.cpu_slices, ns))
        first = ns // 2 - k // 2
        d0 = displaced[0]
        t0 = time.perf_counter()
        srt = orc.sort_atoms(og, d0.x, d0.y, d0.z, d0.Z)
        trans = orc.transmission(og, srt, p.values, p.form, p.i_per_atom, first=first, count=k, nthreads=cores)
        t_pot = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.run_plane(trans, orc.propagator(og))
        t_loop = time.perf_counter() - t0
        cpu = {"value": round(k / (t_pot + t_loop), 4), "unit": "configuration·slices/s", "cores": cores,
               "host_cpus_visible": visible, "kind": "port",
               "sample": f"{k} mid-specimen slices of configuration 0: FP64 C++ binning and full-3D Lobato "
                         f"potentials with band limit on {cores} threads ({t_pot:.1f} s) and the slice loop "
                         f"on one thread ({t_loop:.2f} s)"}
    engine.close()
    line = {
        "metric": "full-3D frozen-phonon CTEM configuration·slices/s at 1024² (30 configurations batched)",
        "value": round(value, 2), "unit": "configuration·slices/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_total / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (amorphous SiO2 by seeded random sequential addition, seeded Philox phonons)",
        "config": {"workload": f"full-3D CTEM, amorphous SiO2 {len(spec)} atoms 60x60x100 A, {g.resolution}^2, "
                               f"{ns} slices, 300 kV, Lobato 20 integrals, {n_cfg} configurations u_rms 0.1 A, "
                               f"phonon-averaged exit intensity", "device": dev},
        "kernel_tuning": tune, "potential_build_s_per_job": round(t_build / args.steps, 3),
        "slice_loop_rate": round(loop_rate, 2),
        "timing": "each step: binning + full-3D potentials + band limit of all configurations, the "
                  "graph-captured slice loop and the FP64 intensity sum",
        "exit_intensity_mean": float(mean.mean() / n_cfg) if mean is not None else None,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_mixed(args):
    from cltem_amd.mixed import run_mixed_cbed
    inst = wl.mixed_cbed()
    log(f"{len(inst)} CBED instances, sizes "
        f"{sorted(set(i.scenario['resolution'] for i in inst))}, {inst[0].n_configs} configurations each")
    concurrent = not args.mixed_sequential
    timing = []
    patterns, groups = run_mixed_cbed(inst, warmup=args.warmup, repeats=args.steps, concurrent=concurrent,
                                      timing=timing)
    tm = timing[-1]
    value = tm.wave_slices / tm.job_s
    per = {str(res): {"instances": len(r.instances), "wave_slices": r.wave_slices,
                      "tuning": {"best_row": r.tuning["best_row"], "best_col": r.tuning["best_col"]}
                      if r.tuning else None}
           for res, r in sorted(groups.items())}
    log(f"mixed: whole job {tm.job_s:.3f} s ({tm.build_s:.3f} s binning + potentials, {tm.loop_s:.3f} s "
        f"probes + slice loops + sums) -> {value:.0f} wave.slices/s; slice loops alone "
        f"{tm.wave_slices / tm.loop_s:.0f}; {per}")
    line = {
        "metric": "mixed-size batched CBED wave·slices/s (48 instances, 256²-768², whole job, "
                  + ("all sizes in one graph launch" if concurrent else "sizes one after another") + ")",
        "value": round(value, 1), "unit": "wave·slices/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * tm.job_s, 3), "higher_is_better": True,
        "timing": "last timed step: binning + potentials of all 240 configurations (displaced coordinates "
                  "generated beforehand, as for the STEM workload), probe set-up, the slice loops of every "
                  "size and the FP64 diffraction sums; set-up and autotuning excluded",
        "build_s": round(tm.build_s, 4), "loop_s": round(tm.loop_s, 4),
        "slice_loop_rate": round(tm.wave_slices / tm.loop_s, 1), "setup_s": round(tm.setup_s, 3),
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (SrTiO3 5x5 cells, seeded thickness/size/convergence, seeded Philox phonons)",
        "config": {"workload": "mixed-size CBED, 48 SrTiO3 [001] instances, 10-80 cells thick, grids "
                               "256/384/512/768, 5-25 mrad, 5 configurations u_rms 0.08 A, FP64 "
                               "diffraction sums"},
        "per_size": per,
        "pattern_totals": [float(p.sum()) for p in patterns[:4]],
    }
    print(json.dumps(line), flush=True)
    return 0


def pixels_within(ax, ay, g, radius=8.0):
    """Per atom: the grid pixels (startx + i px, starty + j px) within `radius` of it, the
    pairs the potential kernels evaluate (their r^2 <= 64 test)."""
    res, px = g.resolution, g.real_scale
    ys = g.sim_y[0] + np.arange(res) * px
    out = np.zeros(len(ax), dtype=np.int64)
    for a0 in range(0, len(ax), 1024):
        x = ax[a0:a0 + 1024, None]
        rem = radius * radius - (ys[None, :] - ay[a0:a0 + 1024, None]) ** 2
        w = np.sqrt(np.maximum(rem, 0.0))
        lo = np.maximum(np.ceil((x - w - g.sim_x[0]) / px), 0)
        hi = np.minimum(np.floor((x + w - g.sim_x[0]) / px), res - 1)
        out[a0:a0 + 1024] = np.where(rem >= 0, np.maximum(hi - lo + 1, 0), 0).sum(axis=1)
    return out


def potential_evaluations(g, spec, z_ids):
    """Algorithmic work of one configuration's potentials: projected, the atom-pixel pairs
    within 8 A in each atom's own slice; full 3D, those pairs for every slice whose load_z
    window holds the atom, times the trapezoid points h = 0..integrals at depth
    rad_z = (z_slice - h dz / integrals) - z <= 3 (kernels/transmission_potentials_full_3d_d.cl)."""
    keep = z_ids >= 0
    npx = pixels_within(spec.x[keep], spec.y[keep], g)
    if not g.full3d:
        return int(npx.sum())
    z, zid = spec.z[keep], z_ids[keep]
    total = 0
    h = np.arange(g.full3d_integrals + 1)
    for k in range(g.n_slices):
        sel = np.abs(zid - k) <= g.load_z
        if not np.any(sel):
            continue
        cz = g.z_lims[0] + (g.n_slices - k) * g.slice_thickness
        rz = (cz - h[None, :] * g.slice_thickness / g.full3d_integrals) - z[sel, None]
        total += int((npx[sel] * (rz <= 3.0).sum(axis=1)).sum())
    return total


def bench_potentials(args):
    """Potential builders alone: the projected Kirkland potentials of the STEM configuration
    (SrTiO3 16x16x40, 512^2, 83 slices, one phonon configuration) and the full-3D Lobato
    potentials of the SiO2 configuration (20,000 atoms, 1024^2, 106 slices, 20 integrals, one
    configuration), as atom-pixel evaluations per second of the potential kernels (hipEvents
    around each launch) and the whole build with the band limit."""
    cases = [("projected Kirkland, SrTiO3 16x16x40, 512^2", wl.stem_srtio3()),
             ("full-3D Lobato 20 integrals, SiO2 20,000 atoms, 1024^2", wl.ctem_amorphous_sio2(n_configs=1))]
    out = {}
    for label, w in cases:
        engine = cltem_amd.Engine(0)
        g, p = setup(engine, w)
        spec = w["specimen"].displaced(w["u_rms"], w["seed"], 0)
        ids = engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z, want_ids=True)
        evals = potential_evaluations(g, spec, ids["z_ids"])
        for _ in range(max(1, args.warmup)):
            engine.build_transmission(0)
        engine.synchronize()
        engine.profile(True)
        engine.profile_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            engine.build_transmission(0)
        engine.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        launches, pot_ms = engine.profile_read(2)
        engine.profile(False)
        pot_s = pot_ms * 1e-3 / args.steps
        rate = evals / pot_s
        log(f"{label}: {evals:.4g} evaluations in {pot_s * 1e3:.2f} ms of potential kernels "
            f"({launches // args.steps} launches) -> {rate:.4g} evaluations/s; whole build {wall * 1e3:.2f} ms")
        out["projected" if not g.full3d else "full3d"] = {
            "case": label, "evaluations": evals, "kernel_ms": round(pot_s * 1e3, 3),
            "launches": launches // args.steps, "evaluations_per_s": round(rate, 1),
            "ns_per_evaluation_per_cu": round(pot_s / evals * 1e9 * engine.device_info()[1], 4),
            "build_ms_with_band_limit": round(wall * 1e3, 3), "slices": g.n_slices,
            "ms_per_slice": round(pot_s * 1e3 / g.n_slices, 4)}
        engine.close()
    line = {"metric": "potential-builder atom·pixel evaluations/s (projected, 512²)",
            "value": out["projected"]["evaluations_per_s"], "unit": "evaluations/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": out["projected"]["build_ms_with_band_limit"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 distances and sums)",
            "data": "synthetic (seeded Philox phonons)", "config": {"workload": "potential builders"},
            "cases": out}
    print(json.dumps(line), flush=True)
    return 0


def bench_adjoint(args):
    """Position gradients of the Au/C adjoint case (2000 atoms, 256^2, 46 slices, projected
    Kirkland, L = sum (I - I_target)^2): one step rebuilds the transmission functions from
    the positions (binning + potentials), runs the slice loop with sqrt(N) checkpoints and
    the adjoint slice loop, as one iteration of a position fit would."""
    w = wl.adjoint_auc()
    spec, tgt = w["specimen"], w["target"]
    engine = cltem_amd.Engine(0)
    dev, cus, _ = engine.device_info()
    g, p = setup(engine, w)
    ns, n_atoms = g.n_slices, len(spec)
    log(f"{dev}: {n_atoms} atoms, {g.resolution}^2, {ns} slices, adjoint position gradients")
    build_config(engine, 0, tgt)
    # the sweep variants for one wave (at one wave the passes are latency-bound, and the
    # variant that spreads a pass over the most blocks need not be the default), or the ones
    # given: under a kernel tracer every launch is slowed, and the tuner's choice with it
    if args.variants:
        rv, cv = (int(x) for x in args.variants.split(","))
        engine.set_kernel_variant(0, rv)
        engine.set_kernel_variant(1, cv)
        log(f"sweep variants given: row {rv}, column {cv}")
    else:
        tune = engine.autotune(0, 1, reps=10)
        log(f"autotune (1 wave): row variant {tune['best_row']}, column variant {tune['best_col']}; "
            f"column us {[round(x, 1) for x in tune['col_us']]}")
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    target = engine.forward_checkpointed(0)[0]

    def forward():
        build_config(engine, 0, spec)
        engine.init_plane_wave(1.0)
        return engine.forward_checkpointed(0)[0]

    def gradient():
        intensity = forward()
        return engine.adjoint_positions((2.0 * (intensity - target))[None], n_atoms)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        engine.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        engine.synchronize()
        return (time.perf_counter() - t0) / args.steps, out

    # the workload's sqrt(n)-slice checkpoints (budget 0), then every state kept (the engine's
    # default whenever the states fit its budget: 46 states of one 256^2 wave take 24 MB)
    engine.set_checkpoint_budget(0)
    t_grad, grad = timed(gradient)
    t_fwd, _ = timed(forward)
    engine.set_checkpoint_budget(1 << 30)
    t_all, grad_all = timed(gradient)
    value = 1.0 / t_grad
    log(f"gradient {1e3 * t_grad:.2f} ms with sqrt(n) checkpoints (forward alone {1e3 * t_fwd:.2f} ms) -> "
        f"{value:.2f} gradients/s; every state kept {1e3 * t_all:.2f} ms -> {1.0 / t_all:.2f} gradients/s")
    assert np.allclose(grad_all, grad, rtol=1e-5, atol=1e-9 * np.abs(grad).max())
    cpu = None
    if not args.no_cpu:
        from oracle import oracle_py as orc
        og = orc.geometry(orc.scenario(**w["scenario"]))
        t0 = time.perf_counter()
        orc.ctem_intensity(og, spec.x, spec.y, spec.z, spec.Z, p.values, p.form, p.i_per_atom, nthreads=16)
        dt = time.perf_counter() - t0
        n_fd = 2 * 3 * n_atoms
        cpu = {"value": 1.0 / (n_fd * dt), "unit": "gradients/s", "cores": 16, "kind": "port",
               "sample": f"one FP64 forward of the oracle (binning, potentials, {ns} slices) in {dt:.2f} s "
                         f"on 16 threads, scaled to the {n_fd} forwards of a central-difference gradient"}
    engine.close()
    line = {
        "metric": "adjoint position gradients/s (2000 Au/C atoms, 256², 46 slices, sqrt(N) checkpoints)",
        "value": round(value, 3), "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_grad, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (Au/C 1:3 by seeded random sequential addition, target displaced by 0.05 A)",
        "config": {"workload": f"CTEM position gradients, {n_atoms} atoms in 40^3 A, {g.resolution}^2, {ns} slices, "
                               f"200 kV, projected Kirkland, L = sum (I - I_target)^2"},
        "forward_ms": round(1e3 * t_fwd, 3), "gradient_over_forward": round(t_grad / t_fwd, 2),
        "checkpoints": f"every {int(np.ceil(np.sqrt(ns)))}th state (sqrt(n) checkpointing, segments recomputed)",
        "all_states_kept": {"gradients_per_s": round(1.0 / t_all, 3), "ms_per_gradient": round(1e3 * t_all, 3),
                            "state_bytes": int(ns * g.resolution * g.resolution * 8)},
        "grad_norm": float(np.linalg.norm(grad)), "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_sweeps(args):
    """Row and column passes alone at several grid sizes: every row and column kernel variant
    timed with hipEvents over 10 back-to-back launches (mse_time_sweeps) on --batch waves of a
    small SrTiO3 specimen (one configuration), against copy kernels of each variant's access
    pattern (mse_bandwidth_probe).  Algorithmic bytes: B*16N + 8N per row pass, B*16N per
    column pass."""
    B = args.batch
    out = {}
    for res in [int(x) for x in args.sizes.split(",")]:
        w = wl.stem_srtio3(nx=4, ny=4, nz=3, resolution=res, pixels=4, window_cells=(1, 2))
        engine = cltem_amd.Engine(0)
        g, p = setup(engine, w)
        spec = w["specimen"]
        engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
        engine.build_transmission(0)
        engine.set_wave_count(B)
        engine.set_wave_configs([0] * B)
        N = res * res
        rowb, colb = B * 16 * N + 8 * N, B * 16 * N
        rows, cols = [], []
        for rv in range(7):
            engine.set_kernel_variant(0, rv)
            rows.append(engine.time_sweeps(0, reps=10)[0])
        engine.set_kernel_variant(0, 4)
        for cv in range(14):
            engine.set_kernel_variant(1, cv)
            cols.append((engine.time_sweeps(0, reps=10)[1], engine.bandwidth_probe(1, B, reps=10, variant=cv)))
        copy_row = engine.bandwidth_probe(0, B, reps=10)
        engine.close()
        br = int(np.argmin(rows))
        bc = int(np.argmin([c[0] for c in cols]))
        out[str(res)] = {
            "row_us": [round(x, 2) for x in rows], "col_us": [round(c[0], 2) for c in cols],
            "col_copy_pattern_us": [round(c[1], 2) for c in cols], "row_copy_pattern_us": round(copy_row, 2),
            "best_row": br, "best_col": bc,
            "row_GBs": round(rowb / rows[br] / 1e3, 1), "col_GBs": round(colb / cols[bc][0] / 1e3, 1),
            "row_frac": round(rowb / rows[br] / 8e6, 4), "col_frac": round(colb / cols[bc][0] / 8e6, 4),
            "row_copy_GBs": round(rowb / copy_row / 1e3, 1), "col_copy_GBs": round(colb / cols[bc][1] / 1e3, 1)}
        log(f"{res}^2, B = {B}: row variant {br} {out[str(res)]['row_GBs']} GB/s "
            f"(copy pattern {out[str(res)]['row_copy_GBs']}), column variant {bc} {out[str(res)]['col_GBs']} GB/s "
            f"(copy pattern {out[str(res)]['col_copy_GBs']})")
    best = max(out.values(), key=lambda d: d["col_frac"])
    line = {"metric": "slice-sweep HBM bandwidth per grid size (best variants)", "value": best["col_GBs"],
            "unit": "GB/s", "n_gpus": 1, "steps": 10, "warmup": 0, "ms_per_step": None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (SrTiO3 4x4x3, one configuration)",
            "config": {"workload": f"row and column passes, B = {B}", "sizes": args.sizes}, "sizes": out}
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=("stem", "sio2", "mixed", "adjoint", "potentials", "sweeps"),
                    default="stem")
    ap.add_argument("--sizes", default="512,768,1024", help="sweeps: grid sizes")
    ap.add_argument("--configs", type=int, default=30, help="sio2: frozen-phonon configurations")
    ap.add_argument("--cpu-slices", type=int, default=2, help="sio2: slices in the CPU baseline sample")
    ap.add_argument("--mixed-sequential", action="store_true",
                    help="mixed: run the size groups one after another instead of concurrently")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed frozen-phonon configurations")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--sweep", default="1,16,64,256", help="batch sizes for the slice-loop rate")
    ap.add_argument("--sweep-probes", type=int, default=256)
    ap.add_argument("--plain-probes", type=int, default=8)
    ap.add_argument("--cpu-probes", type=int, default=16)
    ap.add_argument("--check-probes", type=int, default=8,
                    help="probes of the first timed configuration checked against the FP64 oracle")
    ap.add_argument("--roof-batch", type=int, default=64, help="batch size of the roofline measurement")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variants", default="",
                    help="adjoint: row,column sweep variants to use instead of autotuning them")
    ap.add_argument("--graph", type=int, default=1, help="replay each probe batch as a captured hipGraph")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: one batch of --batch probes through every slice, "
                         "then 5 launches each of the row and column-tile copy kernels "
                         "(the traffic-counter calibration); prints kernel timings only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this build "
              "(probes and phonon configurations would shard across GPUs with one FP64 "
              "all-reduce of the detector or intensity sums at the end)", file=sys.stderr)
        return 2
    if args.workload == "sio2":
        return bench_sio2(args)
    if args.workload == "mixed":
        return bench_mixed(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "potentials":
        return bench_potentials(args)
    if args.workload == "sweeps":
        return bench_sweeps(args)

    w = wl.stem_srtio3()
    spec = w["specimen"]
    engine = cltem_amd.Engine(0)
    dev, cus, _ = engine.device_info()
    g, p = setup(engine, w)
    pos = wl.stem_positions(g)
    probe = cltem_amd.make_probe(w["cond_ap_mrad"])
    dets = cltem_amd.make_detectors(w["detectors"])
    N = g.resolution * g.resolution
    ns = g.n_slices
    log(f"{dev}, {cus} CUs; {len(spec)} atoms, {g.resolution}^2, {ns} slices, {len(pos)} probes")

    if args.profile:
        build_config(engine, 0, spec.displaced(w["u_rms"], w["seed"], 0))
        engine.stem_batch(0, pos[: args.batch], args.batch, probe, dets, use_graph=False)
        engine.synchronize()
        roof = roofline(engine, g, pos, args.batch, probe, dets)
        copy_row = engine.bandwidth_probe(0, args.batch, reps=5)
        copy_col = engine.bandwidth_probe(1, args.batch, reps=5, variant=0)
        print(json.dumps({"profile_run": True, "batch": args.batch, "roofline": roof,
                          "copy_row_us": copy_row, "copy_col_us": copy_col,
                          "copy_bytes_per_launch": args.batch * 16 * N}), flush=True)
        engine.close()
        return 0

    n_cfg = args.warmup + args.steps
    displaced = [spec.displaced(w["u_rms"], w["seed"], c) for c in range(n_cfg)]
    accum = np.zeros((len(pos), len(dets)))
    first_cfg = None
    tune = None
    for c in range(args.warmup):
        build_config(engine, 0, displaced[c])
        if tune is None:
            tune = engine.autotune(0, args.batch, reps=10)  # the column variants lie within a few percent
            log(f"autotune at B={args.batch}: {tune}")
        engine.stem_batch(0, pos, args.batch, probe, dets, use_graph=bool(args.graph))
        log(f"warmup configuration {c} done")
    engine.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        c = args.warmup + k
        build_config(engine, 0, displaced[c])
        engine.stem_batch(0, pos, args.batch, probe, dets, use_graph=bool(args.graph), accum=accum)
        if k == 0:
            engine.synchronize()
            first_cfg = accum.copy()
    engine.synchronize()
    t_total = time.perf_counter() - t0
    value = args.steps * len(pos) * ns / t_total
    log(f"timed {args.steps} configurations in {t_total:.3f} s -> {value:.0f} probe.slices/s")

    # potential build alone (configuration 0 of the timed set)
    engine.synchronize()
    t0 = time.perf_counter()
    build_config(engine, 0, displaced[args.warmup])
    engine.synchronize()
    t_build = time.perf_counter() - t0

    # slice-loop rate against B (transmission functions resident)
    sweep = {}
    sweep_tuning = {}
    for B in [int(b) for b in args.sweep.split(",") if b]:
        npr = max(B, args.sweep_probes if B > 1 else min(args.sweep_probes, 64))
        sub = pos[:npr]
        if B > 1 and B != args.batch:  # kernel variants tuned for this batch size too
            t = engine.autotune(0, B, reps=10)
            sweep_tuning[str(B)] = {"best_row": t["best_row"], "best_col": t["best_col"]}
        engine.stem_batch(0, sub[:B], B, probe, dets)
        engine.synchronize()
        t0 = time.perf_counter()
        engine.stem_batch(0, sub, B, probe, dets, use_graph=bool(args.graph))
        engine.synchronize()
        dt = time.perf_counter() - t0
        sweep[str(B)] = round(len(sub) * ns / dt, 1)
        log(f"slice loop B={B}: {sweep[str(B)]:.0f} probe.slices/s")
    plain, _ = plain_rate(engine, g, pos[: args.plain_probes], probe, dets)
    log(f"plain (B=1, sync per slice): {plain:.0f} probe.slices/s")

    roof = roofline(engine, g, pos, args.roof_batch, probe, dets)
    log(f"roofline: {roof}")

    cpu = None
    parity = None
    if not args.no_cpu:
        gpu_cfg0 = first_cfg
        # probes spread over the raster (every len/n-th), so the check is not one corner
        chk = np.linspace(0, len(pos) - 1, args.check_probes).astype(int)
        cpu, err, nchk = cpu_baseline(w, g, displaced[args.warmup], pos[: args.cpu_probes], pos[chk], gpu_cfg0[chk])
        parity = {"checked": "detector intensities of the first timed phonon configuration against the "
                             "FP64 oracle (same displaced atoms, oracle potentials and slice loop)",
                  "detector_max_rel_err_vs_fp64": err, "probes_checked": nchk, "probe_indices": chk.tolist(),
                  "tolerance": 1e-3, "ok": bool(err <= 1e-3)}
        log(f"cpu baseline: {cpu}; parity {parity}")
    engine.close()

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "probe·slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (SrTiO3 16x16x40, seeded Philox frozen phonons)",
        "config": {
            "workload": "batched HAADF STEM, SrTiO3 16x16x40 [001], 512^2, 83 slices, 200 kV, "
                        "20 mrad, 64x64 raster, 60-200 mrad, projected Kirkland",
            "resolution": g.resolution, "probes": len(pos), "slices": ns, "batch": args.batch,
            "phonon_configs_timed": args.steps, "atoms": len(spec), "device": dev,
        },
        "kernel_tuning": tune,
        "slice_loop_rate_by_batch": sweep, "kernel_tuning_by_batch": sweep_tuning,
        "plain_rate": round(plain, 1),
        # whole job at B (value) over the plain form with a host sync after every slice
        "batched_over_plain": round(value / plain, 2) if plain > 0 else None,
        # like for like: the batched slice loop at B over the graph-replayed B = 1 loop
        "batched_over_plain_graph": (round(sweep[str(args.batch)] / sweep["1"], 2)
                                     if "1" in sweep and str(args.batch) in sweep else None),
        "potential_build_ms_per_config": round(1e3 * t_build, 2),
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(line), flush=True)
    if parity is not None and not parity["ok"]:
        log(f"parity check FAILED: {parity['detector_max_rel_err_vs_fp64']:.3e} > {parity['tolerance']}")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
