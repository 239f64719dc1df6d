"""Mixed-size batched CBED: independent specimens of different thickness, grid size and
probe convergence, each averaged over its own frozen-phonon configurations.

clTEM runs one simulation per job on one grid (SimulationCbed::simulate,
microscope/simulationcbed.cpp:122-210, merged by SimulationManager::updateImages,
simulationmanager.cpp:274-350).  Here every instance of one grid size shares an engine:
the engine's geometry is that of the thickest instance of the size and thinner instances
get their own z geometry (mse_set_config_geometry).  Their waves, ordered by thickness,
go through one graph-captured slice loop in which each wave stops at its own exit
surface.  Each size has its own engine, stream and autotuned kernel variants, and the
sizes run concurrently from one host thread each (the engine calls release the
interpreter lock).
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import Engine, load_params, make_probe, make_scenario, plan_geometry


@dataclass
class CbedInstance:
    specimen: object             # specimens.Specimen, undisplaced
    scenario: dict               # mse_scenario fields (mode 3)
    alpha_mrad: float            # condenser aperture
    position: tuple              # probe position (x, y), Angstrom
    n_configs: int = 5
    u_rms: float = 0.08
    seed: int = 0


@dataclass
class SizeGroupResult:
    resolution: int
    instances: list
    tuning: dict | None
    build_s: float
    loop_s: float                # last timed repetition (probe init, slice loop, FP64 sums)
    wave_slices: int             # slices actually run per repetition, summed over waves
    t_start: float = 0.0         # perf_counter at the start / end of the timed repetitions
    t_end: float = 0.0
    patterns: dict = field(default_factory=dict)  # instance index -> mean diffraction (res, res)


def _run_group(res, items, params_name, device, autotune, graph, warmup, repeats, barrier, out, errors):
    try:
        _run_group_body(res, items, params_name, device, autotune, graph, warmup, repeats, barrier, out)
    except Exception as ex:  # reported by run_mixed_cbed
        errors.append((res, ex))
        if barrier is not None:
            barrier.abort()


def _run_group_body(res, items, params_name, device, autotune, graph, warmup, repeats, barrier, out):
    eng = Engine(device)
    p = load_params(params_name)
    geoms = [plan_geometry(make_scenario(**inst.scenario)) for _, inst in items]
    # waves ordered by thickness: each slice's active waves are then a prefix of the batch
    order = sorted(range(len(items)), key=lambda k: -geoms[k].n_slices)
    items = [items[k] for k in order]
    geoms = [geoms[k] for k in order]
    eng.set_geometry(geoms[0])
    eng.set_parameterisation(p)
    t0 = time.perf_counter()
    wave_cfg, wave_group = [], []
    c = 0
    for gi, ((idx, inst), g) in enumerate(zip(items, geoms)):
        for ph in range(inst.n_configs):
            if gi != 0:
                eng.set_config_geometry(c, g)
            d = inst.specimen.displaced(inst.u_rms, inst.seed, ph)
            eng.sort_atoms(c, d.x, d.y, d.z, d.Z, want_ids=False)
            eng.build_transmission(c)
            wave_cfg.append(c)
            wave_group.append(gi)
            c += 1
    eng.synchronize()
    build_s = time.perf_counter() - t0
    tuning = eng.autotune(0, min(len(wave_cfg), 64), reps=3) if autotune else None
    eng.set_wave_count(len(wave_cfg))
    eng.set_wave_configs(wave_cfg)
    probes = [make_probe(inst.alpha_mrad) for _, inst in items]

    def one_pass():
        w0 = 0
        for (idx, inst), pr in zip(items, probes):
            pos = np.tile(np.asarray(inst.position, dtype=np.float64), (inst.n_configs, 1))
            eng.init_probes_range(w0, pos, pr)
            w0 += inst.n_configs
        eng.multislice(graph=graph)
        return eng.diffraction_sum(wave_group, len(items))

    for _ in range(warmup):
        one_pass()
    eng.synchronize()
    if barrier is not None:
        barrier.wait()
    t_start = time.perf_counter()
    loop_s = 0.0
    for _ in range(repeats):
        t0 = time.perf_counter()
        sums = one_pass()
        loop_s = time.perf_counter() - t0
    t_end = time.perf_counter()
    wave_slices = sum(inst.n_configs * g.n_slices for (_, inst), g in zip(items, geoms))
    r = SizeGroupResult(res, [idx for idx, _ in items], tuning, build_s, loop_s, wave_slices, t_start, t_end)
    for gi, (idx, inst) in enumerate(items):
        r.patterns[idx] = sums[gi] / inst.n_configs
    eng.close()
    out[res] = r


def run_mixed_cbed(instances, params_name="kirkland", device=0, autotune=True, graph=True, warmup=0,
                   repeats=1, concurrent=True):
    """Phonon-averaged CBED patterns |Psi(k)|^2 (fft-shifted, FP64) of every instance.
    Each size group builds its transmission functions, runs `warmup` untimed passes, then
    `repeats` timed passes (probe init, slice loop, FP64 sums): concurrently with the other
    groups after a common barrier, or (concurrent=False) one group after another, each with
    the whole GPU.  Returns (patterns in instance order, per-size results)."""
    groups: dict = {}
    for i, inst in enumerate(instances):
        groups.setdefault(inst.scenario["resolution"], []).append((i, inst))
    out, errors = {}, []
    if concurrent:
        barrier = threading.Barrier(len(groups)) if len(groups) > 1 else None
        threads = [threading.Thread(target=_run_group, args=(res, items, params_name, device, autotune, graph,
                                                              warmup, max(1, repeats), barrier, out, errors))
                   for res, items in sorted(groups.items())]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    else:
        for res, items in sorted(groups.items()):
            _run_group(res, items, params_name, device, autotune, graph, warmup, max(1, repeats), None, out, errors)
            if errors:
                break
    if errors:
        real = [e for e in errors if not isinstance(e[1], threading.BrokenBarrierError)] or errors
        raise RuntimeError(f"size group {real[0][0]} failed: {real[0][1]}") from real[0][1]
    patterns = [None] * len(instances)
    for r in out.values():
        for idx, pat in r.patterns.items():
            patterns[idx] = pat
    return patterns, out
