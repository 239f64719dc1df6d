"""Mixed-size batched CBED: independent specimens of different thickness, grid size and
probe convergence, each averaged over its own frozen-phonon configurations.

clTEM runs one simulation per job on one grid (SimulationCbed::simulate,
microscope/simulationcbed.cpp:122-210, merged by SimulationManager::updateImages,
simulationmanager.cpp:274-350).  Here every instance of one grid size shares an engine:
the engine's geometry is that of the thickest instance of the size, thinner instances
get their own z geometry (mse_set_config_geometry) and run vacuum slices below their
exit surface (free propagation leaves |Psi(k)|^2 unchanged), so all waves of a size go
through one graph-captured slice loop.  Each size has its own engine, stream and
autotuned kernel variants, and the sizes run concurrently from one host thread each
(the engine calls release the interpreter lock).
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
    loop_s: float
    wave_slices: int
    patterns: dict = field(default_factory=dict)  # instance index -> mean diffraction (res, res)


def _run_group(res, items, params_name, device, autotune, graph, out, errors):
    try:
        eng = Engine(device)
        p = load_params(params_name)
        geoms = [plan_geometry(make_scenario(**inst.scenario)) for _, inst in items]
        thick = int(np.argmax([g.n_slices for g in geoms]))
        eng.set_geometry(geoms[thick])
        eng.set_parameterisation(p)
        t0 = time.perf_counter()
        wave_cfg, wave_group = [], []
        c = 0
        for gi, ((idx, inst), g) in enumerate(zip(items, geoms)):
            for ph in range(inst.n_configs):
                if gi != thick:
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
        w0 = 0
        for idx, inst in items:
            pos = np.tile(np.asarray(inst.position, dtype=np.float64), (inst.n_configs, 1))
            eng.init_probes_range(w0, pos, make_probe(inst.alpha_mrad))
            w0 += inst.n_configs
        eng.synchronize()
        t0 = time.perf_counter()
        eng.multislice(graph=graph)
        sums = eng.diffraction_sum(wave_group, len(items))
        loop_s = time.perf_counter() - t0
        r = SizeGroupResult(res, [idx for idx, _ in items], tuning, build_s, loop_s,
                            len(wave_cfg) * geoms[thick].n_slices)
        for gi, (idx, inst) in enumerate(items):
            r.patterns[idx] = sums[gi] / inst.n_configs
        eng.close()
        out[res] = r
    except Exception as ex:  # reported by run_mixed_cbed
        errors.append((res, ex))


def run_mixed_cbed(instances, params_name="kirkland", device=0, autotune=True, graph=True):
    """Phonon-averaged CBED patterns |Psi(k)|^2 (fft-shifted, FP64) of every instance.
    Returns (patterns in instance order, per-size group results)."""
    groups: dict = {}
    for i, inst in enumerate(instances):
        groups.setdefault(inst.scenario["resolution"], []).append((i, inst))
    out, errors = {}, []
    threads = [threading.Thread(target=_run_group, args=(res, items, params_name, device, autotune, graph,
                                                          out, errors))
               for res, items in sorted(groups.items())]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise RuntimeError(f"size group {errors[0][0]} failed: {errors[0][1]}") from errors[0][1]
    patterns = [None] * len(instances)
    for r in out.values():
        for idx, pat in r.patterns.items():
            patterns[idx] = pat
    return patterns, out
