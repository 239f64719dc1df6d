"""Mixed-size batched CBED: independent specimens of different thickness, grid size and
probe convergence, each averaged over its own frozen-phonon configurations.

clTEM runs one simulation per job on one grid (SimulationCbed::simulate,
microscope/simulationcbed.cpp:122-210, merged by SimulationManager::updateImages,
simulationmanager.cpp:274-350).  Here every instance of one grid size shares an engine:
the engine's geometry is that of the thickest instance of the size and thinner instances
get their own z geometry (mse_set_config_geometry); their waves, ordered by thickness, stop
at their own exit surfaces.  The engines of all sizes form one group (mse_group_*), and the
slice loops of every size run from one captured graph launched once per job: each size's
kernels (its own FFT plan and autotuned variants) are a branch of that graph, and the
branches run concurrently on the one GPU.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import Engine, EngineGroup, load_params, make_probe, make_scenario, plan_geometry


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
    wave_slices: int             # slices actually run per job, summed over the size's waves
    patterns: dict = field(default_factory=dict)  # instance index -> mean diffraction (res, res)


@dataclass
class MixedTiming:
    setup_s: float               # engines, geometries, first builds and autotuning (once)
    job_s: float                 # one whole job: binning, potentials, probes, slice loops, sums
    build_s: float               # of which binning and potentials
    loop_s: float                # of which probe set-up, the group's slice loops and FP64 sums
    wave_slices: int             # per job, all sizes


class _SizeEngine:
    """The instances of one grid size on one engine, waves ordered by thickness."""

    def __init__(self, res, items, params, device):
        self.res = res
        self.eng = Engine(device)
        geoms = [plan_geometry(make_scenario(**inst.scenario)) for _, inst in items]
        order = sorted(range(len(items)), key=lambda k: -geoms[k].n_slices)
        self.items = [items[k] for k in order]
        self.geoms = [geoms[k] for k in order]
        self.eng.set_geometry(self.geoms[0])
        self.eng.set_parameterisation(params)
        self.wave_cfg, self.wave_group, self.jobs = [], [], []
        c = 0
        for gi, ((idx, inst), g) in enumerate(zip(self.items, self.geoms)):
            for ph in range(inst.n_configs):
                if gi != 0:
                    self.eng.set_config_geometry(c, g)
                # the displaced coordinates of every configuration, generated once
                self.jobs.append((c, inst.specimen.displaced(inst.u_rms, inst.seed, ph)))
                self.wave_cfg.append(c)
                self.wave_group.append(gi)
                c += 1
        self.probes = [make_probe(inst.alpha_mrad) for _, inst in self.items]
        self.wave_slices = sum(inst.n_configs * g.n_slices for (_, inst), g in zip(self.items, self.geoms))

    def sort(self):
        for c, d in self.jobs:
            self.eng.sort_atoms(c, d.x, d.y, d.z, d.Z, want_ids=False)

    def build(self):
        for c, _ in self.jobs:
            self.eng.build_transmission(c)

    def sort_and_build(self):
        """Each configuration binned, then its potentials queued: a binning's synchronisation
        waits for the previous configuration's build, which keeps the GPU busy meanwhile."""
        for c, d in self.jobs:
            self.eng.sort_atoms(c, d.x, d.y, d.z, d.Z, want_ids=False)
            self.eng.build_transmission(c)
        self.eng.synchronize()

    def set_waves(self):
        self.eng.set_wave_count(len(self.wave_cfg))
        self.eng.set_wave_configs(self.wave_cfg)

    def init_probes(self):
        w0 = 0
        for (idx, inst), pr in zip(self.items, self.probes):
            pos = np.tile(np.asarray(inst.position, dtype=np.float64), (inst.n_configs, 1))
            self.eng.init_probes_range(w0, pos, pr)
            w0 += inst.n_configs

    def sums(self):
        return self.eng.diffraction_sum(self.wave_group, len(self.items))


def run_mixed_cbed(instances, params_name="kirkland", device=0, autotune=True, graph=True, warmup=0,
                   repeats=1, concurrent=True, timing=None):
    """Phonon-averaged CBED patterns |Psi(k)|^2 (fft-shifted, FP64) of every instance.

    After the set-up (one engine per grid size, the displaced coordinates of every
    configuration, first builds, autotuning), `warmup` untimed and `repeats` timed whole jobs
    run: binning and potentials of every instance and configuration, probe set-up, the slice
    loops of all sizes and the FP64 sums.  With concurrent=True each size's binning and
    potentials run from a host thread of their own and the slice loops of all sizes are one
    graph launch (mse_group_multislice);
    otherwise each size runs its own loop after the previous one (graph per size when graph=True,
    eager launches otherwise).  When `timing` is a list, a MixedTiming of the last timed job is
    appended.  Returns (patterns in instance order, per-size results)."""
    t_setup = time.perf_counter()
    by_size: dict = {}
    for i, inst in enumerate(instances):
        by_size.setdefault(inst.scenario["resolution"], []).append((i, inst))
    params = load_params(params_name)
    sizes = [_SizeEngine(res, items, params, device) for res, items in sorted(by_size.items())]
    tuning = {}
    for s in sizes:
        s.sort()
        s.build()
        if autotune:
            tuning[s.res] = s.eng.autotune(0, min(len(s.wave_cfg), 64), reps=3)
        s.set_waves()
    for s in sizes:
        s.eng.synchronize()
    t_setup = time.perf_counter() - t_setup
    group = EngineGroup([s.eng for s in sizes]) if concurrent else None

    pool = ThreadPoolExecutor(len(sizes)) if concurrent else None

    def job():
        t0 = time.perf_counter()
        if pool is not None:
            # one host thread per size: one engine's binning round trips overlap the other
            # engines' potential kernels (the ABI calls release the interpreter lock)
            for f in [pool.submit(s.sort_and_build) for s in sizes]:
                f.result()
        else:
            for s in sizes:
                s.sort_and_build()
        t1 = time.perf_counter()
        for s in sizes:
            s.init_probes()
        if group is not None:
            group.multislice()
        else:
            for s in sizes:
                s.eng.multislice(graph=graph)
                s.eng.synchronize()
        out = [s.sums() for s in sizes]
        t2 = time.perf_counter()
        return out, t2 - t0, t1 - t0, t2 - t1

    try:
        for _ in range(warmup):
            job()
        for _ in range(max(1, repeats)):
            sums, job_s, build_s, loop_s = job()
    finally:
        if pool is not None:
            pool.shutdown()
        if group is not None:
            group.close()
    wave_slices = sum(s.wave_slices for s in sizes)
    if timing is not None:
        timing.append(MixedTiming(t_setup, job_s, build_s, loop_s, wave_slices))
    results, patterns = {}, [None] * len(instances)
    for s, sm in zip(sizes, sums):
        r = SizeGroupResult(s.res, [idx for idx, _ in s.items], tuning.get(s.res), s.wave_slices)
        for gi, (idx, inst) in enumerate(s.items):
            r.patterns[idx] = sm[gi] / inst.n_configs
            patterns[idx] = r.patterns[idx]
        results[s.res] = r
        s.eng.close()
    return patterns, results
