"""Mixed static potentials (useParallelPotentials) against the FP64 oracle: every wave
multiplies, at every slice, the transmission function of one of several phonon
configurations, drawn per (probe, slice) from a seeded counter; the oracle replays the same
draws (cltem_amd.incoherence.mixed_potential_picks)."""
import os
import subprocess

import numpy as np
import pytest

import cltem_amd
from cltem_amd import incoherence as inc
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "cltem_amd", "mse_run")
PARAMS = os.path.join(ROOT, "cltem_amd", "params")


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def oracle_sets(oracle, og, spec, p, n_sets, u_rms, seed):
    out = []
    for j in range(n_sets):
        d = spec.displaced(u_rms, seed, j)
        srt = oracle.sort_atoms(og, d.x, d.y, d.z, d.Z)
        out.append(oracle.transmission(og, srt, p.values, p.form, p.i_per_atom))
    return out


def oracle_mixed_recip(oracle, og, sets, picks, pos, alpha):
    psi = oracle.fft2(oracle.probe_kspace(og, pos[0] - og.sim_x[0], pos[1] - og.sim_y[0], None, alpha), inverse=True)
    P = oracle.propagator(og)
    for sl in range(og.n_slices):
        psi, recip = oracle.slice_step(psi, sets[picks[sl]][sl], P)
    return recip


def test_cbed_waves_draw_sets_per_slice(engine, oracle):
    spec = sp.srtio3(3, 3, 4)
    lo, hi = spec.limits
    cx = 1.5 * 3.905
    scen = dict(mode=3, resolution=256, slice_thickness=1.9525, struct_min=tuple(lo), struct_max=tuple(hi),
                area=(cx, cx, cx, cx))
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    og = oracle.geometry(oracle.scenario(**scen))
    p = cltem_amd.load_params("kirkland")
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    n_sets, seed = 3, 0x5EED
    for j in range(n_sets):
        d = spec.displaced(0.1, 21, j)
        engine.sort_atoms(j, d.x, d.y, d.z, d.Z)
        engine.build_transmission(j)
    pos = np.array([[cx, cx], [cx + 0.5, cx - 0.3], [cx - 0.9, cx + 0.2]])
    engine.set_wave_count(3)
    engine.set_wave_configs([0, 0, 0])
    engine.set_mixed_potentials(0, n_sets, seed)
    engine.init_probes(pos, cltem_amd.make_probe(16.0))
    engine.multislice()
    picks = inc.mixed_potential_picks(seed, [0, 1, 2], g.n_slices, n_sets)
    assert set(np.unique(picks)) == {0, 1, 2}
    sets = oracle_sets(oracle, og, spec, p, n_sets, 0.1, 21)
    for w in range(3):
        ref = oracle.diffraction(oracle_mixed_recip(oracle, og, sets, picks[w], pos[w], 16.0))
        err = rel_l2(engine.diffraction(w), ref)
        print(f"wave {w}: sets {picks[w].tolist()}: rel-L2 {err:.3e}")
        assert err <= 1e-4
    # the graph-captured loop draws the same sets
    engine.init_probes(pos, cltem_amd.make_probe(16.0))
    engine.multislice(graph=True)
    ref = oracle.diffraction(oracle_mixed_recip(oracle, og, sets, picks[2], pos[2], 16.0))
    assert rel_l2(engine.diffraction(2), ref) <= 1e-4
    # position gradients are not defined over mixed sets
    with pytest.raises(cltem_amd.MseError):
        engine.forward_checkpointed(0)
    engine.set_mixed_potentials(0, 0)


def test_stem_batches_draw_by_probe_index(engine, oracle):
    """Probe k of a STEM raster draws with key k in every batch, also in graph replays and
    in the eager final partial batch."""
    w = wl.stem_srtio3(nx=5, ny=5, nz=3, resolution=256, pixels=3, window_cells=(2, 3))
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    p = cltem_amd.load_params("kirkland")
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    n_sets, seed = 4, 99
    for j in range(n_sets):
        d = w["specimen"].displaced(0.08, 3, j)
        engine.sort_atoms(j, d.x, d.y, d.z, d.Z)
        engine.build_transmission(j)
    engine.set_mixed_potentials(0, n_sets, seed)
    pos = wl.stem_positions(g)  # 9 probes: batches of 4, 4 (graph replays) and 1 (eager)
    rows = [(60.0, 200.0, 0.0, 0.0), (0.0, 15.0, 0.0, 0.0)]
    got = engine.stem_batch(0, pos, 4, cltem_amd.make_probe(20.0), cltem_amd.make_detectors(rows), use_graph=True)
    picks = inc.mixed_potential_picks(seed, np.arange(len(pos)), g.n_slices, n_sets)
    sets = oracle_sets(oracle, og, w["specimen"], p, n_sets, 0.08, 3)
    worst = 0.0
    for k in range(len(pos)):
        recip = oracle_mixed_recip(oracle, og, sets, picks[k], pos[k], 20.0)
        for j, (inner, outer, xc, yc) in enumerate(rows):
            want = oracle.stem_pixel(recip, inner, outer, xc, yc, og.inverse_scale_angle, True)
            worst = max(worst, abs(got[k, j] - want) / abs(want))
    print(f"STEM with {n_sets} mixed sets, 9 probes: max detector rel err {worst:.3e}")
    assert worst <= 1e-3
    engine.set_mixed_potentials(0, 0)


@pytest.mark.skipif(not os.path.exists(RUN), reason="mse_run not built")
def test_driver_stem_parallel_potentials_match_oracle(tmp_path, oracle):
    """Three frozen-phonon jobs sharing two potential sets, each job drawing with its own seed."""
    prefix = str(tmp_path / "stem_pp")
    seed = 31
    subprocess.run([RUN, "--mode", "stem", "--srtio3", "5,5,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{2 * 3.905},{3 * 3.905},{2 * 3.905},{3 * 3.905}", "--stem-pixels", "2",
                    "--det", "ADF,60,200", "--alpha", "20", "--batch", "4", "--phonons", "3", "--urms", "0.08",
                    "--seed", str(seed), "--parallel-potentials", "2",
                    "--params", os.path.join(PARAMS, "kirkland.dat"), "--out", prefix], check=True)
    import json
    meta = json.load(open(prefix + ".json"))
    got = np.fromfile(meta["ADF"]["file"], dtype=np.float64)
    w = wl.stem_srtio3(nx=5, ny=5, nz=3, resolution=256, pixels=2, window_cells=(2, 3))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    p = cltem_amd.load_params("kirkland")
    sets = oracle_sets(oracle, og, w["specimen"], p, 2, 0.08, seed)
    pos = wl.stem_positions(og)
    ref = np.zeros(len(pos))
    for job in range(3):
        picks = inc.mixed_potential_picks(inc.mixed_seed(seed, job), np.arange(len(pos)), og.n_slices, 2)
        for k in range(len(pos)):
            recip = oracle_mixed_recip(oracle, og, sets, picks[k], pos[k], 20.0)
            ref[k] += oracle.stem_pixel(recip, 60.0, 200.0, 0.0, 0.0, og.inverse_scale_angle, True)
    ref /= 3
    err = np.max(np.abs(got - ref) / np.abs(ref))
    print(f"host STEM, 3 jobs over 2 shared sets: max rel err {err:.3e}")
    assert err <= 1e-3
