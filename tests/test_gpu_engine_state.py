"""Engine state across calls: a STEM batch on a configuration thinner than the engine
geometry, and the autotuner growing the wave set while a configuration carries its own beam
vector, each checked against the FP64 oracle and followed by ordinary work on the engine."""
import math

import numpy as np
import pytest

import cltem_amd
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def max_rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))))


@pytest.mark.parametrize("graph", [False, True])
def test_stem_batch_on_a_thinner_configuration(engine, oracle, graph):
    """A configuration with fewer slices than the engine geometry (mse_set_config_geometry)
    runs its probes to its own exit surface, with the detectors on its last slice; the
    engine then runs a batch of the thick configuration as usual."""
    thick = wl.stem_srtio3(nx=3, ny=3, nz=6, resolution=256, pixels=3, window_cells=(1, 2),
                           detectors=((60.0, 200.0), (0.0, 20.0)))
    thin = wl.stem_srtio3(nx=3, ny=3, nz=3, resolution=256, pixels=3, window_cells=(1, 2),
                          detectors=((60.0, 200.0), (0.0, 20.0)))
    gk = cltem_amd.plan_geometry(cltem_amd.make_scenario(**thick["scenario"]))
    gn = cltem_amd.plan_geometry(cltem_amd.make_scenario(**thin["scenario"]))
    assert gn.n_slices < gk.n_slices
    p = cltem_amd.load_params("kirkland")
    engine.set_geometry(gk)
    engine.set_parameterisation(p)
    for c, w in ((0, thick), (1, thin)):
        if c == 1:
            engine.set_config_geometry(1, gn)
        s = w["specimen"]
        engine.sort_atoms(c, s.x, s.y, s.z, s.Z)
        engine.build_transmission(c)
    pos = wl.stem_positions(gk)
    probe = cltem_amd.make_probe(20.0)
    dets = cltem_amd.make_detectors(thick["detectors"])
    det4 = np.array([list(d) + [0.0] * (4 - len(d)) for d in thick["detectors"]])
    refs = {}
    for c, w in ((0, thick), (1, thin)):
        og = oracle.geometry(oracle.scenario(**w["scenario"]))
        s = w["specimen"]
        srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
        refs[c] = oracle.run_stem(og, oracle.transmission(og, srt, p.values, p.form, p.i_per_atom),
                                  oracle.propagator(og), pos, det4, cond_ap_mrad=20.0)
    got_thin = engine.stem_batch(1, pos, 4, probe, dets, use_graph=graph)
    err = max_rel(got_thin, refs[1])
    print(f"thin configuration ({gn.n_slices} of {gk.n_slices} slices), graph={graph}: max rel {err:.2e}")
    assert err <= 1e-3
    got_thick = engine.stem_batch(0, pos, 4, probe, dets, use_graph=graph)
    assert max_rel(got_thick, refs[0]) <= 1e-3
    # and back to the thin one (a graph of the other slice count is captured anew)
    assert max_rel(engine.stem_batch(1, pos, 4, probe, dets, use_graph=graph), refs[1]) <= 1e-3


def test_autotune_grows_waves_under_a_configuration_beam(engine, oracle):
    """mse_autotune with more waves than the engine held while configuration 0 has its own
    beam vector: the per-wave propagator table follows the new wave set, so the slice loop
    afterwards propagates every wave with the configuration's tilted propagator."""
    w = wl.ctem_srtio3(3, 3, 3, resolution=256)
    scen = w["scenario"]
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    p = cltem_amd.load_params("kirkland")
    s = w["specimen"]
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    engine.sort_atoms(0, s.x, s.y, s.z, s.Z)
    k0 = g.wavenumber
    tilt = 0.006
    k = np.array([k0 * math.sin(tilt) * math.cos(0.4), k0 * math.sin(tilt) * math.sin(0.4), k0 * math.cos(tilt)])
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.set_config_beam(0, k)
    engine.build_transmission(0)
    engine.autotune(0, 6, reps=1)   # six waves of configuration 0 from here on
    engine.set_kernel_variant(0, 4)
    engine.set_kernel_variant(1, 7)
    engine.init_plane_wave(1.0)
    engine.multislice()
    og0 = oracle.geometry(oracle.scenario(**scen))
    og = oracle.geometry(oracle.scenario(**dict(scen, beam_tilt_mrad=1000.0 * tilt, beam_azimuth=0.4)))
    og.sigma_eff = og0.sigma_eff  # modifyBeamTilt keeps the original interaction constant
    srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
    psi, _ = oracle.run_plane(oracle.transmission(og, srt, p.values, p.form, p.i_per_atom), oracle.propagator(og))
    for i in range(6):
        err = rel_l2(engine.exit_wave(i), psi)
        assert err <= 1e-4, (i, err)
    print(f"6 waves after autotune, tilted configuration: exit waves match the FP64 oracle")
    engine.set_config_beam(0, None)
