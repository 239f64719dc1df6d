"""Batched paths against the FP64 oracle: frozen-phonon CTEM intensity averaged over
configurations in one launch sequence, and CBED over specimens of different thickness and
probes of different convergence in one batch with FP64 diffraction sums."""
import numpy as np
import pytest

import cltem_amd
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_rebuild_after_geometry_change(engine, oracle):
    """A configuration rebuilt under a geometry with more slices gets a larger buffer; the
    sweeps must read the new one (regression: the device pointer table kept the old one).
    Waves of a configuration not rebuilt after a geometry change are refused."""
    p = cltem_amd.load_params("kirkland")
    for nz in (3, 6):  # 11 then 17 slices
        w = wl.ctem_srtio3(3, 3, nz, resolution=256)
        g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
        engine.set_geometry(g)
        engine.set_parameterisation(p)
        for c in (0, 1):
            engine.sort_atoms(c, w["specimen"].x, w["specimen"].y, w["specimen"].z, w["specimen"].Z)
            engine.build_transmission(c)
        engine.set_wave_count(2)
        engine.set_wave_configs([0, 1])
        engine.init_plane_wave(1.0)
        engine.multislice()
        og = oracle.geometry(oracle.scenario(**w["scenario"]))
        srt = oracle.sort_atoms(og, w["specimen"].x, w["specimen"].y, w["specimen"].z, w["specimen"].Z)
        psi, _ = oracle.run_plane(oracle.transmission(og, srt, p.values, p.form, p.i_per_atom), oracle.propagator(og))
        for i in range(2):
            assert rel_l2(engine.exit_wave(i), psi) <= 1e-4, (nz, i)
    w = wl.ctem_srtio3(3, 3, 8, resolution=256)
    engine.set_geometry(cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"])))
    engine.set_wave_count(1)
    with pytest.raises(cltem_amd.MseError) as ex:
        engine.multislice()
    assert ex.value.code == -2


@pytest.mark.parametrize("full3d", [False, True])
def test_ctem_phonon_batch_mean_intensity(engine, oracle, full3d):
    w = wl.ctem_srtio3(3, 3, 4, resolution=256, params="lobato" if full3d else "kirkland")
    scen = dict(w["scenario"], full3d=int(full3d), full3d_integrals=8)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    og = oracle.geometry(oracle.scenario(**scen))
    p = cltem_amd.load_params(w["params"])
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    n_cfg = 4
    ref = np.zeros((256, 256))
    for c in range(n_cfg):
        s = w["specimen"].displaced(0.1, 31, c)
        engine.sort_atoms(c, s.x, s.y, s.z, s.Z)
        engine.build_transmission(c)
        srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
        psi, _ = oracle.run_plane(oracle.transmission(og, srt, p.values, p.form, p.i_per_atom),
                                  oracle.propagator(og))
        ref += np.abs(psi) ** 2
    engine.set_wave_count(n_cfg)
    engine.set_wave_configs(list(range(n_cfg)))
    engine.init_plane_wave(1.0)
    engine.multislice()
    got = engine.exit_intensity_sum()
    err = rel_l2(got / n_cfg, ref / n_cfg)
    print(f"phonon-averaged exit intensity (full3d={full3d}): rel-L2 {err:.3e}")
    assert err <= 1e-4
    # accumulation adds onto the caller's FP64 buffer
    acc = engine.exit_intensity_sum(got.copy())
    np.testing.assert_allclose(acc, 2 * got, rtol=1e-15)


def test_cbed_mixed_thickness_and_convergence(engine, oracle):
    cx = 1.5 * 3.905
    thick = [2, 5, 3]           # unit cells along z per instance; engine geometry = thickest
    alphas = [12.0, 20.0, 8.0]  # mrad per instance
    n_ph = 2
    specs = [sp.srtio3(3, 3, nz) for nz in thick]
    scens = []
    for s in specs:
        lo, hi = s.limits
        scens.append(dict(mode=3, resolution=256, slice_thickness=1.9525, struct_min=tuple(lo),
                          struct_max=tuple(hi), area=(cx, cx, cx, cx), area_padding=4.0))
    imax = int(np.argmax(thick))
    gmax = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scens[imax]))
    p = cltem_amd.load_params("kirkland")
    engine.set_geometry(gmax)
    engine.set_parameterisation(p)
    wave_cfg, wave_group = [], []
    for i, (s, sc) in enumerate(zip(specs, scens)):
        gi = cltem_amd.plan_geometry(cltem_amd.make_scenario(**sc))
        for ph in range(n_ph):
            c = i * n_ph + ph
            if i != imax:
                engine.set_config_geometry(c, gi)
            d = s.displaced(0.08, 5, ph)
            engine.sort_atoms(c, d.x, d.y, d.z, d.Z)
            engine.build_transmission(c)
            wave_cfg.append(c)
            wave_group.append(i)
    nw = len(wave_cfg)
    engine.set_wave_count(nw)
    engine.set_wave_configs(wave_cfg)
    for i in range(len(specs)):
        engine.init_probes_range(i * n_ph, np.tile([cx, cx], (n_ph, 1)), cltem_amd.make_probe(alphas[i]))
    engine.multislice()
    got = engine.diffraction_sum(wave_group, len(specs))
    for i, (s, sc) in enumerate(zip(specs, scens)):
        og = oracle.geometry(oracle.scenario(**sc))
        ref = np.zeros((256, 256))
        for ph in range(n_ph):
            d = s.displaced(0.08, 5, ph)
            srt = oracle.sort_atoms(og, d.x, d.y, d.z, d.Z)
            tr = oracle.transmission(og, srt, p.values, 0, 3)
            psi = oracle.fft2(oracle.probe_kspace(og, cx - og.sim_x[0], cx - og.sim_y[0], None, alphas[i]),
                              inverse=True)
            P_ = oracle.propagator(og)
            for sl in range(og.n_slices):  # the instance's own slices only
                psi, recip = oracle.slice_step(psi, tr[sl], P_)
            ref += oracle.diffraction(recip)
        err = rel_l2(got[i], ref)
        print(f"instance {i}: {og.n_slices} of {gmax.n_slices} slices, alpha {alphas[i]} mrad: rel-L2 {err:.3e}")
        assert err <= 1e-3


def test_graph_slice_loop_replays_match_eager(engine, oracle):
    """mse_multislice_graph: every replay (with eager work and synchronisation between the
    replays) reproduces the eager slice loop bit for bit, for waves of several phonon
    configurations and thicknesses."""
    w = wl.ctem_srtio3(3, 3, 6, resolution=256)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    p = cltem_amd.load_params(w["params"])
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    for c in range(3):
        s = w["specimen"].displaced(0.08, 17, c)
        engine.sort_atoms(c, s.x, s.y, s.z, s.Z)
        engine.build_transmission(c)
    engine.set_wave_count(6)
    engine.set_wave_configs([0, 1, 2, 0, 1, 2])
    engine.init_plane_wave(1.0)
    engine.multislice()
    ref = np.stack([engine.exit_wave(i) for i in range(6)])
    for rep in range(3):
        engine.init_plane_wave(1.0)
        engine.multislice(graph=True)
        got = np.stack([engine.exit_wave(i) for i in range(6)])
        assert np.array_equal(got, ref), f"replay {rep}"
        # eager work between replays: a different slice range, then the full loop again
        engine.init_plane_wave(1.0)
        engine.multislice(0, 3)
        engine.synchronize()
    # a new slice range is captured anew
    engine.init_plane_wave(1.0)
    engine.multislice(0, 5, graph=True)
    part = engine.exit_wave(0)
    engine.init_plane_wave(1.0)
    engine.multislice(0, 5)
    assert np.array_equal(part, engine.exit_wave(0))


def test_amorphous_sio2_full3d_1024(engine, oracle):
    """The full-size amorphous SiO2 case (20000 atoms, 1024^2, 106 slices, full-3D Lobato,
    20 integrals, 300 kV) with 3 frozen-phonon configurations through the graph-captured
    loop: transmission of a mid-specimen slice against the FP64 oracle, graph against eager
    slice loop, and the configuration mean against per-configuration runs."""
    w = wl.ctem_amorphous_sio2(n_configs=3)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    assert g.n_slices == 106 and g.resolution == 1024
    p = cltem_amd.load_params(w["params"])
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    n_cfg = w["n_configs"]
    for c in range(n_cfg):
        s = w["specimen"].displaced(w["u_rms"], w["seed"], c)
        engine.sort_atoms(c, s.x, s.y, s.z, s.Z)
        engine.build_transmission(c)
    # A3 parity at full size on one slice of configuration 0
    s0 = w["specimen"].displaced(w["u_rms"], w["seed"], 0)
    srt = oracle.sort_atoms(og, s0.x, s0.y, s0.z, s0.Z)
    ref_t = oracle.transmission(og, srt, p.values, p.form, p.i_per_atom, first=53, count=1)[0]
    err_t = rel_l2(engine.read_transmission(0, 53), ref_t)
    print(f"slice 53 transmission vs FP64 oracle: rel-L2 {err_t:.3e}")
    assert err_t <= 1e-5
    engine.set_wave_count(n_cfg)
    engine.set_wave_configs(list(range(n_cfg)))
    engine.init_plane_wave(1.0)
    engine.multislice(graph=True)
    mean = engine.exit_intensity_sum() / n_cfg
    per = np.zeros_like(mean)
    engine.set_wave_count(1)
    for c in range(n_cfg):
        engine.set_wave_configs([c])
        engine.init_plane_wave(1.0)
        engine.multislice()
        per += np.abs(engine.exit_wave(0)) ** 2
    err = rel_l2(mean, per / n_cfg)
    print(f"graph-batched configuration mean vs per-configuration eager runs: rel-L2 {err:.3e}")
    assert err <= 1e-6
    assert np.all(np.isfinite(mean)) and 0.5 < mean.mean() < 1.5


@pytest.mark.parametrize("order", ["thin_first", "thick_first"])
def test_thinner_specimens_stop_at_their_exit_surface(engine, oracle, order):
    """Waves of a thinner specimen batched with a thicker one hold their own exit wave (the
    result of a separate run over their own slices), whether the waves are ordered by
    thickness (prefix sets) or not (device wave lists), eagerly and from the graph."""
    thick = wl.ctem_srtio3(3, 3, 6, resolution=256)
    thin = wl.ctem_srtio3(3, 3, 3, resolution=256)
    gk = cltem_amd.plan_geometry(cltem_amd.make_scenario(**thick["scenario"]))
    gn = cltem_amd.plan_geometry(cltem_amd.make_scenario(**thin["scenario"]))
    assert gn.n_slices < gk.n_slices
    p = cltem_amd.load_params("kirkland")
    engine.set_geometry(gk)
    engine.set_parameterisation(p)
    engine.sort_atoms(0, thick["specimen"].x, thick["specimen"].y, thick["specimen"].z, thick["specimen"].Z)
    engine.build_transmission(0)
    engine.set_config_geometry(1, gn)
    engine.sort_atoms(1, thin["specimen"].x, thin["specimen"].y, thin["specimen"].z, thin["specimen"].Z)
    engine.build_transmission(1)
    cfgs = [1, 0, 1] if order == "thin_first" else [0, 1, 1]
    refs = {}
    for c, w in ((0, thick), (1, thin)):
        og = oracle.geometry(oracle.scenario(**w["scenario"]))
        srt = oracle.sort_atoms(og, w["specimen"].x, w["specimen"].y, w["specimen"].z, w["specimen"].Z)
        refs[c], _ = oracle.run_plane(oracle.transmission(og, srt, p.values, p.form, p.i_per_atom),
                                      oracle.propagator(og))
    engine.set_wave_count(3)
    engine.set_wave_configs(cfgs)
    for graph in (False, True, True):
        engine.init_plane_wave(1.0)
        engine.multislice(graph=graph)
        for i, c in enumerate(cfgs):
            err = rel_l2(engine.exit_wave(i), refs[c])
            assert err <= 1e-4, (order, graph, i, err)
    print(f"{order}: thin {gn.n_slices} and thick {gk.n_slices} slices match separate FP64 runs")


def test_mixed_size_cbed_driver(oracle):
    """cltem_amd.mixed: instances of all four grid sizes (256, 384, 512, 768) and several
    thicknesses and convergence angles, the slice loops of every size in one graph launch
    (each size with its own autotuned kernels), two timed whole jobs; every phonon-averaged
    pattern against the FP64 oracle."""
    from cltem_amd.mixed import run_mixed_cbed
    inst = wl.mixed_cbed(n_instances=6, cells_xy=3, thickness_cells=(2, 5), sizes=(256, 384, 512, 768),
                         n_configs=2)
    for i, res in enumerate((256, 384, 512, 768)):
        inst[i].scenario["resolution"] = res
    timing = []
    patterns, groups = run_mixed_cbed(inst, warmup=1, repeats=1, timing=timing)
    assert set(groups) == {256, 384, 512, 768}
    assert timing[0].job_s > 0 and timing[0].wave_slices == sum(r.wave_slices for r in groups.values())
    p = cltem_amd.load_params("kirkland")
    for i, it in enumerate(inst):
        og = oracle.geometry(oracle.scenario(**it.scenario))
        P_ = oracle.propagator(og)
        ref = np.zeros((og.resolution, og.resolution))
        for ph in range(it.n_configs):
            d = it.specimen.displaced(it.u_rms, it.seed, ph)
            srt = oracle.sort_atoms(og, d.x, d.y, d.z, d.Z)
            tr = oracle.transmission(og, srt, p.values, p.form, p.i_per_atom)
            psi = oracle.fft2(oracle.probe_kspace(og, it.position[0] - og.sim_x[0], it.position[1] - og.sim_y[0],
                                                  None, it.alpha_mrad), inverse=True)
            for sl in range(og.n_slices):
                psi, recip = oracle.slice_step(psi, tr[sl], P_)
            ref += oracle.diffraction(recip)
        ref /= it.n_configs
        err = rel_l2(patterns[i], ref)
        print(f"instance {i}: {og.resolution}^2, {og.n_slices} slices, alpha {it.alpha_mrad:.1f} mrad: rel-L2 {err:.3e}")
        assert err <= 1e-3


def test_waves_past_a_stem_batch_stay_consistent(engine):
    """mse_stem_batch with fewer waves than the engine holds: every wave then refers to the
    batch's configuration, and per-wave calls on the waves past the batch stay valid."""
    w = wl.stem_srtio3(nx=3, ny=3, nz=2, resolution=256, pixels=2, window_cells=(1, 2))
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    engine.set_geometry(g)
    engine.set_parameterisation(cltem_amd.load_params("kirkland"))
    s = w["specimen"]
    engine.sort_atoms(0, s.x, s.y, s.z, s.Z)
    engine.build_transmission(0)
    engine.set_wave_count(6)
    engine.set_wave_configs([0] * 6)
    engine.init_plane_wave(1.0)
    got = engine.stem_batch(0, wl.stem_positions(g), 2, cltem_amd.make_probe(20.0),
                            cltem_amd.make_detectors([(60.0, 200.0, 0.0, 0.0)]))
    assert np.all(got > 0)
    d = engine.diffraction(5)  # a plane wave untouched by the batches: one bright pixel
    assert np.isclose(d.sum(), d.max(), rtol=1e-6) and d.max() > 0
    engine.multislice(0, 1)


def test_stem_batch_on_a_fresh_engine_after_autotune(oracle):
    """The benchmark's order on a new engine: autotune (which sizes the waves itself), then
    the batched STEM path, whose probe set-ups also write the per-wave draw keys."""
    w = wl.stem_srtio3(nx=3, ny=3, nz=2, resolution=256, pixels=2, window_cells=(1, 2))
    with cltem_amd.Engine(0) as eng:
        g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
        eng.set_geometry(g)
        p = cltem_amd.load_params("kirkland")
        eng.set_parameterisation(p)
        s = w["specimen"]
        eng.sort_atoms(0, s.x, s.y, s.z, s.Z)
        eng.build_transmission(0)
        eng.autotune(0, 4, reps=1)
        pos = wl.stem_positions(g)
        got = eng.stem_batch(0, pos, 2, cltem_amd.make_probe(20.0),
                             cltem_amd.make_detectors([(60.0, 200.0, 0.0, 0.0)]), use_graph=True)
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
    ref = oracle.run_stem(og, oracle.transmission(og, srt, p.values, p.form, p.i_per_atom), oracle.propagator(og),
                          pos, np.array([[60.0, 200.0, 0.0, 0.0]]), cond_ap_mrad=20.0)
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-3
