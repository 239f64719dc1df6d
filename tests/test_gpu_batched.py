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
