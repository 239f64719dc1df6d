"""Plasmon beam tilt against the FP64 oracle.  A wave's beam vector turns between slices
(modifyBeamTilt): its configuration's later transmission functions and its propagator follow
the new vector, and its diffraction pattern and detector values are read translated back
(translateDiffImage).  Engine-level runs with fixed scatterings, and the C++ host's CBED and
STEM loops with seeded draws replayed through the oracle."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import cltem_amd
from cltem_amd import incoherence as inc
from cltem_amd import specimens as sp

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "cltem_amd", "mse_run")
PARAMS = os.path.join(ROOT, "cltem_amd", "params")
CX = 1.5 * 3.905


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def cbed_scenario(spec, res=256, **kw):
    lo, hi = spec.limits
    return dict(mode=3, resolution=res, slice_thickness=1.9525, struct_min=tuple(lo), struct_max=tuple(hi),
                area=(CX, CX, CX, CX), **kw)


def tilted_geometry(oracle, scen, og0, k):
    """The oracle's geometry for beam vector k.  modifyBeamTilt leaves the interaction
    constant of the potentials at the original beam's, so sigma_eff is kept."""
    tilt = 1000.0 * math.atan2(math.hypot(k[0], k[1]), k[2])
    og = oracle.geometry(oracle.scenario(**dict(scen, beam_tilt_mrad=tilt, beam_azimuth=math.atan2(k[1], k[0]))))
    assert tuple(og.full_x) == tuple(og0.full_x) and og.n_slices == og0.n_slices
    og.sigma_eff = og0.sigma_eff
    return og


def oracle_moved_diffraction(oracle, scen, spec, p, pos, alpha, events, aberr=None):
    """FP64 CBED with the beam vector replaced after the slices in `events` ({slice: k}), and
    the diffraction pattern translated by the final change of the beam vector."""
    og0 = oracle.geometry(oracle.scenario(**scen))
    srt = oracle.sort_atoms(og0, spec.x, spec.y, spec.z, spec.Z)
    tr = oracle.transmission(og0, srt, p.values, p.form, p.i_per_atom)
    P = oracle.propagator(og0)
    psi = oracle.fft2(oracle.probe_kspace(og0, pos[0] - og0.sim_x[0], pos[1] - og0.sim_y[0], aberr, alpha),
                      inverse=True)
    k = np.array(og0.k_vec[:])
    for sl in range(og0.n_slices):
        psi, recip = oracle.slice_step(psi, tr[sl], P)
        if sl in events:
            k = np.asarray(events[sl], dtype=np.float64)
            og = tilted_geometry(oracle, scen, og0, k)
            tr = oracle.transmission(og, srt, p.values, p.form, p.i_per_atom)
            P = oracle.propagator(og)
    diff = oracle.diffraction(recip)
    return oracle.translate_image(diff, k[0] - og0.k_vec[0], k[1] - og0.k_vec[1], og0.inverse_scale), og0


def run_engine_with_events(engine, g, events):
    """Slices in segments; after each event slice configuration 0 takes the new beam vector
    and its remaining transmission functions are rebuilt."""
    n = g.n_slices
    first = 0
    for sl in sorted(events):
        engine.multislice(first, sl + 1 - first)
        first = sl + 1
        engine.set_config_beam(0, events[sl])
        if sl + 1 < n:
            engine.build_transmission(0, sl + 1, n - sl - 1)
    if first < n:
        engine.multislice(first, n - first)


def two_scatterings(g):
    k, y = inc.initial_beam(g.wavenumber, g.beam_tilt_mrad, g.beam_azimuth)
    ev = {}
    for sl, (th, ph) in ((3, (9.0, 0.7)), (7, (6.0, 2.0))):
        k, y = inc.rotate_vector_spherical(k, y, th / 1000.0, ph)
        ev[sl] = k
    return ev


# 384 and 768 points: the column kernels read per-wave propagator rows from memory there
# (more than 8 elements per lane) instead of holding them in registers
@pytest.mark.parametrize("params,full3d,res", [("kirkland", False, 256), ("lobato", True, 256),
                                               ("kirkland", False, 384), ("kirkland", False, 768)])
def test_cbed_two_scatterings_match_oracle(engine, oracle, params, full3d, res):
    spec = sp.srtio3(3, 3, 4)
    scen = cbed_scenario(spec, res=res, full3d=int(full3d), full3d_integrals=4)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    p = cltem_amd.load_params(params)
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    probe = cltem_amd.make_probe(15.0)
    engine.init_probes([[CX, CX]], probe)
    events = two_scatterings(g)
    run_engine_with_events(engine, g, events)
    got = engine.diffraction(0)
    ref, og0 = oracle_moved_diffraction(oracle, scen, spec, p, (CX, CX), 15.0, events)
    err = rel_l2(got, ref)
    k = events[7]
    shift = ((k[0] - og0.k_vec[0]) / og0.inverse_scale, (k[1] - og0.k_vec[1]) / og0.inverse_scale)
    print(f"{params} full3d={full3d} {res}^2: two scatterings, pattern moved by ({shift[0]:.2f}, {shift[1]:.2f}) px, "
          f"rel-L2 {err:.3e}")
    assert err <= 1e-4
    # the scatterings are visible: the untilted run differs
    engine.set_config_beam(0, None)
    engine.build_transmission(0)
    engine.init_probes([[CX, CX]], probe)
    engine.multislice()
    assert rel_l2(engine.diffraction(0), ref) > 0.05


def test_detectors_over_moved_patterns(engine, oracle):
    """STEM detectors of waves with a beam of their own read the translated pattern: with
    the geometry's own vector they equal the fused detector path; with a moved vector they
    equal the annular sums of the translated diffraction patterns."""
    spec = sp.srtio3(3, 3, 3)
    scen = cbed_scenario(spec)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    engine.set_geometry(g)
    engine.set_parameterisation(cltem_amd.load_params("kirkland"))
    for c in (0, 1):
        engine.sort_atoms(c, spec.x, spec.y, spec.z, spec.Z)
        engine.build_transmission(c)
    engine.set_wave_count(3)
    engine.set_wave_configs([0, 1, 1])
    engine.init_probes([[CX, CX], [CX + 0.4, CX], [CX, CX - 0.7]], cltem_amd.make_probe(18.0))
    engine.multislice()
    rows = [(0, 15, 0, 0), (40, 150, 0, 0), (10, 30, 5, -3)]
    dets = cltem_amd.make_detectors(rows)
    fused = engine.stem_detect(dets)
    engine.set_config_beam(1, list(g.k_vec))
    same = engine.stem_detect(dets)
    assert np.max(np.abs(same - fused) / np.abs(fused)) <= 1e-5
    plain = {w: engine.diffraction(w) for w in (1, 2)}  # no move: the geometry's own vector
    k, _ = inc.rotate_vector_spherical(*inc.initial_beam(g.wavenumber, 0.0, 0.0), 0.011, 0.4)
    engine.set_config_beam(1, k)
    moved = engine.stem_detect(dets)
    np.testing.assert_array_equal(moved[0], same[0])  # configuration 0 keeps its beam
    W = g.resolution
    yy, xx = np.mgrid[0:W, 0:W]
    for w in (1, 2):
        d = engine.diffraction(w)
        ref = oracle.translate_image(plain[w], k[0] - g.k_vec[0], k[1] - g.k_vec[1], g.inverse_scale)
        assert rel_l2(d, ref) <= 1e-12
        for j, (inner, outer, xc, yc) in enumerate(rows):
            cx, cy = W / 2 + xc / g.inverse_scale_angle, W / 2 + yc / g.inverse_scale_angle
            r = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2)
            mask = (r <= outer / g.inverse_scale_angle) & (r >= inner / g.inverse_scale_angle)
            want = math.sqrt(2.0) * d[mask].sum()
            assert abs(moved[w, j] - want) <= 1e-9 * abs(want), (w, j)
        assert np.max(np.abs(moved[w] - same[w]) / np.abs(same[w])) > 1e-3
    engine.set_config_beam(1, None)


def test_beam_vector_is_checked(engine):
    spec = sp.srtio3(2, 2, 2)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**cbed_scenario(spec)))
    engine.set_geometry(g)
    with pytest.raises(cltem_amd.MseError):
        engine.set_config_beam(0, [0.0, 0.0, 2 * g.wavenumber])  # not the wavenumber
    with pytest.raises(cltem_amd.MseError):
        engine.set_config_beam(0, [g.wavenumber, 0.0, 0.0])  # kz = 0


def _load(prefix):
    meta = json.load(open(prefix + ".json"))
    return meta, {k: np.fromfile(v["file"], dtype=np.float64) for k, v in meta.items() if isinstance(v, dict)}


def _host_events(g, spec, seed, job, mfp, te, tc, tilt=0.0, az=0.0):
    """The C++ host's draws for one job, from the Python restatement of the sampler."""
    z = spec.z
    depths = inc.plasmon_depths(seed, job, float(z.max() - z.min()), mfp)
    slices = inc.plasmon_events(depths, g.pre_slices, g.slice_thickness, g.n_slices)
    polar, azim = inc.plasmon_angles(seed, job, len(slices), te, tc)
    k, y = inc.initial_beam(g.wavenumber, tilt, az)
    ev = {}
    for sl, th, ph in zip(slices, polar, azim):
        k, y = inc.rotate_vector_spherical(k, y, th / 1000.0, ph)
        ev[sl] = k
    return ev


@pytest.mark.skipif(not os.path.exists(RUN), reason="mse_run not built")
def test_driver_cbed_with_plasmons_matches_oracle(tmp_path, oracle):
    prefix = str(tmp_path / "cbed_pl")
    seed = 5
    subprocess.run([RUN, "--mode", "cbed", "--srtio3", "3,3,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{CX},{CX},{CX},{CX}", "--alpha", "15", "--seed", str(seed),
                    "--plasmon", "0.4,2.0,20.0", "--params", os.path.join(PARAMS, "kirkland.dat"),
                    "--out", prefix], check=True)
    _, out = _load(prefix)
    spec = sp.srtio3(3, 3, 3)
    scen = cbed_scenario(spec)
    og = oracle.geometry(oracle.scenario(**scen))
    ev = _host_events(og, spec, seed, 0, 4.0, 2.0, 20.0)
    assert len(ev) >= 1
    ref, _ = oracle_moved_diffraction(oracle, scen, spec, cltem_amd.load_params("kirkland"), (CX, CX), 15.0, ev)
    err = rel_l2(out["Diff"], ref)
    print(f"host CBED with {len(ev)} plasmon scatterings: rel-L2 {err:.3e}")
    assert err <= 1e-4


@pytest.mark.skipif(not os.path.exists(RUN), reason="mse_run not built")
def test_driver_stem_with_plasmons_matches_oracle(tmp_path, oracle):
    """Every batch of probes is one job with its own scattering history."""
    prefix = str(tmp_path / "stem_pl")
    seed = 8
    x0, x1 = 3.905, 2 * 3.905
    subprocess.run([RUN, "--mode", "stem", "--srtio3", "3,3,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{x0},{x1},{x0},{x1}", "--stem-pixels", "2", "--det", "ADF,40,150",
                    "--det", "BF,0,12", "--alpha", "18", "--batch", "2", "--seed", str(seed),
                    "--plasmon", "0.4,2.0,20.0", "--params", os.path.join(PARAMS, "kirkland.dat"),
                    "--out", prefix], check=True)
    _, out = _load(prefix)
    spec = sp.srtio3(3, 3, 3)
    lo, hi = spec.limits
    scen = dict(mode=2, resolution=256, slice_thickness=1.9525, struct_min=tuple(lo), struct_max=tuple(hi),
                area=(x0, x1, x0, x1), stem_pixels=(2, 2))
    og = oracle.geometry(oracle.scenario(**scen))
    p = cltem_amd.load_params("kirkland")
    W = og.resolution
    yy, xx = np.mgrid[0:W, 0:W]
    masks = []
    for inner, outer in ((40.0, 150.0), (0.0, 12.0)):
        r = np.sqrt((xx - W / 2) ** 2 + (yy - W / 2) ** 2)
        masks.append((r <= outer / og.inverse_scale_angle) & (r >= inner / og.inverse_scale_angle))
    n_events = 0
    for pix in range(4):
        job = pix // 2  # batches of 2 pixels, configuration 0
        ev = _host_events(og, spec, seed, job, 4.0, 2.0, 20.0)
        n_events += len(ev) if pix % 2 == 0 else 0
        pos = (og.stem_start[0] + (pix % 2) * og.stem_scale[0], og.stem_start[1] + (pix // 2) * og.stem_scale[1])
        d, _ = oracle_moved_diffraction(oracle, scen, spec, p, pos, 18.0, ev)
        for name, m in zip(("ADF", "BF"), masks):
            want = math.sqrt(2.0) * d[m].sum()
            got = out[name][pix]
            assert abs(got - want) <= 1e-3 * abs(want), (pix, name, got, want)
    assert n_events >= 1
