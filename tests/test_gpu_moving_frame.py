"""Moving STEM frame on the GPU (clTEM with parallel_stem off): mse_stem_batch_frames against
the FP64 oracle run pixel by pixel as clTEM's jobs are (each pixel's own field, transmission
functions rebuilt for it: SimulationGeneral::initialiseSimulation,
microscope/simulationgeneral.cpp:345-358, and SimulationStem::simulate,
microscope/simulationstem.cpp:143-278), and against mse_stem_batch when every frame is the
parallel field."""
import numpy as np
import pytest

import cltem_amd
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu

DET_TOL = 1e-3
PIXELS = [0, 5, 18, 27, 36, 45, 54, 63]
DETS = ((60.0, 200.0), (0.0, 20.0), (10.0, 45.0, 7.5, -12.5))


def _workload(moving):
    w = wl.stem_srtio3(nx=6, ny=6, nz=6, resolution=256, pixels=8, window_cells=(2, 3), detectors=DETS)
    w["scenario"] = dict(w["scenario"], moving_stem_frame=int(moving), area_padding=1.5)
    return w


def _load(engine, w, spec):
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    p = cltem_amd.load_params(w["params"])
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z, want_ids=False)
    return g, p


def test_moving_frame_matches_oracle_per_pixel(engine, oracle):
    w = _workload(True)
    spec = w["specimen"].displaced(0.08, 20261202, 0)
    g, p = _load(engine, w, spec)
    assert g.moving_frame == 1
    pos = wl.stem_positions(g)[PIXELS]
    frames = cltem_amd.stem_pixel_frames(g, PIXELS)
    assert len(np.unique(frames[:, 0])) > 4
    probe = cltem_amd.make_probe(w["cond_ap_mrad"])
    dets = cltem_amd.make_detectors(DETS)
    got4 = engine.stem_batch_frames(0, pos, frames, 4, probe, dets)
    got8 = engine.stem_batch_frames(0, pos, frames, 8, probe, dets)
    assert np.max(np.abs(got4 - got8) / np.abs(got8)) <= 1e-6
    s = oracle.scenario(**w["scenario"])
    det4 = np.array([list(d) + [0.0] * (4 - len(d)) for d in DETS])
    ref = np.zeros_like(got4)
    for i, px in enumerate(PIXELS):
        gp = oracle.geometry_pixel(s, px)
        assert (gp.sim_x[0], gp.sim_y[0]) == tuple(frames[i])
        srt = oracle.sort_atoms(gp, spec.x, spec.y, spec.z, spec.Z)
        tr = oracle.transmission(gp, srt, p.values, p.form, p.i_per_atom)
        ref[i] = oracle.run_stem(gp, tr, oracle.propagator(gp), pos[i:i + 1], det4,
                                 cond_ap_mrad=w["cond_ap_mrad"])[0]
    err = float(np.max(np.abs(got4 - ref) / np.abs(ref)))
    print(f"moving frame, 8 pixels x 3 detectors vs FP64 per-pixel jobs: max rel {err:.2e}")
    assert err <= DET_TOL
    # the frames matter: the same probes in pixel 0's field alone give other intensities
    same = engine.stem_batch_frames(0, pos, np.repeat(frames[:1], len(PIXELS), axis=0), 8, probe, dets)
    assert np.max(np.abs(same - ref) / np.abs(ref)) > 10 * DET_TOL


def test_frames_of_the_parallel_field_equal_stem_batch(engine):
    w = _workload(False)
    spec = w["specimen"].displaced(0.08, 20261202, 1)
    g, _ = _load(engine, w, spec)
    engine.build_transmission(0)
    pos = wl.stem_positions(g)[:12]
    frames = cltem_amd.stem_pixel_frames(g, range(12))
    probe = cltem_amd.make_probe(w["cond_ap_mrad"])
    dets = cltem_amd.make_detectors(DETS)
    a = engine.stem_batch(0, pos, 4, probe, dets, use_graph=False)
    b = engine.stem_batch_frames(0, pos, frames, 4, probe, dets)
    err = float(np.max(np.abs(a - b) / np.abs(a)))
    print(f"parallel field through the frame path vs mse_stem_batch: max rel {err:.2e}")
    assert err <= 1e-6
