"""The metric's own configuration against the FP64 oracle: batched HAADF STEM on SrTiO3
16x16x40 (51,200 atoms), 512^2, 83 slices, 200 kV, 20 mrad probe, 64x64 raster over one
unit cell, 60-200 mrad detector, frozen phonons of u_rms 0.08 A -- exactly
workloads.stem_srtio3() -- through mse_stem_batch (SimulationStem::simulate,
microscope/simulationstem.cpp:143-278) at B = 16, 64 and 256 with graph replays, and with
the column-sweep variants 7, 10, 11 and the autotuner's pick.

Every run covers the whole 4096-probe raster; 16 probes on a 4x4 sub-grid of the raster
(corners and interior) are checked against the oracle for two phonon configurations
(detector tolerance 1e-3), and all 4096 probes are checked across batch sizes and kernel
variants (1e-5).  The oracle's per-configuration references are computed once per module.
"""
import numpy as np
import pytest

import cltem_amd
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu

SAMPLE = np.array([y * 64 + x for y in (0, 21, 42, 63) for x in (0, 21, 42, 63)])
DET_TOL = 1e-3      # detector intensities vs FP64 (north star)
CROSS_TOL = 1e-5    # the same FP32 arithmetic in another batching or kernel variant


def max_rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))))


class Config2:
    def __init__(self, oracle):
        self.oracle = oracle
        self.w = wl.stem_srtio3()
        self.g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**self.w["scenario"]))
        self.og = oracle.geometry(oracle.scenario(**self.w["scenario"]))
        assert (self.g.resolution, self.g.n_slices) == (512, 83)
        assert len(self.w["specimen"]) == 51200
        self.p = cltem_amd.load_params(self.w["params"])
        self.pos = wl.stem_positions(self.g)
        assert len(self.pos) == 4096
        self.probe = cltem_amd.make_probe(self.w["cond_ap_mrad"])
        self.dets = cltem_amd.make_detectors(self.w["detectors"])
        self.det4 = np.array([list(d) + [0.0] * (4 - len(d)) for d in self.w["detectors"]])
        self._refs = {}

    def specimen(self, c):
        return self.w["specimen"].displaced(self.w["u_rms"], self.w["seed"], c)

    def reference(self, c):
        """FP64 oracle detector values of the 16 sampled probes for phonon configuration c."""
        if c not in self._refs:
            orc = self.oracle
            s = self.specimen(c)
            srt = orc.sort_atoms(self.og, s.x, s.y, s.z, s.Z)
            tr = orc.transmission(self.og, srt, self.p.values, self.p.form, self.p.i_per_atom)
            self._refs[c] = orc.run_stem(self.og, tr, orc.propagator(self.og), self.pos[SAMPLE], self.det4,
                                         cond_ap_mrad=self.w["cond_ap_mrad"])
        return self._refs[c]

    def load(self, engine, c):
        engine.set_geometry(self.g)
        engine.set_parameterisation(self.p)
        s = self.specimen(c)
        engine.sort_atoms(0, s.x, s.y, s.z, s.Z, want_ids=False)
        engine.build_transmission(0)

    def run(self, engine, batch, graph=True):
        return engine.stem_batch(0, self.pos, batch, self.probe, self.dets, use_graph=graph)


@pytest.fixture(scope="module")
def cfg2(oracle):
    return Config2(oracle)


@pytest.mark.parametrize("c", [0, 1])
def test_config2_batches_against_oracle(engine, cfg2, c):
    """B = 16, 64, 256 (graph replays): 16 sampled probes against the FP64 oracle for phonon
    configuration c, and the whole raster identical to 1e-5 across batch sizes."""
    cfg2.load(engine, c)
    engine.set_kernel_variant(0, 4)
    engine.set_kernel_variant(1, 7)
    ref = cfg2.reference(c)
    runs = {}
    for B in (16, 64, 256):
        got = cfg2.run(engine, B)
        assert got.shape == (4096, 1) and np.all(np.isfinite(got)) and np.all(got > 0)
        err = max_rel(got[SAMPLE], ref)
        print(f"config {c}, B={B}: 16 sampled probes vs FP64, max rel {err:.2e}")
        assert err <= DET_TOL, (c, B, err)
        runs[B] = got
    for B in (16, 256):
        cross = max_rel(runs[B], runs[64])
        print(f"config {c}: B={B} vs B=64 over 4096 probes, max rel {cross:.2e}")
        assert cross <= CROSS_TOL, (c, B, cross)
    # the eager path (no graph) over the whole raster equals the replayed graphs
    eager = cfg2.run(engine, 64, graph=False)
    assert max_rel(eager, runs[64]) <= CROSS_TOL


def test_config2_column_variants_and_autotune(engine, cfg2):
    """Column variants 7 (persistent, prefetching), 10 (8-column tiles), 11 (one tile per block,
    two blocks per CU) and the autotuner's pick at B = 64, phonon configuration 0."""
    cfg2.load(engine, 0)
    ref = cfg2.reference(0)
    engine.set_kernel_variant(0, 4)
    base = None
    for v in (7, 10, 11):
        engine.set_kernel_variant(1, v)
        got = cfg2.run(engine, 64)
        err = max_rel(got[SAMPLE], ref)
        print(f"column variant {v}: 16 sampled probes vs FP64, max rel {err:.2e}")
        assert err <= DET_TOL, (v, err)
        if base is None:
            base = got
        else:
            assert max_rel(got, base) <= CROSS_TOL, v
    tune = engine.autotune(0, 64, reps=2)
    print(f"autotune: row {tune['best_row']}, column {tune['best_col']}")
    got = cfg2.run(engine, 64)
    err = max_rel(got[SAMPLE], ref)
    print(f"autotuned variants: 16 sampled probes vs FP64, max rel {err:.2e}")
    assert err <= DET_TOL
    assert max_rel(got, base) <= CROSS_TOL


def test_config2_phonon_accumulation(engine, cfg2):
    """Two configurations accumulated into one FP64 buffer (updateImages): the sum of the
    per-configuration references."""
    acc = np.zeros((4096, 1))
    for c in (0, 1):
        cfg2.load(engine, c)
        engine.stem_batch(0, cfg2.pos, 64, cfg2.probe, cfg2.dets, use_graph=True, accum=acc)
    ref = cfg2.reference(0) + cfg2.reference(1)
    err = max_rel(acc[SAMPLE], ref)
    print(f"accumulated configurations 0+1: max rel {err:.2e}")
    assert err <= DET_TOL


def test_config2_specimen_ctem_exit_wave_512(engine, oracle):
    """The config-2 specimen (SrTiO3 16x16x40) as a 512^2 plane-wave CTEM run over its 83
    slices: exit wave rel-L2 <= 1e-4 and diffraction <= 1e-3 against the FP64 oracle."""
    w = wl.ctem_srtio3(16, 16, 40, resolution=512)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    assert (g.resolution, g.n_slices) == (512, 83)
    p = cltem_amd.load_params(w["params"])
    s = w["specimen"]
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    engine.sort_atoms(0, s.x, s.y, s.z, s.Z, want_ids=False)
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    engine.multislice(graph=True)
    ew = engine.exit_wave(0)
    dif = engine.diffraction(0)
    srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
    psi, recip = oracle.run_plane(oracle.transmission(og, srt, p.values, p.form, p.i_per_atom),
                                  oracle.propagator(og))
    e1 = float(np.linalg.norm(ew - psi) / np.linalg.norm(psi))
    d = oracle.diffraction(recip)
    e2 = float(np.linalg.norm(dif - d) / np.linalg.norm(d))
    print(f"512^2 CTEM, 83 slices: exit wave rel-L2 {e1:.2e}, diffraction rel-L2 {e2:.2e}")
    assert e1 <= 1e-4
    assert e2 <= 1e-3
