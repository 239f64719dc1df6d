"""GPU parity: the HIP path against the FP64 oracle on the same inputs."""
import numpy as np
import pytest

import cltem_amd
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def setup_case(engine, oracle, w, config=0, spec=None):
    scen = w["scenario"]
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    og = oracle.geometry(oracle.scenario(**scen))
    p = cltem_amd.load_params(w["params"])
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    spec = w["specimen"] if spec is None else spec
    return g, og, p, spec


def test_binning_bit_exact_config1(engine, oracle):
    w = wl.ctem_srtio3()
    g, og, p, spec = setup_case(engine, oracle, w)
    r = engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    o = oracle.sort_atoms(og, spec.x, spec.y, spec.z, spec.Z)
    assert r["n_in_range"] == o["n_in_range"]
    np.testing.assert_array_equal(r["block_start"], o["block_start"])
    keep = r["block_ids"] >= 0
    np.testing.assert_array_equal(r["block_ids"][keep], o["block_ids"])
    np.testing.assert_array_equal(r["z_ids"][keep], o["z_ids"])


def test_transmission_config1_slices(engine, oracle):
    w = wl.ctem_srtio3()
    g, og, p, spec = setup_case(engine, oracle, w)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    o = oracle.sort_atoms(og, spec.x, spec.y, spec.z, spec.Z)
    for s in (0, 2, 3, 21, 40):
        ref = oracle.transmission(og, o, p.values, p.form, p.i_per_atom, first=s, count=1)[0]
        got = engine.read_transmission(0, s)
        err = rel_l2(got, ref)
        print(f"slice {s}: rel-L2 {err:.3e}")
        assert err <= 1e-5


def test_ctem_exit_wave_config1(engine, oracle):
    w = wl.ctem_srtio3()
    g, og, p, spec = setup_case(engine, oracle, w)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    engine.multislice()
    ew = engine.exit_wave(0)
    dif = engine.diffraction(0)
    o = oracle.sort_atoms(og, spec.x, spec.y, spec.z, spec.Z)
    trans = oracle.transmission(og, o, p.values, p.form, p.i_per_atom)
    P = oracle.propagator(og)
    psi, recip = oracle.run_plane(trans, P)
    e1 = rel_l2(ew, psi)
    e2 = rel_l2(dif, oracle.diffraction(recip))
    print(f"exit wave rel-L2 {e1:.3e}, diffraction rel-L2 {e2:.3e}")
    assert e1 <= 1e-4
    assert e2 <= 1e-3


def test_stem_small_vs_oracle(engine, oracle):
    w = wl.stem_srtio3(nx=6, ny=6, nz=6, resolution=256, pixels=8, window_cells=(2, 3),
                       detectors=((60.0, 200.0), (0.0, 20.0), (20.0, 60.0, 3.0, -2.0)))
    g, og, p, spec = setup_case(engine, oracle, w)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    pos = wl.stem_positions(g)[::7]
    probe = cltem_amd.make_probe(w["cond_ap_mrad"])
    dets = cltem_amd.make_detectors(w["detectors"])
    got = engine.stem_batch(0, pos, 4, probe, dets)
    o = oracle.sort_atoms(og, spec.x, spec.y, spec.z, spec.Z)
    trans = oracle.transmission(og, o, p.values, p.form, p.i_per_atom)
    P = oracle.propagator(og)
    det4 = np.array([list(d) + [0.0] * (4 - len(d)) for d in w["detectors"]])
    ref = oracle.run_stem(og, trans, P, pos, det4, cond_ap_mrad=w["cond_ap_mrad"])
    err = np.abs(got - ref) / np.abs(ref)
    print("detector rel err max", err.max(), got[:2], ref[:2])
    assert err.max() <= 1e-3
    # split path (init_probes / multislice / stem_detect) agrees with the fused batch
    engine.set_wave_count(len(pos))
    engine.set_wave_configs([0] * len(pos))
    engine.init_probes(pos, probe)
    engine.multislice()
    split = engine.stem_detect(dets)
    assert np.abs(split - got).max() / np.abs(got).max() <= 1e-5


def test_graph_replays_match_eager_and_variants_agree(engine, oracle):
    w = wl.stem_srtio3(nx=6, ny=6, nz=6, resolution=256, pixels=8, window_cells=(2, 3),
                       detectors=((60.0, 200.0), (0.0, 20.0)))
    g, og, p, spec = setup_case(engine, oracle, w)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    pos = wl.stem_positions(g)[:32]
    probe = cltem_amd.make_probe(w["cond_ap_mrad"])
    dets = cltem_amd.make_detectors(w["detectors"])
    eager = engine.stem_batch(0, pos, 8, probe, dets, use_graph=False)
    for rep in range(3):
        # eager work and synchronisation between replays
        engine.set_wave_count(2)
        engine.set_wave_configs([0, 0])
        engine.init_probes(pos[:2], probe)
        engine.multislice(0, 2)
        engine.synchronize()
        replay = engine.stem_batch(0, pos, 8, probe, dets, use_graph=True)
        np.testing.assert_array_equal(replay, eager)
    # every kernel variant computes the same detector values
    for rv in range(7):
        for cv in range(14):
            engine.set_kernel_variant(0, rv)
            engine.set_kernel_variant(1, cv)
            out = engine.stem_batch(0, pos[:8], 8, probe, dets)
            assert np.max(np.abs(out - eager[:8]) / np.abs(eager[:8])) < 1e-5, (rv, cv)
    engine.set_kernel_variant(0, 4)
    engine.set_kernel_variant(1, 7)


@pytest.mark.parametrize("res", [512, 384, 768, 1024])
def test_column_variants_agree(engine, oracle, res):
    """The column kernels (persistent: prefetching, half lanes, 8-column tiles, XCD-aware 8- and
    16-column tiles; lean: two blocks per CU, 512 points) do the same butterflies with the same twiddles as the default; only the compiler's fusing of
    multiplies and adds into FMAs may differ between kernels, so the exit waves of a small
    specimen agree to float rounding (rel-L2 1e-6)."""
    w = wl.stem_srtio3(nx=4, ny=4, nz=3, resolution=res, pixels=4, window_cells=(1, 2))
    g, og, p, spec = setup_case(engine, oracle, w)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    pos = wl.stem_positions(g)[:6]
    probe = cltem_amd.make_probe(w["cond_ap_mrad"])
    engine.set_wave_count(len(pos))
    engine.set_wave_configs([0] * len(pos))
    ref = None
    for cv in (7, 8, 10, 11, 12, 13):
        engine.set_kernel_variant(1, cv)
        engine.init_probes(pos, probe)
        engine.multislice()
        got = np.stack([engine.exit_wave(k) for k in range(len(pos))])
        assert np.all(np.isfinite(got))
        if ref is None:
            ref = got
        else:
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            print(f"{res}^2 column variant {cv}: rel-L2 {err:.2e} against variant 7")
            assert err <= 1e-6, (res, cv)
    engine.set_kernel_variant(1, 7)
