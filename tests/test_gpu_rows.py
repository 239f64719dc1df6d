"""GPU parity per scope row: potentials (all forms, projected and full 3D), beam tilt,
probe aberrations and taper, FFT sizes, and binning across phonon configurations."""
import numpy as np
import pytest

import cltem_amd
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def prepare(engine, oracle, scen, spec, params="kirkland"):
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    og = oracle.geometry(oracle.scenario(**scen))
    p = cltem_amd.load_params(params)
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    o = oracle.sort_atoms(og, spec.x, spec.y, spec.z, spec.Z)
    return g, og, p, o


def ctem_compare(engine, oracle, g, og, p, o, tol=1e-4):
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    engine.multislice()
    ew = engine.exit_wave(0)
    trans = oracle.transmission(og, o, p.values, p.form, p.i_per_atom)
    psi, recip = oracle.run_plane(trans, oracle.propagator(og))
    err = rel_l2(ew, psi)
    print(f"res {g.resolution}: {g.n_slices} slices, exit-wave rel-L2 {err:.3e}")
    assert err <= tol
    return trans


@pytest.mark.parametrize("params", ["peng", "lobato"])
def test_projected_peng_lobato(engine, oracle, params):
    w = wl.ctem_srtio3(3, 3, 4, resolution=256, params=params)
    spec = w["specimen"].displaced(0.08, 11, 0)
    g, og, p, o = prepare(engine, oracle, w["scenario"], spec, params)
    trans = ctem_compare(engine, oracle, g, og, p, o)
    for s in range(0, g.n_slices, 3):
        assert rel_l2(engine.read_transmission(0, s), trans[s]) <= 1e-5


@pytest.mark.parametrize("params", ["kirkland", "lobato", "peng"])
def test_full3d_amorphous(engine, oracle, params):
    spec = sp.random_box({"Si": 60, "O": 120}, (14.0, 14.0, 9.0), 1.5, seed=3)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=300.0, slice_thickness=1.0, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]), full3d=1, full3d_integrals=20)
    g, og, p, o = prepare(engine, oracle, scen, spec, params)
    trans = oracle.transmission(og, o, p.values, p.form, p.i_per_atom)
    errs = [rel_l2(engine.read_transmission(0, s), trans[s]) for s in range(g.n_slices)]
    print(params, "full-3D transmission rel-L2 max", max(errs))
    assert max(errs) <= 1e-5
    ctem_compare(engine, oracle, g, og, p, o)


@pytest.mark.parametrize("params,full3d", [("kirkland", 0), ("peng", 0), ("lobato", 0),
                                           ("kirkland", 1), ("peng", 1), ("lobato", 1)])
def test_tabulated_and_direct_potentials(engine, oracle, params, full3d):
    """The tabulated potentials (default) and the formulas evaluated directly
    (mse_set_potential_tables(e, 0)) give the same transmission functions, and both match the
    FP64 oracle to rel-L2 1e-5 per slice.  The full-3D tables read from LDS (default) and from
    global memory (mse_set_potential_tables(e, 2)) give identical bits."""
    spec = sp.random_box({"Sr": 8, "Ti": 8, "O": 24}, (12.0, 12.0, 6.0), 1.5, seed=21)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.5, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]), full3d=full3d, full3d_integrals=10)
    try:
        g, og, p, o = prepare(engine, oracle, scen, spec, params)
        tab = [engine.read_transmission(0, s) for s in range(g.n_slices)]
        engine.set_potential_tables(True, lds=False)
        engine.build_transmission(0)
        for s in range(g.n_slices):
            np.testing.assert_array_equal(engine.read_transmission(0, s), tab[s])
        engine.set_potential_tables(False)
        engine.build_transmission(0)
        direct = [engine.read_transmission(0, s) for s in range(g.n_slices)]
    finally:
        engine.set_potential_tables(True)
    trans = oracle.transmission(og, o, p.values, p.form, p.i_per_atom)
    e_td = max(rel_l2(a, b) for a, b in zip(tab, direct))
    e_t = max(rel_l2(a, b) for a, b in zip(tab, trans))
    e_d = max(rel_l2(a, b) for a, b in zip(direct, trans))
    print(f"{params} full3d={full3d}: tabulated vs direct {e_td:.2e}, vs FP64 {e_t:.2e} / {e_d:.2e}")
    assert e_td <= 2e-6 and e_t <= 1e-5 and e_d <= 1e-5


def test_beam_tilt(engine, oracle):
    w = wl.ctem_srtio3(3, 3, 5, resolution=256)
    scen = dict(w["scenario"], beam_tilt_mrad=8.0, beam_azimuth=0.7)
    g, og, p, o = prepare(engine, oracle, scen, w["specimen"])
    ctem_compare(engine, oracle, g, og, p, o)


def test_full3d_beam_tilt(engine, oracle):
    spec = sp.random_box({"Si": 30, "O": 60}, (12.0, 12.0, 6.0), 1.6, seed=9)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, slice_thickness=1.0, struct_min=tuple(lo), struct_max=tuple(hi),
                area=(lo[0], hi[0], lo[1], hi[1]), full3d=1, full3d_integrals=10, beam_tilt_mrad=6.0,
                beam_azimuth=1.1)
    g, og, p, o = prepare(engine, oracle, scen, spec, "kirkland")
    ctem_compare(engine, oracle, g, og, p, o)


@pytest.mark.parametrize("taper", [False, True])
def test_probe_aberrations(engine, oracle, taper):
    w = wl.stem_srtio3(nx=4, ny=4, nz=2, resolution=256, pixels=4, window_cells=(1, 2))
    g, og, p, o = prepare(engine, oracle, w["scenario"], w["specimen"])
    ab = {"C10": 40.0, "C12": (15.0, 0.4), "C21": (300.0, 1.0), "C23": (120.0, -0.5), "C30": 2.0e4,
          "C32": (5e3, 0.2), "C34": (3e3, 0.9), "C41": (1e5, 0.1), "C43": (8e4, 0.3),
          "C45": (6e4, 0.5), "C50": 1e6, "C52": (1e5, 0.6), "C54": (2e5, 0.7), "C56": (3e5, 0.8)}
    probe = cltem_amd.make_probe(22.0, smooth_mrad=2.0 if taper else 0.0, taper=taper, aberrations=ab)
    pos = wl.stem_positions(g)[[0, 5, 10]]
    engine.set_wave_count(3)
    engine.set_wave_configs([0, 0, 0])
    engine.init_probes(pos, probe)
    vec = cltem_amd.probe_aberration_vector(probe)
    for i in range(3):
        got = engine.exit_wave(i)
        k = oracle.probe_kspace(og, pos[i, 0] - og.sim_x[0], pos[i, 1] - og.sim_y[0], vec, 22.0,
                                2.0 if taper else 0.0, taper)
        ref = oracle.fft2(k, inverse=True)
        assert rel_l2(got, ref) <= 2e-6


@pytest.mark.parametrize("res", [384, 768, 1024])
def test_other_fft_sizes(engine, oracle, res):
    w = wl.ctem_srtio3(2, 2, 3, resolution=res)
    spec = w["specimen"].displaced(0.05, 5, 1)
    g, og, p, o = prepare(engine, oracle, w["scenario"], spec)
    ctem_compare(engine, oracle, g, og, p, o)


def test_binning_bit_exact_phonon_configs(engine, oracle):
    w = wl.stem_srtio3()
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    engine.set_geometry(g)
    engine.set_parameterisation(cltem_amd.load_params("kirkland"))
    for c in range(10):
        spec = w["specimen"].displaced(w["u_rms"], w["seed"], c)
        r = engine.sort_atoms(c % 3, spec.x, spec.y, spec.z, spec.Z)
        o = oracle.sort_atoms(og, spec.x, spec.y, spec.z, spec.Z)
        assert r["n_in_range"] == o["n_in_range"]
        np.testing.assert_array_equal(r["block_start"], o["block_start"])
        keep = r["block_ids"] >= 0
        np.testing.assert_array_equal(r["block_ids"][keep], o["block_ids"])
        np.testing.assert_array_equal(r["z_ids"][keep], o["z_ids"])


def test_empty_specimen_plane_wave(engine):
    # no atoms: the exit wave stays 1 (P(k=0) = 1, band limit keeps k = 0)
    scen = dict(mode=1, resolution=512, slice_thickness=2.0, struct_min=(0, 0, 0),
                struct_max=(10, 10, 10), area=(0, 10, 0, 10))
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    engine.set_geometry(g)
    engine.set_parameterisation(cltem_amd.load_params("kirkland"))
    engine.sort_atoms(0, [], [], [], [])
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    engine.multislice()
    np.testing.assert_allclose(engine.exit_wave(0), 1.0, atol=2e-6)


def test_errors_are_reported(engine):
    scen = dict(mode=1, resolution=512, slice_thickness=2.0, struct_min=(0, 0, 0),
                struct_max=(10, 10, 10), area=(0, 10, 0, 10))
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    engine.set_geometry(g)
    with pytest.raises(cltem_amd.MseError):
        engine.build_transmission(5)  # configuration never sorted
    with pytest.raises(cltem_amd.MseError):
        engine.sort_atoms(0, [1.0], [1.0], [1.0], [200])  # Z outside the parameterisation
    bad = cltem_amd.plan_geometry(cltem_amd.make_scenario(**dict(scen, resolution=640)))
    with pytest.raises(cltem_amd.MseError):
        engine.set_geometry(bad)
    # device allocation failure is an error, not a crash, and the engine stays usable
    big = cltem_amd.plan_geometry(cltem_amd.make_scenario(**dict(scen, resolution=1024)))
    engine.set_geometry(big)
    with pytest.raises(cltem_amd.MseError) as ex:
        engine.set_wave_count(40000)  # 40000 x 8 MiB = 313 GiB > HBM
    assert ex.value.code == -3
    with pytest.raises(cltem_amd.MseError):
        engine.multislice()  # no waves
    engine.set_wave_count(2)
    engine.init_plane_wave(1.0)
    engine.synchronize()


def test_ctem_image_formation(engine, oracle):
    """CTEM image (objective aperture with smoothed edge, defocus, C30, two-fold
    astigmatism, both coherence envelopes) of two phonon configurations, summed in FP64,
    against the FP64 oracle; the exit waves are left as they were."""
    w = wl.ctem_srtio3(3, 3, 4, resolution=256)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    p = cltem_amd.load_params("kirkland")
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    ab = {"C10": -60.0, "C30": 1.0e7, "C12": (20.0, 0.4)}
    im = cltem_amd.make_imaging(obj_ap_mrad=25.0, obj_smooth_mrad=2.0, beta_mrad=0.4, delta=25.0, aberrations=ab)
    lens = cltem_amd.imaging_vector(im)
    ref = np.zeros((256, 256))
    for c in range(2):
        s = w["specimen"].displaced(0.08, 41, c)
        engine.sort_atoms(c, s.x, s.y, s.z, s.Z)
        engine.build_transmission(c)
        srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
        _, recip = oracle.run_plane(oracle.transmission(og, srt, p.values, p.form, p.i_per_atom),
                                    oracle.propagator(og))
        ref += oracle.ctem_image(og, recip, lens, obj_ap_mrad=25.0, obj_smooth_mrad=2.0, beta_mrad=0.4,
                                 delta=25.0)
    engine.set_wave_count(2)
    engine.set_wave_configs([0, 1])
    engine.init_plane_wave(1.0)
    engine.multislice()
    ew = engine.exit_wave(0)
    got = engine.ctem_image(im)
    err = rel_l2(got, ref)
    print(f"CTEM image (2 configurations): rel-L2 {err:.3e}, contrast {got.std() / got.mean():.3f}")
    assert err <= 1e-4
    assert got.std() / got.mean() > 0.05  # a real image, not a flat field
    assert np.array_equal(engine.exit_wave(0), ew)
    acc = engine.ctem_image(im, got.copy())
    np.testing.assert_allclose(acc, 2 * got, rtol=1e-15)


def test_ccd_dose_matches_oracle_draw_for_draw(engine, oracle):
    """CCD DQE, seeded Poisson dose and NTF on a CTEM image: with unit tables the engine's
    Poisson draws equal the oracle generator's draw for draw (both sampler branches); with
    the K2 tables and 2x binning the whole image matches the FP64 restatement."""
    from cltem_amd.ccd import TABLE_LEN, load_ccd
    from oracle import ccd_oracle as co
    w = wl.ctem_srtio3(3, 3, 4, resolution=256)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    engine.set_geometry(g)
    rng = np.random.default_rng(8)
    x = np.linspace(0, 6 * np.pi, 256)
    img = 1.0 + 0.5 * np.sin(x)[None, :] * np.cos(0.5 * x)[:, None] + 0.01 * rng.standard_normal((256, 256))
    ones = np.ones(TABLE_LEN)
    for dose in (3.0, 40.0):  # means below 10 (inversion) and above (PTRS)
        got = engine.ccd_image(img, ones, ones, binning=1, dose_per_pixel=dose, seed=77)
        counts = np.round(got)
        assert np.abs(got - counts).max() < 1e-2
        want, _ = co.ccd_image(img, ones, ones, 1, dose, seed=77)
        same = np.mean(counts == np.round(want))
        print(f"dose {dose}: {same * 100:.4f}% of pixels draw the same count as the oracle")
        assert same >= 0.999
    k2 = load_ccd("k2")
    got = engine.ccd_image(img, k2.dqe, k2.ntf, binning=2, dose_per_pixel=25.0, conversion=1.5, seed=3)
    want, _ = co.ccd_image(img, k2.dqe, k2.ntf, 2, 25.0, conversion=1.5, seed=3)
    err = rel_l2(got, want)
    print(f"K2, binning 2, 100 e/pixel: rel-L2 {err:.3e}")
    assert err <= 1e-3


def test_full3d_depth_cutoff_exact(engine, oracle):
    """A gold atom placed so that one integration point of a slice lies exactly 3 A above it
    in double (rad_z == 3: the point counts), then at the nearest position where rad_z > 3
    (it does not).  The GPU sum finds an atom's first counted point by search instead of
    testing every point; it must decide the boundary as the FP64 per-point test of the
    oracle does.  The two placements differ by one ulp, so the change they make to the
    slice's phase is that one point's weight (about 8e-6 rad with 2 A slices and 4
    integration points); the GPU's change must match the oracle's (float noise cancels
    between the two GPU runs, while a wrong decision on either side would leave a change of
    zero or double)."""
    spec = sp.random_box({"Au": 6, "C": 20}, (10.0, 10.0, 8.0), 1.5, seed=5)
    lo, hi = spec.limits
    lo, hi = lo - 0.5, hi + 0.5  # fixed limits, so that moving an atom keeps the geometry
    integrals = 4
    scen = dict(mode=1, resolution=256, voltage_kv=300.0, slice_thickness=2.0, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]), full3d=1, full3d_integrals=integrals)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    s, h = g.n_slices // 2, 1
    dz, int_r = g.slice_thickness, 1.0 / integrals
    point_z = (g.z_lims[0] + (g.n_slices - s) * dz) - h * dz * int_r  # the kernels' expression
    z_in = point_z - 3.0
    z_out = z_in
    while point_z - z_out <= 3.0:
        z_out = np.nextafter(z_out, -np.inf)
    assert point_z - z_in == 3.0 and point_z - z_out > 3.0
    au = int(np.flatnonzero(spec.Z == 79)[0])
    gpu, ref = [], []
    for z in (z_in, z_out):
        zs = spec.z.copy()
        zs[au] = z
        placed = sp.Specimen(spec.x.copy(), spec.y.copy(), zs, spec.Z.copy())
        g, og, p, o = prepare(engine, oracle, scen, placed, "lobato")
        gpu.append(engine.read_transmission(0, s))
        ref.append(oracle.transmission(og, o, p.values, p.form, p.i_per_atom, first=s, count=1)[0])
    d_gpu = np.angle(gpu[0] / gpu[1])
    d_ref = np.angle(ref[0] / ref[1])
    err = np.abs(d_gpu - d_ref).max()
    print(f"slice {s}: the point moves the phase by {np.abs(d_ref).max():.3e} rad (oracle), "
          f"{np.abs(d_gpu).max():.3e} rad (GPU); difference {err:.2e} rad")
    assert np.abs(d_ref).max() > 5e-6
    assert err <= 0.05 * np.abs(d_ref).max()


def test_parameterisation_smaller_than_sorted_atoms(engine, oracle):
    """Atoms sorted under one parameterisation, then a parameterisation with fewer elements
    than they use: building refuses (MSE_ERR_STATE) instead of reading past the smaller
    parameter rows and tables, sorting again refuses the atomic numbers outside it, and the
    full parameterisation builds as before."""
    import dataclasses
    w = wl.ctem_srtio3(2, 2, 2, resolution=256)
    g, og, p, o = prepare(engine, oracle, w["scenario"], w["specimen"])
    ref = engine.read_transmission(0, 0)
    small = dataclasses.replace(p, values=p.values[: 22 * p.per_atom])  # up to Ti; Sr is 38
    engine.set_parameterisation(small)
    with pytest.raises(cltem_amd.MseError) as err:
        engine.build_transmission(0)
    assert err.value.code == -2, err.value  # MSE_ERR_STATE
    spec = w["specimen"]
    with pytest.raises(cltem_amd.MseError):
        engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.set_parameterisation(p)
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    np.testing.assert_array_equal(engine.read_transmission(0, 0), ref)


def test_full3d_tables_beyond_the_lds_budget(engine, oracle):
    """Seven elements: their full-3D tables exceed the 32 KB the kernels copy into LDS, so the
    tables are read from global memory (same bits as mse_set_potential_tables(e, 2)); the
    transmission functions match the formulas evaluated directly and the FP64 oracle."""
    spec = sp.random_box({"H": 6, "C": 6, "O": 6, "Si": 6, "Ti": 6, "Sr": 6, "Au": 6}, (12.0, 12.0, 6.0), 1.5,
                         seed=5)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.5, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]), full3d=1, full3d_integrals=10)
    try:
        g, og, p, o = prepare(engine, oracle, scen, spec, "lobato")
        tab = [engine.read_transmission(0, s) for s in range(g.n_slices)]
        engine.set_potential_tables(True, lds=False)
        engine.build_transmission(0)
        for s in range(g.n_slices):
            np.testing.assert_array_equal(engine.read_transmission(0, s), tab[s])
        engine.set_potential_tables(False)
        engine.build_transmission(0)
        direct = [engine.read_transmission(0, s) for s in range(g.n_slices)]
    finally:
        engine.set_potential_tables(True)
    trans = oracle.transmission(og, o, p.values, p.form, p.i_per_atom)
    e_td = max(rel_l2(a, b) for a, b in zip(tab, direct))
    e_t = max(rel_l2(a, b) for a, b in zip(tab, trans))
    print(f"7 elements, full 3D: tabulated vs direct {e_td:.2e}, vs FP64 {e_t:.2e}")
    assert e_td <= 2e-6 and e_t <= 1e-5


@pytest.mark.parametrize("full3d,tilt", [(0, 0.0), (1, 0.0), (1, 6.0)])
def test_dense_rows_span_staging_chunks(engine, oracle, full3d, tilt):
    """The builders stage the atoms of a pixel tile's block rows (and, in full 3D, of every
    slice of its depth window) as one sequence, 256 at a time.  A dense slab puts several
    hundred atoms in each tile's sequence, so chunk boundaries fall inside block rows: the
    full-3D trapezoid must carry p2 across them and restart it only at a row's first atom
    (also with a tilted beam, where no atom is passed over).  Transmission functions against
    the FP64 oracle."""
    if full3d:
        spec = sp.random_box({"C": 420}, (12.0, 12.0, 2.5), 0.6, seed=21)
        extra = dict(slice_thickness=1.0, full3d=1, full3d_integrals=6, beam_tilt_mrad=tilt, beam_azimuth=0.4)
    else:
        spec = sp.random_box({"C": 700}, (12.0, 12.0, 1.5), 0.5, seed=22)
        extra = dict(slice_thickness=2.0)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=300.0, struct_min=tuple(lo), struct_max=tuple(hi),
                area=(lo[0], hi[0], lo[1], hi[1]), **extra)
    g, og, p, o = prepare(engine, oracle, scen, spec, "kirkland")
    # atoms per slice (and per full-3D window of slices s - load_z .. s + load_z)
    nb = g.blocks_x * g.blocks_y
    bs = np.asarray(o["block_start"])
    per_slice = np.array([bs[(s + 1) * nb] - bs[s * nb] for s in range(g.n_slices)])
    if full3d:
        lz = g.load_z
        per_seq = [per_slice[max(0, s - lz): s + lz + 1].sum() for s in range(g.n_slices)]
    else:
        per_seq = list(per_slice)
    print(f"full3d={full3d} tilt={tilt}: {g.n_slices} slices, atoms per slice {per_slice.tolist()}, "
          f"largest staging sequence {max(per_seq)}")
    assert max(per_seq) > 300
    trans = oracle.transmission(og, o, p.values, p.form, p.i_per_atom)
    errs = [rel_l2(engine.read_transmission(0, s), trans[s]) for s in range(g.n_slices)]
    print(f"transmission rel-L2 max {max(errs):.2e}")
    # the phases stay below 1 rad, so float rounding keeps the builders near 1e-7; one trapezoid
    # term lost or doubled per row would cost about 1e-4 of a phase near the atom
    assert max(errs) <= 1e-6
