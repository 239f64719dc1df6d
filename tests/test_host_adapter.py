"""The C++ host side (clTEM-shaped SimulationGeneral/Ctem/Cbed/Stem classes and the mse_run
driver) against the Python input generators (CPU) and the FP64 oracle (GPU)."""
import json
import os
import subprocess

import numpy as np
import pytest

import cltem_amd
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "cltem_amd", "mse_run")
PARAMS = os.path.join(ROOT, "cltem_amd", "params")


def have_driver():
    return os.path.exists(RUN)


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_cpp_phonon_displacements_match_python(tmp_path):
    f = tmp_path / "atoms.bin"
    for cfg in (0, 3):
        subprocess.run([RUN, "--srtio3", "4,3,2", "--urms", "0.08", "--seed", "20261202", "--dump-atoms",
                        str(f), "--dump-config", str(cfg)], check=True)
        raw = f.read_bytes()
        n = 4 * 3 * 2 * 5
        xyz = np.frombuffer(raw[:24 * n], dtype=np.float64).reshape(3, n)
        Z = np.frombuffer(raw[24 * n:], dtype=np.int32)
        ref = sp.srtio3(4, 3, 2).displaced(0.08, 20261202, cfg)
        np.testing.assert_array_equal(Z, ref.Z)
        # libm and numpy may differ in the last ulp of log/cos before the float32 rounding
        for got, exp in zip(xyz, (ref.x, ref.y, ref.z)):
            assert np.max(np.abs(got - exp)) <= 1e-5


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_cpp_incoherence_draws_match_python():
    from cltem_amd import incoherence as inc
    for job in (0, 2):
        r = subprocess.run([RUN, "--chromatic", "1.4,0.4,0.6", "--source-size", "0.8", "--kv", "300", "--seed", "7",
                            "--dump-incoherence", "6", "--dump-config", str(job)],
                           capture_output=True, text=True, check=True)
        vals = np.array([float(v) for v in r.stdout.split()])
        sn, spos = inc.sigmas_from_hwhm(0.4, 0.6)
        df = inc.focus_changes(7, job, np.arange(6), 1.4, sn, spos, 300.0)
        dx, dy = inc.source_offset(7, job, inc.source_sigma_from_fwhm(0.8))
        np.testing.assert_allclose(vals[:6], df, rtol=1e-12)
        np.testing.assert_allclose(vals[6:], [dx, dy], rtol=1e-12)
    # both halves of the asymmetric spread occur, scaled by their own sigma
    df = inc.focus_changes(7, 0, np.arange(4000), 1.4, sn, spos, 300.0)
    frac_neg = np.mean(df < 0)
    assert abs(frac_neg - sn / (sn + spos)) < 0.03
    scale = 10000 * 1.4 / 300.0
    assert abs(np.mean(-df[df < 0]) / scale - sn * np.sqrt(2 / np.pi)) < 0.05 * sn
    assert abs(np.mean(df[df > 0]) / scale - spos * np.sqrt(2 / np.pi)) < 0.05 * spos


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_cpp_plasmon_draws_match_python():
    """The C++ host's PlasmonScattering draws (depths, then polar and azimuthal angles) equal
    cltem_amd.incoherence's, for the combined and the individual-count modes."""
    from cltem_amd import incoherence as inc
    for job, ind in ((0, 0), (5, 0), (2, 3)):
        spec = "8,1.2,15" + (f",{ind}" if ind else "")
        r = subprocess.run([RUN, "--srtio3", "2,2,2", "--seed", "77", "--plasmon", spec, "--dump-plasmons", "250",
                            "--dump-config", str(job)], check=True, capture_output=True, text=True)
        vals = np.array(r.stdout.split(), dtype=np.float64)
        n = int(vals[0])
        ref = inc.plasmon_depths(77, job, 250.0, 80.0, individual=ind)
        assert n == len(ref) and (ind == 0 or n == ind)
        np.testing.assert_allclose(vals[1:n + 1], ref, rtol=1e-13)
        polar, az = inc.plasmon_angles(77, job, n, 1.2, 15.0)
        np.testing.assert_allclose(vals[n + 1::2], polar, rtol=1e-12)
        np.testing.assert_allclose(vals[n + 2::2], az, rtol=1e-13)


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_cpp_mixed_seed_matches_python():
    from cltem_amd import incoherence as inc
    for job in (0, 1, 7):
        r = subprocess.run([RUN, "--seed", "20261202", "--dump-mixed-seed", "--dump-config", str(job)], check=True,
                           capture_output=True, text=True)
        assert int(r.stdout.strip()) == inc.mixed_seed(20261202, job)


def _settings(*args):
    r = subprocess.run([RUN, "--params", os.path.join(PARAMS, "kirkland.dat"), *args], check=True,
                       capture_output=True, text=True)
    return json.loads(r.stdout)


def _at(j, *path):
    for k in path:
        j = j[k]
    return j


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_settings_json_follows_cltem_schema():
    """The settings written beside each image carry the fields clTEM's JsonToManager reads
    (utilities/jsonutils.cpp:15-360) at the same paths, with clTEM's units."""
    st = _settings("--mode", "stem", "--srtio3", "3,3,2", "--res", "256", "--dz", "1.9525", "--kv", "300",
                   "--det", "ADF,60,200", "--stem-pixels", "8", "--batch", "16", "--phonons", "3", "--urms", "0.08",
                   "--parallel-potentials", "2", "--chromatic", "1.2,0.5,0.8", "--source-size", "0.7",
                   "--tilt", "2.5,0.3", "--dump-settings", "ADF")
    for path, want in [(("mode", "id"), 2), (("mode", "name"), "STEM"), (("resolution",), 256),
                       (("slice thickness", "val"), 1.9525), (("slice offset", "val"), 0.0),
                       (("microscope", "voltage", "val"), 300.0),
                       (("microscope", "condenser aperture", "semi-angle"), 20.0),
                       (("microscope", "beam tilt", "inclination", "val"), 2.5),
                       (("stem", "scan", "x", "pixels"), 8), (("stem", "static area", "concurrent pixels"), 16),
                       (("stem", "detectors", "ADF", "radius", "inner"), 60.0),
                       (("stem", "detectors", "ADF", "radius", "outer"), 200.0),
                       (("incoherence", "iterations"), 3), (("incoherence", "probe", "chromatic", "enabled"), True),
                       (("incoherence", "probe", "chromatic", "Cc", "val"), 1.2),
                       (("incoherence", "inelastic scattering", "phonon", "mixed static potentials", "count"), 2),
                       (("default padding", "xy", "val"), 8.0), (("full 3d", "state"), False),
                       (("precalculate transmission",), True), (("intermediate output", "enabled"), False)]:
        assert _at(st, *path) == pytest.approx(want) if isinstance(want, float) else _at(st, *path) == want, path
    assert _at(st, "microscope", "beam tilt", "azimuth", "val") == pytest.approx(0.3 * 180 / np.pi)
    assert _at(st, "incoherence", "probe", "chromatic", "dE", "HWHM -") == pytest.approx(0.5)
    assert _at(st, "incoherence", "probe", "source size", "FWHM", "val") == pytest.approx(0.7)
    assert _at(st, "slice count") > 0 and "aberrations" in st["microscope"]
    for k in ("C10", "C30", "C50"):
        assert "val" in st["microscope"]["aberrations"][k]
    for k in ("C12", "C21", "C23", "C32", "C34", "C41", "C43", "C45", "C52", "C54", "C56"):
        assert set(st["microscope"]["aberrations"][k]) == {"mag", "ang", "units"}
    assert st["stem"]["static area"]["enabled"] is True
    mv = _settings("--mode", "stem", "--srtio3", "3,3,2", "--res", "256", "--det", "ADF,60,200",
                   "--stem-pixels", "8", "--moving-frame", "--dump-settings", "ADF")
    assert mv["stem"]["static area"]["enabled"] is False
    # CBED diffraction: imageReturned drops the objective-lens entries and the aberrations
    cb = _settings("--mode", "cbed", "--srtio3", "2,2,2", "--res", "256", "--area", "3,3,4,4",
                   "--plasmon", "50,1.5,12,2", "--dump-settings", "Diff")
    assert cb["cbed"]["position"]["x"] == 3.0 and cb["cbed"]["position"]["y"] == 4.0
    assert "aberrations" not in cb["microscope"] and "alpha" not in cb["microscope"]
    pl = cb["incoherence"]["inelastic scattering"]["plasmon"]
    assert pl["type"] == "individual" and pl["individual"] == 2 and pl["mean free path"]["value"] == 50.0
    assert cb["precalculate transmission"] is False
    ct = _settings("--mode", "ctem", "--srtio3", "2,2,2", "--res", "256", "--dump-settings", "EW")
    assert ct["mode"]["name"] == "CTEM" and "ctem" in ct and "condenser aperture" not in ct["microscope"]


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_rejects_bad_input(tmp_path):
    r = subprocess.run([RUN, "--mode", "ctem"], capture_output=True, text=True)
    assert r.returncode == 1 and "no structure" in r.stderr
    r = subprocess.run([RUN, "--srtio3", "1,1,1", "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown option" in r.stderr


def _load(prefix):
    meta = json.load(open(prefix + ".json"))
    out = {}
    for k, v in meta.items():
        if isinstance(v, dict):
            out[k] = np.fromfile(v["file"], dtype=np.float64)
    return meta, out


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_ctem_matches_oracle(tmp_path, oracle):
    prefix = str(tmp_path / "ctem")
    subprocess.run([RUN, "--mode", "ctem", "--srtio3", "3,3,4", "--res", "256", "--dz", "1.9525",
                    "--params", os.path.join(PARAMS, "kirkland.dat"), "--out", prefix], check=True)
    meta, out = _load(prefix)
    w = wl.ctem_srtio3(3, 3, 4)
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    s = w["specimen"]
    p = cltem_amd.load_params("kirkland")
    srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
    psi, recip = oracle.run_plane(oracle.transmission(og, srt, p.values, 0, 3), oracle.propagator(og))
    ew = out["EW"].view(np.complex128).reshape(256, 256)
    assert np.linalg.norm(ew - psi) / np.linalg.norm(psi) <= 1e-4
    dif = out["Diff"].reshape(256, 256)
    ref = oracle.diffraction(recip)
    assert np.linalg.norm(dif - ref) / np.linalg.norm(ref) <= 1e-3


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_stem_phonon_average_matches_oracle(tmp_path, oracle):
    prefix = str(tmp_path / "stem")
    x0, x1 = 2 * 3.905, 3 * 3.905
    subprocess.run([RUN, "--mode", "stem", "--srtio3", "5,5,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{x0},{x1},{x0},{x1}", "--stem-pixels", "4", "--det", "ADF,60,200",
                    "--det", "BF,0,15", "--alpha", "20", "--batch", "8", "--phonons", "2",
                    "--urms", "0.08", "--seed", "99", "--params", os.path.join(PARAMS, "kirkland.dat"),
                    "--out", prefix], check=True)
    meta, out = _load(prefix)
    assert meta["jobs"] == 2
    w = wl.stem_srtio3(nx=5, ny=5, nz=3, resolution=256, pixels=4, window_cells=(2, 3))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    p = cltem_amd.load_params("kirkland")
    pos = wl.stem_positions(og)
    acc = np.zeros((16, 2))
    for c in range(2):
        s = w["specimen"].displaced(0.08, 99, c)
        srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
        tr = oracle.transmission(og, srt, p.values, 0, 3)
        acc += oracle.run_stem(og, tr, oracle.propagator(og), pos, np.array([[60, 200, 0, 0], [0, 15, 0, 0.]]))
    acc /= 2
    for d, name in enumerate(("ADF", "BF")):
        got = out[name]
        assert np.max(np.abs(got - acc[:, d]) / np.abs(acc[:, d])) <= 1e-3, name


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_stem_moving_frame_matches_oracle(tmp_path, oracle):
    """The static area off ("static area enabled": false): every pixel job in its own field
    with its own transmission functions (currentAreaBase(pixel), simulationmanager.cpp:485-501),
    averaged over two phonon configurations, against the FP64 oracle run pixel by pixel."""
    prefix = str(tmp_path / "stem_moving")
    x0, x1 = 2 * 3.905, 3 * 3.905
    subprocess.run([RUN, "--mode", "stem", "--srtio3", "5,5,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{x0},{x1},{x0},{x1}", "--stem-pixels", "4", "--det", "ADF,60,200",
                    "--det", "BF,0,15", "--alpha", "20", "--batch", "5", "--phonons", "2", "--moving-frame",
                    "--urms", "0.08", "--seed", "99", "--params", os.path.join(PARAMS, "kirkland.dat"),
                    "--out", prefix], check=True)
    meta, out = _load(prefix)
    assert meta["jobs"] == 2
    w = wl.stem_srtio3(nx=5, ny=5, nz=3, resolution=256, pixels=4, window_cells=(2, 3))
    s_or = oracle.scenario(**dict(w["scenario"], moving_stem_frame=1))
    p = cltem_amd.load_params("kirkland")
    pos = wl.stem_positions(oracle.geometry(s_or))
    acc = np.zeros((16, 2))
    for c in range(2):
        s = w["specimen"].displaced(0.08, 99, c)
        for px in range(16):
            gp = oracle.geometry_pixel(s_or, px)
            srt = oracle.sort_atoms(gp, s.x, s.y, s.z, s.Z)
            tr = oracle.transmission(gp, srt, p.values, 0, 3)
            acc[px] += oracle.run_stem(gp, tr, oracle.propagator(gp), pos[px:px + 1],
                                       np.array([[60, 200, 0, 0], [0, 15, 0, 0.]]))[0]
    acc /= 2
    for d, name in enumerate(("ADF", "BF")):
        got = out[name]
        assert np.max(np.abs(got - acc[:, d]) / np.abs(acc[:, d])) <= 1e-3, name


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_cbed_matches_oracle(tmp_path, oracle):
    prefix = str(tmp_path / "cbed")
    cx = 1.5 * 3.905
    subprocess.run([RUN, "--mode", "cbed", "--srtio3", "3,3,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{cx},{cx},{cx},{cx}", "--alpha", "15",
                    "--params", os.path.join(PARAMS, "kirkland.dat"), "--out", prefix], check=True)
    meta, out = _load(prefix)
    s = sp.srtio3(3, 3, 3)
    lo, hi = s.limits
    scen = dict(mode=3, resolution=256, slice_thickness=1.9525, struct_min=tuple(lo), struct_max=tuple(hi),
                area=(cx, cx, cx, cx))
    og = oracle.geometry(oracle.scenario(**scen))
    p = cltem_amd.load_params("kirkland")
    srt = oracle.sort_atoms(og, s.x, s.y, s.z, s.Z)
    tr = oracle.transmission(og, srt, p.values, 0, 3)
    k = oracle.probe_kspace(og, cx - og.sim_x[0], cx - og.sim_y[0], None, 15.0)
    psi = oracle.fft2(k, inverse=True)
    P_ = oracle.propagator(og)
    for sl in range(og.n_slices):
        psi, recip = oracle.slice_step(psi, tr[sl], P_)
    ref = oracle.diffraction(recip)
    got = out["Diff"].reshape(256, 256)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-3


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_cbed_chromatic_and_source_size_match_oracle(tmp_path, oracle):
    """CBED averaged over 3 jobs, each with its own focus change and source-size offset of
    the potential grid, against FP64 oracle runs with the same per-job draws."""
    import ctypes
    from cltem_amd import incoherence as inc
    prefix = str(tmp_path / "cbed_inc")
    cx = 1.5 * 3.905
    seed = 11
    subprocess.run([RUN, "--mode", "cbed", "--srtio3", "3,3,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{cx},{cx},{cx},{cx}", "--alpha", "15", "--phonons", "3", "--urms", "0",
                    "--seed", str(seed), "--chromatic", "1.2,0.5,0.8", "--source-size", "1.0",
                    "--params", os.path.join(PARAMS, "kirkland.dat"), "--out", prefix], check=True)
    meta, out = _load(prefix)
    assert meta["jobs"] == 3
    s = sp.srtio3(3, 3, 3)
    lo, hi = s.limits
    scen = dict(mode=3, resolution=256, slice_thickness=1.9525, struct_min=tuple(lo), struct_max=tuple(hi),
                area=(cx, cx, cx, cx))
    og = oracle.geometry(oracle.scenario(**scen))
    p = cltem_amd.load_params("kirkland")
    P_ = oracle.propagator(og)
    sn, spos = inc.sigmas_from_hwhm(0.5, 0.8)
    ref = np.zeros((256, 256))
    for job in range(3):
        d = s.displaced(0.0, seed, job)
        dx, dy = inc.source_offset(seed, job, inc.source_sigma_from_fwhm(1.0))
        og_t = type(og).from_buffer_copy(og)
        og_t.sim_x[0] += dx
        og_t.sim_y[0] += dy
        srt = oracle.sort_atoms(og, d.x, d.y, d.z, d.Z)
        tr = oracle.transmission(og_t, srt, p.values, 0, 3)
        ab = np.zeros(25)
        ab[0] = inc.focus_changes(seed, job, [0], 1.2, sn, spos, 200.0)[0]
        psi = oracle.fft2(oracle.probe_kspace(og, cx - og.sim_x[0], cx - og.sim_y[0], ab, 15.0), inverse=True)
        for sl in range(og.n_slices):
            psi, recip = oracle.slice_step(psi, tr[sl], P_)
        ref += oracle.diffraction(recip)
    ref /= 3
    got = out["Diff"].reshape(256, 256)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"CBED with focus spread and source size, 3 jobs: rel-L2 {err:.3e}")
    assert err <= 1e-3
    # the effects are visible: the same run without them differs
    plain = str(tmp_path / "cbed_plain")
    subprocess.run([RUN, "--mode", "cbed", "--srtio3", "3,3,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{cx},{cx},{cx},{cx}", "--alpha", "15", "--params",
                    os.path.join(PARAMS, "kirkland.dat"), "--out", plain], check=True)
    _, outp = _load(plain)
    assert np.linalg.norm(outp["Diff"] - out["Diff"]) / np.linalg.norm(out["Diff"]) > 1e-3


def read_tiff(path):
    """Minimal little-endian baseline TIFF reader: tags and the single float32 strip."""
    import struct
    b = open(path, "rb").read()
    assert b[:4] == b"II*\x00"
    ifd = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, ifd)[0]
    tags = {}
    for i in range(n):
        tid, typ, cnt, val = struct.unpack_from("<HHII", b, ifd + 2 + 12 * i)
        if typ == 3:
            val &= 0xFFFF
        tags[tid] = val
    w, h = tags[256], tags[257]
    data = np.frombuffer(b, dtype="<f4", count=w * h, offset=tags[273]).reshape(h, w)
    return tags, data


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_tiff_writer(tmp_path):
    f = str(tmp_path / "ramp.tif")
    subprocess.run([RUN, "--write-test-tiff", f, "7", "5"], check=True)
    tags, data = read_tiff(f)
    # fileio::SaveTiff<float>: 32-bit IEEE float, one sample, one strip, no compression,
    # min-is-black, bottom-left orientation, contiguous planes
    assert tags[256] == 7 and tags[257] == 5 and tags[258] == 32 and tags[339] == 3
    assert tags[277] == 1 and tags[278] == 5 and tags[279] == 7 * 5 * 4 and tags[259] == 1
    assert tags[262] == 1 and tags[274] == 4 and tags[284] == 1
    x = np.arange(7)[None, :]
    y = np.arange(5)[:, None]
    np.testing.assert_array_equal(data, (x + 1000.0 * y).astype(np.float32))


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_tiff_outputs(tmp_path):
    """--tiff writes clTEM's outputs: EW as amplitude and phase, Diff, each with a settings
    .json; the .tif values equal the FP64 results rounded to float32."""
    prefix = str(tmp_path / "ctem")
    subprocess.run([RUN, "--mode", "ctem", "--srtio3", "2,2,2", "--res", "256", "--dz", "1.9525", "--tiff",
                    "--params", os.path.join(PARAMS, "kirkland.dat"), "--out", prefix], check=True)
    meta, out = _load(prefix)
    ew = out["EW"].reshape(256, 256, 2)
    amp = np.hypot(ew[..., 0], ew[..., 1])
    ph = np.arctan2(ew[..., 1], ew[..., 0])
    for name, want in (("EW_amplitude", amp), ("EW_phase", ph), ("Diff", out["Diff"].reshape(256, 256))):
        tags, data = read_tiff(f"{prefix}_{name}.tif")
        np.testing.assert_allclose(data, want.astype(np.float32), rtol=1e-6, atol=1e-6 * np.abs(want).max())
        settings = json.load(open(f"{prefix}_{name}.json"))
        assert settings["mode"]["name"] == "CTEM" and settings["resolution"] == 256
        assert "aberrations" not in settings["microscope"]  # imageReturned drops them for EW and Diff


def _pm3m_ops():
    """The 48 operations of Pm-3m as CIF strings, generated from (z,x,y), (-y,x,z), (-x,-y,-z)."""
    gens = [np.array(m) for m in ([[0, 0, 1], [1, 0, 0], [0, 1, 0]], [[0, -1, 0], [1, 0, 0], [0, 0, 1]],
                                  [[-1, 0, 0], [0, -1, 0], [0, 0, -1]])]
    ops = {tuple(np.eye(3, dtype=int).ravel())}
    frontier = list(ops)
    while frontier:
        new = []
        for o in frontier:
            for g in gens:
                m = tuple((g @ np.array(o).reshape(3, 3)).ravel())
                if m not in ops:
                    ops.add(m)
                    new.append(m)
        frontier = new
    out = []
    for o in sorted(ops):
        m = np.array(o).reshape(3, 3)
        terms = []
        for r in range(3):
            t = "".join(("+" if m[r, c] > 0 else "-") + "xyz"[c] for c in range(3) if m[r, c] != 0)
            terms.append(t)
        out.append(",".join(terms))
    assert len(out) == 48
    return out


def _write_cif(path, a, sites, ops=None):
    lines = ["data_test", f"_cell_length_a {a}(2)", f"_cell_length_b {a}", f"_cell_length_c {a}",
             "_cell_angle_alpha 90.0", "_cell_angle_beta 90", "_cell_angle_gamma 90.000", "", "loop_",
             "_symmetry_equiv_pos_as_xyz"]
    lines += [f"'{o}'" for o in (ops or ["x,y,z"])]
    lines += ["", "loop_", "_atom_site_label", "_atom_site_type_symbol", "_atom_site_fract_x", "_atom_site_fract_y",
              "_atom_site_fract_z", "_atom_site_occupancy"]
    lines += [f"{lab} {sym} {x} {y} {z} {occ}" for lab, sym, x, y, z, occ in sites]
    open(path, "w").write("\n".join(lines) + "\n")


def _cif_atoms(tmp_path, cif, *args):
    f = tmp_path / "atoms.bin"
    subprocess.run([RUN, "--cif", str(cif), *args, "--dump-atoms", str(f)], check=True)
    raw = f.read_bytes()
    n = len(raw) // 28
    xyz = np.frombuffer(raw[:24 * n], dtype=np.float64).reshape(3, n)
    Z = np.frombuffer(raw[24 * n:], dtype=np.int32)
    return xyz, Z


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_cif_supercell_matches_generator(tmp_path):
    """SrTiO3 from its Pm-3m asymmetric unit (symmetry expansion) and from an explicit P1
    cell: the [001] supercell equals the Python generator's tiling."""
    a = 3.905
    w = 3 * a - 1e-3
    ref = sp.srtio3(3, 3, 3)
    order_ref = np.lexsort((ref.z, ref.y, ref.x, ref.Z))
    for name, sites, ops in (
            ("sto_pm3m.cif", [("Sr1", "Sr2+", 0, 0, 0, 1), ("Ti1", "Ti4+", 0.5, 0.5, 0.5, 1), ("O1", "O2-", 0.5, 0.5, 0, 1)],
             _pm3m_ops()),
            ("sto_p1.cif", [("Sr", "Sr", 0, 0, 0, 1), ("Ti", "Ti", 0.5, 0.5, 0.5, 1), ("O1", "O", 0.5, 0.5, 0, 1),
                            ("O2", "O", 0.5, 0, 0.5, 1), ("O3", "O", 0, 0.5, 0.5, 1)], None)):
        cif = tmp_path / name
        _write_cif(cif, a, sites, ops)
        xyz, Z = _cif_atoms(tmp_path, cif, "--widths", f"{w},{w},{w}")
        assert len(Z) == len(ref.Z), name
        r = np.round(xyz, 6)  # the cell vectors carry cos(90 deg) ~ 6e-17 terms
        o = np.lexsort((r[2], r[1], r[0], Z))
        np.testing.assert_array_equal(Z[o], ref.Z[order_ref])
        for got, exp in zip(xyz, (ref.x, ref.y, ref.z)):
            np.testing.assert_allclose(got[o], exp[order_ref], atol=2e-6)


@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_cif_zone_axis_and_occupancy(tmp_path):
    a = 3.905
    cif = tmp_path / "sto.cif"
    _write_cif(cif, a, [("Sr1", "Sr", 0, 0, 0, 1), ("Ti1", "Ti", 0.5, 0.5, 0.5, 1), ("O1", "O", 0.5, 0.5, 0, 1)],
               _pm3m_ops())
    # [110] along z: the cell is rotated, density and bond lengths are unchanged
    W = 20.0
    xyz, Z = _cif_atoms(tmp_path, cif, "--widths", f"{W},{W},{W}", "--zone", "1,1,0")
    assert abs(len(Z) / W ** 3 - 5 / a ** 3) < 0.08 * 5 / a ** 3
    pts = xyz.T
    ti = pts[Z == 22]
    o = pts[Z == 8]
    d = np.sqrt(((ti[:, None, :] - o[None, :, :]) ** 2).sum(-1)).min(axis=1)
    inner = np.all((ti > 2.5) & (ti < W - 2.5), axis=1)
    np.testing.assert_allclose(d[inner], a / 2, atol=1e-5)
    # rows of Sr along z repeat every a*sqrt(2) for [110]
    sr = pts[Z == 38]
    col = sr[np.all(np.abs(sr[:, :2] - sr[0, :2]) < 1e-6, axis=1)][:, 2]
    assert np.allclose(np.diff(np.sort(col)), a * np.sqrt(2), atol=1e-5)
    # partial occupancy: a half-occupied O site and a 70/30 Sr/Ca site; seeded, one atom per site
    cif2 = tmp_path / "mixed.cif"
    _write_cif(cif2, a, [("Sr1", "Sr", 0, 0, 0, 0.7), ("Ca1", "Ca", 0, 0, 0, 0.3), ("Ti1", "Ti", 0.5, 0.5, 0.5, 1),
                         ("O1", "O", 0.5, 0.5, 0, 0.5)], _pm3m_ops())
    w = 10 * a - 1e-3
    xyz1, Z1 = _cif_atoms(tmp_path, cif2, "--widths", f"{w},{w},{w}", "--seed", "5")
    xyz2, Z2 = _cif_atoms(tmp_path, cif2, "--widths", f"{w},{w},{w}", "--seed", "5")
    np.testing.assert_array_equal(Z1, Z2)
    n_a = np.sum(Z1 == 38) + np.sum(Z1 == 20)
    assert n_a == 1000 and np.sum(Z1 == 22) == 1000
    assert abs(np.sum(Z1 == 20) / 1000 - 0.3) < 0.05
    assert abs(np.sum(Z1 == 8) / 3000 - 0.5) < 0.04
    _, Z3 = _cif_atoms(tmp_path, cif2, "--widths", f"{w},{w},{w}", "--seed", "6")
    assert not np.array_equal(Z1, Z3)


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_intermediate_outputs_match_oracle(tmp_path, oracle):
    """--intermediate 4: EW and diffraction after every 4 slices and at the exit surface
    (SimulationCtem::simulate, simulationctem.cpp:285-400), written as clTEM names the stacks."""
    prefix = str(tmp_path / "ctem_int")
    subprocess.run([RUN, "--mode", "ctem", "--srtio3", "3,3,3", "--res", "256", "--dz", "1.9525",
                    "--intermediate", "4", "--tiff", "--params", os.path.join(PARAMS, "kirkland.dat"),
                    "--out", prefix], check=True)
    meta, out = _load(prefix)
    w = wl.ctem_srtio3(3, 3, 3)
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    s = w["specimen"]
    p = cltem_amd.load_params("kirkland")
    tr = oracle.transmission(og, oracle.sort_atoms(og, s.x, s.y, s.z, s.Z), p.values, 0, 3)
    P = oracle.propagator(og)
    n = og.n_slices
    n_out = -(-n // 4)
    assert meta["EW"]["depth"] == n_out and n % 4 != 0  # the exit surface is an extra output
    ew = out["EW"].view(np.complex128).reshape(n_out, 256, 256)
    dif = out["Diff"].reshape(n_out, 256, 256)
    width = len(str(n - 1))
    for k in range(n_out):
        m = min(4 * (k + 1), n)
        psi, recip = oracle.run_plane(tr[:m], P)
        assert np.linalg.norm(ew[k] - psi) / np.linalg.norm(psi) <= 1e-4, k
        ref = oracle.diffraction(recip)
        assert np.linalg.norm(dif[k] - ref) / np.linalg.norm(ref) <= 1e-3, k
        sid = str(m - 1).zfill(width)  # saveTiffOutput: the slice that ends the output
        for name in (f"EW_amplitude_{sid}", f"EW_phase_{sid}", f"Diff_{sid}"):
            assert os.path.exists(f"{prefix}_{name}.tif"), name
    settings = json.load(open(f"{prefix}_Diff.json"))
    assert settings["intermediate output"] == {"enabled": True, "slice interval": 4}


@pytest.mark.gpu
@pytest.mark.skipif(not have_driver(), reason="mse_run not built")
def test_driver_stem_intermediate_outputs_match_oracle(tmp_path, oracle):
    prefix = str(tmp_path / "stem_int")
    x0, x1 = 2 * 3.905, 3 * 3.905
    subprocess.run([RUN, "--mode", "stem", "--srtio3", "5,5,3", "--res", "256", "--dz", "1.9525",
                    "--area", f"{x0},{x1},{x0},{x1}", "--stem-pixels", "2", "--det", "ADF,60,200",
                    "--alpha", "20", "--batch", "3", "--intermediate", "4",
                    "--params", os.path.join(PARAMS, "kirkland.dat"), "--out", prefix], check=True)
    meta, out = _load(prefix)
    w = wl.stem_srtio3(nx=5, ny=5, nz=3, resolution=256, pixels=2, window_cells=(2, 3))
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    s = w["specimen"]
    p = cltem_amd.load_params("kirkland")
    tr = oracle.transmission(og, oracle.sort_atoms(og, s.x, s.y, s.z, s.Z), p.values, 0, 3)
    P = oracle.propagator(og)
    n = og.n_slices
    n_out = -(-n // 4)
    got = out["ADF"].reshape(n_out, 4)
    pos = wl.stem_positions(og)
    for k in range(n_out):
        m = min(4 * (k + 1), n)
        og_m = type(og).from_buffer_copy(og)  # run_stem takes the slice count from the geometry
        og_m.n_slices = m
        ref = oracle.run_stem(og_m, np.ascontiguousarray(tr[:m]), P, pos, np.array([[60.0, 200.0, 0.0, 0.0]]))[:, 0]
        assert np.max(np.abs(got[k] - ref) / np.abs(ref)) <= 1e-3, k
