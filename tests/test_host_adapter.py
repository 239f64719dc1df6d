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
        assert settings["mode"] == "CTEM" and settings["resolution"] == 256
