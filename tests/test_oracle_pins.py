"""Pins of the FP64 oracle against golden vectors, closed forms and invariants (CPU only).

Each test checks the oracle against something it does not share code with: numpy/scipy
golden vectors, published constants, analytic identities, or hand-derived answers.
"""
import os

import numpy as np
import pytest
import scipy.integrate as si

from cltem_amd import params as P
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- constants and geometry -----------------------------------------------------------
def test_wavelength_sigma_closed_forms(oracle):
    # relativistic wavelength h / sqrt(2 m e V (1 + e V / 2 m c^2)) with clTEM's constants
    h, c, m, e = 6.6262e-034, 299792458.0, 9.10938291e-031, 1.6021773e-019
    for kv in (80.0, 200.0, 300.0):
        V = kv * 1e3
        lam = h / np.sqrt(2 * m * e * V * (1 + e * V / (2 * m * c * c))) * 1e10
        assert oracle.wavelength(kv) == pytest.approx(lam, rel=1e-12)
        # Kirkland eq. 5.6: sigma = 2 pi m e lam / h^2 (relativistic m), in rad/(V A)
        mr = m + e * V / c**2
        sig = 2 * np.pi * mr * e * (lam * 1e-10) / h**2 * 1e-10
        assert oracle.sigma(kv) == pytest.approx(sig, rel=1e-9)
    # published values (Kirkland, Advanced Computing in EM, table 5.2 / eq. 2.5)
    assert oracle.wavelength(200.0) == pytest.approx(0.025079, abs=2e-6)
    assert oracle.wavelength(300.0) == pytest.approx(0.019687, abs=2e-6)
    assert oracle.sigma(200.0) == pytest.approx(7.2884e-4, rel=2e-4)


def test_geometry_config1_hand_derived(oracle):
    w = wl.ctem_srtio3()
    g = oracle.geometry(oracle.scenario(**w["scenario"]))
    # 8 cells -> x in [0, 7.5 a]; field = 7.5 a + 2 * 8 A padding
    L = 7.5 * 3.905 + 16.0
    assert g.full_x[1] - g.full_x[0] == pytest.approx(L, abs=1e-5)
    assert g.real_scale == pytest.approx(L / 256, abs=1e-7)
    assert g.blocks_x == int(np.ceil(L / 8.0)) == 6
    assert g.load_x == 2
    # z: 19.5 a of structure, padded by 2 slices on top and 2 below (>= 3 A each)
    assert g.n_slices == 43
    assert g.pre_slices == 2
    assert g.load_z == 2


def test_geometry_config2_band_limit(oracle):
    w = wl.stem_srtio3()
    g = oracle.geometry(oracle.scenario(**w["scenario"]))
    assert g.real_scale * 512 == pytest.approx(3.905 + 16.0, abs=1e-9)
    assert g.n_slices == 83
    # 2/3 Nyquist in mrad: (2/3)(256/L) lambda
    bl = (2.0 / 3.0) * (256.0 / (3.905 + 16.0)) * g.wavelength * 1e3
    assert g.kmax_band * g.limit_factor * g.wavelength * 1e3 == pytest.approx(bl, rel=1e-12)
    assert bl > 200.0  # the 60-200 mrad annulus lies inside the band limit


def test_product_geometry_matches_oracle(oracle):
    import cltem_amd
    cases = [wl.ctem_srtio3()["scenario"], wl.stem_srtio3()["scenario"],
             wl.stem_srtio3(resolution=256, pixels=8)["scenario"],
             dict(wl.ctem_srtio3(4, 3, 7, resolution=384, slice_thickness=1.3)["scenario"],
                  slice_offset=0.4, beam_tilt_mrad=5.0, beam_azimuth=0.3, area_padding=2.5),
             dict(wl.ctem_srtio3(2, 2, 2, resolution=768)["scenario"], full3d=1, voltage_kv=300.0)]
    for sc in cases:
        a = cltem_amd.plan_geometry(cltem_amd.make_scenario(**sc))
        b = oracle.geometry(oracle.scenario(**sc))
        for f, _ in b._fields_:
            va, vb = getattr(a, f), getattr(b, f)
            try:
                va, vb = list(va), list(vb)
            except TypeError:
                pass
            assert va == vb, f


# ---- binning -----------------------------------------------------------------------------
def _toy_geometry(oracle, res=64, dz=2.0):
    # area [0, 16] with no area padding and zero cell padding -> full limits exactly [0, 16]
    sc = dict(mode=1, resolution=res, slice_thickness=dz, struct_min=(0, 0, 0),
              struct_max=(16, 16, 10), area=(0, 16, 0, 16), pad_xy=(0.0, 0.0), pad_z=(0.0, 0.0))
    return oracle.geometry(oracle.scenario(**sc))


def test_binning_hand_cases(oracle):
    g = _toy_geometry(oracle)
    assert (g.full_x[0], g.full_x[1], g.blocks_x) == (0.0, 16.0, 2)
    assert (g.z_lims[0], g.z_lims[1], g.n_slices) == (0.0, 10.0, 5)
    x = np.array([8.0, 7.999999, 15.999, 0.5, 16.0, 0.0, 4.0, 4.0, 4.0, 4.0])
    y = np.array([1.0, 1.0, 15.999, 8.0, 1.0, 1.0, 4.0, 4.0, 4.0, 4.0])
    z = np.array([5.0, 5.0, 9.999, 0.001, 5.0, 5.0, 8.0, 2.0, 0.0, 10.0])
    Z = np.full(len(x), 8)
    r = oracle.sort_atoms(g, x, y, z, Z)
    # atoms 4 (x == max), 5 (x == min), 8 (z == min), 9 (z == max) fail the strict host filter
    assert r["n_filtered"] == 6 and r["n_in_range"] == 6
    # x = 8 -> floor(0.5 * 2) = 1 ; x = 7.999999 -> 0 ; y = 8 -> block row 1
    # z: zid = floor((10 - z) / 2): 5 -> 2, 9.999 -> 0, 0.001 -> 4, 8 -> 1, 2 -> 4
    exp_bid = [1, 0, 3, 2, 0, 0]
    exp_zid = [2, 2, 0, 4, 1, 4]
    np.testing.assert_array_equal(r["block_ids"], exp_bid)
    np.testing.assert_array_equal(r["z_ids"], exp_zid)


def test_binning_invariants_config1_displaced(oracle):
    w = wl.ctem_srtio3()
    g = oracle.geometry(oracle.scenario(**w["scenario"]))
    spec = w["specimen"].displaced(0.08, 7, 3)
    r = oracle.sort_atoms(g, spec.x, spec.y, spec.z, spec.Z)
    bs = r["block_start"]
    assert np.all(np.diff(bs) >= 0) and bs[-1] == r["n_in_range"] and bs[0] == 0
    nb = g.blocks_x * g.blocks_y
    bx = (g.full_x[1] - g.full_x[0]) / g.blocks_x
    # every sorted atom lies inside its bin (edges inclusive) and bins keep input order
    keep = ((spec.x > g.full_x[0]) & (spec.x < g.full_x[1]) & (spec.y > g.full_y[0])
            & (spec.y < g.full_y[1]) & (spec.z > g.z_lims[0]) & (spec.z < g.z_lims[1]))
    orig_index = np.nonzero(keep)[0]
    for key in range(len(bs) - 1):
        lo, hi = bs[key], bs[key + 1]
        if hi == lo:
            continue
        s_, b_ = divmod(key, nb)
        j, i = divmod(b_, g.blocks_x)
        xs, ys, zs = r["x"][lo:hi], r["y"][lo:hi], r["z"][lo:hi]
        assert np.all(xs >= g.full_x[0] + i * bx - 1e-9) and np.all(xs <= g.full_x[0] + (i + 1) * bx + 1e-9)
        assert np.all(ys >= g.full_y[0] + j * bx - 1e-9) and np.all(ys <= g.full_y[0] + (j + 1) * bx + 1e-9)
        top = g.z_lims[1] - s_ * g.slice_thickness
        assert np.all(zs <= top + 1e-9) and np.all(zs >= top - g.slice_thickness - 1e-9)
        # stability: the sorted run is the input order restricted to this bin
        idx = [orig_index[k] for k in range(len(orig_index))
               if r["block_ids"][k] == b_ and r["z_ids"][k] == s_]
        np.testing.assert_array_equal(r["x"][lo:hi], spec.x[idx])


def test_binning_numpy_restatement(oracle):
    """The atom_sort_d formula evaluated by numpy (IEEE double, same operation order)."""
    w = wl.stem_srtio3()
    g = oracle.geometry(oracle.scenario(**w["scenario"]))
    spec = w["specimen"].displaced(0.08, 20261202, 1)
    r = oracle.sort_atoms(g, spec.x, spec.y, spec.z, spec.Z)
    keep = ((spec.x > g.full_x[0]) & (spec.x < g.full_x[1]) & (spec.y > g.full_y[0])
            & (spec.y < g.full_y[1]) & (spec.z > g.z_lims[0]) & (spec.z < g.z_lims[1]))
    x, y, z = spec.x[keep], spec.y[keep], spec.z[keep]
    bidx = np.floor((x - g.full_x[0]) * (1.0 / (g.full_x[1] - g.full_x[0])) * g.blocks_x).astype(int)
    bidy = np.floor((y - g.full_y[0]) * (1.0 / (g.full_y[1] - g.full_y[0])) * g.blocks_y).astype(int)
    zid = np.floor((g.z_lims[1] - z) * (1.0 / g.slice_thickness)).astype(int)
    zid -= zid == g.n_slices
    bidx -= bidx == g.blocks_x
    bidy -= bidy == g.blocks_y
    np.testing.assert_array_equal(r["block_ids"], bidx + g.blocks_x * bidy)
    np.testing.assert_array_equal(r["z_ids"], zid)


# ---- Bessel functions and potentials -------------------------------------------------
def test_bessel_vs_scipy(oracle):
    d = np.load(os.path.join(GOLD, "bessel.npz"))
    L = oracle.lib()
    k0 = np.array([L.orc_bessk0(v) for v in d["x"]])
    k1 = np.array([L.orc_bessk1(v) for v in d["x"]])
    np.testing.assert_allclose(k0, d["k0"], rtol=1e-12)
    np.testing.assert_allclose(k1, d["k1"], rtol=1e-12)


@pytest.mark.parametrize("name", ["kirkland", "peng", "lobato"])
def test_projected_potentials_vs_scipy(oracle, name):
    d = np.load(os.path.join(GOLD, "vproj_scipy.npz"))
    import ctypes as C
    p = P.load(name)
    L = oracle.lib()
    vp = np.ascontiguousarray(p.values)
    ptr = vp.ctypes.data_as(C.POINTER(C.c_double))
    for i, Z in enumerate(d["Z"]):
        got = np.array([L.orc_vproj(p.form, ptr, p.i_per_atom, int(Z), r) for r in d["r"]])
        np.testing.assert_allclose(got, d[name][i], rtol=1e-11)


@pytest.mark.parametrize("name,Z", [("kirkland", 8), ("kirkland", 79), ("peng", 38), ("peng", 14)])
def test_projection_identity_3d(oracle, name, Z):
    """Integral over z of the full-3D form (kernels/transmission_potentials_full_3d_d.cl:57-134)
    against its analytic projection: int exp(-a r)/r dz = 2 K0(a rho) and
    int exp(-pi^2 r^2/d) dz = sqrt(d/pi) exp(-pi^2 rho^2/d).  With the reference's constants
    the Lorentzian term projects onto the projected kernel's 300.8242834 exactly, while the
    Gaussian prefactor 266.5157269 projects to 266.5157269/sqrt(pi) = 150.3654, 3.1e-4 below
    the projected kernel's 150.4121417 (a mismatch of the reference, reproduced as written)."""
    import ctypes as C
    import scipy.special as ss
    p = P.load(name)
    vp = np.ascontiguousarray(p.values)
    ptr = vp.ctypes.data_as(C.POINTER(C.c_double))
    L = oracle.lib()
    c = p.values[(Z - 1) * p.per_atom:Z * p.per_atom]
    g3 = 266.5157269 / np.sqrt(np.pi)
    for rho in (0.05, 0.3, 1.0, 2.5):
        f = lambda zz: L.orc_v3d(p.form, ptr, p.i_per_atom, Z, float(np.hypot(rho, zz)))
        integral = 2 * si.quad(f, 0.0, 60.0, limit=400, epsabs=0, epsrel=1e-12, points=[rho])[0]
        if p.form == 0:
            a, b, cc, d = c[0:6:2], c[1:6:2], c[6:12:2], c[7:12:2]
            expect = (2 * 150.4121417 * np.sum(a * ss.k0(2 * np.pi * rho * np.sqrt(b)))
                      + g3 * np.sum(cc / d * np.exp(-(np.pi * rho) ** 2 / d)))
            proj = L.orc_vproj(0, ptr, 3, Z, rho)
            lor = 300.8242834 * np.sum(a * ss.k0(2 * np.pi * rho * np.sqrt(b)))
            # the projected kernel = same Lorentzian term + Gaussian term with 150.4121417
            assert proj - lor == pytest.approx(150.4121417 / g3 * (expect - lor), rel=1e-9)
        else:
            a, b = c[:5], c[5:]
            expect = g3 * np.sum(a / b * np.exp(-(np.pi * rho) ** 2 / b))
            assert L.orc_vproj(1, ptr, 5, Z, rho) == pytest.approx(150.4121417 / g3 * expect, rel=1e-12)
        assert integral == pytest.approx(expect, rel=1e-9)


def test_lobato_3d_projection(oracle):
    """The full-3D Lobato form integrates to 945.09 sum a b^-3/2 (sqrt(b)/pi K0 + rho K1); the
    reference's projected Lobato kernel omits the sqrt(b)/pi factor on K0 (reproduced as is)."""
    import ctypes as C
    import scipy.special as ss
    p = P.load("lobato")
    vp = np.ascontiguousarray(p.values)
    ptr = vp.ctypes.data_as(C.POINTER(C.c_double))
    L = oracle.lib()
    Z = 14
    c = p.values[(Z - 1) * 10:Z * 10]
    a, b = c[:5], c[5:]
    for rho in (0.2, 1.0, 3.0):
        f = lambda zz: L.orc_v3d(2, ptr, 5, Z, float(np.hypot(rho, zz)))
        integral = 2 * si.quad(f, 0.0, 60.0, limit=400, epsrel=1e-12, points=[rho])[0]
        x = 2 * np.pi * rho / np.sqrt(b)
        expect = 945.090144399935 * np.sum(a * b**-1.5 * (np.sqrt(b) / np.pi * ss.k0(x) + rho * ss.k1(x)))
        assert integral == pytest.approx(expect, rel=1e-7)


def _one_atom(oracle, x, y, z, Zn=38, res=64, dz=2.0):
    g = _toy_geometry(oracle, res=res, dz=dz)
    r = oracle.sort_atoms(g, [x], [y], [z], [Zn])
    return g, r


def test_single_atom_rmin_and_cutoff(oracle):
    import ctypes as C
    p = P.load("kirkland")
    g, r = _one_atom(oracle, 4.0, 4.0, 5.0)
    px = g.real_scale  # 0.25 A
    V = oracle.potential_slice(g, r, p.values, 0, 3, slice_=2, nthreads=1)
    vp = np.ascontiguousarray(p.values)
    ptr = vp.ctypes.data_as(C.POINTER(C.c_double))
    L = oracle.lib()
    # the atom sits exactly on pixel (16, 16): r = 0 -> clamped to 0.25 px
    assert V[16, 16] == L.orc_vproj(0, ptr, 3, 38, 0.25 * px)
    yy, xx = np.mgrid[0:64, 0:64]
    rr = np.hypot(xx * px - 4.0, yy * px - 4.0)
    assert np.all(V[rr > 8.0] == 0.0)
    inside = (rr <= 8.0) & (rr >= 0.25 * px)
    ref = np.array([L.orc_vproj(0, ptr, 3, 38, v) for v in rr[inside]])
    np.testing.assert_allclose(V[inside], ref, rtol=1e-13)
    # other slices are empty
    assert np.all(oracle.potential_slice(g, r, p.values, 0, 3, slice_=0, nthreads=1) == 0)


def test_superposition(oracle):
    p = P.load("kirkland")
    g = _toy_geometry(oracle)
    r1 = oracle.sort_atoms(g, [3.1], [5.2], [5.0], [22])
    r2 = oracle.sort_atoms(g, [9.7], [6.3], [5.5], [8])
    r12 = oracle.sort_atoms(g, [3.1, 9.7], [5.2, 6.3], [5.0, 5.5], [22, 8])
    V1 = oracle.potential_slice(g, r1, p.values, 0, 3, 2, 1)
    V2 = oracle.potential_slice(g, r2, p.values, 0, 3, 2, 1)
    V12 = oracle.potential_slice(g, r12, p.values, 0, 3, 2, 1)
    np.testing.assert_allclose(V12, V1 + V2, rtol=1e-15, atol=1e-12)


# ---- FFT --------------------------------------------------------------------------------
def test_fft_golden_vectors(oracle):
    d = np.load(os.path.join(GOLD, "fft_ortho.npz"))
    for k in d.files:
        if not k.startswith("in_"):
            continue
        shape = k[3:]
        a = d[k]
        np.testing.assert_allclose(oracle.fft2(a), d["fwd_" + shape], rtol=0, atol=1e-13)
        np.testing.assert_allclose(oracle.fft2(a, inverse=True), d["inv_" + shape], rtol=0, atol=1e-13)


@pytest.mark.parametrize("shape", [(2, 4), (8, 8), (16, 64), (512, 512), (1024, 256)])
def test_fast_fft_matches(oracle, shape):
    """The CPU baseline's Stockham transform (power-of-two sizes) is the oracle's FFT: numpy's
    orthonormal transform both ways, and the oracle's own recursive one."""
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    f = oracle.fft2_fast(a)
    np.testing.assert_allclose(f, np.fft.fft2(a, norm="ortho"), rtol=0, atol=1e-13)
    np.testing.assert_allclose(oracle.fft2_fast(a, inverse=True), np.fft.ifft2(a, norm="ortho"), rtol=0,
                               atol=1e-13)
    np.testing.assert_allclose(f, oracle.fft2(a), rtol=0, atol=1e-13)


def test_fast_fft_refuses_other_sizes(oracle):
    a = np.zeros((384, 384), complex)
    assert oracle.lib().orc_fft2_fast(oracle._dp(a.view(np.float64)), 384, 384, 0) == -1


@pytest.mark.parametrize("n", [256, 384, 768])
def test_fft_identities(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    f = oracle.fft2(a)
    assert np.sum(np.abs(f) ** 2) == pytest.approx(np.sum(np.abs(a) ** 2), rel=1e-12)  # Parseval
    np.testing.assert_allclose(oracle.fft2(f, inverse=True), a, atol=1e-12)
    d = np.zeros((n, n), complex)
    d[0, 0] = 1.0
    np.testing.assert_allclose(oracle.fft2(d), np.full((n, n), 1.0 / n), atol=1e-15)  # impulse
    yy, xx = np.mgrid[0:n, 0:n]
    pw = np.exp(2j * np.pi * (5 * xx + 7 * yy) / n)  # plane wave -> one bin
    f = oracle.fft2(pw)
    assert abs(f[7, 5]) == pytest.approx(n, rel=1e-12)
    f[7, 5] = 0
    assert np.abs(f).max() < 1e-9


# ---- propagator, band limit, probe --------------------------------------------------
def _g512(oracle):
    return oracle.geometry(oracle.scenario(**wl.stem_srtio3()["scenario"]))


def test_propagator_closed_form(oracle):
    g = _g512(oracle)
    P_ = oracle.propagator(g)
    kx, ky = oracle.kgrid(g)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    assert P_[0, 0] == 1.0
    ii = np.where(np.arange(512) < 256, np.arange(512), np.arange(512) - 512)
    n2 = ii[None, :] ** 2 + ii[:, None] ** 2
    keep = n2 < (512 / 3.0) ** 2   # 512^2/9 is not an integer: no ties
    assert np.array_equal(np.abs(P_) > 0, keep)
    ref = np.exp(-1j * np.pi * g.wavelength * k2 * g.slice_thickness)
    np.testing.assert_allclose(P_[keep], ref[keep], atol=2e-12)


def test_band_limit_mask(oracle):
    g = _g512(oracle)
    t = np.ones((512, 512), complex)
    out = oracle.band_limit(g, t)
    ii = np.where(np.arange(512) < 256, np.arange(512), np.arange(512) - 512)
    keep = ii[None, :] ** 2 + ii[:, None] ** 2 <= (512 / 3.0) ** 2
    assert np.array_equal(out != 0, keep)


def test_probe_aperture_and_shift(oracle):
    g = _g512(oracle)
    kx, ky = oracle.kgrid(g)
    k = np.hypot(kx[None, :], ky[:, None])
    pr = oracle.probe_kspace(g, 0.0, 0.0, None, 20.0)
    inside = k < 0.020 / g.wavelength
    np.testing.assert_allclose(np.abs(pr[inside]), 1.0, atol=1e-14)
    assert np.all(pr[~inside] == 0)
    # unitary inverse FFT keeps the norm: sum |psi|^2 = pixels inside the aperture
    psi = oracle.fft2(pr, inverse=True)
    assert np.sum(np.abs(psi) ** 2) == pytest.approx(inside.sum(), rel=1e-12)
    # translation is a phase ramp exp(-2 pi i k.r)
    pr2 = oracle.probe_kspace(g, 1.3, -0.7, None, 20.0)
    ramp = np.exp(-2j * np.pi * (kx[None, :] * 1.3 + ky[:, None] * -0.7))
    np.testing.assert_allclose(pr2[inside], ramp[inside], atol=1e-11)


def test_probe_defocus_and_astigmatism(oracle):
    g = _g512(oracle)
    kx, ky = oracle.kgrid(g)
    lam = g.wavelength
    ab = np.zeros(25)
    ab[0] = 50.0                      # C10 (A)
    ab[1], ab[2] = 12.0, -5.0         # C12 as (re, im)
    ab[7] = 1.0e4                     # C30
    pr = oracle.probe_kspace(g, 0.0, 0.0, ab, 20.0)
    w = lam * (kx[None, :] + 1j * ky[:, None])
    wc = np.conj(w)
    chi = 2 * np.pi / lam * (0.5 * ab[0] * np.abs(w) ** 2 + np.real(0.5 * (ab[1] + 1j * ab[2]) * wc**2)
                             + 0.25 * ab[7] * np.abs(w) ** 4)
    inside = np.abs(pr) > 0
    np.testing.assert_allclose(pr[inside], np.exp(-1j * chi[inside]), atol=1e-10)


# ---- whole multislice identities ---------------------------------------------------
def test_empty_specimen_plane_wave_is_unity(oracle):
    g = _g512(oracle)
    trans = np.ones((5, 512, 512), complex)
    P_ = oracle.propagator(g)
    psi, recip = oracle.run_plane(trans, P_)
    np.testing.assert_allclose(psi, 1.0, atol=1e-12)


def test_empty_specimen_bright_field(oracle):
    w = wl.stem_srtio3(resolution=256, pixels=4)
    g = oracle.geometry(oracle.scenario(**w["scenario"]))
    trans = np.ones((g.n_slices, 256, 256), complex)
    P_ = oracle.propagator(g)
    pos = wl.stem_positions(g)[:3]
    det = np.array([[0.0, 30.0, 0, 0], [30.0, 100.0, 0, 0]])
    out = oracle.run_stem(g, trans, P_, pos, det, cond_ap_mrad=20.0)
    kx, ky = oracle.kgrid(g)
    n_ap = np.sum(np.hypot(kx[None, :], ky[:, None]) < 0.020 / g.wavelength)
    np.testing.assert_allclose(out[:, 0], np.sqrt(2) * n_ap, rtol=1e-12)
    np.testing.assert_allclose(out[:, 1], 0.0, atol=1e-9)


def test_multislice_vs_numpy_restatement(oracle):
    """A 64^2 toy run of the oracle against an independent numpy multislice loop."""
    g = _toy_geometry(oracle, res=64, dz=2.0)
    p = P.load("kirkland")
    r = oracle.sort_atoms(g, [4.0, 11.0, 8.3], [5.0, 9.0, 12.1], [7.0, 3.0, 5.2], [79, 8, 22])
    trans = oracle.transmission(g, r, p.values, 0, 3)
    P_ = oracle.propagator(g)
    psi, recip = oracle.run_plane(trans, P_)
    ref = np.ones((64, 64), complex)
    for s in range(g.n_slices):
        rec = np.fft.fft2(trans[s] * ref, norm="ortho") * P_
        ref = np.fft.ifft2(rec, norm="ortho")
    np.testing.assert_allclose(psi, ref, atol=1e-12)
    np.testing.assert_allclose(recip, rec, atol=1e-12)


# ---- full-3D slice sums (kernels/transmission_potentials_full_3d_d.cl:239-279) ----------
def _toy3d(oracle, dz=1.0, integrals=20, tilt_mrad=0.0, azimuth=0.0):
    """64^2 toy field [0, 16]^2 (0.25 A pixels, two 8 A blocks, load_x 1), z in [0, 10] with
    no padding: slices of dz = 1 A with tops at z = 10 - s, full 3D with `integrals` points."""
    sc = dict(mode=1, resolution=64, slice_thickness=dz, struct_min=(0, 0, 0), struct_max=(16, 16, 10),
              area=(0, 16, 0, 16), pad_xy=(0.0, 0.0), pad_z=(0.0, 0.0), full3d=1,
              full3d_integrals=integrals, beam_tilt_mrad=tilt_mrad, beam_azimuth=azimuth)
    g = oracle.geometry(oracle.scenario(**sc))
    assert (g.z_lims[0], g.z_lims[1], g.n_slices, g.load_z) == (0.0, 10.0, int(round(10 / dz)), 3)
    return g


def _v3d_fn(oracle, name, Z):
    import ctypes as C
    p = P.load(name)
    vp = np.ascontiguousarray(p.values)
    ptr = vp.ctypes.data_as(C.POINTER(C.c_double))
    L = oracle.lib()
    return p, (lambda r: L.orc_v3d(p.form, ptr, p.i_per_atom, Z, float(r))), vp


@pytest.mark.parametrize("name,Z", [("kirkland", 38), ("peng", 22), ("lobato", 8)])
@pytest.mark.parametrize("z0", [4.0, 5.0, 6.5])
def test_full3d_slice_sum_integrates_isolated_atom(oracle, name, Z, z0):
    """For an isolated atom whose 3 A window edge z0 + 3 falls on a slice top (z0 = 4, 5: an
    atom on a slice boundary at 5), summing sst * sumz over every slice gives the trapezoid
    rule, step dz/integrals, of V3D(sqrt(rho^2 + (z - z0)^2)) over the admitted range
    [bottom of slice zid + load_z, z0 + 3]: signed window (points below the atom are kept,
    points more than 3 A above are not), each slice's own trapezoid with no double-counted
    boundary point.  Checked against scipy quad plus the Euler-Maclaurin term
    h^2/12 (f'(b) - f'(a)), to O(h^4).  z0 = 6.5 puts the edge inside a slice: the admitted
    points then start at the first grid point at or below z0 + 3 and the first one carries an
    extra half step (its stale p2 is 0), i.e. the range ends half a step above it."""
    g = _toy3d(oracle)
    p, f3, _ = _v3d_fn(oracle, name, Z)
    x0, y0 = 6.1, 7.3
    r = oracle.sort_atoms(g, [x0], [y0], [z0], [Z])
    zid = int(r["z_ids"][0])
    assert zid == int(np.floor(10.0 - z0))
    h = g.slice_thickness / g.full3d_integrals
    total = np.zeros((64, 64))
    for s in range(g.n_slices):
        total += h * oracle.potential_slice(g, r, p.values, p.form, p.i_per_atom, s, nthreads=1)
    z_bot = 10.0 - (min(zid + 3, g.n_slices - 1) + 1) * g.slice_thickness
    top_slice = max(zid - 3, 0)
    z_top_slice = 10.0 - top_slice * g.slice_thickness
    if z0 + 3.0 >= z_top_slice:
        z_top, extra = z_top_slice, 0.0
    else:  # the first admitted grid point below the edge, plus the lone half step
        k = np.ceil((z_top_slice - (z0 + 3.0)) / h - 1e-9)
        z_top, extra = z_top_slice - k * h, 0.5 * h
    # pixels 0.6 to 3 A from the atom's axis: the integrand is smooth on the step's scale there,
    # so the Euler-Maclaurin expansion holds to O(h^4)
    for (xi, yi) in ((26, 31), (30, 30), (20, 40), (34, 22)):
        rho = float(np.hypot(xi * 0.25 - x0, yi * 0.25 - y0))
        assert rho <= 8.0
        f = lambda zz: f3(np.hypot(rho, zz - z0))
        a, b = z_bot, z_top
        integral = si.quad(f, a, b, limit=500, epsabs=0, epsrel=1e-13, points=[z0] if a < z0 < b else None)[0]
        d = 1e-5
        fp = lambda zz: (f(zz + d) - f(zz - d)) / (2 * d)
        em = h * h / 12.0 * (fp(b) - fp(a))
        got = total[yi, xi]
        if extra == 0.0:
            print(f"rho {rho:.2f}: sum {got:.10g}, quad + EM {integral + em:.10g}")
            assert got == pytest.approx(integral + em, rel=1e-9, abs=1e-9), (rho, got, integral)
        else:
            # the lone half step: [z_top, z_top + h/2] by its left value, error <= h^2/8 |f'|
            lone = extra * f(z_top)
            print(f"rho {rho:.2f}: sum - lone {got - lone:.10g}, quad + EM {integral + em:.10g}")
            assert got - lone == pytest.approx(integral + em, rel=1e-9, abs=1e-9)
            half = si.quad(f, z_top, z_top + extra, epsabs=0, epsrel=1e-13)[0]
            assert abs(got - integral - half - em) <= h * h / 8 * abs(fp(z_top)) * 1.05 + 1e-9 * abs(got)


@pytest.mark.parametrize("order", ["a_first", "b_first"])
def test_full3d_carried_p2_between_atoms_of_a_block_row(oracle, order):
    """Two atoms in one (slice, block row) bin, both within 8 A of the pixel.  In slice 3
    (z in [6, 7]) atom A at z = 4 admits every point (7 - 4 = 3 <= 3); atom B at z = 3.83 first
    admits h0 = 4 (7 - 0.2 - 3.83 <= 3).  The reference keeps p2 across atoms of a row, so
    B's first admitted point is paired with the last point of the atom before it.  Expected
    sumz from point values of V3D with explicit trapezoid weights; the atoms' order in the row
    is their input order (stable binning), so swapping them changes the sum."""
    name, Zn = "kirkland", 38
    g = _toy3d(oracle)
    p, f3, _ = _v3d_fn(oracle, name, Zn)
    A = (2.5, 2.0, 4.0)
    B = (1.5, 2.5, 3.83)
    atoms = [A, B] if order == "a_first" else [B, A]
    r = oracle.sort_atoms(g, [a[0] for a in atoms], [a[1] for a in atoms], [a[2] for a in atoms], [Zn, Zn])
    assert list(r["z_ids"]) == [6, 6] and list(r["block_ids"]) == [0, 0]
    xi, yi = 8, 8  # pixel at (2, 2)
    V = oracle.potential_slice(g, r, p.values, p.form, p.i_per_atom, 3, nthreads=1)
    I, step = g.full3d_integrals, g.slice_thickness / g.full3d_integrals

    def pts(atom):
        rho = np.hypot(xi * 0.25 - atom[0], yi * 0.25 - atom[1])
        zs = 7.0 - np.arange(I + 1) * step
        return np.array([f3(np.hypot(rho, z - atom[2])) for z in zs]), (zs - atom[2]) <= 3.0

    pa, ka = pts(A)
    pb, kb = pts(B)
    assert ka.all() and (not kb[:4].any()) and kb[4:].all()
    trap_a = sum(0.5 * (pa[h] + pa[h - 1]) for h in range(1, I + 1))
    trap_b = sum(0.5 * (pb[h] + pb[h - 1]) for h in range(5, I + 1))
    if order == "a_first":
        expect = trap_a + trap_b + 0.5 * (pb[4] + pa[I])   # B's first point pairs with A's last
    else:
        expect = trap_b + 0.5 * (pb[4] + 0.0) + trap_a      # B first in its row: p2 starts at 0
    assert V[yi, xi] == pytest.approx(expect, rel=1e-13)
    # and the two orders differ by exactly the carried value
    assert abs(0.5 * pa[I]) > 1e-3 * abs(expect)


def test_full3d_tilt_shift_is_cumulative(oracle):
    """With a beam tilt each integration point moves the lateral distance by one more step:
    point h of an atom uses rad - (h + 1) * (kx, ky) / kz * dz / integrals
    (kernels/transmission_potentials_full_3d_d.cl:245-246), so the slice sum of a single atom
    equals the trapezoid over points at those shifted distances."""
    tilt, az = 30.0, 0.7   # mrad, rad
    g = _toy3d(oracle, tilt_mrad=tilt, azimuth=az)
    name, Zn = "peng", 22
    p, f3, _ = _v3d_fn(oracle, name, Zn)
    atom = (6.0, 7.0, 4.0)  # window edge at slice 3's top: every point of slices 3..9 admitted
    r = oracle.sort_atoms(g, [atom[0]], [atom[1]], [atom[2]], [Zn])
    I, step = g.full3d_integrals, g.slice_thickness / g.full3d_integrals
    k = np.array(g.k_vec[:])
    sx, sy = k[0] / k[2] * step, k[1] / k[2] * step
    assert abs(np.hypot(k[0], k[1]) / k[2] - np.tan(tilt * 1e-3)) < 1e-12
    xi, yi = 25, 27
    for s in (3, 5, 9):
        V = oracle.potential_slice(g, r, p.values, p.form, p.i_per_atom, s, nthreads=1)
        top = 10.0 - s
        vals = []
        for h in range(I + 1):
            rx = xi * 0.25 - atom[0] - (h + 1) * sx
            ry = yi * 0.25 - atom[1] - (h + 1) * sy
            vals.append(f3(np.sqrt(rx * rx + ry * ry + (top - h * step - atom[2]) ** 2)))
        expect = sum(0.5 * (vals[h] + vals[h - 1]) for h in range(1, I + 1))
        assert V[yi, xi] == pytest.approx(expect, rel=1e-12), s
        # without the cumulative shift (a single shift per atom) the sum is different
        alt = []
        for h in range(I + 1):
            rx = xi * 0.25 - atom[0] - sx
            ry = yi * 0.25 - atom[1] - sy
            alt.append(f3(np.sqrt(rx * rx + ry * ry + (top - h * step - atom[2]) ** 2)))
        assert abs(sum(0.5 * (alt[h] + alt[h - 1]) for h in range(1, I + 1)) - expect) > 1e-9 * abs(expect)


# ---- off-centre detector (kernels/band_pass_d.cl:15-38, simulationstem.cpp:98-131) -----
@pytest.mark.parametrize("det", [(5.0, 25.0, 10.0, -5.0), (0.0, 12.0, -14.0, 9.0), (8.0, 30.0, 3.0, 17.0)])
def test_off_centre_annulus_counts_pixels(oracle, det):
    """Empty specimen: |Psi(k)| = 1 inside the condenser aperture (k < alpha / lambda) and 0
    outside, so a detector reads sqrt(2) x the number of aperture pixels inside its annulus.
    The annulus is centred at (W/2 + xc, H/2 + yc) pixels of the fft-shifted image (xc, yc in
    mrad over the angular pixel scale), counted here in numpy from the k-grid."""
    w = wl.stem_srtio3(resolution=256, pixels=4)
    g = oracle.geometry(oracle.scenario(**w["scenario"]))
    trans = np.ones((g.n_slices, 256, 256), complex)
    P_ = oracle.propagator(g)
    pos = wl.stem_positions(g)[:2]
    out = oracle.run_stem(g, trans, P_, pos, np.array([det]), cond_ap_mrad=20.0)
    # independent count: shifted pixel (xs, ys) holds unshifted ((xs + W/2) % W, (ys + H/2) % H)
    n = 256
    kx, ky = np.fft.fftfreq(n, d=g.real_scale), np.fft.fftfreq(n, d=g.real_scale)
    us = (np.arange(n) + n // 2) % n
    kk = np.hypot(kx[us][None, :], ky[us][:, None])
    aperture = kk < 0.020 / g.wavelength
    sc = g.inverse_scale_angle
    ys, xs = np.mgrid[0:n, 0:n]
    rr = np.hypot(xs - (n / 2 + det[2] / sc), ys - (n / 2 + det[3] / sc))
    ring = (rr >= det[0] / sc) & (rr <= det[1] / sc)
    count = int(np.sum(aperture & ring))
    assert 0 < count < int(np.sum(aperture))
    np.testing.assert_allclose(out[:, 0], np.sqrt(2) * count, rtol=1e-12)


@pytest.mark.parametrize("det", [(5.0, 25.0, 10.0, -5.0), (0.0, 12.0, -14.0, 9.0), (8.0, 30.0, 3.0, 17.0)])
def test_off_centre_annulus_direction(oracle, det):
    """The aperture disc above is symmetric, so it cannot tell the offset's direction: on a
    random unshifted Psi(k) the detector is sqrt(2) x the sum of |Psi|^2 over the ring pixels
    of the fft-shifted image centred at (W/2 + xc, H/2 + yc), and the mirrored centres read
    other values."""
    n = 128
    rng = np.random.default_rng(11)
    recip = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    sc = 0.9  # mrad per pixel
    got = oracle.stem_pixel(recip, det[0], det[1], det[2], det[3], sc, sqrt2=True)
    shifted = np.fft.fftshift(np.abs(recip) ** 2)
    ys, xs = np.mgrid[0:n, 0:n]

    def ring(cx, cy):
        rr = np.hypot(xs - (n / 2 + cx / sc), ys - (n / 2 + cy / sc))
        return np.sqrt(2) * float(np.sum(shifted[(rr >= det[0] / sc) & (rr <= det[1] / sc)]))

    assert got == pytest.approx(ring(det[2], det[3]), rel=1e-12)
    for mx, my in ((-1, 1), (1, -1), (-1, -1)):
        assert abs(ring(mx * det[2], my * det[3]) - got) > 1e-6 * got
