"""Pins of the oracle's CTEM image formation (ctem_image + inverse FFT + sqabs)."""
import numpy as np
import pytest


def geom(oracle, res=64):
    scen = dict(mode=1, resolution=res, voltage_kv=200.0, slice_thickness=1.0,
                struct_min=(0.0, 0.0, 0.0), struct_max=(20.0, 20.0, 5.0), area=(0.0, 20.0, 0.0, 20.0))
    return oracle.geometry(oracle.scenario(**scen))


def test_aberration_free_wide_aperture_is_exit_intensity(oracle):
    g = geom(oracle)
    rng = np.random.default_rng(3)
    recip = rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))
    img = oracle.ctem_image(g, recip, obj_ap_mrad=1e4)
    ref = np.abs(np.fft.ifft2(recip, norm="ortho")) ** 2
    np.testing.assert_allclose(img, ref, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("c10,c30,beta,delta", [(-120.0, 0.0, 0.0, 0.0), (80.0, 1.2e7, 0.5, 30.0),
                                               (0.0, -5e6, 1.0, 0.0)])
def test_two_beam_closed_form(oracle, c10, c30, beta, delta):
    """Psi = a0 at k = 0 and a1 at one reflection: the image is
    |a0 + a1 T(k1) exp(2 pi i k1.r)|^2 / N with T = E_t E_s exp(-i chi),
    chi = pi lambda C10 k^2 + (pi / 2) C30 lambda^3 k^4,
    E_t = exp(-(pi delta lambda k^2)^2 / 4), E_s = exp(-(pi alpha k)^2 (C10 + C30 lambda^2 k^2)^2)."""
    g = geom(oracle)
    res = g.resolution
    kx, ky = oracle.kgrid(g)
    lam = g.wavelength
    i1, j1 = 5, 3
    a0, a1 = 0.9 + 0.1j, 0.2 - 0.3j
    recip = np.zeros((res, res), complex)
    recip[0, 0] = a0
    recip[j1, i1] = a1
    lens = np.zeros(25)
    lens[0] = c10
    lens[7] = c30
    img = oracle.ctem_image(g, recip, lens, obj_ap_mrad=1e4, beta_mrad=beta, delta=delta)
    k2 = kx[i1] ** 2 + ky[j1] ** 2
    chi = np.pi * lam * c10 * k2 + 0.5 * np.pi * c30 * lam ** 3 * k2 ** 2
    alpha = beta * 1e-3
    e_t = np.exp(-0.25 * (np.pi * delta * lam * k2) ** 2)
    e_s = np.exp(-(np.pi * alpha) ** 2 * k2 * (c10 + c30 * lam ** 2 * k2) ** 2)
    T1 = e_t * e_s * np.exp(-1j * chi)
    x = np.arange(res)
    ph = np.exp(2j * np.pi * (i1 * x[None, :] + j1 * x[:, None]) / res)
    ref = np.abs(a0 + a1 * T1 * ph) ** 2 / (res * res)
    np.testing.assert_allclose(img, ref, rtol=1e-10, atol=1e-16)


def test_objective_aperture_edge(oracle):
    """A reflection outside the aperture is removed; one inside the smoothing band is scaled
    by 1 - smoothstep(ap - r, ap + r, k)."""
    g = geom(oracle)
    res = g.resolution
    kx, ky = oracle.kgrid(g)
    lam = g.wavelength
    i1 = 6
    k = abs(kx[i1])
    ap = k * lam * 1e3 * 0.98        # aperture just inside k ...
    sm = k * lam * 1e3 * 0.05        # ... with a smoothing band that contains it
    recip = np.zeros((res, res), complex)
    recip[0, 0] = 1.0
    recip[0, i1] = 0.5
    img = oracle.ctem_image(g, recip, obj_ap_mrad=ap, obj_smooth_mrad=sm)
    e0, e1 = (ap - sm) * 1e-3 / lam, (ap + sm) * 1e-3 / lam
    t = np.clip((k - e0) / (e1 - e0), 0, 1)
    edge = 1 - t * t * (3 - 2 * t)
    x = np.arange(res)
    ref = np.abs(1.0 + 0.5 * edge * np.exp(2j * np.pi * i1 * x[None, :] / res)) ** 2 / res ** 2
    np.testing.assert_allclose(img, np.broadcast_to(ref, (res, res)), rtol=1e-12)
    img2 = oracle.ctem_image(g, recip, obj_ap_mrad=k * lam * 1e3 * 0.9)
    np.testing.assert_allclose(img2, np.full((res, res), 1.0 / res ** 2), rtol=1e-12)
