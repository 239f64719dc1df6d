"""Pins of the plasmon-tilt pieces on the CPU: the oracle's diffraction translation
(translateDiffImage + bilinear_translate_d) and spherical rotation (rotateVectorSpherical),
and the seeded plasmon samplers of cltem_amd.incoherence (PlasmonScattering)."""
import numpy as np
import pytest
from scipy import stats
from scipy.spatial.transform import Rotation

from cltem_amd import incoherence as inc


@pytest.fixture(scope="module")
def img():
    rng = np.random.default_rng(5)
    return rng.random((32, 32))


@pytest.mark.parametrize("ix,iy", [(0, 0), (3, -2), (-7, 5), (31, 1)])
def test_translate_whole_pixels_is_roll(oracle, img, ix, iy):
    # inverse_scale 1: the shift in pixels is exactly dk
    out = oracle.translate_image(img, float(ix), float(iy), 1.0)
    np.testing.assert_array_equal(out, np.roll(img, (iy, ix), axis=(0, 1)))


def test_translate_fraction_is_bilinear_blend_of_rolls(oracle, img):
    sx, sy = 0.25, 0.625
    out = oracle.translate_image(img, 2 + sx, -3 + sy, 1.0)
    r = lambda a, b: np.roll(img, (b, a), axis=(0, 1))  # noqa: E731
    ref = ((1 - sx) * (1 - sy) * r(2, -3) + sx * (1 - sy) * r(3, -3) + (1 - sx) * sy * r(2, -2)
           + sx * sy * r(3, -2))
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-15)
    assert abs(out.sum() - img.sum()) <= 1e-12 * img.sum()


def test_translate_scales_by_inverse_scale(oracle, img):
    s = 0.0371
    out = oracle.translate_image(img, 4 * s + 0.5 * s, 0.0, s)
    ref = 0.5 * np.roll(img, 4, axis=1) + 0.5 * np.roll(img, 5, axis=1)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12)


def test_translate_wraps_only_once(oracle, img):
    n = img.shape[0]
    out = oracle.translate_image(img, float(n + 2), 0.0, 1.0)
    # a pixel taken from x - n - 2 is wrapped once: x >= 2 reads x - 2, x < 2 stays outside
    np.testing.assert_array_equal(out[:, 2:], img[:, :-2])
    assert not out[:, :2].any()


def test_rotation_gives_wavevector_and_matches_scipy(oracle):
    k = 39.87
    for theta_mrad, phi in [(0.0, 0.0), (5.0, 0.3), (12.0, -2.2), (40.0, 3.0)]:
        z, y = oracle.rotate_vector_spherical([0.0, 0.0, k], [0.0, 1.0, 0.0], theta_mrad / 1000, phi)
        t = theta_mrad / 1000
        # MicroscopeParameters::Wavevector (utilities/commonstructs.h:282-289)
        np.testing.assert_allclose(z, k * np.array([np.sin(t) * np.cos(phi), np.sin(t) * np.sin(phi), np.cos(t)]),
                                   rtol=0, atol=1e-12)
        np.testing.assert_allclose(y, [-np.sin(phi), np.cos(phi), 0.0], rtol=0, atol=1e-15)
    # a second rotation: about the current y by theta, then about the previous z by phi
    z0, y0 = oracle.rotate_vector_spherical([0.0, 0.0, k], [0.0, 1.0, 0.0], 0.01, 0.7)
    z1, y1 = oracle.rotate_vector_spherical(z0, y0, 0.004, 2.1)
    r = (Rotation.from_rotvec(2.1 * z0 / np.linalg.norm(z0)) *
         Rotation.from_rotvec(0.004 * y0 / np.linalg.norm(y0)))
    np.testing.assert_allclose(z1, r.apply(z0), rtol=0, atol=1e-12)
    np.testing.assert_allclose(y1, Rotation.from_rotvec(2.1 * z0 / np.linalg.norm(z0)).apply(y0), atol=1e-14)
    assert abs(np.linalg.norm(z1) - k) <= 1e-12 and abs(np.dot(z1, y1)) <= 1e-12


def test_python_rotation_matches_oracle(oracle):
    z, y = [0.1, -0.2, 39.8], [0.0, 1.0, 0.0]
    for theta, phi in [(0.003, 1.0), (0.02, -0.4)]:
        zp, yp = inc.rotate_vector_spherical(z, y, theta, phi)
        zo, yo = oracle.rotate_vector_spherical(z, y, theta, phi)
        np.testing.assert_allclose(zp, zo, rtol=0, atol=1e-13)
        np.testing.assert_allclose(yp, yo, rtol=0, atol=1e-15)
        z, y = zp, yp


def test_depths_are_a_poisson_process():
    # exponential free paths cut at the thickness: the number of scatterings is Poisson(T / mfp)
    # and, pooled, the depths are uniform over the thickness
    T, mfp = 300.0, 100.0
    counts, depths = [], []
    for job in range(3000):
        d = inc.plasmon_depths(3, job, T, mfp)
        counts.append(len(d))
        depths.extend(d)
    counts = np.array(counts)
    kmax = 9
    obs = np.bincount(np.minimum(counts, kmax), minlength=kmax + 1)
    p = stats.poisson.pmf(np.arange(kmax + 1), T / mfp)
    p[-1] = stats.poisson.sf(kmax - 1, T / mfp)
    assert stats.chisquare(obs, p * len(counts)).pvalue > 1e-3
    assert stats.kstest(np.array(depths) / T, "uniform").pvalue > 1e-3


def test_depths_stay_inside_and_individual_count(monkeypatch):
    for job in range(50):
        d = inc.plasmon_depths(9, job, 300.0, 100.0)
        assert all(0 < a < b for a, b in zip(d, d[1:])) and (not d or d[-1] <= 300.0)
        d2 = inc.plasmon_depths(9, job, 300.0, 100.0, individual=2)
        assert len(d2) == 2 and d2[-1] <= 300.0
    monkeypatch.setattr(inc, "_DEPTH_ATTEMPTS", 8)
    with pytest.raises(ValueError):
        inc.plasmon_depths(9, 0, 1.0, 1e6, individual=3)  # three scatterings in 1 A: none found


def test_polar_angles_follow_the_truncated_lorentzian():
    te, tc = 0.8, 12.0
    polar, az = inc.plasmon_angles(4, 1, 4000, te, tc)
    assert np.all(polar <= tc + 1e-12)
    # P(theta <= t) = ln(1 + t^2/te^2) / ln(1 + tc^2/te^2) (inverse of getScatteringPolar)
    cdf = lambda t: np.log1p((t / te) ** 2) / np.log1p((tc / te) ** 2)  # noqa: E731
    assert stats.kstest(polar, cdf).pvalue > 1e-3
    assert stats.kstest(az / (2 * np.pi), "uniform").pvalue > 1e-3


def test_events_one_per_slice_and_dropped_past_the_exit():
    # pre_slices 2, dz 2: slice i ends (i - 1) * 2 below the top; depths 0.5 and 0.7 both lie
    # in slice 2 but the second waits for slice 3; 100 A is below the exit surface
    ev = inc.plasmon_events([0.5, 0.7, 6.0, 100.0], pre_slices=2, dz=2.0, n_slices=10)
    assert ev == [2, 3, 4]
