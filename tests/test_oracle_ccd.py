"""Pins of the CCD restatement (oracle/ccd_oracle.py) and of the CCD table reader."""
import numpy as np
import pytest
from scipy import stats

from cltem_amd.ccd import TABLE_LEN, load_ccd
from oracle import ccd_oracle as co


def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        assert tuple(int(v) for v in co.philox4x32_10([ctr], key)[0]) == want


def test_ccd_tables_read_like_a_stream():
    for name, first, last in (("k2", 0.72963, 0.03842), ("orius", 0.40000001, 0.00227227)):
        c = load_ccd(name)
        assert len(c.dqe) == len(c.ntf) == TABLE_LEN
        assert c.dqe[0] == pytest.approx(first) and c.dqe[-1] == pytest.approx(last)
    # entries that run together in the file ("0.7056940.701858") are two values
    k2 = load_ccd("k2")
    assert k2.dqe[6] == pytest.approx(0.705694) and k2.dqe[7] == pytest.approx(0.701858)


def test_radial_filter_closed_form():
    res, binning = 64, 2
    table = np.arange(10, dtype=np.float64)  # f(r) = r / binning until the clamp at 9
    f = co.radial_filter(res, table, binning, take_sqrt=False)
    x = np.arange(res)
    xp = np.where(x < res // 2, x, x - res)
    rad = np.sqrt(xp[None, :] ** 2 + xp[:, None] ** 2)
    want = np.minimum(rad / binning, 9.0)
    inside = rad / binning < 8.999
    np.testing.assert_allclose(f[inside], want[inside], rtol=1e-12, atol=1e-12)
    assert np.all(f[rad / binning >= 10] == 9.0)
    np.testing.assert_allclose(co.radial_filter(res, table + 4, binning, True) ** 2,
                               co.radial_filter(res, table + 4, binning, False), rtol=1e-12)


@pytest.mark.parametrize("lam", [0.3, 3.0, 9.9, 10.0, 47.0, 2.0e4])
def test_poisson_moments(lam):
    n = 200_000
    k = co.poisson(np.full(n, lam), np.arange(n), seed=12345)
    assert np.all(k == np.floor(k)) and np.all(k >= 0)
    assert abs(k.mean() - lam) < 5 * np.sqrt(lam / n)
    # sample variance of a Poisson sample: sd ~ lam * sqrt(2 / n) for large lam
    assert abs(k.var() - lam) < 6 * np.sqrt((lam + 2 * lam * lam) / n)


@pytest.mark.parametrize("lam", [2.5, 30.0])
def test_poisson_distribution_chi_square(lam):
    n = 200_000
    k = co.poisson(np.full(n, lam), np.arange(n) + 7, seed=99).astype(int)
    lo, hi = int(stats.poisson.ppf(1e-4, lam)), int(stats.poisson.ppf(1 - 1e-4, lam))
    obs = np.array([np.sum(k == v) for v in range(lo, hi + 1)])
    exp = n * stats.poisson.pmf(np.arange(lo, hi + 1), lam)
    chi2 = np.sum((obs - exp) ** 2 / exp)
    assert stats.chi2.sf(chi2, len(obs) - 1) > 1e-4


def test_unit_tables_and_dose_limit():
    rng = np.random.default_rng(0)
    img = 1.0 + 0.2 * rng.standard_normal((32, 32))
    ones = np.ones(TABLE_LEN)
    out, lam = co.ccd_image(img, ones, ones, 1, 1e3, seed=5)
    np.testing.assert_allclose(lam, 1e3 * img, rtol=1e-12)
    np.testing.assert_allclose(out, np.round(out), atol=1e-9)  # unit NTF keeps integer counts
    # infinite-dose limit: counts / dose -> NTF(DQE-filtered image)
    c = load_ccd("k2")
    out, _ = co.ccd_image(img, c.dqe, c.ntf, 2, 1e10, seed=5)
    z = co.apply_filter(img.astype(complex), co.radial_filter(32, c.dqe, 2, True)).real
    want = co.apply_filter(z.astype(complex), co.radial_filter(32, c.ntf, 2, False)).real
    np.testing.assert_allclose(out / (4e10), want, rtol=0, atol=1e-4 * np.abs(want).max())
