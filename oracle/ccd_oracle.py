"""CPU restatement of clTEM's CCD response and dose step -- TEST INFRASTRUCTURE ONLY.

Follows SimulationCtem::simulateImageDose (src/simulation/microscope/simulationctem.cpp:174-263)
with ccd_dqe_d / ccd_ntf_d (kernels/ccd_dqe_d.cl, kernels/ccd_ntf_d.cl) in float64 numpy,
and restates the engine's seeded Poisson generator (Philox4x32-10 uniforms; inversion below a
mean of 10, Hormann's PTRS transformed rejection above) so draws can be compared one by one.

Pins (tests/test_oracle_ccd.py): Philox known answers; the radial filter on linear tables
(closed form, clamp); sampler mean/variance and chi-square against scipy.stats.poisson; the
whole step with unit tables and an infinite-dose limit.
"""
from __future__ import annotations

import numpy as np
from scipy.special import gammaln

_MASK = 0xFFFFFFFF


def philox4x32_10(counter, key):
    """Philox4x32-10 on (n, 4) counters and a (2,) key; returns (n, 4) uint32."""
    c = np.array(counter, dtype=np.uint64).reshape(-1, 4)
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    c0, c1, c2, c3 = (c[:, i].copy() for i in range(4))
    for r in range(10):
        if r > 0:
            k0 = np.uint64((int(k0) + 0x9E3779B9) & _MASK)
            k1 = np.uint64((int(k1) + 0xBB67AE85) & _MASK)
        p0 = c0 * np.uint64(0xD2511F53)
        p1 = c2 * np.uint64(0xCD9E8D57)
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(_MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(_MASK)
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return np.stack([c0, c1, c2, c3], axis=1).astype(np.uint32)


def _unit(u):
    return (u.astype(np.float64) + 0.5) / 4294967296.0


def poisson(lam, ids, seed):
    """Poisson deviates of means `lam` for streams `ids` (the engine's generator)."""
    lam = np.asarray(lam, dtype=np.float64)
    ids = np.asarray(ids, dtype=np.uint64)
    key = (seed & _MASK, (seed >> 32) & _MASK)
    out = np.zeros(lam.shape)
    small = (lam > 0) & (lam < 10)
    if small.any():
        r = philox4x32_10(np.stack([ids[small], 0 * ids[small], 0 * ids[small], 0 * ids[small]], axis=1), key)
        u = _unit(r[:, 0])
        ls = lam[small]
        p = np.exp(-ls)
        F = p.copy()
        k = np.zeros(ls.shape)
        active = u > F
        for _ in range(200):
            if not active.any():
                break
            k = np.where(active, k + 1, k)
            p = np.where(active, p * ls / np.maximum(k, 1), p)
            F = np.where(active, F + p, F)
            active = active & (u > F)
        out[small] = k
    big = lam >= 10
    if big.any():
        lb = lam[big]
        idb = ids[big]
        slam, loglam = np.sqrt(lb), np.log(lb)
        b = 0.931 + 2.53 * slam
        a = -0.059 + 0.02483 * b
        inv_alpha = 1.1239 + 1.1328 / (b - 3.4)
        vr = 0.9277 - 3.6224 / (b - 2.0)
        res = np.floor(lb + 0.5)
        todo = np.ones(lb.shape, bool)
        for it in range(64):
            if not todo.any():
                break
            r = philox4x32_10(np.stack([idb, np.full(idb.shape, it + 1, np.uint64), 0 * idb, 0 * idb], axis=1), key)
            U = _unit(r[:, 0]) - 0.5
            V = _unit(r[:, 1])
            us = 0.5 - np.abs(U)
            kk = np.floor((2.0 * a / us + b) * U + lb + 0.43)
            acc1 = (us >= 0.07) & (V <= vr)
            skip = (kk < 0) | ((us < 0.013) & (V > us))
            with np.errstate(invalid="ignore", divide="ignore"):
                acc2 = ~skip & (np.log(V) + np.log(inv_alpha) - np.log(a / (us * us) + b)
                                <= -lb + kk * loglam - gammaln(np.maximum(kk, 0) + 1.0))
            acc = todo & (acc1 | acc2)
            res = np.where(acc, kk, res)
            todo = todo & ~acc
        out[big] = res
    return out


def radial_filter(res, table, binning, take_sqrt):
    """ccd_dqe_d / ccd_ntf_d factors over the unshifted res x res spectrum."""
    t = np.asarray(table, dtype=np.float64)
    x = np.arange(res)
    xp = np.where(x < res // 2, x, x - res).astype(np.float64)
    rad = np.sqrt(xp[None, :] ** 2 + xp[:, None] ** 2)
    i0 = np.floor(rad / binning).astype(int)
    v0 = t[np.minimum(i0, len(t) - 1)]
    v1 = t[np.minimum(i0 + 1, len(t) - 1)]
    interp = rad / binning - np.floor(rad / binning)
    f = interp * v1 + (1 - interp) * v0
    return np.sqrt(f) if take_sqrt else f


def apply_filter(z, filt):
    """FFT -> x filt -> IFFT with clFFT's 1/sqrt(N) scaling both ways."""
    return np.fft.ifft2(np.fft.fft2(z, norm="ortho") * filt, norm="ortho")


def ccd_image(image, dqe, ntf, binning, dose_per_pixel, conversion=1.0, seed=0, lam_override=None):
    """The whole step; returns (output image, Poisson means used).  lam_override replaces the
    means (to draw from exactly the means another implementation used)."""
    res = image.shape[0]
    z = apply_filter(image.astype(np.complex128), radial_filter(res, dqe, binning, True))
    lam = dose_per_pixel * binning * binning * z.real if lam_override is None else lam_override
    counts = conversion * poisson(lam.ravel(), np.arange(lam.size), seed).reshape(lam.shape)
    out = apply_filter(counts.astype(np.complex128), radial_filter(res, ntf, binning, False)).real
    return out, lam
