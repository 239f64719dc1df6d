"""Incoherent effects drawn per job, as the C++ host classes draw them.

ChromaticAberration::getFocusChange (incoherence/chromaticaberration.cpp:43-58, 99-106) and
ProbeSourceSize::getOffset (incoherence/probesourcesize.cpp:29-31) with Philox4x32-10 keyed
by a seed in place of clTEM's clock-seeded generators: counter (job, 0xC0, probe, 1) for the
focus change of one probe, (job, 0x50, 0, 2) for a job's source offset; uniforms
(u + 0.5) / 2^32, Box-Muller normals.
"""
from __future__ import annotations

import numpy as np

from .specimens import philox4x32

_FWHM = 2.0 * np.sqrt(2.0 * np.log(2.0))


def _uniforms(seed: int, ctr) -> np.ndarray:
    key = np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], dtype=np.uint64)
    r = philox4x32(np.asarray(ctr, dtype=np.uint64).reshape(-1, 4), key).astype(np.float64)
    return (r + 0.5) / 4294967296.0


def sigmas_from_hwhm(hwhm_neg: float, hwhm_pos: float):
    return abs(2.0 * hwhm_neg / _FWHM), abs(2.0 * hwhm_pos / _FWHM)


def focus_changes(seed: int, job: int, probes, cc_mm: float, sigma_neg: float, sigma_pos: float,
                  kv: float) -> np.ndarray:
    """Focus change (Angstrom) of each probe index in `probes` for one job."""
    probes = np.asarray(probes, dtype=np.uint64)
    n = len(probes)
    u = _uniforms(seed, np.stack([np.full(n, job), np.full(n, 0xC0), probes, np.ones(n)], axis=1))
    tot = sigma_neg + sigma_pos
    ratio = sigma_neg / tot if tot > 0 else 0.0
    mag = np.abs(np.sqrt(-2.0 * np.log(u[:, 1])) * np.cos(2.0 * np.pi * u[:, 2]))
    de = np.where(u[:, 0] >= ratio, mag * sigma_pos, -mag * sigma_neg)
    return 10000.0 * cc_mm * de / kv


def source_offset(seed: int, job: int, sigma: float):
    """(dx, dy) of the potential grid origin for one job (Angstrom)."""
    u = _uniforms(seed, [[job, 0x50, 0, 2]])[0]
    r = np.sqrt(-2.0 * np.log(u[0]))
    return r * np.cos(2.0 * np.pi * u[1]) * sigma, r * np.sin(2.0 * np.pi * u[1]) * sigma


def source_sigma_from_fwhm(fwhm: float) -> float:
    return fwhm / _FWHM


# ---- plasmon scattering (incoherence/inelastic/plasmon.cpp) --------------------------------
# Counters: (job, 0x9A + 256 * attempt, k, 4) for the k-th free path of an attempt; (job, 0xA5,
# n, 5) for the polar (word 0) and azimuthal (word 1) angles of the n-th scattering.
_DEPTH_ATTEMPTS = 1 << 20


def plasmon_depths(seed: int, job: int, thickness: float, mean_free_path: float, individual: int = 0):
    """generateScatteringDepths: depths (Angstrom below the specimen top) of one job's plasmon
    scatterings, each free path -mfp ln(u) (getScatteringDistance, plasmon.cpp:9-12), while the
    running depth stays <= thickness (getDistancesforCombined, plasmon.cpp:100-127).  With
    individual = n > 0 the draw is repeated until exactly n scatterings fall inside the specimen
    (getDistancesForIndividual); raises ValueError when no attempt gives that."""
    for attempt in range(_DEPTH_ATTEMPTS if individual > 0 else 1):
        depths = []
        k = 0
        t = 0.0
        while True:
            u = _uniforms(seed, [[job, 0x9A + 256 * attempt, k, 4]])[0][0]
            t += -mean_free_path * np.log(u)
            k += 1
            if t > thickness:
                break
            depths.append(t)
            if individual > 0 and len(depths) > individual:
                break
        if individual <= 0 or len(depths) == individual:
            return depths
    raise ValueError("could not generate a plasmon configuration with the requested scattering count")


def plasmon_angles(seed: int, job: int, n: int, characteristic_mrad: float, critical_mrad: float):
    """Polar angle (mrad) and azimuth (rad) of scatterings 0..n-1 of a job: getScatteringPolar
    theta^2 = theta_E^2 (((theta_c^2 + theta_E^2) / theta_E^2)^u - 1) and getScatteringAzimuth
    2 pi u (plasmon.cpp:14-29)."""
    if n <= 0:
        return np.zeros(0), np.zeros(0)
    u = _uniforms(seed, np.stack([np.full(n, job), np.full(n, 0xA5), np.arange(n), np.full(n, 5)], axis=1))
    ce2 = characteristic_mrad ** 2
    cc2 = critical_mrad ** 2
    polar = np.sqrt(ce2 * (np.power((cc2 + ce2) / ce2, u[:, 0]) - 1.0))
    return polar, 2.0 * np.pi * u[:, 1]


def rotation_about(axis, angle):
    """generateRotationAroundVector (utilities/vectorutils.cpp:8-18)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return np.eye(3) + c * np.sin(angle) + c @ c * (1.0 - np.cos(angle))


def rotate_vector_spherical(z, y, theta, phi):
    """rotateVectorSpherical (utilities/vectorutils.cpp:21-31): returns the new (z, y)."""
    r_theta = rotation_about(y, theta)
    r_phi = rotation_about(z, phi)
    z = r_phi @ (r_theta @ np.asarray(z, dtype=np.float64))
    y = r_phi @ np.asarray(y, dtype=np.float64)
    return z, y


def initial_beam(wavenumber: float, tilt_mrad: float, azimuth: float):
    """(k_vec, y axis) at the top of the specimen (simulationcbed.cpp:152-162)."""
    return rotate_vector_spherical([0.0, 0.0, wavenumber], [0.0, 1.0, 0.0], tilt_mrad / 1000.0, azimuth)


def plasmon_events(depths, pre_slices: int, dz: float, n_slices: int):
    """Slice after which each scattering is applied: the first slice i whose end, (i + 1 -
    pre_slices) dz below the top, reaches the scattering depth, at most one per slice
    (simulationcbed.cpp:171-188).  Scatterings past the last slice are dropped."""
    out = []
    j = 0
    for i in range(n_slices):
        if j < len(depths) and (i + 1 - pre_slices) * dz >= depths[j]:
            out.append(i)
            j += 1
    return out


# ---- mixed static potentials (microscope/simulationgeneral.cpp:731-745) ---------------------

def _key(seed: int) -> np.ndarray:
    return np.array([seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF], dtype=np.uint64)


def mixed_potential_picks(seed: int, keys, n_slices: int, n_sets: int) -> np.ndarray:
    """Set drawn by each wave key at each slice, as the engine's row pass draws it:
    floor(n_sets r / 2^32), r the first word of Philox4x32-10 at (key, slice, 0x3A, 6)."""
    keys = np.asarray(keys, dtype=np.uint64)
    k, s = np.meshgrid(keys, np.arange(n_slices, dtype=np.uint64), indexing="ij")
    ctr = np.stack([k.ravel(), s.ravel(), np.full(k.size, 0x3A, np.uint64), np.full(k.size, 6, np.uint64)], axis=1)
    r = philox4x32(ctr, _key(seed))[:, 0].astype(np.uint64)
    return ((r * np.uint64(n_sets)) >> np.uint64(32)).astype(np.int64).reshape(len(keys), n_slices)


def mixed_seed(seed: int, job: int, batch: int = 0) -> int:
    """Seed of one job's set draws (the C++ host's SimulationManager::mixedSeed; `batch` for
    probes set up one batch at a time): words 0 and 1 of Philox4x32-10 keyed by the manager's
    seed at (job, 0x3B, batch, 7)."""
    r = philox4x32(np.array([[job, 0x3B, batch, 7]], dtype=np.uint64), _key(seed))[0]
    return int(r[0]) | (int(r[1]) << 32)
