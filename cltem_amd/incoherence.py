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
