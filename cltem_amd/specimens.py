"""Seeded synthetic specimens and frozen-phonon displacements.

clTEM reads atoms from .xyz/.cif files and displaces them with a clock-seeded
mt19937 (incoherence/inelastic/phonon.cpp:9,98-129), so its runs cannot be
repeated.  Here specimens are generated in memory and displaced by a
counter-based Philox4x32-10 generator keyed by (seed, configuration, atom), so
every configuration is reproducible and identical on every device.  All
coordinates are rounded to float32-representable doubles (what the .xyz reader
would produce through std::stof, structure/crystalstructure.cpp:156-158).

Displacement semantics follow SimulationGeneral::sortAtoms
(microscope/simulationgeneral.cpp:188-205): each axis of an orthogonal cell gets
sqrt(<u^2>) * N(0, 1).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ELEMENTS = {"H": 1, "C": 6, "O": 8, "Si": 14, "Ti": 22, "Sr": 38, "Au": 79}

# Philox4x32 constants (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC11)
_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32(counter: np.ndarray, key: np.ndarray, rounds: int = 10) -> np.ndarray:
    """Vectorised Philox4x32-R.  counter: (n, 4) uint32, key: (n, 2) or (2,) uint32."""
    c = np.array(counter, dtype=np.uint64).reshape(-1, 4)
    k = np.broadcast_to(np.array(key, dtype=np.uint64), (c.shape[0], 2)).copy()
    c0, c1, c2, c3 = (c[:, i].copy() for i in range(4))
    k0, k1 = k[:, 0].copy(), k[:, 1].copy()
    for r in range(rounds):
        if r > 0:
            k0 = (k0 + _W0) & _MASK
            k1 = (k1 + _W1) & _MASK
        p0 = c0 * _M0
        p1 = c2 * _M1
        hi0, lo0 = p0 >> 32, p0 & _MASK
        hi1, lo1 = p1 >> 32, p1 & _MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return np.stack([c0, c1, c2, c3], axis=1).astype(np.uint32)


def gaussian_triples(seed: int, config: int, n: int) -> np.ndarray:
    """(n, 3) standard normal deviates for atoms 0..n-1 of one configuration.

    Counter = (atom, config, 0, 0), key = (seed low 32 bits, seed high 32 bits);
    Box-Muller on the four 32-bit outputs mapped to (0, 1) by (u + 0.5) / 2^32.
    """
    ctr = np.zeros((n, 4), dtype=np.uint64)
    ctr[:, 0] = np.arange(n, dtype=np.uint64)
    ctr[:, 1] = config
    key = np.array([seed & _MASK, (seed >> 32) & _MASK], dtype=np.uint64)
    r = philox4x32(ctr, key).astype(np.float64)
    u = (r + 0.5) / 4294967296.0
    rad0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    rad1 = np.sqrt(-2.0 * np.log(u[:, 2]))
    n0 = rad0 * np.cos(2.0 * np.pi * u[:, 1])
    n1 = rad0 * np.sin(2.0 * np.pi * u[:, 1])
    n2 = rad1 * np.cos(2.0 * np.pi * u[:, 3])
    return np.stack([n0, n1, n2], axis=1)


def f32_exact(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


@dataclass
class Specimen:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    Z: np.ndarray

    def __len__(self) -> int:
        return len(self.x)

    @property
    def limits(self):
        """CrystalStructure::limitsX/Y/Z: min/max of the atom coordinates."""
        return (np.array([self.x.min(), self.y.min(), self.z.min()]),
                np.array([self.x.max(), self.y.max(), self.z.max()]))

    def displaced(self, u_rms, seed: int, config: int) -> "Specimen":
        """One frozen-phonon configuration; u_rms is a scalar or per-atom array (Angstrom)."""
        g = gaussian_triples(seed, config, len(self))
        u = np.broadcast_to(np.asarray(u_rms, dtype=np.float64), (len(self),))
        return Specimen(f32_exact(self.x + u * g[:, 0]), f32_exact(self.y + u * g[:, 1]),
                        f32_exact(self.z + u * g[:, 2]), self.Z.copy())


def srtio3(nx: int, ny: int, nz: int, a: float = 3.905) -> Specimen:
    """Cubic perovskite SrTiO3 tiled nx*ny*nz along [001]: Sr (0,0,0), Ti (1/2,1/2,1/2),
    O (1/2,1/2,0), (1/2,0,1/2), (0,1/2,1/2)."""
    basis = [(38, 0.0, 0.0, 0.0), (22, 0.5, 0.5, 0.5), (8, 0.5, 0.5, 0.0), (8, 0.5, 0.0, 0.5),
             (8, 0.0, 0.5, 0.5)]
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    xs, ys, zs, Zs = [], [], [], []
    for Zb, bx, by, bz in basis:
        xs.append((cells[:, 0] + bx) * a)
        ys.append((cells[:, 1] + by) * a)
        zs.append((cells[:, 2] + bz) * a)
        Zs.append(np.full(len(cells), Zb, dtype=np.int32))
    # order: cell-major, basis-minor
    x = np.stack(xs, axis=1).ravel()
    y = np.stack(ys, axis=1).ravel()
    z = np.stack(zs, axis=1).ravel()
    Z = np.stack(Zs, axis=1).ravel()
    return Specimen(f32_exact(x), f32_exact(y), f32_exact(z), Z.astype(np.int32))


def random_box(counts: dict, box, min_sep: float, seed: int, max_tries: int = 200) -> Specimen:
    """Random sequential addition in an orthogonal box with a minimum separation.

    counts maps element symbol -> number of atoms; box = (Lx, Ly, Lz).  Atoms are
    placed in a seeded order, each trial position drawn from Philox keyed by
    (seed, trial, atom); a cell list keeps the rejection test local.
    """
    symbols = []
    for sym, n in counts.items():
        symbols += [ELEMENTS[sym]] * int(n)
    n = len(symbols)
    order = np.argsort(philox4x32(np.stack([np.arange(n), np.full(n, 7), np.zeros(n),
                                            np.zeros(n)], axis=1), [seed & _MASK, 0])[:, 0],
                       kind="stable")
    Zs = np.array(symbols, dtype=np.int32)[order]
    L = np.asarray(box, dtype=np.float64)
    cell = max(min_sep, 1e-6)
    ncell = np.maximum((L // cell).astype(int), 1)
    grid: dict = {}
    pos = np.zeros((n, 3))
    sep2 = min_sep * min_sep
    for i in range(n):
        for trial in range(max_tries):
            ctr = np.array([[i, trial, 1, 0]], dtype=np.uint64)
            u = (philox4x32(ctr, [seed & _MASK, (seed >> 32) & _MASK])[0, :3].astype(np.float64)
                 + 0.5) / 4294967296.0
            p = u * L
            ci = np.minimum((p // cell).astype(int), ncell - 1)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dzc in (-1, 0, 1):
                        for j in grid.get((ci[0] + dx, ci[1] + dy, ci[2] + dzc), ()):
                            d = pos[j] - p
                            if d @ d < sep2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[i] = p
                grid.setdefault(tuple(ci), []).append(i)
                break
        else:
            raise RuntimeError("random_box: could not place atom %d" % i)
    return Specimen(f32_exact(pos[:, 0]), f32_exact(pos[:, 1]), f32_exact(pos[:, 2]), Zs)
