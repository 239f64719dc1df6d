"""Pins of the oracle's finite-difference loss gradient (the checker of the adjoint).

orc_ctem_loss_fd rebuilds only the slices a perturbed atom touches; here it is checked
against brute force (the whole forward rerun from numpy-perturbed coordinates through
orc_ctem_intensity) and for the O(h^2) convergence of a central difference."""
import numpy as np
import pytest

from cltem_amd import params as prm
from cltem_amd import specimens as sp


def tiny_case(oracle, full3d=False, tilt=0.0):
    spec = sp.random_box({"Au": 4, "C": 12}, (9.0, 9.0, 5.0), 1.2, seed=5)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=64, voltage_kv=200.0, slice_thickness=1.0, pad_z=(0.0, 0.0),
                struct_min=tuple(lo), struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]),
                full3d=int(full3d), full3d_integrals=4, beam_tilt_mrad=tilt, beam_azimuth=0.3)
    g = oracle.geometry(oracle.scenario(**scen))
    return spec, g


def loss(oracle, g, p, x, y, z, Z, target):
    I = oracle.ctem_intensity(g, x, y, z, Z, p.values, p.form, p.i_per_atom, nthreads=1)
    return float(np.sum((I - target) ** 2))


@pytest.mark.parametrize("full3d,tilt", [(False, 0.0), (False, 5.0), (True, 0.0)])
def test_fd_matches_brute_force(oracle, full3d, tilt):
    spec, g = tiny_case(oracle, full3d, tilt)
    p = prm.load("kirkland")
    tgt = spec.displaced(0.05, 9, 0)
    target = oracle.ctem_intensity(g, tgt.x, tgt.y, tgt.z, tgt.Z, p.values, p.form, p.i_per_atom)
    atoms = np.array([0, 3, 7, 11, 2])
    axes = np.array([0, 1, 0, 1, 2])
    h = 1e-3
    fd, L0 = oracle.ctem_loss_fd(g, spec.x, spec.y, spec.z, spec.Z, p.values, p.form, p.i_per_atom, target,
                                 atoms, axes, h=h)
    assert L0 == pytest.approx(loss(oracle, g, p, spec.x, spec.y, spec.z, spec.Z, target), rel=1e-13)
    for k, (a, ax) in enumerate(zip(atoms, axes)):
        Ls = []
        for sgn in (1.0, -1.0):
            c = [spec.x.copy(), spec.y.copy(), spec.z.copy()]
            c[ax][a] += sgn * h
            Ls.append(loss(oracle, g, p, c[0], c[1], c[2], spec.Z, target))
        brute = (Ls[0] - Ls[1]) / (2 * h)
        assert fd[k] == pytest.approx(brute, rel=1e-9, abs=1e-12), (a, ax)
    # projected potentials do not depend on z inside a slice
    if not full3d:
        assert fd[4] == 0.0


def test_fd_converges_as_h_squared(oracle):
    spec, g = tiny_case(oracle)
    p = prm.load("lobato")
    tgt = spec.displaced(0.05, 9, 0)
    target = oracle.ctem_intensity(g, tgt.x, tgt.y, tgt.z, tgt.Z, p.values, p.form, p.i_per_atom)
    atoms, axes = np.array([1, 1, 6]), np.array([0, 1, 0])
    d = [oracle.ctem_loss_fd(g, spec.x, spec.y, spec.z, spec.Z, p.values, p.form, p.i_per_atom, target,
                             atoms, axes, h=h)[0] for h in (4e-3, 2e-3, 1e-3)]
    e1 = np.abs(d[0] - d[2])
    e2 = np.abs(d[1] - d[2])
    # error ~ c h^2: halving h from 4e-3 to 2e-3 cuts the distance to the 1e-3 value ~4x (15/3 = 5)
    assert np.all(e2 <= 0.35 * e1 + 1e-9 * np.abs(d[2]))
    assert np.all(np.abs(d[2]) > 0)
