"""Named workloads: specimen + scenario fields for the engine (and, identically, the oracle).

Each function returns a dict with
  scenario: keyword fields of mse_scenario (see cltem_amd.make_scenario)
  specimen: cltem_amd.specimens.Specimen (undisplaced)
  params:   parameterisation name
plus mode-specific entries (probe positions, detectors, phonon settings).
"""
from __future__ import annotations

import numpy as np

from . import specimens as sp

SRTIO3_A = 3.905
PHONON_SEED = 20261202


def _limits(spec):
    lo, hi = spec.limits
    return dict(struct_min=tuple(lo), struct_max=tuple(hi))


def ctem_srtio3(nx=8, ny=8, nz=20, resolution=256, slice_thickness=1.9525, voltage_kv=200.0,
                params="kirkland"):
    """Plain CTEM exit wave: SrTiO3 [001] tiled nx*ny*nz, plane wave, no phonons
    (8x8x20, 256^2, 1.9525 A slices, 200 kV is the CPU-runnable reference case)."""
    spec = sp.srtio3(nx, ny, nz, SRTIO3_A)
    lim = _limits(spec)
    scen = dict(mode=1, resolution=resolution, voltage_kv=voltage_kv,
                slice_thickness=slice_thickness, slice_offset=0.0,
                area=(lim["struct_min"][0], lim["struct_max"][0], lim["struct_min"][1],
                      lim["struct_max"][1]), **lim)
    return dict(name=f"ctem_srtio3_{nx}x{ny}x{nz}_{resolution}", scenario=scen, specimen=spec,
                params=params)


def stem_srtio3(nx=16, ny=16, nz=40, resolution=512, pixels=64, window_cells=(8, 9),
                slice_thickness=SRTIO3_A / 2, voltage_kv=200.0, cond_ap_mrad=20.0,
                detectors=((60.0, 200.0, 0.0, 0.0),), u_rms=0.08, params="kirkland"):
    """Batched HAADF STEM on SrTiO3 [001] tiled nx*ny*nz.

    The scan window is one unit cell (cells window_cells[0]..[1] in x and y).  With
    clTEM's xy padding of 8 A each side the field is 3.905 + 16 = 19.905 A, so at 512^2
    the 2/3 band limit is 215 mrad and the 60-200 mrad annulus lies inside it.  (A
    scan over the whole 16x16-cell specimen would give a 76.5 A field whose band limit,
    55.9 mrad, is below the detector's inner angle.)
    """
    spec = sp.srtio3(nx, ny, nz, SRTIO3_A)
    lim = _limits(spec)
    x0, x1 = window_cells[0] * SRTIO3_A, window_cells[1] * SRTIO3_A
    scen = dict(mode=2, resolution=resolution, voltage_kv=voltage_kv,
                slice_thickness=slice_thickness, slice_offset=0.0, area=(x0, x1, x0, x1),
                stem_pixels=(pixels, pixels), **lim)
    return dict(name=f"stem_srtio3_{nx}x{ny}x{nz}_{resolution}_{pixels}px", scenario=scen,
                specimen=spec, params=params, cond_ap_mrad=cond_ap_mrad,
                detectors=[tuple(d) for d in detectors], u_rms=u_rms, seed=PHONON_SEED)


def stem_positions(geom, pixels=None):
    """Probe positions of SimulationStem::simulate (microscope/simulationstem.cpp:156-187):
    pixel p -> (start_x + (p % nx) * step_x, start_y + (p // nx) * step_y)."""
    nx, ny = (geom.stem_pixels[0], geom.stem_pixels[1]) if pixels is None else pixels
    p = np.arange(nx * ny)
    xi = p % nx
    yi = p // nx
    x = geom.stem_start[0] + xi * geom.stem_scale[0]
    y = geom.stem_start[1] + yi * geom.stem_scale[1]
    return np.stack([x, y], axis=1)


def adjoint_auc(n_atoms=2000, box=(40.0, 40.0, 40.0), resolution=256, slice_thickness=1.0,
                voltage_kv=200.0, min_sep=1.0, seed=PHONON_SEED, noise=0.05, params="kirkland"):
    """Adjoint (position-gradient) run: random Au/C (1:3) in a box, projected potential,
    plane-wave CTEM in 1 A slices (2000 atoms, 40^3 A, 256^2, 200 kV: 40 slices of
    specimen plus clTEM's default +-3 A z padding, 46 slices).  `target` holds positions
    displaced by Gaussian noise (sigma = noise A) from which the target intensity is
    simulated."""
    n_au = n_atoms // 4
    spec = sp.random_box({"Au": n_au, "C": n_atoms - n_au}, box, min_sep, seed)
    lim = _limits(spec)
    scen = dict(mode=1, resolution=resolution, voltage_kv=voltage_kv,
                slice_thickness=slice_thickness, slice_offset=0.0,
                area=(lim["struct_min"][0], lim["struct_max"][0], lim["struct_min"][1],
                      lim["struct_max"][1]), **lim)
    tgt = spec.displaced(noise, seed ^ 0x5A5A, 0)
    return dict(name=f"adjoint_auc_{n_atoms}_{resolution}", scenario=scen, specimen=spec,
                params=params, target=tgt)


def ctem_amorphous_sio2(n_atoms=20000, box=(60.0, 60.0, 100.0), resolution=1024, slice_thickness=1.0,
                        voltage_kv=300.0, min_sep=1.5, n_configs=30, u_rms=0.1, integrals=20,
                        seed=PHONON_SEED, params="lobato"):
    """Full-3D CTEM on an amorphous SiO2-like cell: Si:O = 1:2 by random sequential addition
    in a box with a minimum separation, plane wave, frozen-phonon configurations batched
    (20000 atoms in 60x60x100 A, 1024^2, 1 A slices (106 with clTEM's z padding), 300 kV,
    Lobato with 20 integrals, 30 configurations of u_rms 0.1 A).  The output is the
    phonon-averaged exit-wave intensity."""
    n_si = n_atoms // 3
    spec = sp.random_box({"Si": n_si, "O": n_atoms - n_si}, box, min_sep, seed)
    lim = _limits(spec)
    scen = dict(mode=1, resolution=resolution, voltage_kv=voltage_kv,
                slice_thickness=slice_thickness, slice_offset=0.0, full3d=1, full3d_integrals=integrals,
                area=(lim["struct_min"][0], lim["struct_max"][0], lim["struct_min"][1],
                      lim["struct_max"][1]), **lim)
    return dict(name=f"ctem_sio2_{n_atoms}_{resolution}_{n_configs}cfg", scenario=scen, specimen=spec,
                params=params, n_configs=n_configs, u_rms=u_rms, seed=seed)


def mixed_cbed(n_instances=48, cells_xy=5, thickness_cells=(10, 80), sizes=(256, 384, 512, 768),
               alpha_mrad=(5.0, 25.0), n_configs=5, u_rms=0.08, slice_thickness=SRTIO3_A / 2,
               voltage_kv=200.0, seed=PHONON_SEED):
    """Mixed-size batched CBED: independent SrTiO3 [001] instances (cells_xy^2 cells wide)
    with seeded thickness U{10..80} cells, grid size from `sizes`, convergence U(5, 25)
    mrad and n_configs frozen-phonon configurations each; the probe sits at the centre of
    the specimen (48 instances, 200 kV, 5 configurations of u_rms 0.08 A)."""
    from .mixed import CbedInstance
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_instances):
        nz = int(rng.integers(thickness_cells[0], thickness_cells[1] + 1))
        res = int(sizes[int(rng.integers(0, len(sizes)))])
        alpha = float(rng.uniform(*alpha_mrad))
        spec = sp.srtio3(cells_xy, cells_xy, nz, SRTIO3_A)
        lo, hi = spec.limits
        cx, cy = 0.5 * cells_xy * SRTIO3_A, 0.5 * cells_xy * SRTIO3_A
        scen = dict(mode=3, resolution=res, voltage_kv=voltage_kv, slice_thickness=slice_thickness,
                    slice_offset=0.0, struct_min=tuple(lo), struct_max=tuple(hi), area=(cx, cx, cy, cy),
                    area_padding=0.5 * cells_xy * SRTIO3_A)
        out.append(CbedInstance(specimen=spec, scenario=scen, alpha_mrad=alpha, position=(cx, cy),
                                n_configs=n_configs, u_rms=u_rms, seed=seed + i))
    return out
