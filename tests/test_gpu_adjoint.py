"""Adjoint slice loop: position gradients through the torch.autograd binding, checked
against central finite differences (h = 1e-3 A) of the FP64 oracle's loss."""
import numpy as np
import pytest
import torch

import cltem_amd
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl
from cltem_amd.autograd import exit_intensity

pytestmark = pytest.mark.gpu


def setup(engine, scen, params):
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**scen))
    p = cltem_amd.load_params(params)
    engine.set_geometry(g)
    engine.set_parameterisation(p)
    return g, p


def gpu_loss_and_grad(engine, spec, target_I):
    pos = torch.tensor(np.stack([spec.x, spec.y, spec.z], axis=1), dtype=torch.float64, requires_grad=True)
    I = exit_intensity(pos, engine, spec.Z)
    L = ((I - torch.from_numpy(target_I)) ** 2).sum()
    L.backward()
    return float(L), pos.grad.numpy(), I.detach().numpy()


def away_from_slice_edges(g, z, margin):
    d = (g.z_lims[1] - z) / g.slice_thickness
    frac = d - np.floor(d)
    return (frac > margin / g.slice_thickness) & (frac < 1 - margin / g.slice_thickness)


def compare(grad_sel, fd, label):
    err = np.abs(grad_sel - fd)
    rel_norm = float(np.linalg.norm(grad_sel - fd) / np.linalg.norm(fd))
    scale = np.max(np.abs(fd))
    for k in np.argsort(-np.abs(fd))[:8]:
        print(f"  {label} fd {fd[k]: .6e} adjoint {grad_sel[k]: .6e} rel {err[k] / max(abs(fd[k]), 1e-300):.2e}")
    print(f"{label}: {len(fd)} coordinates, rel-L2 {rel_norm:.3e}, max |err| / max |fd| {err.max() / scale:.3e}")
    return rel_norm, err, scale


def test_adjoint_auc_matches_fd(engine, oracle):
    """2000 Au/C atoms, 256^2, 46 slices of 1 A, loss sum (I - I_target)^2, sqrt(46)
    checkpointing; 64 sampled coordinates against FP64 central differences."""
    w = wl.adjoint_auc()
    spec, tgt = w["specimen"], w["target"]
    g, p = setup(engine, w["scenario"], w["params"])
    og = oracle.geometry(oracle.scenario(**w["scenario"]))
    target_I = oracle.ctem_intensity(og, tgt.x, tgt.y, tgt.z, tgt.Z, p.values, p.form, p.i_per_atom)
    L, grad, I = gpu_loss_and_grad(engine, spec, target_I)
    assert grad.shape == (len(spec.Z), 3)
    assert np.all(np.isfinite(grad))
    assert np.all(grad[:, 2] == 0.0)  # projected potential: no depth dependence within a slice

    rng = np.random.default_rng(4)
    au = np.flatnonzero(spec.Z == 79)
    c = np.flatnonzero(spec.Z == 6)
    atoms = np.concatenate([rng.choice(au, 16, replace=False), rng.choice(c, 16, replace=False)])
    sel_atoms = np.repeat(atoms, 2)
    sel_axes = np.tile([0, 1], len(atoms))
    zc = [a for a in rng.permutation(len(spec.Z)) if away_from_slice_edges(g, spec.z[a], 0.01)][:8]
    sel_atoms = np.concatenate([sel_atoms, zc])
    sel_axes = np.concatenate([sel_axes, np.full(len(zc), 2)])
    fd, L0 = oracle.ctem_loss_fd(og, spec.x, spec.y, spec.z, spec.Z, p.values, p.form, p.i_per_atom, target_I,
                                 sel_atoms, sel_axes, h=1e-3)
    print(f"loss: GPU {L:.8e}, FP64 oracle {L0:.8e}, rel {abs(L - L0) / L0:.2e}")
    assert abs(L - L0) / L0 <= 1e-4
    got = grad[sel_atoms, sel_axes]
    xy = sel_axes < 2
    rel_norm, err, scale = compare(got[xy], fd[xy], "x/y")
    assert rel_norm <= 1e-3
    assert np.all(err <= 1e-3 * np.abs(fd[xy]) + 1e-3 * scale)
    assert np.all(fd[~xy] == 0.0) and np.all(got[~xy] == 0.0)
    # most of the remaining difference is the O(h^2) error of the h = 1e-3 difference itself:
    # against the Richardson combination of h and h/2 the adjoint agrees far more closely
    fd_half, _ = oracle.ctem_loss_fd(og, spec.x, spec.y, spec.z, spec.Z, p.values, p.form, p.i_per_atom,
                                     target_I, sel_atoms[xy], sel_axes[xy], h=5e-4)
    fd_r = (4.0 * fd_half - fd[xy]) / 3.0
    rel_r, _, _ = compare(got[xy], fd_r, "x/y vs Richardson")
    assert rel_r <= 2e-4


@pytest.mark.parametrize("params,tilt", [("lobato", 0.0), ("peng", 6.0), ("kirkland", 6.0)])
def test_adjoint_forms_and_tilt_match_fd(engine, oracle, params, tilt):
    """Every parameterisation, with and without beam tilt, on a 0.125 A pixel grid.  Lobato's
    gold core (b5 = 3.8e-4 A^2) varies on a 0.02 A scale, where a plain h = 1e-3 A central
    difference is still 5e-3 from its limit; the reference here is the Richardson
    combination (4 D(h/2) - D(h)) / 3 of the FP64 central differences at h = 1e-3 and
    5e-4 A, whose error is O(h^4)."""
    spec = sp.random_box({"Au": 20, "C": 60}, (16.0, 16.0, 8.0), 1.2, seed=8)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.0, pad_z=(0.0, 0.0),
                struct_min=tuple(lo), struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]),
                beam_tilt_mrad=tilt, beam_azimuth=0.7)
    g, p = setup(engine, scen, params)
    og = oracle.geometry(oracle.scenario(**scen))
    tgt = spec.displaced(0.05, 21, 0)
    target_I = oracle.ctem_intensity(og, tgt.x, tgt.y, tgt.z, tgt.Z, p.values, p.form, p.i_per_atom)
    L, grad, _ = gpu_loss_and_grad(engine, spec, target_I)
    atoms = np.arange(0, len(spec.Z), 5)
    sel_atoms = np.repeat(atoms, 2)
    sel_axes = np.tile([0, 1], len(atoms))
    fd = {}
    for h in (1e-3, 5e-4):
        fd[h], L0 = oracle.ctem_loss_fd(og, spec.x, spec.y, spec.z, spec.Z, p.values, p.form, p.i_per_atom,
                                        target_I, sel_atoms, sel_axes, h=h)
    assert abs(L - L0) / L0 <= 1e-4
    fd_r = (4.0 * fd[5e-4] - fd[1e-3]) / 3.0
    got = grad[sel_atoms, sel_axes]
    print(f"plain h=1e-3 central difference vs Richardson: rel-L2 "
          f"{np.linalg.norm(fd[1e-3] - fd_r) / np.linalg.norm(fd_r):.3e}")
    rel_norm, err, scale = compare(got, fd_r, f"{params} tilt {tilt}")
    assert rel_norm <= 1e-3
    assert np.all(err <= 1e-3 * np.abs(fd_r) + 1e-3 * scale)


def test_checkpointed_forward_equals_multislice_and_waves_add(engine, oracle):
    spec = sp.random_box({"Au": 10, "C": 30}, (12.0, 12.0, 9.0), 1.2, seed=2)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.0, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]))
    g, p = setup(engine, scen, "kirkland")
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    engine.multislice()
    ref = np.abs(engine.exit_wave(0)) ** 2
    engine.init_plane_wave(1.0)
    I1 = engine.forward_checkpointed(0)
    assert np.abs(I1[0] - ref).max() <= 1e-6 * ref.max()
    w = np.random.default_rng(1).standard_normal(I1.shape)
    g1 = engine.adjoint_positions(w, len(spec.Z))
    # two identical waves with the same weights: twice the gradient (per-wave g_t summed)
    engine.set_wave_count(2)
    engine.set_wave_configs([0, 0])
    engine.init_plane_wave(1.0)
    I2 = engine.forward_checkpointed(0)
    assert np.array_equal(I2[0], I1[0]) and np.array_equal(I2[1], I1[0])
    g2 = engine.adjoint_positions(np.concatenate([w, w]), len(spec.Z))
    assert np.allclose(g2, 2 * g1, rtol=1e-5, atol=1e-9 * np.abs(g1).max())
    # a weight on one wave only gives that wave's gradient
    g3 = engine.adjoint_positions(np.concatenate([w, np.zeros_like(w)]), len(spec.Z))
    assert np.allclose(g3, g1, rtol=1e-5, atol=1e-9 * np.abs(g1).max())


def test_all_states_and_sqrt_checkpoints_agree(engine):
    """forward_checkpointed keeps every forward state when they fit the checkpoint budget,
    else sqrt(n) checkpoints whose segments the adjoint recomputes.  Both routes give the
    intensities of multislice and the same gradients, to float rounding (a slice can run out of
    place on one route and in place on the other)."""
    spec = sp.random_box({"Au": 10, "C": 30}, (12.0, 12.0, 9.0), 1.2, seed=2)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.0, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]))
    g, p = setup(engine, scen, "kirkland")
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    engine.set_wave_count(2)
    engine.set_wave_configs([0, 0])
    engine.init_plane_wave(1.0)
    engine.multislice()
    ref = np.abs(engine.exit_wave(0)) ** 2
    w = np.random.default_rng(5).standard_normal((2, 256, 256))
    out = {}
    try:
        for budget in (1 << 30, 0):
            engine.set_checkpoint_budget(budget)
            engine.init_plane_wave(1.0)
            I = engine.forward_checkpointed(0)
            out[budget] = (I, engine.adjoint_positions(w, len(spec.Z)))
    finally:
        engine.set_checkpoint_budget(1 << 30)
    print(f"{g.n_slices} slices")
    for budget, (I, _) in out.items():
        for wv in (0, 1):
            assert np.abs(I[wv] - ref).max() <= 1e-6 * ref.max(), (budget, wv)
    g_all, g_ck = out[1 << 30][1], out[0][1]
    rel = np.linalg.norm(g_all - g_ck) / np.linalg.norm(g_ck)
    print(f"every state kept vs sqrt(n) checkpoints: gradient rel-L2 {rel:.2e}")
    assert np.abs(g_ck).max() > 0 and rel <= 1e-5


def test_adjoint_with_wave_beam_tables(engine):
    """The backward column pass is the forward propagator pass on conjugated px, py tables, or,
    where waves carry beam tables of their own, a pass that conjugates each wave's factors.  A
    configuration beam equal to the geometry's own vector takes the second route with the same
    propagator, so intensities and gradients agree with the first."""
    spec = sp.random_box({"Au": 10, "C": 30}, (12.0, 12.0, 9.0), 1.2, seed=2)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.0, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]))
    g, p = setup(engine, scen, "kirkland")
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    w = np.random.default_rng(3).standard_normal((1, 256, 256))
    grads, ints = [], []
    for beam in (None, list(g.k_vec)):
        engine.set_config_beam(0, beam)
        engine.build_transmission(0)
        engine.init_plane_wave(1.0)
        ints.append(engine.forward_checkpointed(0))
        grads.append(engine.adjoint_positions(w, len(spec.Z)))
    engine.set_config_beam(0, None)
    I_err = np.abs(ints[1] - ints[0]).max() / ints[0].max()
    g_err = np.linalg.norm(grads[1] - grads[0]) / np.linalg.norm(grads[0])
    print(f"own beam as a configuration beam: intensity max err {I_err:.2e}, gradient rel-L2 {g_err:.2e}")
    assert np.abs(grads[0]).max() > 0
    assert I_err <= 1e-5 and g_err <= 1e-4


def test_adjoint_errors(engine):
    spec = sp.random_box({"C": 20}, (10.0, 10.0, 6.0), 1.2, seed=4)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.0, struct_min=tuple(lo),
                struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]))
    g, p = setup(engine, scen, "kirkland")
    engine.sort_atoms(0, spec.x, spec.y, spec.z, spec.Z)
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    I = engine.forward_checkpointed(0)
    engine.build_transmission(0)  # invalidates the checkpoints
    with pytest.raises(cltem_amd.MseError) as ex:
        engine.adjoint_positions(np.ones_like(I), len(spec.Z))
    assert ex.value.code == -2


def away_from_depth_cutoffs(g, z, margin):
    """The full-3D sum takes integration point h of slice s when its depth minus the atom's
    is at most 3 A; those depths lie on z_min + m dz / integrals, so the potential jumps as an
    atom crosses z = z_min + m dz / integrals - 3.  A central difference must not straddle one."""
    step = g.slice_thickness / g.full3d_integrals
    u = (z - g.z_lims[0] + 3.0) / step
    frac = u - np.floor(u)
    return (frac > margin / step) & (frac < 1 - margin / step)


@pytest.mark.parametrize("params,tilt", [("kirkland", 0.0), ("lobato", 6.0), ("peng", 0.0)])
def test_adjoint_full3d_matches_fd(engine, oracle, params, tilt):
    """Full-3D potentials (8 integration points per 1 A slice, +-3 slices of neighbours):
    x, y and z gradients against the Richardson combination of FP64 central differences at
    h = 1e-3 and 5e-4 A.  Tilt shifts every integration point sideways (the shifted path of
    the sum).  Atoms are kept 3e-3 A away from slice edges and from the 3 A depth cutoffs,
    where the potential is discontinuous."""
    spec = sp.random_box({"Au": 8, "C": 24}, (12.0, 12.0, 5.0), 1.2, seed=11)
    lo, hi = spec.limits
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.0, pad_z=(0.0, 0.0),
                struct_min=tuple(lo), struct_max=tuple(hi), area=(lo[0], hi[0], lo[1], hi[1]),
                beam_tilt_mrad=tilt, beam_azimuth=0.7, full3d=1, full3d_integrals=8)
    g, p = setup(engine, scen, params)
    og = oracle.geometry(oracle.scenario(**scen))
    tgt = spec.displaced(0.05, 21, 0)
    target_I = oracle.ctem_intensity(og, tgt.x, tgt.y, tgt.z, tgt.Z, p.values, p.form, p.i_per_atom)
    L, grad, _ = gpu_loss_and_grad(engine, spec, target_I)
    assert np.all(np.isfinite(grad))
    ok = away_from_slice_edges(g, spec.z, 3e-3) & away_from_depth_cutoffs(g, spec.z, 3e-3)
    atoms = np.flatnonzero(ok)[:10]
    assert len(atoms) >= 6
    sel_atoms = np.repeat(atoms, 3)
    sel_axes = np.tile([0, 1, 2], len(atoms))
    fd = {}
    for h in (1e-3, 5e-4):
        fd[h], L0 = oracle.ctem_loss_fd(og, spec.x, spec.y, spec.z, spec.Z, p.values, p.form, p.i_per_atom,
                                        target_I, sel_atoms, sel_axes, h=h)
    print(f"loss: GPU {L:.8e}, FP64 oracle {L0:.8e}, rel {abs(L - L0) / L0:.2e}")
    assert abs(L - L0) / L0 <= 1e-4
    fd_r = (4.0 * fd[5e-4] - fd[1e-3]) / 3.0
    got = grad[sel_atoms, sel_axes]
    z = sel_axes == 2
    assert np.abs(fd_r[z]).max() > 0.05 * np.abs(fd_r).max()  # depth matters in full 3D
    rel_norm, err, scale = compare(got, fd_r, f"full 3D {params} tilt {tilt}")
    assert rel_norm <= 1e-3
    assert np.all(err <= 1e-3 * np.abs(fd_r) + 1e-3 * scale)


@pytest.mark.parametrize("full3d", [False, True])
def test_adjoint_without_atoms_in_range(engine, full3d):
    """Atoms outside the padded limits are filtered out of every slice: the exit intensity is
    that of vacuum and every gradient is exactly zero."""
    scen = dict(mode=1, resolution=256, voltage_kv=200.0, slice_thickness=1.0, struct_min=(0.0, 0.0, 0.0),
                struct_max=(10.0, 10.0, 6.0), area=(0.0, 10.0, 0.0, 10.0), full3d=int(full3d),
                full3d_integrals=4)
    g, p = setup(engine, scen, "kirkland")
    x = np.array([500.0, -400.0, 5.0])
    y = np.array([5.0, 5.0, 700.0])
    z = np.array([3.0, 2.0, 1.0])
    Z = np.array([79, 6, 14], dtype=np.int32)
    r = engine.sort_atoms(0, x, y, z, Z)
    assert r["n_in_range"] == 0
    engine.build_transmission(0)
    engine.set_wave_count(1)
    engine.set_wave_configs([0])
    engine.init_plane_wave(1.0)
    I = engine.forward_checkpointed(0)
    np.testing.assert_allclose(I[0], 1.0, atol=1e-5)
    grad = engine.adjoint_positions(np.ones_like(I), len(Z))
    assert grad.shape == (3, 3) and np.all(grad == 0.0)
