"""Host-side inputs: Philox displacements, specimens, parameter tables (CPU only)."""
import numpy as np
import pytest

from cltem_amd import params as P
from cltem_amd import specimens as sp
from cltem_amd import workloads as wl


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32-10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        got = sp.philox4x32(np.array([ctr], dtype=np.uint64), np.array(key, dtype=np.uint64))[0]
        assert tuple(int(v) for v in got) == out


def test_gaussian_displacements_are_seeded_normals():
    g = sp.gaussian_triples(20261202, 0, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    assert np.array_equal(g[:1000], sp.gaussian_triples(20261202, 0, 1000))
    assert not np.allclose(g[:1000], sp.gaussian_triples(20261202, 1, 1000))


def test_srtio3_specimen():
    s = sp.srtio3(8, 8, 20)
    assert len(s) == 6400
    assert np.bincount(s.Z)[[8, 22, 38]].tolist() == [3840, 1280, 1280]
    lo, hi = s.limits
    np.testing.assert_allclose(hi, [29.2875, 29.2875, 76.1475], atol=1e-5)
    assert np.all(s.x.astype(np.float32).astype(np.float64) == s.x)  # float32-exact
    d = s.displaced(0.08, 20261202, 3)
    dz = d.z - s.z
    assert 0.07 < dz.std() < 0.09 and np.all(d.Z == s.Z)


def test_random_box_min_separation():
    s = sp.random_box({"Si": 40, "O": 80}, (12.0, 12.0, 12.0), 1.5, seed=5)
    pts = np.stack([s.x, s.y, s.z], 1)
    d = np.linalg.norm(pts[:, None] - pts[None], axis=-1) + np.eye(len(pts)) * 99
    assert d.min() >= 1.5 - 1e-6
    assert np.bincount(s.Z)[[8, 14]].tolist() == [80, 40]


def test_parameter_tables():
    k = P.load("kirkland")
    assert (k.form, k.i_per_atom, k.max_atomic_number) == (0, 3, 103)
    assert (P.load("lobato").max_atomic_number, P.load("peng").max_atomic_number) == (103, 98)
    assert P.load("peng").form == 1 and P.load("lobato").form == 2


def test_stem_positions_follow_the_raster():
    w = wl.stem_srtio3(pixels=4)
    import cltem_amd
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**w["scenario"]))
    pos = wl.stem_positions(g)
    assert pos.shape == (16, 2)
    step = 3.905 / 4
    np.testing.assert_allclose(pos[5], [8 * 3.905 + step, 8 * 3.905 + step])
