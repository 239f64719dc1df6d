"""Moving STEM frame (clTEM with parallel_stem off): each pixel runs in a field of its own
(SimulationManager::currentAreaBase, simulationmanager.cpp:485-501; StemArea::getPixelSimArea,
utilities/commonstructs.h:379-396).  CPU only: the geometry of the product and of the oracle
against hand-derived numbers and each other, and the oracle's per-pixel potentials against a
translation identity."""
import numpy as np
import pytest

import cltem_amd
from cltem_amd import params as P
from cltem_amd import workloads as wl


def _scenario(area_padding=0.0, res=256):
    # a 4 A x 4 A scan area at (10..14, 20..24), 8 x 8 pixels, cell padding +-8 A
    return dict(mode=2, resolution=res, slice_thickness=2.0, struct_min=(8.0, 16.0, 0.0),
                struct_max=(12.0, 24.0, 2.0), area=(10.0, 14.0, 20.0, 24.0), area_padding=area_padding,
                stem_pixels=(8, 8), moving_stem_frame=1)


# Hand derivation, area padding 0: the full field is [2, 22] x [12, 32] (22 - 2 = 20 A), so
# frame origins snap to rs = 20/256 = 0.078125 A.  Pixel 43 = (3, 5): scan offsets 1.5 and
# 2.5 A -> floor(19.2) = 19 and floor(32) = 32 steps -> 1.484375 and 2.5 A -> positions
# (11.484375, 22.5); a zero-width area padded by 8 A starts at (3.484375, 14.5).  Pixel 0's
# field [2, 18] x [12, 28] is 16 A wide: real scale 16/256 = 0.0625 A.
# Area padding 2: the corrected scan area is [9, 15] (6 A), the full field 22 A, rs =
# 22/256 = 0.0859375; pixel 43: floor(17.45...) = 17 and floor(29.09...) = 29 steps ->
# (11.4609375, 22.4921875); the frame is 1 A wider each side: start (2.4609375, 13.4921875);
# pixel 0's field [1, 19]: real scale 18/256 = 0.0703125.
HAND = {
    0.0: dict(rs=0.078125, p43=(3.484375, 14.5), p0=(2.0, 18.0, 12.0, 28.0), real=0.0625),
    2.0: dict(rs=0.0859375, p43=(2.4609375, 13.4921875), p0=(1.0, 19.0, 11.0, 29.0), real=0.0703125),
}


@pytest.mark.parametrize("pad", [0.0, 2.0])
def test_moving_frame_geometry_hand_derived(oracle, pad):
    h = HAND[pad]
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**_scenario(pad)))
    assert g.moving_frame == 1
    assert g.frame_step == h["rs"]
    assert (g.sim_x[0], g.sim_x[1], g.sim_y[0], g.sim_y[1]) == h["p0"]
    assert g.real_scale == h["real"]
    assert tuple(cltem_amd.stem_pixel_frames(g, [43])[0]) == h["p43"]
    assert tuple(cltem_amd.stem_pixel_frames(g, [0])[0]) == (h["p0"][0], h["p0"][2])
    # the full field (binning, blocks) is that of parallel STEM
    gp = cltem_amd.plan_geometry(cltem_amd.make_scenario(**dict(_scenario(pad), moving_stem_frame=0)))
    assert list(g.full_x) == list(gp.full_x) and list(g.full_y) == list(gp.full_y)
    assert (g.blocks_x, g.load_x) == (gp.blocks_x, gp.load_x)
    # the oracle's own restatement
    s = oracle.scenario(**_scenario(pad))
    og = oracle.geometry_pixel(s, 43)
    assert (og.sim_x[0], og.sim_y[0]) == h["p43"]
    assert og.real_scale == h["real"]
    og0 = oracle.geometry(s)
    assert (og0.sim_x[0], og0.sim_x[1], og0.sim_y[0], og0.sim_y[1]) == h["p0"]


def test_moving_frame_product_matches_oracle_every_pixel(oracle):
    """A SrTiO3 scan window whose numbers are not dyadic: every pixel's frame and every
    geometry field agree exactly."""
    sc = dict(wl.stem_srtio3(resolution=512, pixels=16)["scenario"], moving_stem_frame=1, area_padding=1.3)
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**sc))
    s = oracle.scenario(**sc)
    og = oracle.geometry(s)
    for f, _ in og._fields_:
        va, vb = getattr(g, f), getattr(og, f)
        try:
            va, vb = list(va), list(vb)
        except TypeError:
            pass
        assert va == vb, f
    frames = cltem_amd.stem_pixel_frames(g, range(256))
    for p in range(256):
        gp = oracle.geometry_pixel(s, p)
        assert (gp.sim_x[0], gp.sim_y[0]) == tuple(frames[p]), p
    # frames are distinct along a row and snap to the full field's pixel size
    assert len(np.unique(frames[:16, 0])) == 16
    steps = (frames[:, 0] - frames[0, 0]) / g.frame_step
    assert np.allclose(steps, np.round(steps), atol=1e-6)


def test_parallel_stem_ignores_frames():
    sc = wl.stem_srtio3(resolution=256, pixels=8)["scenario"]
    g = cltem_amd.plan_geometry(cltem_amd.make_scenario(**sc))
    assert g.moving_frame == 0
    fr = cltem_amd.stem_pixel_frames(g, [0, 17, 63])
    assert np.all(fr == np.array([g.sim_x[0], g.sim_y[0]]))


def test_moving_frame_rejects_missing_pixels():
    sc = dict(_scenario(), stem_pixels=(0, 0))
    with pytest.raises(cltem_amd.MseError):
        cltem_amd.plan_geometry(cltem_amd.make_scenario(**sc))


def test_moving_frame_potential_is_translated(oracle):
    """Pixels (0, y) and (2, y) of the hand scenario have frames 12 snapping steps = 0.9375 A =
    15 pixels of the 0.0625 A field apart: the oracle's potential of the second frame is the
    first one's shifted by 15 columns, bit for bit (dyadic grid, same atoms in the same order)."""
    s = oracle.scenario(**_scenario())
    g0 = oracle.geometry_pixel(s, 8 * 3 + 0)
    g2 = oracle.geometry_pixel(s, 8 * 3 + 2)
    assert g2.sim_x[0] - g0.sim_x[0] == 0.9375 and g2.sim_y[0] == g0.sim_y[0]
    rng = np.random.default_rng(5)
    n = 12
    x = rng.uniform(9.0, 11.0, n)
    y = rng.uniform(17.0, 23.0, n)
    z = rng.uniform(0.1, 1.9, n)
    Z = rng.choice([8, 22, 38], n)
    tab = P.load("kirkland")
    srt = oracle.sort_atoms(g0, x, y, z, Z)
    assert srt["n_in_range"] == n
    peak = 0.0
    for sl in range(g0.n_slices):
        v0 = oracle.potential_slice(g0, srt, tab.values, tab.form, tab.i_per_atom, sl)
        v2 = oracle.potential_slice(g2, srt, tab.values, tab.form, tab.i_per_atom, sl)
        assert np.array_equal(v2[:, :256 - 15], v0[:, 15:]), sl
        peak = max(peak, float(np.max(v0)))
    assert peak > 1.0
