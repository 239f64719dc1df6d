"""The engine's tabulated potentials (cubic pieces in r^2, indexed by the float bits of r^2)
against the FP64 oracle's restatement of the reference formulas
(kernels/transmission_potentials_projected_d.cl:71-181,
kernels/transmission_potentials_full_3d_d.cl:57-134).  CPU only: the table is built on the
host and evaluated here exactly as the kernels evaluate it (float32 r^2 and coefficients)."""
import ctypes as C

import numpy as np
import pytest

import cltem_amd
from cltem_amd import params as P

CASES = [("kirkland", False, (1, 8, 22, 38, 79)), ("kirkland", True, (6, 8, 79)),
         ("peng", False, (6, 14)), ("peng", True, (8, 14)),
         ("lobato", False, (8, 14, 38)), ("lobato", True, (8, 14, 79))]


def _oracle_v(oracle, p, Z, r, full3d):
    L = oracle.lib()
    v = np.ascontiguousarray(p.values, dtype=np.float64)
    ptr = v.ctypes.data_as(C.POINTER(C.c_double))
    f = L.orc_v3d if full3d else L.orc_vproj
    return np.array([f(p.form, ptr, p.i_per_atom, int(Z), float(x)) for x in r])


@pytest.mark.parametrize("name,full3d,zs", CASES)
def test_table_matches_reference_formulas(oracle, name, full3d, zs):
    """From r = 0.002 A (the r_min of 0.008 A pixels) to 8 A (projected) or 10 A (full 3D):
    |V_table - V| <= 1e-6 |V| + 1e-7 |V(1 A)|, i.e. relative 1e-6 wherever V matters and
    nothing above 1e-7 of the potential's scale in the far tail where V itself is negligible."""
    p = P.load(name)
    rng = np.random.default_rng(7)
    for Z in zs:
        tab = cltem_amd.potential_table(p, Z, full3d)
        r = np.exp(rng.uniform(np.log(0.002), np.log(10.0 if full3d else 8.0), 3000))
        r2 = np.float32(r) ** 2
        got = cltem_amd.potential_table_eval(tab, r2)
        exact = _oracle_v(oracle, p, Z, np.sqrt(r2.astype(np.float64)), full3d)
        scale = abs(_oracle_v(oracle, p, Z, [1.0], full3d)[0])
        err = np.abs(got.astype(np.float64) - exact)
        bound = 1e-6 * np.abs(exact) + 1e-7 * scale
        worst = int(np.argmax(err / bound))
        assert np.all(err <= bound), (name, full3d, Z, float(np.sqrt(r2[worst])), float(err[worst]),
                                      float(exact[worst]))


def test_table_bucket_lookup_hand_cases():
    """r^2 = 1 A^2 starts bucket (0 - (-20)) * 32 = 640 at t = 0, so V(1) is its c0; r^2 = 2^-20
    starts bucket 0; the largest r^2 below 2^8 falls in the last bucket (895) at t -> 1."""
    p = P.load("kirkland")
    tab = cltem_amd.potential_table(p, 8)
    assert cltem_amd.potential_table_eval(tab, [1.0])[0] == np.float32(tab[640, 0])
    assert cltem_amd.potential_table_eval(tab, [2.0 ** -20])[0] == np.float32(tab[0, 0])
    top = np.nextafter(np.float32(256.0), np.float32(0.0))
    c = tab[895].astype(np.float32)
    t = np.float32(((1 << 18) - 1) / (1 << 18))
    assert cltem_amd.potential_table_eval(tab, [top])[0] == ((c[3] * t + c[2]) * t + c[1]) * t + c[0]


def test_table_rejects_bad_input():
    p = P.load("kirkland")
    with pytest.raises(cltem_amd.MseError):
        cltem_amd.potential_table(p, 0)
    with pytest.raises(cltem_amd.MseError):
        cltem_amd.potential_table(p, 10_000)


def test_projected_cutoff_threshold():
    """The projected kernels test r2 <= 64 + 2^-46 for the reference's
    r2 <= 64.0000000001 && sqrt(r2) <= 8.0 (correctly rounded double sqrt): the two agree on
    every double from 64 - 2^-40 up to 64.0000000001 and on both sides of it."""
    t = float.fromhex("0x1.0000000000001p+6")
    xs = [64.0]
    for _ in range(64):
        xs.append(np.nextafter(xs[-1], 100.0))
    xs += [np.nextafter(64.0, 0.0), 64.0 - 2.0 ** -40, 64.0000000001, np.nextafter(64.0000000001, 100.0), 65.0]
    for x in xs:
        ref = x <= 64.0000000001 and np.sqrt(x) <= 8.0
        assert (x <= t) == ref, x


def _device_slope(tab, r2):
    """dV/d(r2) as the tabulated position-gradient kernel computes it: the coefficients in the
    low mantissa bits m (c_k 2^(-18 k), float), dV/dm = (3 c3 m + 2 c2) m + c1 in float, times
    dm/d(r2) = 2^(23 - e), the float of biased exponent 277 - (bits >> 23)."""
    r2 = np.asarray(r2, dtype=np.float32)
    bits = r2.view(np.uint32)
    c = tab[(bits >> 18).astype(np.int64) - ((127 - 20) << 5)].astype(np.float64)
    d = np.stack([c[:, 1] * 2.0 ** -18, c[:, 2] * 2.0 ** -36, c[:, 3] * 2.0 ** -54], 1).astype(np.float32)
    m = (bits & ((1 << 18) - 1)).astype(np.float32)
    dv_dm = (np.float32(3) * d[:, 2] * m + np.float32(2) * d[:, 1]) * m + d[:, 0]
    scale = ((277 - (bits >> 23)).astype(np.uint32) << 23).view(np.float32)
    return dv_dm.astype(np.float64) * scale.astype(np.float64)


@pytest.mark.parametrize("name,Z", [("kirkland", 79), ("kirkland", 6), ("peng", 14), ("lobato", 8)])
def test_table_slope_matches_formula_derivative(oracle, name, Z):
    """The projected gradient kernel takes v'(r) = 2 r dV/d(r2) from the slope of the table's
    cubic.  Against a central difference of the FP64 formulas from 0.05 to 8 A: within 1e-5 of
    max |v'| (a wrong exponent or a dropped term of the slope misses by orders of magnitude),
    and on the cubic's own terms it is the exact derivative: the slope at a bucket start is c1
    times 2^(23 - e) / 2^18."""
    p = P.load(name)
    tab = cltem_amd.potential_table(p, Z, False)
    r = np.exp(np.linspace(np.log(0.05), np.log(7.99), 1500))
    r2 = np.float32(r) ** 2
    rr = np.sqrt(r2.astype(np.float64))
    h = 1e-5 * rr
    dv = (_oracle_v(oracle, p, Z, rr + h, False) - _oracle_v(oracle, p, Z, rr - h, False)) / (2 * h)
    got = 2 * rr * _device_slope(tab, r2)
    assert np.max(np.abs(got - dv)) <= 1e-5 * np.max(np.abs(dv))
    # r2 = 1 starts bucket 640 (m = 0): the slope is c1 / width, width = 2^0 / 32
    assert _device_slope(tab, [1.0])[0] == pytest.approx(float(np.float32(tab[640, 1] * 2.0 ** -18)) * 2.0 ** 18 * 32,
                                                          rel=1e-7)
