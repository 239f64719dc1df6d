"""Writes the golden vectors under tests/golden/ (run in the build container).

Everything here comes from numpy/scipy, independent of the oracle and the product:
  fft_ortho.npz      seeded complex arrays and numpy.fft.fft2/ifft2(norm="ortho")
  bessel.npz         scipy.special.k0/k1 on a grid spanning both rational-form branches
  vproj_scipy.npz    projected Kirkland/Peng/Lobato potentials of a few elements, as
                     written in kernels/transmission_potentials_projected_d.cl:104-181,
                     evaluated with scipy's exact K0/K1 instead of the NR approximations
Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.special as ss

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cltem_amd import params as P  # noqa: E402


def fft_vectors():
    rng = np.random.default_rng(20261202)
    out = {}
    for (h, w) in [(16, 16), (24, 24), (48, 48), (32, 64)]:
        a = rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))
        out[f"in_{h}x{w}"] = a
        out[f"fwd_{h}x{w}"] = np.fft.fft2(a, norm="ortho")
        out[f"inv_{h}x{w}"] = np.fft.ifft2(a, norm="ortho")
    np.savez_compressed(os.path.join(HERE, "fft_ortho.npz"), **out)


def bessel():
    x = np.concatenate([np.geomspace(1e-4, 1.0, 60), np.linspace(1.0 + 1e-9, 80.0, 120)])
    np.savez_compressed(os.path.join(HERE, "bessel.npz"), x=x, k0=ss.k0(x), k1=ss.k1(x))


def vproj_tables():
    r = np.concatenate([np.geomspace(0.01, 1.0, 25), np.linspace(1.05, 8.0, 40)])
    elems = [6, 8, 14, 22, 38, 79]
    out = {"r": r, "Z": np.array(elems)}
    for name in ("kirkland", "peng", "lobato"):
        p = P.load(name)
        vals = np.zeros((len(elems), len(r)))
        for i, Z in enumerate(elems):
            c = p.values[(Z - 1) * p.per_atom:(Z * p.per_atom)]
            if p.form == 0:  # a1 b1 a2 b2 a3 b3 c1 d1 c2 d2 c3 d3
                a, b = c[0:6:2], c[1:6:2]
                cc, d = c[6:12:2], c[7:12:2]
                for k in range(3):
                    vals[i] += 300.8242834 * a[k] * ss.k0(2 * np.pi * r * np.sqrt(b[k]))
                    vals[i] += 150.4121417 * (cc[k] / d[k]) * np.exp(-(np.pi * r) ** 2 / d[k])
            elif p.form == 1:  # peng: a1..a5 b1..b5
                a, b = c[:5], c[5:]
                for k in range(5):
                    vals[i] += 150.4121417 * a[k] / b[k] * np.exp(-(np.pi * r) ** 2 / b[k])
            else:  # lobato
                a, b = c[:5], c[5:]
                for k in range(5):
                    x = 2 * np.pi * r / np.sqrt(b[k])
                    vals[i] += 945.090144399935 * a[k] * b[k] ** -1.5 * (ss.k0(x) + r * ss.k1(x))
        out[name] = vals
    np.savez_compressed(os.path.join(HERE, "vproj_scipy.npz"), **out)


if __name__ == "__main__":
    fft_vectors()
    bessel()
    vproj_tables()
    print("golden vectors written to", HERE)
