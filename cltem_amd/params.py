"""Scattering-factor parameterisations (Kirkland, Peng, Lobato).

The tables in cltem_amd/params/*.dat are the published parameterisations in the
file format clTEM reads (utilities/stringutils.cpp:149-184): line 1 the name,
line 2 the form ("kirkland", "peng" or "lobato"), line 3 the number of terms per
atom, then whitespace-separated values, one row per Z starting at Z = 1.
Forms map to 0/1/2 as in structure/structureparameters.h:16-20; parameters per
term are 4 (Kirkland: a,b pairs then c,d pairs) or 2 (Peng, Lobato: a_i then b_i)
(structure/structureparameters.h:65-75).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

PARAM_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params")

FORMS = {"kirkland": 0, "peng": 1, "lobato": 2}
PER_TERM = {0: 4, 1: 2, 2: 2}


@dataclass
class Parameterisation:
    name: str
    form: int
    i_per_atom: int
    values: np.ndarray  # float64, (max_Z * per_atom,)

    @property
    def per_atom(self) -> int:
        return self.i_per_atom * PER_TERM[self.form]

    @property
    def max_atomic_number(self) -> int:
        return len(self.values) // self.per_atom


def load(name: str) -> Parameterisation:
    path = name if os.path.isfile(name) else os.path.join(PARAM_DIR, name + ".dat")
    with open(path) as fh:
        pname = fh.readline().strip()
        form_name = fh.readline().strip()
        per = int(fh.readline().strip())
        values = np.array(fh.read().split(), dtype=np.float64)
    if form_name not in FORMS:
        raise ValueError(f"Unknown parameterisation form {form_name} in file {path}")
    p = Parameterisation(pname, FORMS[form_name], per, values)
    if len(values) % p.per_atom != 0:
        raise ValueError(f"Parameterisation {pname} does not contain correct parameters per atom")
    return p
