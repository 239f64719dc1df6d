"""CCD response tables (DQE and NTF) in clTEM's ccds/*.dat format.

Format (read as Utils::ccdToDqeNtf does, src/simulation/utilities/stringutils.cpp:186-245):
line 1 the name, then a "DQE" line followed by one line of numbers and an "NTF" line followed
by one line of numbers; numbers are read one after another as a stream would (some entries
in the published files run together, e.g. "0.7056940.701858" is 0.705694 then 0.701858), and
each table must hold 725 values.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass

import numpy as np

CCD_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ccds")
TABLE_LEN = 725
_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


@dataclass
class CcdParams:
    name: str
    dqe: np.ndarray
    ntf: np.ndarray


def load_ccd(name_or_path: str) -> CcdParams:
    path = name_or_path if os.path.exists(name_or_path) else os.path.join(CCD_DIR, name_or_path + ".dat")
    with open(path) as f:
        lines = f.read().splitlines()
    name = lines[0].strip()
    tables = {}
    i = 1
    while i < len(lines):
        head = "".join(lines[i].split())
        if head in ("DQE", "NTF") and i + 1 < len(lines):
            tables[head] = np.array([float(t) for t in _NUM.findall(lines[i + 1])])
            i += 2
        else:
            i += 1
    for k in ("DQE", "NTF"):
        if k not in tables or len(tables[k]) != TABLE_LEN:
            raise ValueError(f"{path}: {k} table missing or not {TABLE_LEN} values")
    return CcdParams(name, tables["DQE"], tables["NTF"])
