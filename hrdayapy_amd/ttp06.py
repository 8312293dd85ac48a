"""TTP06 constants shared by the host side of the package.

State column order and resting values of the reference
(hrdayapy/simulation/functions/ionic_models/gpu_ttp06.py:150-176) and the
phenotype conductances (ttp06_phenotype_constants.py:54-61).  The kernels
carry the same numbers in hrdayapy_amd/csrc/ttp06_cell.h.
"""
from __future__ import annotations

from typing import NamedTuple

STATE_NAMES = ["Ki", "Nai", "Cai", "Xr1", "Xr2", "Xs", "m", "h", "j", "d", "f", "f2", "fCass",
               "s", "r", "CaSR", "Rprime", "Cass"]
N_STATES = len(STATE_NAMES)
INIT = {"Ki": 138.4, "Nai": 10.355, "Cai": 0.00013, "Xr1": 0.00448, "Xr2": 0.476, "Xs": 0.0087,
        "m": 0.00155, "h": 0.7573, "j": 0.7225, "d": 3.164e-5, "f": 0.8009, "f2": 0.9778,
        "fCass": 0.9953, "s": 0.3212, "r": 2.235e-8, "CaSR": 3.715, "Rprime": 0.9068,
        "Cass": 0.00036}
V_REST = -86.709


class TTP06PhenotypeConductances(NamedTuple):
    g_Ks: float
    g_to: float
    g_Kr: float
    g_CaL: float
    endo_s: bool


TTP06_PHENOTYPES = {
    "endo": TTP06PhenotypeConductances(0.392, 0.073, 0.153, 3.98e-5, True),
    "mid": TTP06PhenotypeConductances(0.098, 0.294, 0.153, 3.98e-5, True),
    "epi": TTP06PhenotypeConductances(0.392, 0.294, 0.153, 3.98e-5, False),
}
CELLTYPE = {"endo": 0, "mid": 1, "epi": 2}
