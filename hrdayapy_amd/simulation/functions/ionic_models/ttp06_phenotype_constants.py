"""Phenotype conductances (ttp06_phenotype_constants.py:41-67 of the reference)."""
from ....ttp06 import TTP06PhenotypeConductances, TTP06_PHENOTYPES  # noqa: F401

PHENOTYPE_LABELS = {"endo": "Endocardial", "mid": "Mid-myocardial (M-cell)", "epi": "Epicardial"}
