"""CPU restatement of the reference's heart-to-torso registration (test
infrastructure: imported only by tests).

Reference: hrdayapy/ecg/functions/register_heart_to_torso.py
  * the 48 signed permutation matrices   :92-106 (permutation-major, signs
                                           s0, s1, s2 with s2 fastest)
  * _voxel_dice                           :109-138 (lattice anchored at the
                                           common minimum, floor(x / vox))
  * register_heart_to_torso               :146-224 (median per-axis span
                                           ratio, bounding-box centres,
                                           seeded subsample of <= 50,000
                                           target points, first best kept)
Pinned by golden vectors written by running the reference's own functions
(tests/golden/make_torso_vectors.py -> torso_golden.npz), and by brute force
on tiny clouds (the Dice of identical clouds is 1, of disjoint ones 0, and a
cloud rotated by a known signed permutation is recovered).
"""
from __future__ import annotations

from itertools import permutations

import numpy as np


def signed_permutations():
    mats = []
    for perm in permutations([0, 1, 2]):
        for s0 in (1, -1):
            for s1 in (1, -1):
                for s2 in (1, -1):
                    R = np.zeros((3, 3))
                    for i, (p, s) in enumerate(zip(perm, (s0, s1, s2))):
                        R[i, p] = s
                    mats.append(R)
    return mats


def voxel_dice(a, b, vox):
    """Dice of the occupied-voxel sets of two clouds on the lattice anchored at
    their common minimum."""
    mn = np.vstack([a, b]).min(axis=0)
    ia = {tuple(v) for v in np.floor((a - mn) / vox).astype(np.int64)}
    ib = {tuple(v) for v in np.floor((b - mn) / vox).astype(np.int64)}
    den = len(ia) + len(ib)
    return float(2 * len(ia & ib) / den) if den else 0.0


def register(src, tgt, vox=3.0):
    """(R, t, scale, best dice, the 48 dice values)."""
    src = np.asarray(src, dtype=np.float64)
    tgt = np.asarray(tgt, dtype=np.float64)
    scale = float(np.median((tgt.max(0) - tgt.min(0)) / (src.max(0) - src.min(0))))
    s = src * scale
    c_src = (s.max(0) + s.min(0)) / 2.0
    c_tgt = (tgt.max(0) + tgt.min(0)) / 2.0
    sc = s - c_src
    rng = np.random.default_rng(0)
    n_sub = min(len(tgt), 50_000)
    sub = tgt[rng.choice(len(tgt), n_sub, replace=False)]
    dice = []
    best, best_R = -1.0, np.eye(3)
    for R in signed_permutations():
        d = voxel_dice((R @ sc.T).T + c_tgt, sub, vox)
        dice.append(d)
        if d > best:
            best, best_R = d, R
    return best_R, c_tgt - best_R @ c_src, scale, best, np.asarray(dice)
