"""Write golden vectors of the reference's torso-grid build (sharp and smoothed
boundary) and heart-to-torso registration by running the reference's own
functions.

Run in the development container only (the reference tree is not present on
GPU machines); the output ``torso_golden.npz`` is committed next to this
script, and the tests read it (tests/test_oracle_torso.py).

Imported by file path (both modules need only numpy and scipy; no package
``__init__`` runs):
  * hrdayapy/ecg/functions/build_torso_grid.py: _resample_labels,
    compute_smoothed_domain, extend_conductivity, _assemble_stiffness (with
    and without psi), _find_surface_ids;
  * hrdayapy/ecg/functions/register_heart_to_torso.py:
    register_heart_to_torso and _voxel_dice (the 48 Dice values of its
    search, evaluated as its loop does).
Inputs, stored too: a nested-ellipsoid torso segmentation (labels 1 blood,
2 lungs, 3 muscle, 4 bulk) at fine spacing (0.75, 0.75, 1.0) mm resampled to
2 mm; a BiV-shell heart cloud and a target made from it by a known signed
permutation, scale and shift (the cloud carries an off-axis block so that
the permutation is the clear best; one target of 61,000 points, so the search
draws its seeded 50,000-point subsample).
"""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/hrdayapy")
HERE = Path(__file__).resolve().parent
OUT = HERE / "torso_golden.npz"
sys.path.insert(0, str(HERE.parent.parent))

from hrdayapy_amd import geometry as G  # noqa: E402


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def fine_torso():
    """Nested ellipsoids on a 64x56x40 grid at (0.75, 0.75, 1.0) mm."""
    sp = np.array([0.75, 0.75, 1.0])
    shape = (64, 56, 40)
    x, y, z = np.meshgrid(*[(np.arange(n) + 0.5) * h for n, h in zip(shape, sp)], indexing="ij")
    c = np.array([24.0, 21.0, 20.0])
    r2 = lambda a, b, cc, o=(0, 0, 0): (((x - c[0] - o[0]) / a) ** 2 + ((y - c[1] - o[1]) / b) ** 2  # noqa: E731
                                          + ((z - c[2] - o[2]) / cc) ** 2)
    lab = np.zeros(shape, dtype=np.int32)
    lab[r2(20.0, 17.0, 15.0) <= 1] = 4
    lab[r2(16.0, 13.5, 12.0) <= 1] = 3
    lab[r2(5.0, 4.0, 6.0, (-7, 0, 0)) <= 1] = 2
    lab[r2(5.0, 4.0, 6.0, (7, 0, 0)) <= 1] = 2
    lab[r2(3.0, 3.0, 3.0, (0, 5, 2)) <= 1] = 1
    return lab, sp


def main():
    tg = _load("ref_torso", REF / "ecg/functions/build_torso_grid.py")
    rg = _load("ref_register", REF / "ecg/functions/register_heart_to_torso.py")
    out = {}
    # --- torso grid: sharp and SBM ---------------------------------------
    lab, sp = fine_torso()
    dx = 2.0
    sigma_map = {1: 0.7, 2: 0.05, 3: 0.2, 4: 0.2}
    coarse = tg._resample_labels(lab, sp, dx)
    psi = tg.compute_smoothed_domain(lab, sp, dx, coarse.shape, interface_width_mm=3.0,
                                     psi_cutoff=1e-3)
    sigma = np.zeros(coarse.shape, dtype=np.float64)
    for k, v in sigma_map.items():
        sigma[coarse == k] = v
    tissue = coarse != 0
    h_m = dx * 1e-3
    for tag, use_sbm in (("sharp", False), ("sbm", True)):
        if use_sbm:
            active = psi > 0.0
            sig = tg.extend_conductivity(sigma, tissue, active)
            pv = psi
        else:
            active, sig, pv = tissue, sigma, None
        act_flat = np.where(active.ravel())[0].astype(np.int32)
        id_map = np.full(sig.size, -1, dtype=np.int32)
        id_map[act_flat] = np.arange(len(act_flat), dtype=np.int32)
        A = tg._assemble_stiffness(sig, h_m, active, act_flat, id_map, psi_vol=pv).tocsr()
        A.sort_indices()
        out[f"{tag}_sigma"] = sig
        out[f"{tag}_active_ids"] = act_flat
        out[f"{tag}_A_data"], out[f"{tag}_A_indices"], out[f"{tag}_A_indptr"] = (
            A.data, A.indices, A.indptr)
    sf, _ = tg._find_surface_ids(coarse, tissue, np.full(coarse.size, 0, dtype=np.int32))
    out.update(fine_labels=lab, fine_spacing=sp, dx_coarse=np.float64(dx), coarse_labels=coarse,
               psi=psi, surface_flat=sf,
               sigma_keys=np.array(list(sigma_map)), sigma_vals=np.array(list(sigma_map.values())))
    # --- registration ------------------------------------------------------
    geo = G.biv_shell(40, 2.0)
    src = np.argwhere(geo["S"]).astype(np.float64) * 2.0 + np.array([3.0, -4.0, 10.0])
    # an off-axis block breaks the shell's near symmetries, so one signed
    # permutation is the clear best
    blk = np.argwhere(np.ones((6, 5, 4), bool)).astype(np.float64) * 2.0
    src = np.vstack([src, blk + src.min(axis=0) + np.array([0.0, 0.0, 30.0])])
    R_true = np.zeros((3, 3))
    R_true[0, 2], R_true[1, 0], R_true[2, 1] = -1.0, 1.0, -1.0
    rng = np.random.default_rng(7)
    tgt = (R_true @ (src * 1.3).T).T + np.array([120.0, 80.0, -30.0])
    tgt = np.repeat(tgt, int(np.ceil(61_000 / len(tgt))), axis=0)[:61_000]
    # float32-representable values (stored as float32, exact in float64)
    tgt = (tgt + rng.uniform(-1.2, 1.2, size=tgt.shape)).astype(np.float32).astype(np.float64)
    R, t, scale, dice = rg.register_heart_to_torso(src, tgt, dice_vox_mm=3.0, verbose=False)
    # the 48 Dice values of the search, evaluated as its loop does
    s = src * scale
    c_src = (s.max(axis=0) + s.min(axis=0)) / 2.0
    c_tgt = (tgt.max(axis=0) + tgt.min(axis=0)) / 2.0
    sub = tgt[np.random.default_rng(0).choice(len(tgt), min(len(tgt), 50_000), replace=False)]
    all_d = [rg._voxel_dice((Rm @ (s - c_src).T).T + c_tgt, sub, 3.0)
             for Rm in rg._all_signed_permutation_matrices()]
    out.update(reg_src=src, reg_tgt=tgt.astype(np.float32), reg_R=R, reg_t=t, reg_scale=np.float64(scale),
               reg_dice=np.float64(dice), reg_all_dice=np.asarray(all_d),
               reg_R_true=R_true)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: coarse {coarse.shape}, sharp rows {len(out['sharp_active_ids'])}, "
          f"SBM rows {len(out['sbm_active_ids'])}, registration dice {dice:.4f}")


if __name__ == "__main__":
    main()
