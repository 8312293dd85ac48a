"""Rigid heart-to-torso registration with the rotation search on the GPU.

Drop-in for hrdayapy/ecg/functions/register_heart_to_torso.py:
  * _all_signed_permutation_matrices  :92-106 (the 48 matrices, same order)
  * register_heart_to_torso            :146-224 (bounding-box scale and
                                                 centres, seeded target
                                                 subsample, Dice search,
                                                 first best kept)
  * run_registration                   :231-278
  * apply_registration                 :281-297
The 48 voxel-Dice evaluations (_voxel_dice, :109-138) run as one pass of
hp_register_dice_f64: a bitmap pair per permutation, every (point,
permutation) pair a thread.  The lattice of permutation r is anchored at the
minimum over the candidate and target clouds; a signed permutation moves each
coordinate of the centred source exactly, so that minimum and the largest
index are taken from the per-axis extremes here, and the device reproduces
the host's voxel indices bit for bit.
"""
from __future__ import annotations

from itertools import permutations
from pathlib import Path

import numpy as np

from ... import _native

N_PERM = 48
MAX_LATTICE_BITS = 1 << 34  # 2 GiB of bitmap words per call


def _perm_sign_table():
    """(48, 6) int32 rows {perm_0, perm_1, perm_2, s_0, s_1, s_2} in the order of
    _all_signed_permutation_matrices: R[i, perm_i] = s_i."""
    rows = []
    for perm in permutations([0, 1, 2]):
        for s0 in (1, -1):
            for s1 in (1, -1):
                for s2 in (1, -1):
                    rows.append([*perm, s0, s1, s2])
    return np.asarray(rows, dtype=np.int32)


def _all_signed_permutation_matrices():
    mats = []
    for p0, p1, p2, s0, s1, s2 in _perm_sign_table():
        R = np.zeros((3, 3))
        for i, (p, s) in enumerate(zip((p0, p1, p2), (s0, s1, s2))):
            R[i, p] = s
        mats.append(R)
    return mats


def dice_counts(src_c, c_tgt, tgt_sub, vox_mm, device="cuda"):
    """Occupied-voxel counts (|A|, |B|, |A and B|) of the 48 candidates
    s_i * src_c[:, perm_i] + c_tgt_i against tgt_sub, on the device: (48, 3) int64."""
    import torch
    src_c = np.ascontiguousarray(src_c, dtype=np.float64)
    tgt_sub = np.ascontiguousarray(tgt_sub, dtype=np.float64)
    c_tgt = np.asarray(c_tgt, dtype=np.float64)
    if src_c.ndim != 2 or src_c.shape[1] != 3 or tgt_sub.ndim != 2 or tgt_sub.shape[1] != 3:
        raise ValueError("point clouds must be (N, 3)")
    if len(src_c) == 0 or len(tgt_sub) == 0:
        raise ValueError("empty point cloud")
    if not np.all(np.isfinite(src_c)) or not np.all(np.isfinite(tgt_sub)):
        raise ValueError("point clouds must be finite")
    vox = float(vox_mm)
    if not vox > 0.0:
        raise ValueError("dice_vox_mm must be positive")
    ps = _perm_sign_table()
    s_lo, s_hi = src_c.min(axis=0), src_c.max(axis=0)
    t_lo, t_hi = tgt_sub.min(axis=0), tgt_sub.max(axis=0)
    mn = np.empty((N_PERM, 3), dtype=np.float64)
    dims = np.empty((N_PERM, 3), dtype=np.int64)
    for r in range(N_PERM):
        p, s = ps[r, :3], ps[r, 3:]
        # fl(s x + c) is monotone in x: the candidate's extremes are those of
        # the source column, mapped
        lo = np.where(s > 0, s_lo[p], -s_hi[p]) + c_tgt
        hi = np.where(s > 0, s_hi[p], -s_lo[p]) + c_tgt
        m = np.minimum(lo, t_lo)
        mn[r] = m
        top = np.maximum(np.floor((hi - m) / vox), np.floor((t_hi - m) / vox))
        dims[r] = top.astype(np.int64) + 1
    bits = np.prod(dims, axis=1)
    if np.any(bits > MAX_LATTICE_BITS):
        raise ValueError(f"Dice lattice of {int(bits.max()):,} voxels is too fine; "
                         "raise dice_vox_mm")
    if np.any(dims > np.iinfo(np.int32).max):
        raise ValueError("Dice lattice side exceeds int32")
    words = (bits + 31) // 32
    word_off = np.zeros(N_PERM + 1, dtype=np.int64)
    word_off[1:] = np.cumsum(2 * words)
    dev = torch.device(device)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)  # noqa: E731
    d_src, d_tgt = T(src_c, np.float64), T(tgt_sub, np.float64)
    d_ps, d_c = T(ps, np.int32), T(c_tgt, np.float64)
    d_mn, d_dims, d_off = T(mn, np.float64), T(dims, np.int32), T(word_off, np.int64)
    d_bits = torch.empty(int(word_off[-1]), dtype=torch.int32, device=dev)
    d_counts = torch.empty(3 * N_PERM + 1, dtype=torch.int64, device=dev)
    P = _native.ptr
    _native.call("hp_register_dice_f64", P(d_src), len(src_c), P(d_tgt), len(tgt_sub), P(d_ps),
                 N_PERM, P(d_c), P(d_mn), P(d_dims), P(d_off), vox, P(d_bits), int(word_off[-1]),
                 P(d_counts), _native.stream_handle(dev))
    counts = d_counts.cpu().numpy()
    if counts[-1] != 0:
        raise RuntimeError(f"hp_register_dice_f64: {int(counts[-1])} points fell outside "
                           "their lattice")
    return counts[:-1].reshape(N_PERM, 3)


def register_heart_to_torso(src_pts, tgt_pts, dice_vox_mm=3.0, verbose=True, device="cuda"):
    """R, t, scale, dice with (R @ (src * scale).T).T + t ~ tgt (:146-224)."""
    src_pts = np.asarray(src_pts, dtype=np.float64)
    tgt_pts = np.asarray(tgt_pts, dtype=np.float64)
    src_span = src_pts.max(axis=0) - src_pts.min(axis=0)
    tgt_span = tgt_pts.max(axis=0) - tgt_pts.min(axis=0)
    per_axis = tgt_span / src_span
    scale = float(np.median(per_axis))
    if verbose:
        print(f"  Snapshot bbox spans : {np.round(src_span, 1)} mm")
        print(f"  NRRD myo bbox spans : {np.round(tgt_span, 1)} mm")
        print(f"  Isotropic scale     : {scale:.4f}")
    src_scaled = src_pts * scale
    c_src = (src_scaled.max(axis=0) + src_scaled.min(axis=0)) / 2.0
    c_tgt = (tgt_pts.max(axis=0) + tgt_pts.min(axis=0)) / 2.0
    src_c = src_scaled - c_src
    rng = np.random.default_rng(0)
    n_sub = min(len(tgt_pts), 50_000)
    tgt_sub = tgt_pts[rng.choice(len(tgt_pts), n_sub, replace=False)]
    counts = dice_counts(src_c, c_tgt, tgt_sub, dice_vox_mm, device)
    mats = _all_signed_permutation_matrices()
    best_dice, best_R = -1.0, np.eye(3)
    for R, (na, nb, nab) in zip(mats, counts):
        denom = int(na) + int(nb)
        d = float(2 * int(nab) / denom) if denom > 0 else 0.0
        if d > best_dice:
            best_dice, best_R = d, R
    t = c_tgt - best_R @ c_src
    if verbose:
        print(f"  Best Dice : {best_dice:.4f}  (48 signed permutations on the GPU, "
              f"target subsample {n_sub:,})")
        if best_dice < 0.5:
            print("  [warn] Dice < 0.5 — inspect registration visually.")
    return best_R, t, scale, best_dice


def heart_points(label_volume, heart_label=2, spacing=(1.0, 1.0, 1.0), origin_mm=(0.0, 0.0, 0.0)):
    """Physical mm coordinates of the voxels whose label is in heart_label
    (_load_nrrd_heart_pts, :44-85, from an in-memory label volume)."""
    vol = np.asarray(label_volume)
    ijk = np.argwhere(np.isin(vol, np.atleast_1d(heart_label))).astype(np.float64)
    return ijk * np.asarray(spacing, np.float64)[None, :] + np.asarray(origin_mm, np.float64)[None, :]


def run_registration(vm_snapshots_path, torso_nrrd_path=None, heart_label=2, dice_vox_mm=3.0,
                     verbose=True, *, label_volume=None, spacing=(1.0, 1.0, 1.0),
                     origin_mm=(0.0, 0.0, 0.0), device="cuda"):
    """Register the snapshot cloud (coords_mm) to the torso's heart voxels
    (:231-278).  The NRRD route needs pynrrd; label_volume (with spacing and
    origin_mm) passes the segmentation directly."""
    snap = np.load(str(vm_snapshots_path))
    src_pts = snap["coords_mm"].astype(np.float64)
    if label_volume is None:
        if torso_nrrd_path is None:
            raise ValueError("torso_nrrd_path or label_volume is required")
        try:
            import nrrd
        except ImportError as e:
            raise ImportError("reading a torso NRRD needs pynrrd; pass label_volume= instead") from e
        vol, hdr = nrrd.read(str(Path(torso_nrrd_path)))
        if "space directions" in hdr:
            sd = np.asarray(hdr["space directions"], dtype=np.float64)
            spacing = np.array([np.linalg.norm(sd[i]) for i in range(3)])
            origin_mm = np.asarray(hdr.get("space origin", [0.0, 0.0, 0.0]), np.float64)
        elif "spacings" in hdr:
            spacing, origin_mm = np.asarray(hdr["spacings"], np.float64), np.zeros(3)
        else:
            spacing, origin_mm = np.ones(3), np.zeros(3)
        label_volume = vol
    tgt_pts = heart_points(label_volume, heart_label, spacing, origin_mm)
    if len(tgt_pts) == 0:
        raise ValueError(f"no voxel with label {heart_label} in the torso segmentation")
    R, t, scale, dice = register_heart_to_torso(src_pts, tgt_pts, dice_vox_mm=dice_vox_mm,
                                                verbose=verbose, device=device)
    return dict(R=R, t=t, scale=scale, dice=dice)


def apply_registration(coords_mm, R, t, scale):
    """torso_coords = R @ (coords_mm * scale) + t (:281-297)."""
    return (np.asarray(R) @ (np.asarray(coords_mm) * scale).T).T + np.asarray(t)


def projection_rows(coords_mm, R, t, scale, origin_mm, dx_mm, id_map, shape, device="cuda"):
    """Heart node -> torso row (-1 if outside), as build_projection_map
    (compute_bspm.py:126-157), by hp_project_rows_f64: (n,) int64 numpy."""
    import torch
    xyz = np.ascontiguousarray(coords_mm, dtype=np.float64)
    if xyz.ndim != 2 or xyz.shape[1] != 3:
        raise ValueError("coords_mm must be (N, 3)")
    ni, nj, nk = (int(v) for v in shape)
    id_map = np.ascontiguousarray(id_map, dtype=np.int32).ravel()
    if id_map.size != ni * nj * nk:
        raise ValueError("active_id_map does not match the grid shape")
    dev = torch.device(device)
    Rt = np.concatenate([np.asarray(R, np.float64).ravel(), np.asarray(t, np.float64).ravel()])
    if Rt.size != 12:
        raise ValueError("R must be 3x3 and t of length 3")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    d_xyz, d_rt = T(xyz), T(Rt)
    d_org, d_map = T(np.asarray(origin_mm, np.float64).ravel()), T(id_map)
    d_proj = torch.empty(len(xyz), dtype=torch.int64, device=dev)
    P = _native.ptr
    _native.call("hp_project_rows_f64", P(d_xyz), len(xyz), P(d_rt), float(scale), P(d_org),
                 float(dx_mm), P(d_map), ni, nj, nk, P(d_proj), _native.stream_handle(dev))
    return d_proj.cpu().numpy()
