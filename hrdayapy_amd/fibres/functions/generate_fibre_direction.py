"""Rule-based myocardial fibre field on the GPU.

Drop-in for hrdayapy.fibres.functions.generate_fibre_direction (reference
file generate_fibre_direction.py:135-281): same arguments, the same .npy cache
behaviour and the same result.  The crop to the mask's bounding box (+
bbox_pad) and the nearest-value fill of phi and psi into a 2-voxel halo
(`_fill_nan_nearest`, :90-111: scipy's distance transform with indices) run
on the host as in the reference; the gradients, the local frame
(e_t, e_l, e_c), the helix rotation and the normalisations run in one HIP
kernel over the box (hp_fibre_rule_f32), in FP32 in the reference's order of
operations.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import torch
from scipy.ndimage import binary_dilation, distance_transform_edt

from ... import _native as nat


def _fill_nan_nearest(field, S, halo=2):
    out = field.copy()
    inv = ~S
    need = binary_dilation(S, iterations=halo) & inv
    if need.any():
        idx = distance_transform_edt(inv, return_distances=False, return_indices=True)
        out[need] = field[tuple(idx[:, need])]
        del idx
    out[inv & ~need] = 0.0
    return np.nan_to_num(out, nan=0.0).astype(np.float32, copy=False)


def generate_fibre_direction(S, phi, psi, spacing_mm, *, alpha_endo_deg: float = 60.0,
                             alpha_epi_deg: float = -60.0, bbox_pad: int = 4, verbose=True,
                             save_path=None, device="cuda"):
    if save_path is not None:
        save_path = Path(save_path).with_suffix(".npy")
        if save_path.exists():
            if verbose:
                print(f"  [fibres] Loading cached fibre field from {save_path}")
            return np.load(save_path)
    nat.load()
    S = np.asarray(S).astype(bool)
    h = [float(v) for v in np.asarray(spacing_mm, dtype=float)]
    idx = np.argwhere(S)
    if len(idx) == 0:
        raise ValueError("generate_fibre_direction: empty mask")
    lo = np.maximum(idx.min(axis=0) - bbox_pad, 0)
    hi = np.minimum(idx.max(axis=0) + bbox_pad + 1, S.shape)
    box = tuple(slice(int(a), int(b)) for a, b in zip(lo, hi))
    Sc = S[box]
    Fp = _fill_nan_nearest(np.asarray(phi, np.float32)[box], Sc)
    Fs = _fill_nan_nearest(np.asarray(psi, np.float32)[box], Sc)
    if min(Sc.shape) < 2:
        raise ValueError("generate_fibre_direction: the cropped box needs every side >= 2")
    dev = torch.device(device)
    tp = torch.from_numpy(Fp).to(dev)
    ts = torch.from_numpy(Fs).to(dev)
    out = torch.empty(Sc.shape + (3,), dtype=torch.float32, device=dev)
    nat.call("hp_fibre_rule_f32", nat.ptr(tp), nat.ptr(ts), *Sc.shape, *h,
             math.radians(alpha_endo_deg), math.radians(alpha_epi_deg), nat.ptr(out),
             nat.stream_handle())
    f = np.zeros(S.shape + (3,), dtype=np.float32)
    f[box] = out.cpu().numpy()
    if save_path is not None:
        save_path.parent.mkdir(parents=True, exist_ok=True)
        np.save(save_path, f)
    return f
