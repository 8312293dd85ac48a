"""compute_torso_grid / load_torso_grid.

Drop-in for hrdayapy/ecg/torso_grid.py:26-103 and build_torso_grid
(hrdayapy/ecg/functions/build_torso_grid.py:495-691), sharp and smoothed
(SBM) boundary, built on the device (ecg/functions/torso_build.py).  The
NRRD route needs pynrrd; ``fine_label_volume`` (with ``fine_spacing``)
passes the segmentation in memory and is resampled to dx_coarse_mm as the
NRRD would be; ``label_volume`` passes an already coarse label volume
(sharp boundary only: the SBM's domain function needs the fine
segmentation).  The direct factorisation behind solve_fn is replaced by the
device's multigrid-preconditioned solve (TorsoGrid).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .functions.torso import TorsoGrid

_SAVED = ("shape", "dx_mm", "origin_mm", "sigma_vol", "active_mask", "tissue_mask", "active_ids",
          "active_id_map", "surface_flat", "surface_rows", "surface_xyz", "pin_row", "use_sbm")


def compute_torso_grid(torso_nrrd_path=None, sigma_map=None, *, dx_coarse_mm=2.0, verbose=True,
                       save_path=None, use_sbm=False, sbm_interface_width_mm=3.0,
                       sbm_psi_cutoff=1e-3, label_volume=None, origin_mm=(0.0, 0.0, 0.0),
                       fine_label_volume=None, fine_spacing=None, device="cuda"):
    if sigma_map is None:
        raise ValueError("sigma_map is required")
    n_src = sum(x is not None for x in (torso_nrrd_path, label_volume, fine_label_volume))
    if n_src != 1:
        raise ValueError("pass exactly one of torso_nrrd_path, fine_label_volume, label_volume")
    if label_volume is not None:
        if use_sbm:
            raise ValueError("use_sbm needs the fine segmentation (torso_nrrd_path or "
                             "fine_label_volume); label_volume is already coarse")
        g = TorsoGrid.from_labels(np.asarray(label_volume), sigma_map, dx_coarse_mm, origin_mm,
                                  verbose=verbose, device=device)
    else:
        if torso_nrrd_path is not None:
            fine, fine_spacing, origin_mm = _load_nrrd(Path(torso_nrrd_path))
        else:
            if fine_spacing is None:
                raise ValueError("fine_label_volume needs fine_spacing")
            fine = np.asarray(fine_label_volume)
        g = TorsoGrid.from_segmentation(fine, np.asarray(fine_spacing, np.float64), sigma_map,
                                        dx_coarse_mm, origin_mm, use_sbm=use_sbm,
                                        sbm_interface_width_mm=sbm_interface_width_mm,
                                        sbm_psi_cutoff=sbm_psi_cutoff, verbose=verbose,
                                        device=device)
    if save_path is not None:
        kw = {k: g[k] for k in _SAVED}
        kw["shape"] = np.array(g["shape"], np.int32)
        kw["dx_mm"] = np.float64(g["dx_mm"])
        kw["sigma_vol"] = g["sigma_vol"].astype(np.float32)
        kw["pin_row"] = np.int32(g["pin_row"])
        kw["surface_xyz"] = np.asarray(g["surface_xyz"], np.float32)
        if g["use_sbm"]:
            kw["psi_vol"] = g["psi_vol"].astype(np.float32)
        np.savez(str(save_path), **kw)
    return g


def load_torso_grid(save_path, torso_nrrd_path=None, sigma_map=None, *, dx_coarse_mm=2.0,
                    verbose=True, device="cuda"):
    """Rebuild the grid from a saved torso_grid.npz (torso_grid.py:86-103): the
    labels are not stored, so the system is reassembled on the device from the
    saved sigma volume, active and tissue masks (and psi for SBM grids)."""
    z = np.load(str(Path(save_path).with_suffix(".npz")))
    sigma = z["sigma_vol"].astype(np.float64)
    tissue = z["tissue_mask"].astype(bool)
    use_sbm = bool(z["use_sbm"]) if "use_sbm" in z.files else False
    psi = z["psi_vol"].astype(np.float64) if use_sbm else None
    # one label per distinct conductivity reproduces the saved sigma on the tissue;
    # band cells of an SBM grid get their extended sigma back from the nearest tissue
    vals, lab_t = np.unique(sigma[tissue], return_inverse=True)
    lab = np.zeros(sigma.shape, np.int32)
    lab[tissue] = lab_t + 1
    smap = {i + 1: float(v) for i, v in enumerate(vals)}
    return TorsoGrid.from_labels(lab, smap, float(z["dx_mm"]), z["origin_mm"], verbose=verbose,
                                 psi_vol=psi, device=device)


def _load_nrrd(path):
    try:
        import nrrd
    except ImportError as e:
        raise ImportError("reading a torso NRRD needs pynrrd; pass fine_label_volume= "
                          "(with fine_spacing=) or label_volume= instead") from e
    data, hdr = nrrd.read(str(path))
    if "space directions" in hdr:
        sd = np.asarray(hdr["space directions"], dtype=np.float64)
        spacing = np.array([np.linalg.norm(sd[i]) for i in range(3)])
        origin = np.asarray(hdr.get("space origin", [0.0, 0.0, 0.0]), np.float64)
    elif "spacings" in hdr:
        spacing, origin = np.asarray(hdr["spacings"], np.float64), np.zeros(3)
    else:
        spacing, origin = np.ones(3), np.zeros(3)
    return np.asarray(data, dtype=np.int32), spacing, origin
