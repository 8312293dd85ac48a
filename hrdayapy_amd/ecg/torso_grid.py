"""compute_torso_grid / load_torso_grid.

Drop-in for hrdayapy/ecg/torso_grid.py:26-103 and build_torso_grid
(hrdayapy/ecg/functions/build_torso_grid.py:495-691) for the sharp-boundary
path.  The NRRD route needs pynrrd; ``label_volume`` (with ``origin_mm``)
passes a coarse label volume directly.  The direct factorisation
(solve_fn) and the Smoothed Boundary Method are not provided: the solve is
the GPU PCG of hrdayapy_amd.ecg.functions.forward.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .functions.torso import TorsoGrid


def compute_torso_grid(torso_nrrd_path=None, sigma_map=None, *, dx_coarse_mm=2.0, verbose=True,
                       save_path=None, use_sbm=False, sbm_interface_width_mm=3.0,
                       sbm_psi_cutoff=1e-3, label_volume=None, origin_mm=(0.0, 0.0, 0.0)):
    if use_sbm:
        raise NotImplementedError("the Smoothed Boundary Method is not part of the GPU path")
    if sigma_map is None:
        raise ValueError("sigma_map is required")
    if label_volume is None:
        try:
            import nrrd  # noqa: F401
        except ImportError as e:
            raise ImportError("reading a torso NRRD needs pynrrd; pass label_volume= instead") from e
        lab, spacing, origin = _load_and_resample(Path(torso_nrrd_path), dx_coarse_mm)
        origin_mm = origin
    else:
        lab = np.asarray(label_volume)
    g = TorsoGrid.from_labels(lab, sigma_map, dx_coarse_mm, origin_mm, verbose=verbose)
    if save_path is not None:
        np.savez(str(save_path), shape=np.array(g["shape"], np.int32), dx_mm=np.float64(g["dx_mm"]),
                 origin_mm=g["origin_mm"], sigma_vol=g["sigma_vol"].astype(np.float32),
                 active_mask=g["active_mask"], tissue_mask=g["tissue_mask"],
                 active_ids=g["active_ids"], active_id_map=g["active_id_map"],
                 surface_flat=g["surface_flat"], surface_rows=g["surface_rows"],
                 surface_xyz=g["surface_xyz"], pin_row=np.int32(g["pin_row"]), use_sbm=False)
    return g


def _load_and_resample(path, dx):
    import nrrd
    from scipy import ndimage as ndi
    data, hdr = nrrd.read(str(path))
    dirs = np.asarray(hdr.get("space directions", np.eye(3)), dtype=float)
    spacing = np.linalg.norm(dirs, axis=1)
    origin = np.asarray(hdr.get("space origin", np.zeros(3)), dtype=float)
    zoom = spacing / dx
    return ndi.zoom(data, zoom, order=0), spacing, origin
