"""Per-node region labels (drop-in for
hrdayapy/simulation/functions/ionic_models/region_map.py:51-153): the same
label constants, ``region_map_from_segmentation``, ``make_transmural_region_map``
(phi < 0.35 endo, phi >= 0.65 epi, else mid) and ``validate_region_map``
(a UserWarning names the labels that will run SilentCell).
"""
from __future__ import annotations

import warnings

import numpy as np

LABEL_INFARCT = 0
LABEL_ENDO = 1
LABEL_MCELL = 2
LABEL_EPI = 3
LABEL_BORDER = 4

PHI_ENDO_MAX = 0.35
PHI_EPI_MIN = 0.65


def region_map_from_segmentation(seg_mask: np.ndarray, vox_idx: np.ndarray) -> np.ndarray:
    """Label of every tissue node, in the order of vox_idx (np.argwhere(S))."""
    v = np.asarray(vox_idx)
    return np.asarray(seg_mask)[v[:, 0], v[:, 1], v[:, 2]].astype(np.int32)


def make_transmural_region_map(phi_nodes: np.ndarray) -> np.ndarray:
    """Three transmural layers from phi (0 = endocardium, 1 = epicardium)."""
    phi = np.asarray(phi_nodes)
    rm = np.full(phi.shape, LABEL_MCELL, dtype=np.int32)
    rm[phi < PHI_ENDO_MAX] = LABEL_ENDO
    rm[phi >= PHI_EPI_MIN] = LABEL_EPI
    tot = max(1, len(rm))
    n = [int((rm == k).sum()) for k in (LABEL_ENDO, LABEL_MCELL, LABEL_EPI)]
    print(f"[region_map] Transmural 3-layer split:  endo={n[0]} ({100 * n[0] / tot:.1f}%)  "
          f"mcell={n[1]} ({100 * n[1] / tot:.1f}%)  epi={n[2]} ({100 * n[2] / tot:.1f}%)")
    return rm


def validate_region_map(region_map: np.ndarray, model_registry: dict,
                        warn_unlabelled: bool = True) -> None:
    labels = [int(u) for u in np.unique(region_map)]
    unmapped = [u for u in labels if u not in model_registry]
    if unmapped and warn_unlabelled:
        warnings.warn(
            f"[region_map] The following region labels have no entry in model_registry and will "
            f"use SilentCell (zero ionic current): {unmapped}.  If this is intentional (e.g. "
            f"infarct core), suppress this warning with warn_unlabelled=False.", UserWarning,
            stacklevel=2)
    mapped = [u for u in labels if u in model_registry]
    print(f"[region_map] Validation:  mapped={mapped}  silent (no model)={unmapped}")
