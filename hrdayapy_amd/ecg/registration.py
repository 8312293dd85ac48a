"""compute_registration / load_registration.

Drop-in for hrdayapy/ecg/registration.py:21-79: rigid heart-to-torso
registration of the snapshot cloud, with the 48-permutation Dice search on
the device (ecg/functions/registration.py).  The NRRD route needs pynrrd;
``label_volume`` (with ``spacing`` and ``origin_mm``) passes the torso
segmentation in memory.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .functions.registration import run_registration


def compute_registration(vm_snapshots_path, torso_nrrd_path=None, *, heart_label=2,
                         dice_vox_mm=3.0, verbose=True, save_path=None, label_volume=None,
                         spacing=(1.0, 1.0, 1.0), origin_mm=(0.0, 0.0, 0.0), device="cuda"):
    result = run_registration(Path(vm_snapshots_path),
                              None if torso_nrrd_path is None else Path(torso_nrrd_path),
                              heart_label=heart_label, dice_vox_mm=dice_vox_mm, verbose=verbose,
                              label_volume=label_volume, spacing=spacing, origin_mm=origin_mm,
                              device=device)
    if save_path is not None:
        save_path = Path(save_path).with_suffix(".npz")
        save_path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(str(save_path), R=result["R"], t=result["t"], scale=result["scale"],
                 dice=result["dice"])
    return result


def load_registration(path):
    """A saved registration, no recomputation (registration.py:70-79)."""
    path = Path(path).with_suffix(".npz")
    if not path.exists():
        raise FileNotFoundError(f"No saved registration at {path}")
    d = np.load(str(path))
    return {"R": d["R"], "t": d["t"], "scale": float(d["scale"]), "dice": float(d["dice"])}
