"""Write golden vectors of the reference's transmural coordinate and rule-based
fibre field by running the reference's own functions.

Run in the development container only (the reference tree is not present on
GPU machines); the output ``coord_golden.npz`` is committed next to this
script, and the tests read it (tests/test_oracle_coords.py).

Imported by file path (both modules need only numpy, scipy, torch, tqdm; no
package ``__init__`` runs):
  * hrdayapy/coordinates/functions/generate_transmural_coordinate.py:
    generate_transmural_coordinate (CPU device, tol 1e-6);
  * hrdayapy/fibres/functions/generate_fibre_direction.py:
    generate_fibre_direction (verbose off).
Inputs, stored too: a 48^3 synthetic BiV shell at 2 mm (13,100 wall voxels)
with its +1/-1/0 surface labels; for the fibres, that phi and an apicobasal
coordinate psi = (z - z_apex)/(z_base - z_apex), at spacing 2 mm and at an
anisotropic spacing (1.0, 1.5, 2.0) mm.
"""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np
import torch

REF = Path("/root/reference/hrdayapy")
HERE = Path(__file__).resolve().parent
OUT = HERE / "coord_golden.npz"
sys.path.insert(0, str(HERE.parent.parent))

from hrdayapy_amd import geometry as G  # noqa: E402


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    torch.set_num_threads(4)
    tc = _load("ref_transmural", REF / "coordinates/functions/generate_transmural_coordinate.py")
    fb = _load("ref_fibres", REF / "fibres/functions/generate_fibre_direction.py")
    geo = G.biv_shell(48, 2.0)
    S = geo["S"].astype(np.uint8)
    lab = G.biv_surface_labels(geo)
    phi = tc.generate_transmural_coordinate(S, lab, device="cpu", tol=1e-6, maxiter=10_000)
    z = np.nonzero(S.any(axis=(0, 1)))[0]
    kk = np.arange(S.shape[2], dtype=np.float32)[None, None, :] * np.ones(S.shape, np.float32)
    psi = np.where(S.astype(bool), (kk - z.min()) / (z.max() - z.min()), np.nan).astype(np.float32)
    f_iso = fb.generate_fibre_direction(S, phi, psi, np.array([2.0, 2.0, 2.0]), verbose=False)
    f_ani = fb.generate_fibre_direction(S, phi, psi, np.array([1.0, 1.5, 2.0]), verbose=False,
                                        alpha_endo_deg=50.0, alpha_epi_deg=-70.0)
    np.savez_compressed(OUT, S=S, surface_label=lab, phi=phi, psi=psi, fibre_iso=f_iso,
                        fibre_aniso=f_ani)
    print(f"wrote {OUT}: {int(S.sum())} voxels, phi in [{np.nanmin(phi):.3f}, {np.nanmax(phi):.3f}]")


if __name__ == "__main__":
    main()
