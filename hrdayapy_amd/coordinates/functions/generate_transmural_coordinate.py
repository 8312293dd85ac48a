"""Transmural coordinate phi in [0, 1] (0 = endocardium, 1 = epicardium) by
a Laplace solve on the GPU.

Drop-in for hrdayapy.coordinates.functions.generate_transmural_coordinate
(reference file generate_transmural_coordinate.py:190-286): same arguments,
the same .npy cache behaviour and the same result (NaN outside the mask,
Dirichlet values re-applied, clipped to [0, 1]).  The system is the
reference's (:85-183: identity rows where surface_label is +1 or -1, the
6-neighbour graph Laplacian with Neumann faces elsewhere) and so is the
solver's rule (plain CG in FP32 from the distance-ratio start,
||r||/||r0|| < tol, :35-78); the matrix is never formed: the product is the
brick-layout stencil of the HIP kernels (hp_laplace_cg_*), with the CG's
scalars kept on the device.  The distance transforms of the start are
scipy's, on the host, as in the reference.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
from scipy.ndimage import distance_transform_edt

from ..laplace import LaplaceSolver


def generate_transmural_coordinate(S, surface_label, device="cuda", tol: float = 1e-6,
                                   maxiter: int = 10_000, save_path=None, verbose=True):
    if save_path is not None:
        save_path = Path(save_path).with_suffix(".npy")
        if save_path.exists():
            if verbose:
                print(f"  [transmural] Loading cached phi from {save_path}")
            return np.load(save_path)
    S = np.asarray(S)
    lab = np.asarray(surface_label)
    if S.shape != lab.shape:
        raise ValueError(f"S {S.shape} and surface_label {lab.shape} differ in shape")
    Sb = S == 1
    vox = np.argwhere(Sb)
    lv = lab[tuple(vox.T)]
    # distance-ratio start (reference :233-245)
    d_endo = distance_transform_edt(lab != -1)
    d_epi = distance_transform_edt(lab != 1)
    den = d_endo + d_epi
    start = np.where(den > 0, d_endo / np.where(den > 0, den, 1.0), 0.5).astype(np.float32)
    start[lab == -1] = 0.0
    start[lab == 1] = 1.0
    x0 = start[tuple(vox.T)]
    solver = LaplaceSolver(Sb, device)
    x = solver.solve(lv != 0, x0, tol=tol, maxiter=maxiter)
    st = solver.last
    if verbose:
        if st["converged"]:
            print(f"  [transmural] CG converged in {st['iterations']} / {maxiter} iterations  "
                  f"(rel. residual {st['rel_res']:.2e})")
        else:
            print(f"  [transmural] WARNING: CG did NOT converge in {maxiter} iterations  "
                  f"(rel. residual {st['rel_res']:.2e} > tol {tol:.2e})")
    x[lv == -1] = 0.0
    x[lv == 1] = 1.0
    phi = np.full(S.shape, np.nan, dtype=np.float32)
    phi[tuple(vox.T)] = x
    np.clip(phi, 0.0, 1.0, out=phi)
    if save_path is not None:
        save_path.parent.mkdir(parents=True, exist_ok=True)
        np.save(save_path, phi)
    return phi
