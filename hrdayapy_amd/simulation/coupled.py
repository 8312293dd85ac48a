"""compute_coupled / load_coupled (isotropic).

Drop-in for hrdayapy/simulation/coupled.py:33-131: same arguments and
defaults (including the settable A_P = 133.33 and A_M = 0.14 of the isotropic
path).  The isotropic 6-point operator w = sigma_M*h is the 19-point kernel
with sigma_l = sigma_t = sigma_M and fibre-free tensor (the edge-diagonal
terms are exactly zero), so it runs on the same HIP kernels.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from .._spacing_resolve import resolve_voxel_size
from ..anisotropic_simulation.functions.purkinje_myocardium_anisotropic_solver import (
    purkinje_myocardium_anisotropic_solver)


def compute_coupled(nodes, elements, act_times, S, Z, phi, *, anatomy=None, stim_protocol=None,
                    ectopic_region=None, voxel_size=0.4, dx_p=0.2, sigma_P=0.45, sigma_M=1.31e-4,
                    Cm=0.01, A_P=133.33, A_M=0.14, R_P=0.015, dt=0.1, T=400.0, theta=0.5,
                    stim_len_mm=2.0, c_pmj=0.05, n_pmj=1, cg_tol=1e-5, cg_max_iter=500,
                    n_frames=100, device=None, phi_endo_max=0.35, phi_epi_min=0.65,
                    save_path=None, vm_save_dt=None, vm_save_path=None, vm_tmp_dir=None,
                    dtype=torch.float32, scheme="cn", cn_solver="bicgstab"):
    voxel_size = resolve_voxel_size(voxel_size, anatomy, param_name="voxel_size")
    return purkinje_myocardium_anisotropic_solver(
        nodes=nodes, elements=elements, activation_times=act_times, S=S, Z=Z, phi=phi,
        fibre_dir=None, voxel_size=voxel_size, dx_p=dx_p, sigma_P=sigma_P, sigma_l=sigma_M,
        sigma_t=sigma_M, Cm=Cm, A_P=A_P, A_M=A_M, R_P=R_P, dt=dt, T=T, theta=theta,
        stim_protocol=stim_protocol, ectopic_region=ectopic_region, stim_len_mm=stim_len_mm,
        c_pmj=c_pmj, n_pmj=n_pmj, cg_tol=cg_tol, cg_max_iter=cg_max_iter, n_frames=n_frames,
        device=device, out_npz=save_path, phi_endo_max=phi_endo_max, phi_epi_min=phi_epi_min,
        vm_save_dt=vm_save_dt, out_vm_npz=vm_save_path, vm_tmp_dir=vm_tmp_dir, dtype=dtype,
        scheme=scheme, cn_solver=cn_solver, _sigma_M=sigma_M)


def load_coupled(path):
    path = Path(path).with_suffix(".npz")
    if not path.exists():
        raise FileNotFoundError(f"No saved coupled simulation at {path}")
    z = np.load(path)
    return {k: z[k] for k in z.files}
