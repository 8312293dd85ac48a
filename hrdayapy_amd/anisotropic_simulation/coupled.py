"""compute_anisotropic_coupled / load_anisotropic_coupled.

Drop-in for hrdayapy/anisotropic_simulation/coupled.py:35-153: same
arguments and defaults, forwarded to the GPU solver
(purkinje_myocardium_anisotropic_solver).  ``dtype`` (default float32)
selects the arithmetic of the HIP kernels.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from .._spacing_resolve import resolve_voxel_size
from .functions import purkinje_myocardium_anisotropic_solver


def compute_anisotropic_coupled(nodes, elements, act_times, S, Z, phi, fibre_dir, *, anatomy=None,
                                stim_protocol=None, ectopic_region=None, stim_regions=None,
                                voxel_size=0.4, dx_p=0.2, sigma_P=0.45, sigma_l=3.0 * 1.31e-4,
                                sigma_t=1.31e-4, Cm=0.01, A_M=140.0, R_P=0.015, dt=0.1, T=400.0,
                                theta=0.5, stim_len_mm=2.0, c_pmj=0.05, n_pmj=1, cg_tol=1e-5,
                                cg_max_iter=500, n_frames=100, frame_save_dt=None,
                                anim_tmp_dir=None, device=None, phi_endo_max=0.35,
                                phi_epi_min=0.65, save_path=None, vm_save_dt=None,
                                vm_save_path=None, vm_tmp_dir=None, dtype=torch.float32,
                                scheme="cn", cn_solver="bicgstab"):
    voxel_size = resolve_voxel_size(voxel_size, anatomy, param_name="voxel_size")
    return purkinje_myocardium_anisotropic_solver(
        nodes=nodes, elements=elements, activation_times=act_times, S=S, Z=Z, phi=phi,
        fibre_dir=fibre_dir, voxel_size=voxel_size, dx_p=dx_p, sigma_P=sigma_P, sigma_l=sigma_l,
        sigma_t=sigma_t, Cm=Cm, A_M=A_M, R_P=R_P, dt=dt, T=T, theta=theta,
        stim_protocol=stim_protocol, ectopic_region=ectopic_region, stim_regions=stim_regions,
        stim_len_mm=stim_len_mm, c_pmj=c_pmj, n_pmj=n_pmj, cg_tol=cg_tol,
        cg_max_iter=cg_max_iter, n_frames=n_frames, frame_save_dt=frame_save_dt,
        anim_tmp_dir=anim_tmp_dir, device=device, out_npz=save_path,
        phi_endo_max=phi_endo_max, phi_epi_min=phi_epi_min, vm_save_dt=vm_save_dt,
        out_vm_npz=vm_save_path, vm_tmp_dir=vm_tmp_dir, dtype=dtype,
        scheme=scheme, cn_solver=cn_solver)


def load_anisotropic_coupled(path):
    path = Path(path).with_suffix(".npz")
    if not path.exists():
        raise FileNotFoundError(f"No saved anisotropic coupled simulation at {path}")
    z = np.load(path)
    return {k: z[k] for k in z.files}
