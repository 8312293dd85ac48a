"""compute_bspm / load_bspm on the GPU.

Drop-in for hrdayapy/ecg/bspm.py:33-145 and run_bspm_loop
(hrdayapy/ecg/functions/compute_bspm.py:445-670): same arguments, same
(N_surf, n_frames) mV result and the same bspm.npz keys.  Per frame the Vm
snapshot is uploaded into the heart's brick layout, the source is built on
the device and the torso potential is solved by warm-started Jacobi PCG with
the reference's stopping rule (cg_metric 'rms', 'p99' or 'linf' of
z = M^-1 r below cg_tol_mV, tested on the device after every iteration), or
-- with ``refine_tol`` -- by FP32 PCG inside FP64 iterative refinement to a
relative residual.
"""
from __future__ import annotations

import time
from pathlib import Path

import numpy as np
import torch

from .._spacing_resolve import resolve_voxel_size
from ..layout import BrickLayout
from .functions.forward import HeartSource, PoissonSolver

DEFAULT_SIGMA_I = 0.2625
DEFAULT_H_FINE_MM = 0.4


def compute_bspm(vm_snapshots_path, grid, registration, *, anatomy=None, sigma_i=DEFAULT_SIGMA_I,
                 h_fine_mm=DEFAULT_H_FINE_MM, cg_tol_mV=0.01, cg_max_iter=500, cg_metric="rms",
                 t_window=None, verbose=True, save_path=None, full_field_frames=None,
                 full_field_save_path=None, refine_tol=None, device="cuda"):
    if save_path is None:
        raise ValueError("compute_bspm requires save_path (run_bspm_loop writes results "
                         "incrementally as it solves).")
    if cg_metric not in ("rms", "p99", "linf"):
        raise ValueError(f"metric must be one of 'rms', 'p99', 'linf', got {cg_metric!r}")
    h_fine_mm = resolve_voxel_size(h_fine_mm, anatomy, param_name="h_fine_mm")
    dev = torch.device(device)
    snap = np.load(str(Path(vm_snapshots_path)))
    Vm_all = snap["Vm"].astype(np.float32)
    coords = snap["coords_mm"].astype(np.float64)
    vox = snap["vox_idx"].astype(np.int64)
    time_ms = snap["time_vm"].astype(np.float64)
    if t_window is not None:
        keep = (time_ms >= t_window[0]) & (time_ms <= t_window[1])
        if not keep.any():
            raise ValueError(f"t_window={t_window} selected 0 frames out of "
                             f"[{time_ms.min():.1f}, {time_ms.max():.1f}] ms available.")
        Vm_all, time_ms = Vm_all[keep], time_ms[keep]
    n_frames, N_myo = Vm_all.shape
    lo = vox.min(axis=0)
    S = np.zeros(tuple((vox.max(axis=0) - lo + 1).tolist()), dtype=bool)
    S[tuple((vox - lo).T)] = True
    lay = BrickLayout(S, (8, 8, 4))
    # the layout numbers voxels in raster (argwhere) order; snapshots written by
    # the solver use the same order, other producers are mapped explicitly
    if np.array_equal(lay.vox_idx, vox - lo):
        slot_of_col, coords_lay = lay.slot_of_node, coords
    else:
        lin_snap = np.ravel_multi_index(tuple((vox - lo).T), S.shape)
        pos = np.searchsorted(np.ravel_multi_index(tuple(lay.vox_idx.T), S.shape), lin_snap)
        slot_of_col = lay.slot_of_node[pos]
        coords_lay = np.empty_like(coords)
        coords_lay[pos] = coords
    src = HeartSource(lay, coords_lay, grid, registration, sigma_i, h_fine_mm, dev)
    ps = PoissonSolver(grid, dev)
    surf_cells = torch.from_numpy(grid.cell_of_row[grid["surface_rows"]]).to(dev)
    slot_t = torch.from_numpy(slot_of_col.astype(np.int64)).to(dev)
    Vs = torch.zeros(lay.n_slots, dtype=torch.float32, device=dev)
    b = torch.zeros(ps.n, dtype=torch.float64, device=dev)
    x = torch.zeros(ps.n, dtype=torch.float64, device=dev)
    bspm = np.zeros((len(grid["surface_rows"]), n_frames), dtype=np.float32)
    ff = sorted(set(int(f) for f in full_field_frames)) if full_field_frames is not None else []
    if ff and full_field_save_path is None:
        raise ValueError("full_field_frames given but full_field_save_path is None")
    full = {}
    iters = []
    t0 = time.perf_counter()
    for f in range(n_frames):
        Vs[slot_t] = torch.from_numpy(Vm_all[f]).to(dev)
        src(Vs, b, ps.ws)
        if refine_tol is not None:
            ps.solve_refined(b, x, tol=refine_tol)
        else:
            ps.solve_metric(b, x, tol_V=cg_tol_mV * 1e-3, max_iter=cg_max_iter,
                            metric=cg_metric)
        iters.append(ps.last_stats.get("inner", 0))
        bspm[:, f] = (x[surf_cells] * 1e3).float().cpu().numpy()
        if f in ff:
            full[f] = (x[torch.from_numpy(grid.cell_of_row).to(dev)] * 1e3).float().cpu().numpy()
    wall = time.perf_counter() - t0
    if verbose:
        print(f"  Done in {wall:.1f} s ({wall / max(1, n_frames):.3f} s/frame) "
              f"avg CG iters={np.mean(iters):.1f}")
    if ff:
        np.savez(str(full_field_save_path), phi_full=np.stack([full[f] for f in ff], axis=1),
                 frame_indices=np.array(ff, np.int64), time_ms=time_ms[ff],
                 active_ids=np.asarray(grid["active_ids"], np.int32),
                 shape=np.array(grid["shape"], np.int32), dx_mm=np.float64(grid["dx_mm"]),
                 origin_mm=np.asarray(grid["origin_mm"], np.float64))
    np.savez(str(save_path), surface_xyz=np.asarray(grid["surface_xyz"], np.float32),
             surface_flat=np.asarray(grid["surface_flat"], np.int32),
             dx_mm=np.float64(grid["dx_mm"]), bspm_signal=bspm, time_ms=time_ms,
             reg_R=np.asarray(registration["R"], np.float64),
             reg_t=np.asarray(registration["t"], np.float64),
             reg_scale=np.float64(registration["scale"]),
             reg_dice=np.float64(float(registration.get("dice", np.nan))),
             sigma_i=np.float64(sigma_i), h_fine_mm=np.float64(h_fine_mm))
    return bspm


def load_bspm(path):
    p = Path(path).with_suffix(".npz")
    if not p.exists():
        raise FileNotFoundError(f"No saved BSPM at {p}")
    z = np.load(str(p))
    return {k: z[k] for k in z.files}
