"""Coarse torso grid built on the device.

Reference: hrdayapy/ecg/functions/build_torso_grid.py
  * _resample_labels (nearest-neighbour zoom)     :260-269
  * sigma volume and tissue mask                  :554-576
  * compute_smoothed_domain (SBM psi)             :130-186
  * extend_conductivity                           :189-206
  * _assemble_stiffness (sharp and SBM)           :276-397
  * pinned cell (middle tissue voxel, flat order) :616-623
  * _find_surface_ids                             :469-488
Every step is a HIP kernel of csrc/torso_build.hip: the label lookup, the
feature transforms behind the Euclidean distances, the tanh profile and
the linear zoom, the per-column runs and their scans, the row numbering,
the surface and pin selection, and the FP64 face conductances (stored FP32
for the solver, FP64 for the exact matrix).  The host only sizes the
outputs from the counts the device returns.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ... import _native as nat

FT_STACK_ENTRIES = 1 << 28  # 1 GiB of line stacks per feature-transform pass


def _i32_host(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


def _dbl3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def zoom_shape(shape, spacing, dx_mm):
    """Output shape of a zoom by spacing / dx (round half to even, as Python's round)."""
    z = np.asarray(spacing, dtype=np.float64) / float(dx_mm)
    return tuple(int(round(n * f)) for n, f in zip(shape, z))


def resample_labels(labels, spacing, dx_mm, device="cuda"):
    """Nearest-neighbour resample of a label volume to dx_mm (int32 on the device)."""
    dev = torch.device(device)
    lab = torch.as_tensor(np.ascontiguousarray(labels, dtype=np.int32)).to(dev)
    shp = zoom_shape(lab.shape, spacing, dx_mm)
    out = torch.empty(shp, dtype=torch.int32, device=dev)
    nat.call("hp_zoom_nearest_i32", lab.data_ptr(), *lab.shape, _i32_host(shp), out.data_ptr(),
             nat.stream_handle(dev))
    return out


def feature_transform(mask_u8, feature_value, spacing, device="cuda"):
    """(3, n0, n1, n2) int32: coordinates of the nearest voxel with mask ==
    feature_value (-1 if there is none)."""
    dev = torch.device(device)
    m = mask_u8.to(dev).contiguous()
    n0, n1, n2 = (int(x) for x in m.shape)
    N = n0 * n1 * n2
    ft = torch.empty((3, n0, n1, n2), dtype=torch.int32, device=dev)
    tmp = torch.empty_like(ft)
    lmax = max(n0, n1, n2)
    stack_len = int(min(max(N, lmax), FT_STACK_ENTRIES))
    stack = torch.empty(stack_len, dtype=torch.int32, device=dev)
    nat.call("hp_feature_transform", m.data_ptr(), int(feature_value), n0, n1, n2,
             _dbl3(spacing), ft.data_ptr(), tmp.data_ptr(), stack.data_ptr(), stack_len,
             nat.stream_handle(dev))
    return ft


def smoothed_domain(fine_labels, spacing_fine, dx_mm, coarse_shape, interface_width_mm=3.0,
                    psi_cutoff=1e-3, device="cuda"):
    """SBM domain function on the coarse grid (FP64 on the device) and its
    positive set (uint8): compute_smoothed_domain (:130-186)."""
    dev = torch.device(device)
    lab = torch.as_tensor(np.ascontiguousarray(fine_labels, dtype=np.int32)).to(dev)
    inside = (lab != 0).to(torch.uint8)
    n_in = int(inside.sum())
    if n_in == 0 or n_in == inside.numel():
        raise ValueError("the fine segmentation needs both tissue and air voxels")
    del lab
    ft_out = feature_transform(inside, 0, spacing_fine, dev)   # nearest outside voxel
    ft_in = feature_transform(inside, 1, spacing_fine, dev)    # nearest inside voxel
    psi_f = torch.empty(inside.shape, dtype=torch.float64, device=dev)
    nat.call("hp_sbm_psi_f64", inside.data_ptr(), ft_out.data_ptr(), ft_in.data_ptr(),
             *inside.shape, _dbl3(spacing_fine), float(interface_width_mm), psi_f.data_ptr(),
             nat.stream_handle(dev))
    del ft_out, ft_in
    oshape = zoom_shape(inside.shape, spacing_fine, dx_mm)
    psi = torch.empty(tuple(coarse_shape), dtype=torch.float64, device=dev)
    pos = torch.empty(tuple(coarse_shape), dtype=torch.uint8, device=dev)
    nat.call("hp_zoom_linear_f64", psi_f.data_ptr(), *inside.shape, _i32_host(oshape),
             _i32_host(coarse_shape), float(psi_cutoff), psi.data_ptr(), pos.data_ptr(),
             nat.stream_handle(dev))
    return psi, pos


def label_sigma(labels, sigma_map, device="cuda"):
    """(sigma FP64, tissue uint8) of a label volume on the device."""
    dev = torch.device(device)
    lab = labels if isinstance(labels, torch.Tensor) else torch.as_tensor(
        np.ascontiguousarray(labels, dtype=np.int32))
    lab = lab.to(dev, torch.int32).contiguous()
    keys = torch.tensor([int(k) for k in sigma_map], dtype=torch.int32, device=dev)
    vals = torch.tensor([float(v) for v in sigma_map.values()], dtype=torch.float64, device=dev)
    sigma = torch.empty(lab.shape, dtype=torch.float64, device=dev)
    tissue = torch.empty(lab.shape, dtype=torch.uint8, device=dev)
    nat.call("hp_label_sigma_f64", lab.data_ptr(), lab.numel(), keys.data_ptr(), vals.data_ptr(),
             len(keys), sigma.data_ptr(), tissue.data_ptr(), nat.stream_handle(dev))
    return sigma, tissue


def extend_conductivity(sigma, tissue, active, device="cuda"):
    """sigma with the band cells (active, not tissue) taking their nearest
    tissue cell's value (:189-206)."""
    dev = torch.device(device)
    ft = feature_transform(tissue, 1, (1.0, 1.0, 1.0), dev)
    out = torch.empty_like(sigma)
    nat.call("hp_extend_sigma_f64", ft.data_ptr(), tissue.data_ptr(), active.data_ptr(),
             sigma.data_ptr(), *sigma.shape, out.data_ptr(), nat.stream_handle(dev))
    return out


def _column_list(flag, dev, want_id_map=False, k0=None, col_off=None):
    """Flat indices of the set flags in flat order (int32), optionally the
    flat -> row map and row -> stored cell."""
    ni, nj, nk = (int(x) for x in flag.shape)
    s = nat.stream_handle(dev)
    cnt = torch.empty(ni * nj, dtype=torch.int32, device=dev)
    nat.call("hp_column_stats", flag.data_ptr(), ni, nj, nk, None, None, cnt.data_ptr(), s)
    base = torch.empty(ni * nj + 1, dtype=torch.int64, device=dev)
    nat.call("hp_scan_i32", cnt.data_ptr(), ni * nj, base.data_ptr(), s)
    n = int(base[-1])
    lst = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    id_map = torch.empty(ni * nj * nk, dtype=torch.int32, device=dev) if want_id_map else None
    cor = torch.empty(max(n, 1), dtype=torch.int64, device=dev) if col_off is not None else None
    nat.call("hp_column_emit", flag.data_ptr(), ni, nj, nk, base.data_ptr(), lst.data_ptr(),
             nat.ptr(id_map), nat.ptr(k0), nat.ptr(col_off), nat.ptr(cor), s)
    return lst[:n], id_map, (None if cor is None else cor[:n]), base


def assemble(sigma, tissue, active, dx_mm, psi=None, device="cuda"):
    """Column-run stiffness and numbering of a coarse grid held on the device.
    Returns a dict of device tensors and scalars (see TorsoGrid.from_labels)."""
    dev = torch.device(device)
    s = nat.stream_handle(dev)
    ni, nj, nk = (int(x) for x in sigma.shape)
    n_cols = ni * nj
    if ni * nj * nk >= 2 ** 31:
        raise ValueError("torso grid too large for 32-bit flat indices")
    # runs over the active cells
    k0 = torch.empty(n_cols, dtype=torch.int32, device=dev)
    run = torch.empty(n_cols, dtype=torch.int32, device=dev)
    nat.call("hp_column_stats", active.data_ptr(), ni, nj, nk, k0.data_ptr(), run.data_ptr(),
             None, s)
    col_off = torch.empty(n_cols + 1, dtype=torch.int64, device=dev)
    nat.call("hp_scan_i32", run.data_ptr(), n_cols, col_off.data_ptr(), s)
    n_cells = int(col_off[-1])
    if n_cells >= 2 ** 31:
        raise ValueError("torso grid too large for 32-bit cell indices")
    if n_cells == 0:
        raise ValueError("the torso grid has no active cell")
    # rows: active cells in flat order
    active_ids, id_map, cell_of_row, _ = _column_list(active, dev, True, k0, col_off)
    # pinned cell: the middle tissue voxel in flat order
    tissue_ids_cnt = torch.empty(n_cols, dtype=torch.int32, device=dev)
    nat.call("hp_column_stats", tissue.data_ptr(), ni, nj, nk, None, None,
             tissue_ids_cnt.data_ptr(), s)
    tscan = torch.empty(n_cols + 1, dtype=torch.int64, device=dev)
    nat.call("hp_scan_i32", tissue_ids_cnt.data_ptr(), n_cols, tscan.data_ptr(), s)
    n_tissue = int(tscan[-1])
    if n_tissue == 0:
        raise ValueError("the torso grid has no tissue voxel")
    pin_out = torch.empty(1, dtype=torch.int64, device=dev)
    nat.call("hp_select_rank", tissue.data_ptr(), tscan.data_ptr(), n_cols, nk, n_tissue // 2,
             pin_out.data_ptr(), s)
    pin_flat = int(pin_out[0])
    # faces and diagonal
    f32 = dict(dtype=torch.float32, device=dev)
    gx, gy, gz, diag = (torch.empty(n_cells, **f32) for _ in range(4))
    g64 = torch.empty((4, n_cells), dtype=torch.float64, device=dev)
    bad = torch.empty(1, dtype=torch.int64, device=dev)
    nat.call("hp_torso_faces_f64", sigma.data_ptr(), nat.ptr(psi), active.data_ptr(), ni, nj, nk,
             float(dx_mm) * 1e-3, k0.data_ptr(), run.data_ptr(), col_off.data_ptr(), pin_flat,
             gx.data_ptr(), gy.data_ptr(), gz.data_ptr(), diag.data_ptr(), g64.data_ptr(),
             n_cells, bad.data_ptr(), s)
    if int(bad[0]) != 0:
        raise ValueError(f"{int(bad[0])} faces conduct into cells outside the system "
                         "(a conductivity on air?)")
    # body surface
    surf = torch.empty_like(tissue)
    nat.call("hp_surface_flags", tissue.data_ptr(), ni, nj, nk, surf.data_ptr(), s)
    surface_flat, _, _, _ = _column_list(surf, dev)
    return dict(k0=k0, run=run, col_off=col_off[:-1], n_cells=n_cells, active_ids=active_ids,
                id_map=id_map, cell_of_row=cell_of_row, pin_flat=pin_flat, gx=gx, gy=gy, gz=gz,
                diag=diag, g64=g64, surface_flat=surface_flat)
