"""A TorsoGrid assembled on the host from the oracle's stiffness matrix (test
infrastructure): the column-run layout the solver kernels consume, with
each face conductance read off the oracle's FP64 matrix.  CPU tests use it
where the product builds the grid on the device; GPU tests compare the
device build with it."""
import numpy as np

from hrdayapy_amd.ecg.functions.torso import TorsoGrid
from oracle import ecg as OE


def host_grid(lab, smap, dx_mm, origin_mm=(0.0, 0.0, 0.0), psi_vol=None):
    lab = np.asarray(lab)
    shape = lab.shape
    ni, nj, nk = shape
    sigma = np.zeros(shape)
    for k, v in smap.items():
        sigma[lab == k] = v
    tissue = lab != 0
    if psi_vol is not None:
        active = psi_vol > 0
        sigma = OE.extend_conductivity(sigma, tissue, active)
    else:
        active = tissue
    A, active_ids, id_map = OE.assemble_stiffness(sigma, dx_mm * 1e-3, active, psi_vol)
    A = A.tocsr()
    any_col = active.any(axis=2)
    k0 = np.where(any_col, active.argmax(axis=2), 0).ravel()
    k1 = np.where(any_col, nk - 1 - active[:, :, ::-1].argmax(axis=2), -1).ravel()
    col_len = np.where(any_col.ravel(), k1 - k0 + 1, 0).astype(np.int64)
    col_off = np.zeros(ni * nj, dtype=np.int64)
    col_off[1:] = np.cumsum(col_len)[:-1]
    n_cells = int(col_len.sum())
    cols = np.repeat(np.arange(ni * nj), col_len)
    flat_of_cell = cols * nk + k0[cols] + (np.arange(n_cells) - col_off[cols])
    cell_of_flat = np.full(lab.size, -1, dtype=np.int64)
    cell_of_flat[flat_of_cell] = np.arange(n_cells)
    act = active.ravel()
    ci, cj, ck = flat_of_cell // (nj * nk), (flat_of_cell // nk) % nj, flat_of_cell % nk
    g64 = np.zeros((4, n_cells))
    for q, (off, ok) in enumerate(((nj * nk, ci + 1 < ni), (nk, cj + 1 < nj), (1, ck + 1 < nk))):
        c = np.nonzero(ok)[0]
        f, nb = flat_of_cell[c], flat_of_cell[c] + off
        both = act[f] & act[nb]
        c, f, nb = c[both], f[both], nb[both]
        g64[q, c] = -np.asarray(A[id_map[f], id_map[nb]]).ravel()
    rows = id_map[flat_of_cell]
    g64[3] = 1.0
    on = rows >= 0
    g64[3, on] = A.diagonal()[rows[on]]
    g64[3, ~(g64[3] > 0)] = 1.0
    pin_row = OE.pin_row_of(tissue, id_map)
    pin_cell = int(cell_of_flat[active_ids[pin_row]])
    diag32 = g64[3].astype(np.float32)
    diag32[pin_cell] = 1.0  # the solver's pinned row; g64 keeps the unpinned matrix
    sf = OE.surface_flat(lab, tissue).astype(np.int32)
    sidx = np.stack(np.unravel_index(sf, shape), axis=1)
    g = TorsoGrid(shape=shape, dx_mm=float(dx_mm), origin_mm=np.asarray(origin_mm, np.float64),
                  sigma_vol=sigma, active_mask=active, tissue_mask=tissue,
                  active_ids=active_ids.astype(np.int32), active_id_map=id_map.astype(np.int32),
                  surface_flat=sf, surface_rows=id_map[sf].astype(np.int32),
                  surface_xyz=(np.asarray(origin_mm) + sidx * dx_mm).astype(np.float32),
                  pin_row=pin_row, use_sbm=psi_vol is not None)
    if psi_vol is not None:
        g["psi_vol"] = np.asarray(psi_vol, dtype=np.float64)
    g.layout = dict(col_k0=k0.astype(np.int32), col_len=col_len.astype(np.int32),
                    col_off=col_off.astype(np.int32), gx=g64[0].astype(np.float32),
                    gy=g64[1].astype(np.float32), gz=g64[2].astype(np.float32),
                    diag=diag32, pin=pin_cell, n_cells=n_cells)
    g.g64 = g64
    g.cell_of_row = cell_of_flat[active_ids].astype(np.int64)
    g.flat_of_cell = flat_of_cell
    g.psi_cell = None
    return g
