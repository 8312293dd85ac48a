"""Coarse torso grid: conductivities, the column-run layout of the 7-point
operator, the pinned reference cell and the body-surface cells.

Reference: hrdayapy/ecg/functions/build_torso_grid.py
  * sigma volume / tissue mask        :554-576 (active = label != 0)
  * 7-point harmonic-mean stiffness   :276-397 (sharp boundary)
  * pinned cell                       :616-623 (middle tissue voxel in flat order)
  * surface cells                     :469-488 (tissue voxels face-adjacent to air)
  * returned dict keys                :672-691
The operator coefficients are computed vectorised on the host once and held
in HBM in the column-run layout consumed by the HIP kernels (no CSR matrix is
formed); ``stiffness_csr`` builds the equivalent scipy matrix on request.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage as ndi


def _harm(a, b):
    s = a + b
    return np.where(s > 1e-30, 2.0 * a * b / np.where(s > 1e-30, s, 1.0), 0.0)


def _face_conductances(sigma, tissue, flat_of_cell, dx_mm):
    """FP64 conductance of the +x/+y/+z face of every stored cell:
    harm(sigma_c, sigma_nb)/h^2 between two active cells, 0 otherwise
    (build_torso_grid.py:325-390, sharp boundary)."""
    ni, nj, nk = sigma.shape
    sflat, act = sigma.ravel(), tissue.ravel()
    inv_h2 = 1.0 / (dx_mm * 1e-3) ** 2
    n_cells = len(flat_of_cell)
    ci, cj, kk = flat_of_cell // (nj * nk), (flat_of_cell // nk) % nj, flat_of_cell % nk

    def face(off_flat, valid):
        g = np.zeros(n_cells, dtype=np.float64)
        f = flat_of_cell[valid]
        nb = f + off_flat
        both = act[f] & act[nb]
        gv = np.zeros(len(f))
        gv[both] = _harm(sflat[f[both]], sflat[nb[both]]) * inv_h2
        g[valid] = gv
        return g

    return face(nj * nk, ci + 1 < ni), face(nk, cj + 1 < nj), face(1, kk + 1 < nk)


class TorsoGrid(dict):
    """dict with the reference's torso-grid keys plus the GPU layout arrays.

    ``grid["A_scipy"]`` (the unpinned FP64 stiffness as CSR, rows in active
    order) and ``grid["solve_fn"]`` (b -> phi for the pinned system,
    build_torso_grid.py:404-449) are built on first access: the CSR matrix of
    a full torso takes gigabytes of host memory that the GPU path never needs.
    """

    def __missing__(self, key):
        if key == "A_scipy":
            val = self.stiffness_csr()
        elif key == "solve_fn":
            val = self._solve_fn()
        else:
            raise KeyError(key)
        self[key] = val
        return val

    def _solve_fn(self, device="cuda", tol=1e-10):
        """solve(b) -> phi with b, phi in active-row order (numpy FP64).  As in the
        reference, b[pin_row] is set to 0 in place.  The solve is the multigrid
        PCG inside FP64 iterative refinement on the device."""
        import torch
        from .forward import PoissonSolver
        ps = PoissonSolver(self, device, precond="mg")
        cells = torch.from_numpy(self.cell_of_row).to(ps.dev)
        pin_row = int(self["pin_row"])

        def solve_fn(b):
            b[pin_row] = 0.0
            bc = torch.zeros(ps.n, dtype=torch.float64, device=ps.dev)
            bc[cells] = torch.as_tensor(np.asarray(b, dtype=np.float64), device=ps.dev)
            x = ps.solve_refined(bc, tol=tol)
            return x[cells].cpu().numpy()

        return solve_fn

    @classmethod
    def from_labels(cls, label_vol, sigma_map, dx_mm, origin_mm=(0.0, 0.0, 0.0), verbose=False):
        lab = np.asarray(label_vol)
        shape = lab.shape
        ni, nj, nk = shape
        sigma = np.zeros(shape, dtype=np.float64)
        for k, v in sigma_map.items():
            sigma[lab == k] = v
        tissue = lab != 0
        active_flat = np.flatnonzero(tissue.ravel()).astype(np.int64)
        n_act = len(active_flat)
        id_map = np.full(lab.size, -1, dtype=np.int32)
        id_map[active_flat] = np.arange(n_act, dtype=np.int32)
        # column runs over the active cells
        any_col = tissue.any(axis=2)
        k_first = np.where(any_col, tissue.argmax(axis=2), 0)
        k_last = np.where(any_col, nk - 1 - tissue[:, :, ::-1].argmax(axis=2), -1)
        col_len = np.where(any_col, k_last - k_first + 1, 0).astype(np.int64)
        col_off = np.zeros(ni * nj, dtype=np.int64)
        col_off[1:] = np.cumsum(col_len.ravel())[:-1]
        n_cells = int(col_len.sum())
        if n_cells >= 2 ** 31:
            raise ValueError("torso grid too large for 32-bit cell indices")
        # stored cell -> flat index
        cols = np.repeat(np.arange(ni * nj), col_len.ravel())
        dk = np.arange(n_cells) - col_off[cols]
        kk = k_first.ravel()[cols] + dk
        flat_of_cell = cols * nk + kk
        gx, gy, gz = _face_conductances(sigma, tissue, flat_of_cell, dx_mm)
        act = tissue.ravel()
        ci, cj = cols // nj, cols % nj
        # diag = sum of the six faces: + faces of this cell and + faces of the - neighbours
        cell_of_flat = np.full(lab.size, -1, dtype=np.int64)
        cell_of_flat[flat_of_cell] = np.arange(n_cells)
        diag = gx + gy + gz
        for g, off, valid in ((gx, nj * nk, ci > 0), (gy, nk, cj > 0), (gz, 1, kk > 0)):
            nb = cell_of_flat[flat_of_cell[valid] - off]
            ok = nb >= 0
            idx = np.nonzero(valid)[0][ok]
            diag[idx] += g[nb[ok]]
        hole = ~act[flat_of_cell]
        zero_row = (diag <= 0.0) & ~hole
        diag[hole | zero_row] = 1.0
        # pinned reference cell, build_torso_grid.py:621-623
        interior = np.flatnonzero(tissue.ravel())
        pin_flat = int(interior[len(interior) // 2])
        pin_row = int(id_map[pin_flat])
        pin_cell = int(cell_of_flat[pin_flat])
        diag[pin_cell] = 1.0
        # body surface, build_torso_grid.py:469-488
        air = ~tissue
        surf = tissue & ndi.binary_dilation(air, structure=ndi.generate_binary_structure(3, 1))
        surface_flat = np.flatnonzero(surf.ravel()).astype(np.int32)
        surface_rows = id_map[surface_flat]
        sidx = np.stack(np.unravel_index(surface_flat, shape), axis=1)
        surface_xyz = (np.asarray(origin_mm) + sidx * dx_mm).astype(np.float32)
        g = cls(shape=shape, dx_mm=float(dx_mm), origin_mm=np.asarray(origin_mm, dtype=np.float64),
                sigma_vol=sigma, active_mask=tissue, tissue_mask=tissue,
                active_ids=active_flat.astype(np.int32), active_id_map=id_map,
                surface_flat=surface_flat, surface_rows=surface_rows,
                surface_xyz=surface_xyz, pin_row=pin_row, use_sbm=False)
        g.layout = dict(col_k0=k_first.ravel().astype(np.int32),
                        col_len=col_len.ravel().astype(np.int32),
                        col_off=col_off.astype(np.int32), gx=gx.astype(np.float32),
                        gy=gy.astype(np.float32), gz=gz.astype(np.float32),
                        diag=diag.astype(np.float32), pin=pin_cell, n_cells=n_cells)
        g.cell_of_row = cell_of_flat[active_flat].astype(np.int64)
        g.flat_of_cell = flat_of_cell
        if verbose:
            print(f"  torso {shape} @ {dx_mm} mm: {n_act:,} active, {n_cells:,} stored cells, "
                  f"{len(surface_flat):,} surface, pin row {pin_row}")
        return g

    def stiffness_csr(self):
        """The (unpinned) FP64 stiffness in row numbering, for small grids."""
        import scipy.sparse as sp
        L = self.layout
        cor = self.cell_of_row
        n = len(cor)
        row_of_cell = np.full(L["n_cells"], -1, dtype=np.int64)
        row_of_cell[cor] = np.arange(n)
        ni, nj, nk = self["shape"]
        flat = self.flat_of_cell
        cell_of_flat = np.full(ni * nj * nk, -1, dtype=np.int64)
        cell_of_flat[flat] = np.arange(L["n_cells"])
        rows, cols, vals = [], [], []
        g64 = _face_conductances(np.asarray(self["sigma_vol"], dtype=np.float64),
                                 np.asarray(self["tissue_mask"]), flat, self["dx_mm"])
        for g, off in zip(g64, (nj * nk, nk, 1)):
            c = np.nonzero(g > 0)[0]
            nb = cell_of_flat[flat[c] + off]
            a, b = row_of_cell[c], row_of_cell[nb]
            w = g[c]
            rows += [a, b, a, b]
            cols += [b, a, a, b]
            vals += [-w, -w, w, w]
        A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(n, n)).tocsr()
        return A
