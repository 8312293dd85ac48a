"""Coarse torso grid: conductivities, the column-run layout of the 7-point
operator, the pinned reference cell and the body-surface cells.

Reference: hrdayapy/ecg/functions/build_torso_grid.py
  * sigma volume / tissue mask        :554-576 (active = label != 0)
  * 7-point harmonic-mean stiffness   :276-397 (sharp and smoothed boundary)
  * pinned cell                       :616-623 (middle tissue voxel in flat order)
  * surface cells                     :469-488 (tissue voxels face-adjacent to air)
  * returned dict keys                :672-691
The grid is built by HIP kernels (ecg/functions/torso_build.py) and held in
HBM in the column-run layout consumed by the solver kernels (no CSR matrix is
formed); ``stiffness_csr`` builds the equivalent scipy matrix on request from
the device's FP64 face conductances.
"""
from __future__ import annotations

import numpy as np


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
    def from_labels(cls, label_vol, sigma_map, dx_mm, origin_mm=(0.0, 0.0, 0.0), verbose=False,
                    *, psi_vol=None, device="cuda"):
        """Build the grid of a coarse label volume on the device (the kernels
        of ecg/functions/torso_build.py).  psi_vol (coarse, FP64; optional)
        selects the smoothed-boundary system: active = psi > 0, sigma extended
        through the band, conductances scaled by the harmonic mean of psi."""
        import torch
        from . import torso_build as TB
        dev = torch.device(device)
        if isinstance(label_vol, torch.Tensor):
            lab = label_vol.to(dev, torch.int32).contiguous()
        else:
            lab = torch.as_tensor(np.ascontiguousarray(label_vol, dtype=np.int32)).to(dev)
        shape = tuple(int(x) for x in lab.shape)
        ni, nj, nk = shape
        sigma, tissue = TB.label_sigma(lab, sigma_map, dev)
        psi = None
        if psi_vol is not None:
            psi = torch.as_tensor(psi_vol, dtype=torch.float64).to(dev).contiguous()
            if tuple(psi.shape) != shape:
                raise ValueError(f"psi_vol shape {tuple(psi.shape)} != label shape {shape}")
            active = (psi > 0.0).to(torch.uint8)
            sigma = TB.extend_conductivity(sigma, tissue, active, dev)
        else:
            active = tissue
        out = TB.assemble(sigma, tissue, active, dx_mm, psi, dev)
        h = lambda t: t.cpu().numpy()  # noqa: E731
        n_cells = out["n_cells"]
        col_k0, col_len, col_off = h(out["k0"]), h(out["run"]), h(out["col_off"])
        cols = np.repeat(np.arange(ni * nj, dtype=np.int64), col_len.astype(np.int64))
        flat_of_cell = cols * nk + col_k0[cols] + (np.arange(n_cells) - col_off[cols])
        active_ids = h(out["active_ids"])
        id_map = h(out["id_map"])
        pin_flat = out["pin_flat"]
        pin_row = int(id_map[pin_flat])
        c = pin_flat // nk
        pin_cell = int(col_off[c] + (pin_flat % nk) - col_k0[c])
        surface_flat = h(out["surface_flat"])
        surface_rows = id_map[surface_flat]
        sidx = np.stack(np.unravel_index(surface_flat, shape), axis=1)
        surface_xyz = (np.asarray(origin_mm) + sidx * dx_mm).astype(np.float32)
        tissue_h = h(tissue).astype(bool)
        g = cls(shape=shape, dx_mm=float(dx_mm), origin_mm=np.asarray(origin_mm, dtype=np.float64),
                sigma_vol=h(sigma), active_mask=h(active).astype(bool), tissue_mask=tissue_h,
                active_ids=active_ids, active_id_map=id_map, surface_flat=surface_flat,
                surface_rows=surface_rows, surface_xyz=surface_xyz, pin_row=pin_row,
                use_sbm=psi is not None)
        if psi is not None:
            g["psi_vol"] = h(psi)
        g.layout = dict(col_k0=col_k0.astype(np.int32), col_len=col_len.astype(np.int32),
                        col_off=col_off.astype(np.int32), gx=h(out["gx"]), gy=h(out["gy"]),
                        gz=h(out["gz"]), diag=h(out["diag"]), pin=pin_cell, n_cells=n_cells)
        g.g64 = out["g64"]
        g.cell_of_row = h(out["cell_of_row"]).astype(np.int64)
        g.flat_of_cell = flat_of_cell
        # psi of every stored cell (the source's row scale), 0 on cells outside the system
        g.psi_cell = None if psi is None else psi.reshape(-1)[
            torch.from_numpy(flat_of_cell).to(dev)].contiguous()
        if verbose:
            kind = "SBM" if psi is not None else "sharp"
            print(f"  torso {shape} @ {dx_mm} mm ({kind}, built on the device): "
                  f"{len(active_ids):,} active, {n_cells:,} stored cells, "
                  f"{len(surface_flat):,} surface, pin row {pin_row}")
        return g

    @classmethod
    def from_segmentation(cls, fine_labels, spacing, sigma_map, dx_mm, origin_mm=(0.0, 0.0, 0.0),
                          *, use_sbm=False, sbm_interface_width_mm=3.0, sbm_psi_cutoff=1e-3,
                          verbose=False, device="cuda"):
        """Resample a fine segmentation to dx_mm (nearest neighbour) and build
        its grid; with use_sbm the domain function comes from the fine
        segmentation (build_torso_grid.py:549-596)."""
        from . import torso_build as TB
        coarse = TB.resample_labels(fine_labels, spacing, dx_mm, device)
        psi = None
        if use_sbm:
            psi, _ = TB.smoothed_domain(fine_labels, spacing, dx_mm, tuple(coarse.shape),
                                        sbm_interface_width_mm, sbm_psi_cutoff, device)
        return cls.from_labels(coarse, sigma_map, dx_mm, origin_mm, verbose, psi_vol=psi,
                               device=device)

    def stiffness_csr(self):
        """The (unpinned) FP64 stiffness in row numbering (off-diagonals from the
        FP64 face conductances, diagonal from the kernel's six-face sum)."""
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
        g64 = self.g64.cpu().numpy() if hasattr(self.g64, "cpu") else np.asarray(self.g64)
        rows, cols, vals = [np.arange(n)], [np.arange(n)], [g64[3][cor]]
        for g, off in zip(g64[:3], (nj * nk, nk, 1)):
            c = np.nonzero(g > 0)[0]
            nb = cell_of_flat[flat[c] + off]
            a, b = row_of_cell[c], row_of_cell[nb]
            w = g[c]
            rows += [a, b]
            cols += [b, a]
            vals += [-w, -w]
        A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(n, n)).tocsr()
        return A
