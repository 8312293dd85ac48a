"""CPU restatement of the reference's ECG forward solve (test infrastructure:
imported only by tests, smoke() and bench.py's CPU baseline).

Reference: hrdayapy/ecg/functions/
  * _assemble_stiffness (sharp and SBM)    build_torso_grid.py:276-397
  * compute_smoothed_domain, extend_conductivity (SBM psi and sigma)
                                            build_torso_grid.py:130-206
  * _resample_labels (nearest-neighbour zoom) build_torso_grid.py:260-269
  * pin row/column, Jacobi inverse          compute_bspm.py:327-336
  * pinned cell, surface cells              build_torso_grid.py:616-623, 469-488
  * build_heart_laplacian                   compute_bspm.py:88-119
  * build_projection_map                    compute_bspm.py:126-157
  * solve_frame source term                 compute_bspm.py:381-419
  * _pcg_gpu (rms / p99 / linf metrics)     compute_bspm.py:164-263
Parity pins (tests/test_oracle_ecg.py): the assembled matrix is symmetric
with zero row sums before pinning; PCG agrees with scipy's direct solve of
the pinned system; Frank's closed-form sphere-dipole potential
(experiments/experiment_4/sphere_dipole_validation.py:173-183) within the
reference's acceptance RDM < 0.05 at its resolution ratio; a uniform Vm
gives a zero source; the source is linear in Vm.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy import ndimage as ndi


def assemble_stiffness(sigma_vol, h_m, active_mask, psi_vol=None):
    """7-point FDM stiffness with harmonic-mean face conductivity, Neumann at the
    box edge and at air, in the reference's row order (active voxels in flat order).
    With psi_vol (SBM) each face conductance is scaled by the harmonic mean of
    psi on its two cells; faces whose cells both have psi <= 0 are skipped."""
    ni, nj, nk = sigma_vol.shape
    act = active_mask.ravel()
    active_ids = np.flatnonzero(act)
    N = len(active_ids)
    id_map = np.full(act.size, -1, dtype=np.int64)
    id_map[active_ids] = np.arange(N)
    s = sigma_vol.ravel().astype(np.float64)
    psi = None if psi_vol is None else np.asarray(psi_vol, dtype=np.float64).ravel()
    i, j, k = np.unravel_index(active_ids, (ni, nj, nk))
    diag = np.zeros(N)
    rows, cols, vals = [], [], []
    for (di, dj, dk), stride in (((1, 0, 0), nj * nk), ((-1, 0, 0), -nj * nk), ((0, 1, 0), nk),
                                 ((0, -1, 0), -nk), ((0, 0, 1), 1), ((0, 0, -1), -1)):
        ok = ((i + di >= 0) & (i + di < ni) & (j + dj >= 0) & (j + dj < nj) & (k + dk >= 0)
              & (k + dk < nk))
        r = np.nonzero(ok)[0]
        f = active_ids[r]
        nb = f + stride
        sk, sn = s[f], s[nb]
        pair = sk + sn >= 1e-30
        if psi_vol is not None:
            pk, pn = psi[f], psi[nb]
            pair &= ~((pk <= 0.0) & (pn <= 0.0))
            pk, pn = pk[pair], pn[pair]
        r, nb, sk, sn = r[pair], nb[pair], sk[pair], sn[pair]
        if psi_vol is None:
            cond = (2.0 * sk * sn) / (sk + sn) / h_m ** 2
        else:
            ps = pk + pn
            pf = np.where(ps < 1e-30, 0.0, (2.0 * pk * pn) / np.where(ps < 1e-30, 1.0, ps))
            cond = pf * ((2.0 * sk * sn) / (sk + sn)) / h_m ** 2
        np.add.at(diag, r, cond)
        inside = act[nb]
        rows.append(r[inside])
        cols.append(id_map[nb[inside]])
        vals.append(-cond[inside])
    rows.append(np.arange(N))
    cols.append(np.arange(N))
    vals.append(diag)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    return A, active_ids, id_map


def resample_labels(volume, spacing, dx_coarse_mm):
    """Nearest-neighbour resample of a label volume to dx_coarse_mm."""
    zoom = np.asarray(spacing, dtype=np.float64) / dx_coarse_mm
    return ndi.zoom(np.asarray(volume).astype(np.float32), zoom, order=0).astype(np.int32)


def smoothed_domain(volume_fine, spacing_fine, dx_coarse_mm, coarse_shape,
                    interface_width_mm=3.0, psi_cutoff=1e-3):
    """SBM domain function on the coarse grid: psi_fine = (1 + tanh(3 d / w)) / 2
    of the signed distance d (mm, + inside) of the fine segmentation, linearly
    zoomed to the coarse spacing, cropped/padded with 0 to coarse_shape, and 0
    where it falls below psi_cutoff."""
    m = (np.asarray(volume_fine) != 0).astype(np.float64)
    sp_f = np.asarray(spacing_fine, dtype=np.float64)
    d = (ndi.distance_transform_edt(m, sampling=sp_f)
         - ndi.distance_transform_edt(1.0 - m, sampling=sp_f))
    psi_f = 0.5 * (1.0 + np.tanh(3.0 * d / interface_width_mm))
    psi_c = ndi.zoom(psi_f, sp_f / dx_coarse_mm, order=1)
    out = np.zeros(coarse_shape, dtype=np.float64)
    sl = tuple(slice(0, min(a, b)) for a, b in zip(psi_c.shape, coarse_shape))
    out[sl] = psi_c[sl]
    out[out < psi_cutoff] = 0.0
    return out


def extend_conductivity(sigma_vol, tissue_mask, sbm_mask):
    """sigma of the nearest tissue voxel (Euclidean, scipy's feature transform)
    on the band cells (sbm_mask and not tissue)."""
    band = sbm_mask & ~tissue_mask
    out = np.array(sigma_vol, dtype=np.float64, copy=True)
    if not band.any():
        return out
    _, idx = ndi.distance_transform_edt(~tissue_mask, return_indices=True)
    out[band] = sigma_vol[tuple(ix[band] for ix in idx)]
    return out


def pin_row_of(tissue_mask, id_map):
    interior = np.flatnonzero(tissue_mask.ravel())
    return int(id_map[int(interior[len(interior) // 2])])


def pin(A, row):
    """Identity row/column at `row` (compute_bspm.py:327-331)."""
    A = A.tolil()
    A[row, :] = 0.0
    A[:, row] = 0.0
    A[row, row] = 1.0
    return A.tocsr()


def surface_flat(label_vol, active_mask):
    air = label_vol == 0
    dil = ndi.binary_dilation(air, structure=ndi.generate_binary_structure(3, 1))
    return np.flatnonzero((active_mask & dil).ravel())


def heart_laplacian(vox_idx):
    """6-neighbour unit-weight graph Laplacian K (positive form: K V ~ h^2 lap V)."""
    ijk = vox_idx.astype(np.int64) - vox_idx.min(axis=0)
    shape = tuple((ijk.max(axis=0) + 2).tolist())
    lut = np.full(shape, -1, dtype=np.int64)
    n = len(ijk)
    lut[ijk[:, 0], ijk[:, 1], ijk[:, 2]] = np.arange(n)
    rows, cols = [], []
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        q = ijk + np.array(d)
        inb = np.all((q >= 0) & (q < np.array(shape)), axis=1)
        nb = np.full(n, -1)
        nb[inb] = lut[q[inb, 0], q[inb, 1], q[inb, 2]]
        ok = nb >= 0
        rows.append(np.nonzero(ok)[0])
        cols.append(nb[ok])
    r, c = np.concatenate(rows), np.concatenate(cols)
    K = sp.csr_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    return K - sp.diags(np.asarray(K.sum(axis=1)).ravel())


def projection(coords_mm, R, t, scale, origin, dx, id_map, shape):
    torso = (np.asarray(R) @ (coords_mm * scale).T).T + np.asarray(t)
    ijk = np.floor((torso - origin) / dx).astype(np.int64)
    ni, nj, nk = shape
    inb = np.all((ijk >= 0) & (ijk < np.array(shape)), axis=1)
    proj = np.full(len(coords_mm), -1, dtype=np.int64)
    flat = (ijk[inb, 0] * nj + ijk[inb, 1]) * nk + ijk[inb, 2]
    proj[np.nonzero(inb)[0]] = id_map[flat]
    return proj


def source(Vm_mV, K, proj, sigma_i, h_fine_m, dx_coarse_mm, n_rows, pin_row, psi_row=None):
    """RHS of one frame (compute_bspm.py:381-419), FP64; with psi_row (SBM) the
    RHS is scaled by psi row-wise before the pinned row is cleared."""
    valid = proj >= 0
    lap = K @ (np.asarray(Vm_mV, dtype=np.float64) * 1e-3)
    Iv = lap[valid] * (sigma_i / h_fine_m ** 2)
    src = Iv - Iv.mean()
    b = np.zeros(n_rows)
    np.add.at(b, proj[valid], src)
    b /= (dx_coarse_mm / (h_fine_m * 1e3)) ** 3
    if psi_row is not None:
        b *= np.asarray(psi_row, dtype=np.float64)
    b[pin_row] = 0.0
    return b


def pcg(A, b, M_inv, x0=None, tol_abs_mV=0.01, max_iter=500, metric="rms", dtype=np.float64):
    """Jacobi PCG with the reference's per-iteration stopping statistic of z = M^-1 r."""
    A = A.astype(dtype)
    b = b.astype(dtype)
    M_inv = M_inv.astype(dtype)
    N = len(b)
    tol = tol_abs_mV * 1e-3

    def stat(z):
        if metric == "rms":
            return np.linalg.norm(z) / np.sqrt(N)
        if metric == "linf":
            return np.abs(z).max()
        return np.quantile(np.abs(z), 0.99)

    x = x0.astype(dtype).copy() if x0 is not None else np.zeros(N, dtype)
    r = b - A @ x if x0 is not None else b.copy()
    z = M_inv * r
    p = z.copy()
    rz = r @ z
    for i in range(max_iter):
        Ap = A @ p
        pAp = p @ Ap
        if abs(pAp) < 1e-30:
            break
        alpha = rz / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        z = M_inv * r
        e = stat(z)
        if e < tol:
            return x, i + 1, float(e * 1e3)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iter, float(stat(M_inv * r) * 1e3)


def frank_sphere(p_mag, sigma, R_mm, xyz_mm, centre_mm, axis):
    """Potential of a dipole at the centre of an insulated homogeneous sphere."""
    R_m = R_mm * 1e-3
    rv = np.asarray(xyz_mm) - np.asarray(centre_mm)
    r_mm = np.linalg.norm(rv, axis=-1)
    r_m = r_mm * 1e-3
    ct = (rv @ axis) / np.clip(r_mm, 1e-12, None)
    return (p_mag / (4 * np.pi * sigma)) * (ct / r_m ** 2 + 2 * r_m * ct / R_m ** 3)


def sphere_labels(R_mm, dx_mm, margin=1.3):
    n = int(np.ceil(2 * margin * R_mm / dx_mm)) + 1
    origin = -np.full(3, (n - 1) * dx_mm / 2)
    ax = origin[0] + np.arange(n) * dx_mm
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    return (X ** 2 + Y ** 2 + Z ** 2 <= R_mm ** 2).astype(np.uint8), origin
