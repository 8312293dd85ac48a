"""CPU restatement of the reference's transmural coordinate and rule-based
fibre field (test infrastructure: imported only by tests; never by the
product path).

Reference:
  * hrdayapy/coordinates/functions/generate_transmural_coordinate.py
      - `_build_laplacian_transmural` :85-183: 6-neighbour graph Laplacian on
        the myocardial voxels, identity rows (Dirichlet) where the surface
        label is +1 (phi = 1, epicardium) or -1 (phi = 0, endocardium), the
        diagonal counting only tissue neighbours (Neumann at the mask boundary);
      - `_conjugate_gradient` :35-78: plain CG in FP32 from x0, stopping when
        ||r|| / ||r0|| < tol;
      - `generate_transmural_coordinate` :190-286: the EDT warm start
        d_endo / (d_endo + d_epi), the Dirichlet values re-applied after the
        solve, NaN outside the mask, clip to [0, 1].
  * hrdayapy/fibres/functions/generate_fibre_direction.py
      - `_fill_nan_nearest` :90-111 (the nearest in-mask value in a 2-voxel
        halo, 0 beyond it), `_normalize3` :114-120, `_gradient_field` :123-128;
      - `generate_fibre_direction` :135-281: crop to the mask's bounding box
        + 4, e_t = grad(phi)/|.|, e_l = grad(psi) minus its e_t part,
        normalised, e_c = e_t x e_l, alpha = alpha_endo + (alpha_epi -
        alpha_endo) phi, f = cos(alpha) e_c + sin(alpha) e_l, normalised; all
        in FP32; the crop written back into a zero volume.
Pinned by tests/test_oracle_coords.py against the reference's own outputs
(tests/golden/coord_golden.npz) and by the invariants f . e_t = 0, |f| = 1 and
the helix-angle law.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
from scipy.ndimage import binary_dilation, distance_transform_edt

F32 = np.float32
EPS = F32(1e-8)


def transmural_system(S, surface_label):
    """(A csr FP32, b FP32, vox_idx, labels at the voxels) of the Laplace
    problem, rows in np.argwhere(S) order."""
    S = np.asarray(S).astype(bool)
    shape = S.shape
    vox = np.argwhere(S)
    n = len(vox)
    node_of = -np.ones(S.size, dtype=np.int64)
    node_of[np.ravel_multi_index(tuple(vox.T), shape)] = np.arange(n)
    lab = np.asarray(surface_label)[tuple(vox.T)]
    dirichlet = lab != 0
    b = np.where(lab == 1, 1.0, 0.0).astype(F32)
    rows, cols, vals = [np.nonzero(dirichlet)[0]], [np.nonzero(dirichlet)[0]], [np.ones(int(dirichlet.sum()), F32)]
    inner = np.nonzero(~dirichlet)[0]
    deg = np.zeros(n, F32)
    for ax in range(3):
        for d in (-1, 1):
            c = vox[inner].copy()
            c[:, ax] += d
            ok = (c[:, ax] >= 0) & (c[:, ax] < shape[ax])
            nb = np.full(len(inner), -1, dtype=np.int64)
            nb[ok] = node_of[np.ravel_multi_index(tuple(c[ok].T), shape)]
            has = nb >= 0
            rows.append(inner[has])
            cols.append(nb[has])
            vals.append(np.full(int(has.sum()), -1.0, F32))
            deg[inner[has]] += 1.0
    rows.append(inner)
    cols.append(inner)
    vals.append(deg[inner])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n), dtype=F32)
    return A, b, vox, lab


def edt_start(S, surface_label, vox):
    """d_endo / (d_endo + d_epi) at the voxels (0.5 where both vanish), the
    Dirichlet values set."""
    lab = np.asarray(surface_label)
    d_endo = distance_transform_edt(lab != -1)
    d_epi = distance_transform_edt(lab != 1)
    den = d_endo + d_epi
    v = np.where(den > 0, d_endo / np.where(den > 0, den, 1.0), 0.5).astype(F32)
    v[lab == -1] = 0.0
    v[lab == 1] = 1.0
    return v[tuple(vox.T)]


def cg(A, b, x0, tol=1e-6, maxiter=10_000):
    """The reference's CG recurrence in FP32 arithmetic (FP64 inner products);
    returns (x, iterations, final relative residual)."""
    x = x0.astype(F32).copy()
    r = (b - A @ x).astype(F32)
    p = r.copy()
    rs = float(np.dot(r.astype(np.float64), r))
    r0 = math.sqrt(rs)
    rel = 1.0
    it = 0
    for it in range(maxiter):
        Ap = (A @ p).astype(F32)
        alpha = F32(rs / float(np.dot(p.astype(np.float64), Ap)))
        x = x + alpha * p
        r = r - alpha * Ap
        rs_new = float(np.dot(r.astype(np.float64), r))
        rel = math.sqrt(rs_new) / (r0 + 1e-30)
        if rel < tol:
            return x, it + 1, rel
        p = r + F32(rs_new / rs) * p
        rs = rs_new
    return x, it + 1, rel


def transmural_coordinate(S, surface_label, tol=1e-6, maxiter=10_000):
    """phi volume (NaN outside S) and the CG iteration count."""
    A, b, vox, lab = transmural_system(S, surface_label)
    x0 = edt_start(S, surface_label, vox)
    x, it, _ = cg(A, b, x0, tol, maxiter)
    x[lab == -1] = 0.0
    x[lab == 1] = 1.0
    phi = np.full(np.asarray(S).shape, np.nan, dtype=F32)
    phi[tuple(vox.T)] = x
    return np.clip(phi, 0.0, 1.0), it


def fill_nan_nearest(field, S, halo=2):
    """The nearest in-mask value on voxels within `halo` steps of S, 0 beyond."""
    out = field.copy()
    inv = ~S
    need = binary_dilation(S, iterations=halo) & inv
    if need.any():
        idx = distance_transform_edt(inv, return_distances=False, return_indices=True)
        out[need] = field[tuple(idx[:, need])]
    out[inv & ~need] = 0.0
    return np.nan_to_num(out, nan=0.0).astype(F32, copy=False)


def _unit(x, y, z):
    n = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    np.maximum(n, EPS, out=n)
    return (x / n).astype(F32), (y / n).astype(F32), (z / n).astype(F32)


def bbox_crop(S, pad=4):
    idx = np.argwhere(S)
    lo = np.maximum(idx.min(axis=0) - pad, 0)
    hi = np.minimum(idx.max(axis=0) + pad + 1, S.shape)
    return tuple(slice(int(a), int(b)) for a, b in zip(lo, hi))


def fibre_direction(S, phi, psi, spacing, alpha_endo_deg=60.0, alpha_epi_deg=-60.0, pad=4):
    """(Nx,Ny,Nz,3) FP32 fibre field."""
    S = np.asarray(S).astype(bool)
    box = bbox_crop(S, pad)
    Sc = S[box]
    h = [float(v) for v in spacing]
    Fp = fill_nan_nearest(np.asarray(phi, F32)[box], Sc)
    Fs = fill_nan_nearest(np.asarray(psi, F32)[box], Sc)
    gx, gy, gz = (g.astype(F32) for g in np.gradient(Fp, h[0], h[1], h[2]))
    tx, ty, tz = _unit(gx, gy, gz)
    sx, sy, sz = (g.astype(F32) for g in np.gradient(Fs, h[0], h[1], h[2]))
    d = sx * tx + sy * ty + sz * tz
    lx, ly, lz = _unit(sx - d * tx, sy - d * ty, sz - d * tz)
    cx, cy, cz = _unit(ty * lz - tz * ly, tz * lx - tx * lz, tx * ly - ty * lx)
    a0, a1 = math.radians(alpha_endo_deg), math.radians(alpha_epi_deg)
    a = (a0 + (a1 - a0) * Fp).astype(F32)
    ca, sa = np.cos(a).astype(F32), np.sin(a).astype(F32)
    f = np.stack([ca * cx + sa * lx, ca * cy + sa * ly, ca * cz + sa * lz], axis=-1).astype(F32)
    n = np.sqrt(f[..., 0] ** 2 + f[..., 1] ** 2 + f[..., 2] ** 2)
    np.maximum(n, EPS, out=n)
    f /= n[..., None]
    out = np.zeros(S.shape + (3,), F32)
    out[box] = f
    return out
