"""Seeded synthetic inputs for the benchmark workloads (no patient data needed).

* ``slab``       -- rectangular all-myocardium slab (the reference's CPU case).
* ``biv_shell``  -- biventricular shell from two truncated confocal ellipsoids,
                    transmural coordinate phi, rule-based helix fibres
                    (+60 deg endo -> -60 deg epi), and a seeded fractal
                    Purkinje tree on the endocardium, returned in the
                    reference's input conventions: S/phi/fibre_dir in raw
                    array-axis order, tree nodes in voxel units with the
                    (col,row,depth) convention (purkinje_myocardium_anisotropic_solver.py:1057-1068).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage as ndi


def slab(nx=64, ny=64, nz=16):
    S = np.ones((nx, ny, nz), dtype=np.uint8)
    phi = np.zeros(S.shape, dtype=np.float32)
    f = np.zeros(S.shape + (3,), dtype=np.float32)
    f[..., 0] = 1.0
    return S, phi, f


def stub_tree(shape):
    """Inert 2-node Purkinje stub outside the grid (used with c_pmj = 0)."""
    Nx, Ny, Nz = shape
    nodes = np.array([[Ny + 5.0, Nx + 5.0, Nz + 5.0], [Ny + 5.0, Nx + 5.0, Nz + 10.0]])
    return nodes, np.array([[0, 1]]), np.array([0.0, 1.0])


def _ellipsoid(X, Y, Z, c, ax):
    return ((X - c[0]) / ax[0]) ** 2 + ((Y - c[1]) / ax[1]) ** 2 + ((Z - c[2]) / ax[2]) ** 2 <= 1.0


def biv_shell(n=256, h=0.4, seed=0, wall_mm=(11.5, 5.0), coords="distance",
              device="cuda"):
    """Biventricular shell on an n^3 grid of spacing h (mm).

    LV: confocal prolate shell (cavity semi-axes a_i, c_i; epicardium with the
    same focal length, thickness wall_mm[0] at the equator); RV: a second,
    wider confocal shell shifted along x, wrapping the LV, free wall
    wall_mm[1]; both truncated by a base plane.  Returns dict(S, phi,
    fibre_dir, cavity_lv, cavity_rv, h).

    coords="distance": phi is the distance ratio d_endo/(d_endo + d_epi) and
    the fibres follow `rule_based_fibres` (host only).  coords="laplace": phi
    is the reference's Laplace transmural coordinate and the fibres its
    rule-based field from phi and an apicobasal coordinate psi (a Laplace
    solve between the apex and base regions of `biv_apex_base_masks`), all on
    the GPU (`hrdayapy_amd.coordinates`, `hrdayapy_amd.fibres`); the result
    then also holds psi.
    """
    if coords not in ("distance", "laplace"):
        raise ValueError(f"coords must be 'distance' or 'laplace' (got {coords!r})")
    L = n * h
    ax = (np.arange(n) + 0.5) * h - L / 2
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    scale = L / 102.4
    a_i, c_i = 22.0 * scale, 42.0 * scale
    f2 = c_i ** 2 - a_i ** 2
    a_o = a_i + wall_mm[0] * scale
    c_o = np.sqrt(a_o ** 2 + f2)
    zc = 8.0 * scale                      # long axis along z, base at the top
    lv_out = _ellipsoid(X, Y, Z, (0, 0, zc), (a_o, a_o, c_o))
    lv_cav = _ellipsoid(X, Y, Z, (0, 0, zc), (a_i, a_i, c_i))
    rvc = (18.0 * scale, 0.0, zc + 4.0 * scale)
    b_i = (30.0 * scale, 24.0 * scale, 38.0 * scale)
    rv_cav = _ellipsoid(X, Y, Z, rvc, b_i) & ~lv_out
    b_o = tuple(b + wall_mm[1] * scale for b in b_i)
    rv_out = _ellipsoid(X, Y, Z, rvc, b_o)
    base = Z <= zc + 18.0 * scale
    S = (lv_out | rv_out) & ~lv_cav & ~rv_cav & base
    # keep the largest connected piece
    lab, nl = ndi.label(S)
    if nl > 1:
        sizes = ndi.sum(S, lab, range(1, nl + 1))
        S = lab == (1 + int(np.argmax(sizes)))
    cav = (lv_cav | rv_cav) & base
    # transmural coordinate: distance to cavity (endo) vs to the outside (epi)
    outside = ~(S | cav)
    d_endo = ndi.distance_transform_edt(~cav)
    d_epi = ndi.distance_transform_edt(~outside)
    phi = np.where(S, d_endo / np.maximum(d_endo + d_epi, 1e-6), np.nan).astype(np.float32)
    phi = np.clip(phi, 0.0, 1.0)
    geo = dict(S=S.astype(np.uint8), phi=phi, cavity_lv=lv_cav & base, cavity_rv=rv_cav & base,
               h=h)
    if coords == "distance":
        geo["fibre_dir"] = rule_based_fibres(S, phi)
        return geo
    from .coordinates import LaplaceSolver, generate_transmural_coordinate
    from .fibres import generate_fibre_direction
    geo["phi"] = generate_transmural_coordinate(geo["S"], biv_surface_labels(geo), device=device,
                                                verbose=False)
    apex, basal = biv_apex_base_masks(geo)
    vox = np.argwhere(S)
    dirich = (apex | basal)[tuple(vox.T)]
    z = vox[:, 2].astype(np.float32)
    x0 = np.where(dirich, basal[tuple(vox.T)].astype(np.float32),
                  (z - z.min()) / max(1.0, float(z.max() - z.min()))).astype(np.float32)
    solver = LaplaceSolver(S, device)
    psi = np.full(S.shape, np.nan, np.float32)
    psi[tuple(vox.T)] = np.clip(solver.solve(dirich, x0, tol=1e-6, maxiter=20_000), 0.0, 1.0)
    geo["psi"] = psi
    geo["fibre_dir"] = generate_fibre_direction(S, geo["phi"], psi, (h, h, h), verbose=False,
                                                device=device)
    return geo


def biv_surface_labels(geo):
    """Transmural boundary labels of a BiV shell in the reference's convention
    (+1 epicardium, -1 endocardium, 0 elsewhere): wall voxels face-adjacent to
    a cavity are endocardial, wall voxels face-adjacent to the outside below
    the base plane epicardial; the cut base itself is neither (Neumann)."""
    S = geo["S"].astype(bool)
    cav = geo["cavity_lv"] | geo["cavity_rv"]
    k = S.shape[2]
    rows = np.nonzero(S.any(axis=(0, 1)))[0]
    top = int(rows.max()) if len(rows) else k - 1
    below = np.zeros(S.shape, bool)
    below[:, :, :top + 1] = True          # the base plane is the wall's top layer
    outside = ~(S | cav) & below
    st = ndi.generate_binary_structure(3, 1)
    lab = np.zeros(S.shape, np.int8)
    lab[S & ndi.binary_dilation(outside, st)] = 1
    lab[S & ndi.binary_dilation(cav, st)] = -1
    return lab


def biv_apex_base_masks(geo, apex_frac=0.02):
    """Dirichlet regions of the apicobasal coordinate: the wall voxels of the
    lowest apex_frac of the wall's height (apex) and of its top layer (base)."""
    S = geo["S"].astype(bool)
    z = np.nonzero(S.any(axis=(0, 1)))[0]
    z0, z1 = int(z.min()), int(z.max())
    kk = np.arange(S.shape[2])[None, None, :]
    apex = S & (kk <= z0 + max(1, int(round(apex_frac * (z1 - z0)))))
    base = S & (kk == z1)
    return apex, base


def rule_based_fibres(S, phi, alpha_endo=60.0, alpha_epi=-60.0):
    """Helix fibres: alpha(phi) rotation of the circumferential direction about
    the transmural gradient (Bayer-style rule, long axis = array axis 2)."""
    p = np.nan_to_num(phi, nan=0.0).astype(np.float64)
    p = ndi.gaussian_filter(p, 1.0)
    g = np.stack(np.gradient(p), axis=-1)
    t = g / np.maximum(np.linalg.norm(g, axis=-1, keepdims=True), 1e-9)
    ez = np.array([0.0, 0.0, 1.0])
    l = ez - (t @ ez)[..., None] * t
    l /= np.maximum(np.linalg.norm(l, axis=-1, keepdims=True), 1e-9)
    c = np.cross(l, t)
    a = np.deg2rad(alpha_endo + (alpha_epi - alpha_endo) * np.nan_to_num(phi, nan=0.5))[..., None]
    f = np.cos(a) * c + np.sin(a) * l
    f /= np.maximum(np.linalg.norm(f, axis=-1, keepdims=True), 1e-9)
    f[~S.astype(bool)] = 0.0
    return f.astype(np.float32)


def purkinje_tree(geo, n_pmj=800, seed=0, step_mm=0.55, root_frac=0.85):
    """Seeded fractal Purkinje tree on the endocardial surfaces.

    Terminals are n_pmj seeded endocardial voxels (LV and RV in proportion to
    their endocardial area); they are attached, nearest-first, to the growing
    tree through straight paths with nodes every ``step_mm``.  Returns
    (nodes [voxel units, (col,row,depth)], elements (E,2), activation_times).
    """
    rng = np.random.default_rng(seed)
    S = geo["S"].astype(bool)
    h = geo["h"]
    cav = geo["cavity_lv"] | geo["cavity_rv"]
    endo = cav & ndi.binary_dilation(S)  # cavity voxels touching the wall
    pts = np.argwhere(endo).astype(np.float64)
    if len(pts) == 0:
        raise ValueError("no endocardial surface")
    # root: high on the septum (top of the LV cavity, toward the RV)
    lv = np.argwhere(geo["cavity_lv"] & endo)
    top = lv[:, 2] >= np.quantile(lv[:, 2], root_frac)
    root = lv[top][np.argmax(lv[top][:, 0])].astype(np.float64)
    pick = rng.choice(len(pts), size=min(n_pmj, len(pts)), replace=False)
    targets = pts[pick]
    order = np.argsort(np.linalg.norm(targets - root, axis=1))
    nodes = [root]
    elements = []
    dist = [0.0]
    open_ = [True]  # terminals are not attachment points, so each stays a PMJ
    step = step_mm / h
    for k, ti in enumerate(order):
        tgt = targets[ti]
        arr = np.asarray(nodes)
        d = np.linalg.norm(arr - tgt, axis=1)
        d[~np.asarray(open_)] = np.inf
        j = int(np.argmin(d))
        nseg = max(1, int(np.ceil(d[j] / step)))
        prev = j
        for s in range(1, nseg + 1):
            p = arr[j] + (tgt - arr[j]) * (s / nseg)
            nodes.append(p)
            open_.append(s < nseg)
            elements.append([prev, len(nodes) - 1])
            dist.append(dist[prev] + d[j] / nseg)
            prev = len(nodes) - 1
    nodes = np.asarray(nodes)[:, [1, 0, 2]]  # (row,col,depth) -> (col,row,depth)
    act = np.asarray(dist) * h / 2.5          # ~2.5 mm/ms conduction
    return nodes, np.asarray(elements, dtype=np.int64), act


TORSO_SIGMA = {1: 0.2, 2: 0.2, 3: 0.05, 4: 0.7}  # bulk, skeletal muscle, lungs, blood [S/m]


def torso(shape=(384, 384, 256), dx_mm=1.5, heart_centre_frac=(0.58, 0.42, 0.55)):
    """Synthetic torso label volume: nested ellipsoids of bulk tissue (1),
    a skeletal-muscle rim (2), two lungs (3) and a blood pool (4) at the heart
    position.  Returns (labels, origin_mm, heart_centre_mm); the outer
    ellipsoid is inscribed in the box (about 52 % of the cells are tissue)."""
    ni, nj, nk = shape
    ext = np.array(shape, float) * dx_mm
    c = ext / 2.0
    ax = [(np.arange(n) + 0.5) * dx_mm for n in shape]
    X, Y, Z = np.meshgrid(*ax, indexing="ij", sparse=True)
    semi = 0.49 * ext

    def ell(cen, s):
        return (((X - cen[0]) / s[0]) ** 2 + ((Y - cen[1]) / s[1]) ** 2
                + ((Z - cen[2]) / s[2]) ** 2) <= 1.0

    lab = np.zeros(shape, dtype=np.uint8)
    lab[ell(c, semi)] = 2
    lab[ell(c, semi * 0.93)] = 1
    for sgn in (-1, 1):
        lab[ell(c + np.array([sgn * 0.22 * ext[0], 0.05 * ext[1], 0.08 * ext[2]]),
                np.array([0.14, 0.2, 0.26]) * ext)] = 3
    hc = np.asarray(heart_centre_frac) * ext
    lab[ell(hc, np.array([0.05, 0.05, 0.07]) * ext)] = 4
    return lab, np.zeros(3), hc


def apical_stimulus_region(geo, radius_mm=3.0):
    """Myocardial voxels within radius_mm of the LV endocardial apex: an
    ectopic site that starts the ventricular wave (the thin Purkinje cable of
    the reference's parameters cannot drive the tissue load through the PMJs)."""
    S = geo["S"].astype(bool)
    lv = np.argwhere(geo["cavity_lv"])
    apex = lv[np.argmin(lv[:, 2])].astype(float)
    vox = np.argwhere(S)
    d = np.linalg.norm(vox - apex, axis=1) * geo["h"]
    region = np.zeros(S.shape, dtype=bool)
    sel = vox[d <= radius_mm]
    region[sel[:, 0], sel[:, 1], sel[:, 2]] = True
    return region
