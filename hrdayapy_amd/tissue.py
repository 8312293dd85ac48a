"""Host-side set-up of the coupled tissue model (one-time, numpy/scipy).

Builds what the engine consumes from the reference's inputs:
  * conductivity tensor per myocardial voxel
    (purkinje_myocardium_anisotropic_solver.py:856-883);
  * phenotype per node from phi (hetero_solver.py:92-101; NaN -> 0.5 as in
    purkinje_myocardium_anisotropic_solver.py:1548-1551);
  * Purkinje cable discretisation (same file :113-136), root stimulus zone
    (:1491-1502), terminals and PMJ patches (:1145-1177, :1566-1578);
  * derived constants w_pp, Ai_P, Ai_M (:1506-1538);
  * stimulus sites (:1204-1287).
"""
from __future__ import annotations

import math

import numpy as np


def conductivity_D6(fibre_dir, vox_idx, sigma_l, sigma_t, scale=None):
    """(6, N) D00,D11,D22,D01,D02,D12 of D = st*I + (sl-st) f f^T."""
    f = np.asarray(fibre_dir)[vox_idx[:, 0], vox_idx[:, 1], vox_idx[:, 2], :].astype(np.float64)
    ds = sigma_l - sigma_t
    D = np.empty((6, len(vox_idx)))
    D[0] = sigma_t + ds * f[:, 0] * f[:, 0]
    D[1] = sigma_t + ds * f[:, 1] * f[:, 1]
    D[2] = sigma_t + ds * f[:, 2] * f[:, 2]
    D[3] = ds * f[:, 0] * f[:, 1]
    D[4] = ds * f[:, 0] * f[:, 2]
    D[5] = ds * f[:, 1] * f[:, 2]
    if scale is not None:
        D *= np.asarray(scale)[None, :]
    return D


def explicit_dt_limit(D6, h, A_M, Cm, cable_edges=None, w_pp=0.0, Ai_P=1.0):
    """Largest forward-Euler diffusion step (ms) that a spectral bound proves stable.

    Forward Euler V += (dt/Cm) L V is stable while dt * rho(L) <= 2 Cm.
    Myocardium (Gershgorin): a node's 18 stencil weights sum to at most
    h*(2*tr D + 2*sum|D_offdiag|) (faces harm(D_kk)*h, edge diagonals
    |D_ab|*h/2 each, four per plane); divided by Ai_M = A_M*h^3, twice that
    bounds rho.  Cable (Anderson-Morley, uniform weights): rho <= (w_pp/Ai_P)
    * max over edges of deg(a) + deg(b).  The PMJ exchange is integrated
    exactly and sets no limit.
    """
    D = np.abs(np.asarray(D6, dtype=np.float64))
    row = 2.0 * (D[0] + D[1] + D[2]) + 2.0 * (D[3] + D[4] + D[5])
    lim = np.inf
    if row.size and row.max() > 0:
        lim = A_M * h * h * Cm / row.max()
    e = None if cable_edges is None else np.asarray(cable_edges, dtype=np.int64).reshape(-1, 2)
    if e is not None and len(e) and w_pp > 0:
        deg = np.bincount(e.ravel())
        lim = min(lim, 2.0 * Ai_P * Cm / (w_pp * float((deg[e[:, 0]] + deg[e[:, 1]]).max())))
    return lim


def isotropic_D6(n, sigma):
    D = np.zeros((6, n))
    D[:3] = sigma
    return D


def phenotype_from_phi(phi, vox_idx, phi_endo_max=0.35, phi_epi_min=0.65):
    ph = np.asarray(phi)[vox_idx[:, 0], vox_idx[:, 1], vox_idx[:, 2]].astype(np.float32)
    ph = np.where(np.isfinite(ph), ph, np.float32(0.5))
    ct = np.ones(len(ph), dtype=np.uint8)
    ct[ph < phi_endo_max] = 0
    ct[ph >= phi_epi_min] = 2
    return ct


def discretise_tree(nodes_mm, elements, dx):
    """Split every element into max(1, round(len/dx)) segments; original nodes keep their
    index, interior points are appended.  Returns (comp_nodes, comp_edges, branch_map)."""
    nodes_mm = np.asarray(nodes_mm, dtype=np.float32)
    elements = np.asarray(elements, dtype=np.int64).reshape(-1, 2)
    n0 = len(nodes_mm)
    p0 = nodes_mm[elements[:, 0]]
    p1 = nodes_mm[elements[:, 1]]
    nseg = np.maximum(1, np.round(np.linalg.norm(p1 - p0, axis=1) / dx).astype(np.int64))
    n_extra = int((nseg - 1).sum())
    extra = np.empty((n_extra, 3), dtype=np.float32)
    edges = np.empty((int(nseg.sum()), 2), dtype=np.int32)
    bmaps = []
    nxt, k, ke = n0, 0, 0
    for e in range(len(elements)):
        m = int(nseg[e])
        ids = [int(elements[e, 0])] + list(range(nxt, nxt + m - 1)) + [int(elements[e, 1])]
        if m > 1:
            # the reference interpolates in FP32: both weights rounded to FP32,
            # FP32 products and sum (purkinje_myocardium_anisotropic_solver.py:124-125)
            t = np.arange(1, m) / m
            w0 = (1.0 - t).astype(np.float32)[:, None]
            w1 = t.astype(np.float32)[:, None]
            extra[k:k + m - 1] = w0 * p0[e] + w1 * p1[e]
            k += m - 1
            nxt += m - 1
        edges[ke:ke + m, 0] = ids[:-1]
        edges[ke:ke + m, 1] = ids[1:]
        ke += m
        bmaps.append(ids)
    return np.concatenate([nodes_mm, extra]), edges, bmaps


def root_zone(nodes_mm, act_times, comp_nodes, T, stim_len_mm):
    act = np.array(act_times, dtype=np.float32)
    fin = np.isfinite(act)
    act[~fin] = T
    lo = float(act[fin].min())
    root = np.nonzero(np.isclose(act, lo, atol=1e-6))[0]
    centre = np.asarray(nodes_mm)[root].mean(axis=0)
    d = np.linalg.norm(comp_nodes - centre[None, :], axis=1)
    mask = d <= stim_len_mm
    if not mask.any():
        mask[int(d.argmin())] = True
    return mask


def pmj_sites(elements, branch_map, n_nodes, Np):
    """Cable node of every terminal branch end (terminal = degree-1 node other than 0)."""
    elements = np.asarray(elements, dtype=np.int64).reshape(-1, 2)
    deg = np.bincount(elements.ravel(), minlength=n_nodes)
    term = np.nonzero(deg == 1)[0]
    term = term[term != 0]
    if term.size == 0:
        term = np.array([n_nodes - 1])
    is_term = np.zeros(n_nodes, dtype=bool)
    is_term[term] = True
    ids = sorted({bm[-1] for i, bm in enumerate(branch_map) if is_term[elements[i, 1]]})
    return ids if ids else [Np - 1]


def pmj_patches(comp_nodes, ids, myo_nodes_mm, n_pmj=1):
    from scipy.spatial import cKDTree
    k = min(n_pmj, len(myo_nodes_mm))
    _, nn = cKDTree(myo_nodes_mm).query(np.asarray(comp_nodes)[ids], k=k)
    nn = np.asarray(nn).reshape(len(ids), k)
    return [row.tolist() for row in nn]


def purkinje_constants(R_P, sigma_P, dx_p, A_M, voxel_size, A_P=None):
    S_P = math.pi * R_P ** 2
    A_P = 2.0 / R_P if A_P is None else A_P
    return dict(w_pp=sigma_P * S_P / dx_p, Ai_P=A_P * S_P * dx_p, Ai_M=A_M * voxel_size ** 3)


def stim_sites(stim_protocol, root_mask, ectopic_region, vox_idx, Np, stim_regions=None):
    """Per-site global node ids and timing, the schema of the reference's stimulus protocol."""
    root_ids = np.nonzero(root_mask)[0]

    def in_mask(mask):
        m = np.asarray(mask)[vox_idx[:, 0], vox_idx[:, 1], vox_idx[:, 2]].astype(bool)
        return Np + np.nonzero(m)[0]

    ect = in_mask(ectopic_region) if ectopic_region is not None else np.zeros(0, np.int64)
    regions = {k: in_mask(v) for k, v in (stim_regions or {}).items()}
    out = []
    for ev in stim_protocol:
        target = ev.get("target", "root")
        if target == "root":
            ids = root_ids
        elif target == "ectopic":
            ids = ect
            if ids.size == 0:
                print(f"  WARNING: stimulus '{ev.get('name', target)}' targets 'ectopic' but no "
                      "myocardial nodes fall inside ectopic_region.")
        elif target == "region":
            rname = ev.get("region")
            if rname not in regions:
                raise ValueError(f"Stimulus '{ev.get('name', target)}' has target='region' with "
                                 f"region={rname!r}, but stim_regions only defines {sorted(regions)}.")
            ids = regions[rname]
        elif target == "custom":
            ids = np.asarray(ev["node_ids"], dtype=np.int64)
        else:
            raise ValueError(f"Unknown stimulus target '{target}' (expected 'root', 'ectopic', "
                             "'region', or 'custom').")
        out.append(dict(name=ev.get("name", target), target=target, node_ids=np.asarray(ids),
                        n_nodes=int(np.asarray(ids).size), amp=float(ev["amp"]),
                        dur=float(ev["dur"]), onset=float(ev.get("onset", 0.0)),
                        period=ev.get("period", None) or None, n_pulses=ev.get("n_pulses", None)))
    return out
