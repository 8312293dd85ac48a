"""CPU restatement of the reference's myocardium mesh, global Laplacian, CN
system and BiCGSTAB (test infrastructure: imported only by tests, smoke() and
bench.py's CPU baseline; never by the product path).

Reference: hrdayapy/anisotropic_simulation/functions/
purkinje_myocardium_anisotropic_solver.py (alias ``aniso.py``)
  * _discretise_tree                aniso.py:113-136
  * _build_global_laplacian_gpu     aniso.py:606-746
  * _build_cn_matrix                aniso.py:757-783
  * _bicgstab                       aniso.py:786-849
  * _build_conductivity_tensor      aniso.py:856-883
  * build_myo_mesh_anisotropic      aniso.py:886-1074 (edge list and weights)
  * map_pmj / _find_terminals       aniso.py:1145-1177, PMJ ids :1566-1573
  * root stimulus zone              aniso.py:1491-1502
  * _resolve_stim_sites/_stim_active aniso.py:1204-1301
  * isotropic mesh                  hrdayapy/simulation/functions/
                                    purkinje_myocardium_pipeline.py:761-833

Parity pins (tests/test_oracle_mesh.py): row sums of K are zero; K is
symmetric; sigma_l == sigma_t reduces the 19-point stencil to the 7-point one
with weight sigma*h; on a uniform tensor field the stencil is exact on
quadratics (L u = div(D grad u)/A_M); the harmonic face mean reproduces the
exact series-resistance flux of a two-layer 1-D medium; BiCGSTAB's answer
agrees with scipy's direct solve; _stim_active on hand-checked times.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp


# ---------------------------------------------------------------------------
# Myocardium
# ---------------------------------------------------------------------------
def conductivity_tensor(fibre_dir, vox_idx, sigma_l, sigma_t):
    """D = sigma_t I + (sigma_l - sigma_t) f f^T per node (aniso.py:856-883)."""
    f = fibre_dir[vox_idx[:, 0], vox_idx[:, 1], vox_idx[:, 2], :].astype(np.float64)
    ds = sigma_l - sigma_t
    return (sigma_t + ds * f[:, 0] ** 2, sigma_t + ds * f[:, 1] ** 2, sigma_t + ds * f[:, 2] ** 2,
            ds * f[:, 0] * f[:, 1], ds * f[:, 0] * f[:, 2], ds * f[:, 1] * f[:, 2])


FACE_OFFSETS = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
# axis pair -> (co-signed offset, counter-signed offset), aniso.py:1016-1020
DIAG_OFFSETS = {(0, 1): ((1, 1, 0), (1, -1, 0)),
                (0, 2): ((1, 0, 1), (1, 0, -1)),
                (1, 2): ((0, 1, 1), (0, 1, -1))}


def _pairs(vox_idx, flat_to_node, shape, off):
    Nx, Ny, Nz = shape
    ni, nj, nk = (vox_idx[:, 0] + off[0], vox_idx[:, 1] + off[1], vox_idx[:, 2] + off[2])
    inb = (ni >= 0) & (ni < Nx) & (nj >= 0) & (nj < Ny) & (nk >= 0) & (nk < Nz)
    nf = np.where(inb, ni * Ny * Nz + nj * Nz + nk, 0).astype(np.int64)
    nb = flat_to_node[nf]
    valid = inb & (nb >= 0)
    return np.nonzero(valid)[0].astype(np.int64), nb[valid].astype(np.int64)


def build_myo_mesh(S, voxel_size, fibre_dir, sigma_l, sigma_t):
    """Edge list and signed weights of the 18-neighbour mesh (aniso.py:886-1074).

    Returns dict(vox_idx, edges (E,2), edge_w (E,), nodes_mm, N_myo, D6).
    Edge order: 3 face blocks (+x,+y,+z), then per axis pair the co-signed and
    counter-signed diagonal blocks, as the reference builds them.
    """
    shape = S.shape
    Nx, Ny, Nz = shape
    vox_idx = np.argwhere(S == 1).astype(np.int32)
    N = len(vox_idx)
    flat_to_node = -np.ones(Nx * Ny * Nz, dtype=np.int64)
    flat_to_node[(vox_idx[:, 0].astype(np.int64) * Ny + vox_idx[:, 1]) * Nz + vox_idx[:, 2]] = np.arange(N)
    D = conductivity_tensor(fibre_dir, vox_idx, sigma_l, sigma_t)
    rows, cols, ws = [], [], []
    for axis, off in enumerate(FACE_OFFSETS):
        a, b = _pairs(vox_idx, flat_to_node, shape, off)
        Da, Db = D[axis][a], D[axis][b]
        ws.append(2.0 * Da * Db / np.maximum(Da + Db, 1e-30) * voxel_size)
        rows.append(a)
        cols.append(b)
    comp = {(0, 1): D[3], (0, 2): D[4], (1, 2): D[5]}
    for pair, (same, diff) in DIAG_OFFSETS.items():
        Dij = comp[pair]
        for off, sgn in ((same, 1.0), (diff, -1.0)):
            a, b = _pairs(vox_idx, flat_to_node, shape, off)
            ws.append(sgn * 0.5 * (Dij[a] + Dij[b]) * voxel_size / 2.0)
            rows.append(a)
            cols.append(b)
    edges = np.column_stack([np.concatenate(rows), np.concatenate(cols)])
    return dict(vox_idx=vox_idx, edges=edges, edge_w=np.concatenate(ws).astype(np.float64),
                nodes_mm=vox_idx[:, [1, 0, 2]].astype(np.float32) * voxel_size, N_myo=N,
                D6=np.stack(D), shape=shape)


def build_myo_mesh_iso(S, voxel_size, sigma_M):
    """6-connected isotropic mesh, w = sigma_M*h (purkinje_myocardium_pipeline.py:761-833, :1164)."""
    shape = S.shape
    Nx, Ny, Nz = shape
    vox_idx = np.argwhere(S == 1).astype(np.int32)
    N = len(vox_idx)
    flat_to_node = -np.ones(Nx * Ny * Nz, dtype=np.int64)
    flat_to_node[(vox_idx[:, 0].astype(np.int64) * Ny + vox_idx[:, 1]) * Nz + vox_idx[:, 2]] = np.arange(N)
    rows, cols = [], []
    for off in FACE_OFFSETS:
        a, b = _pairs(vox_idx, flat_to_node, shape, off)
        rows.append(a)
        cols.append(b)
    edges = np.column_stack([np.concatenate(rows), np.concatenate(cols)])
    return dict(vox_idx=vox_idx, edges=edges, edge_w=np.full(len(edges), sigma_M * voxel_size),
                nodes_mm=vox_idx[:, [1, 0, 2]].astype(np.float32) * voxel_size, N_myo=N, shape=shape)


# ---------------------------------------------------------------------------
# Purkinje tree
# ---------------------------------------------------------------------------
def discretise_tree(nodes_mm, elements, dx):
    """Subdivide each element into round(len/dx) segments (aniso.py:113-136)."""
    n0 = len(nodes_mm)
    extra, edges, bmaps = [], [], []
    nxt = n0
    for e in elements:
        p0, p1 = nodes_mm[e[0]], nodes_mm[e[1]]
        nseg = max(1, int(np.round(float(np.linalg.norm(p1 - p0)) / dx)))
        bm = [int(e[0])]
        for k in range(1, nseg):
            t = k / nseg
            extra.append(((1 - t) * p0 + t * p1).tolist())
            bm.append(nxt)
            nxt += 1
        bm.append(int(e[1]))
        bmaps.append(bm)
        edges.extend([bm[i], bm[i + 1]] for i in range(len(bm) - 1))
    comp_nodes = np.array(nodes_mm.tolist() + extra, dtype=np.float32)
    return comp_nodes, np.array(edges, dtype=np.int32).reshape(-1, 2), bmaps


def find_terminals(elements, n_nodes):
    """Degree-1 nodes other than node 0 (aniso.py:1172-1177)."""
    deg = np.zeros(n_nodes, dtype=np.int32)
    np.add.at(deg, elements[:, 0], 1)
    np.add.at(deg, elements[:, 1], 1)
    t = [i for i in range(1, n_nodes) if deg[i] == 1]
    return t if t else [n_nodes - 1]


def pmj_comp_ids(elements, branch_map, n_nodes, Np):
    """aniso.py:1566-1573."""
    term = set(find_terminals(elements, n_nodes))
    ids = sorted({bm[-1] for i, bm in enumerate(branch_map) if elements[i][1] in term})
    return ids if ids else [Np - 1]


def map_pmj(comp_nodes, term_ids, nodes_mm, n_pmj=1):
    """n_pmj nearest myocardial nodes of each terminal (aniso.py:1145-1166)."""
    from scipy.spatial import cKDTree
    k = min(n_pmj, len(nodes_mm))
    _, nn = cKDTree(nodes_mm).query(comp_nodes[term_ids], k=k)
    if k == 1:
        nn = nn[:, None]
    return [row.tolist() for row in nn]


def root_stim_mask(nodes_mm, act_times, comp_nodes, T, stim_len_mm):
    """Purkinje nodes within stim_len_mm of the earliest-activated node (aniso.py:1491-1502)."""
    act = np.array(act_times, dtype=np.float32)
    fin = np.isfinite(act)
    act[~fin] = T
    mn = float(act[fin].min())
    root = np.nonzero(np.isclose(act, mn, atol=1e-6))[0]
    pos = nodes_mm[root].mean(axis=0)
    d = np.linalg.norm(comp_nodes - pos[None, :], axis=1)
    m = d <= stim_len_mm
    if not m.any():
        m[int(d.argmin())] = True
    return m


# ---------------------------------------------------------------------------
# Global Laplacian L = D^-1 K, CN matrix, BiCGSTAB
# ---------------------------------------------------------------------------
def global_laplacian(Np, N_myo, comp_edges_p, w_pp, Ai_P, myo_edges, myo_w, Ai_M,
                     pmj_ids, pmj_patches, c_pmj):
    """aniso.py:606-746 as a scipy CSR matrix (FP64) plus its diagonal."""
    N = Np + N_myo
    rows, cols, vals = [], [], []
    ce = np.asarray(comp_edges_p, dtype=np.int64).reshape(-1, 2)
    if len(ce):
        rows += [ce[:, 0], ce[:, 1]]
        cols += [ce[:, 1], ce[:, 0]]
        vals += [np.full(len(ce), w_pp), np.full(len(ce), w_pp)]
    for k, p in enumerate(pmj_ids):
        patch = np.asarray(pmj_patches[k], dtype=np.int64) + Np
        w = c_pmj / len(patch)
        rows += [np.full(len(patch), p), patch]
        cols += [patch, np.full(len(patch), p)]
        vals += [np.full(len(patch), w), np.full(len(patch), w)]
    me = np.asarray(myo_edges, dtype=np.int64)
    mw = np.asarray(myo_w, dtype=np.float64)
    rows += [me[:, 0] + Np, me[:, 1] + Np]
    cols += [me[:, 1] + Np, me[:, 0] + Np]
    vals += [mw, mw]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    K = sp.coo_matrix((v, (r, c)), shape=(N, N)).tocsr()
    deg = np.asarray(K.sum(axis=1)).ravel()
    K = K - sp.diags(deg)
    Ai = np.empty(N)
    Ai[:Np] = Ai_P
    Ai[Np:] = Ai_M
    L = sp.diags(1.0 / Ai) @ K
    return L.tocsr(), -deg / Ai


def cn_matrix(L, diag_L, scale, theta):
    """A = (Cm/dt) I - theta L and the Jacobi inverse (aniso.py:757-783)."""
    A = (sp.identity(L.shape[0]) * scale - theta * L).tocsr()
    return A, 1.0 / np.maximum(scale - theta * diag_L, 1e-30)


def bicgstab(A, M_inv, b, x0, tol=1e-5, maxiter=500, return_iters=False, dot=None):
    """Jacobi-preconditioned BiCGSTAB with the reference's stopping rules (aniso.py:786-849).
    With return_iters, also the number of iterations begun (0 when ||b|| = 0).
    ``dot`` replaces the inner product (and the norms, as sqrt(dot(a, a))):
    tests use it to measure how far the iterates move when the sums are taken
    in another order."""
    if dot is not None:
        vdot = dot

        class _LA:
            @staticmethod
            def norm(a):
                return float(np.sqrt(vdot(a, a)))
    else:
        def vdot(a, c):
            return a @ c
        _LA = np.linalg
    x = x0.copy()
    r = b - A @ x
    r_hat = r.copy()
    rho_prev = alpha = omega = 1.0
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    b_norm = float(_LA.norm(b))
    if b_norm == 0.0:
        return (x, 0) if return_iters else x
    it = -1
    for it in range(maxiter):
        rho = float(vdot(r_hat, r))
        if abs(rho) < 1e-300:
            r_hat = r.copy()
            rho = float(vdot(r_hat, r))
        if it == 0:
            p = r.copy()
        else:
            beta = (rho / rho_prev) * (alpha / omega)
            p = r + beta * (p - omega * v)
        p_hat = M_inv * p
        v = A @ p_hat
        da = float(vdot(r_hat, v))
        if abs(da) < 1e-300:
            break
        alpha = rho / da
        s = r - alpha * v
        if _LA.norm(s) < tol * b_norm:
            x = x + alpha * p_hat
            break
        s_hat = M_inv * s
        t = A @ s_hat
        dw = float(vdot(t, t))
        omega = float(vdot(t, s)) / dw if abs(dw) > 1e-300 else 0.0
        x = x + alpha * p_hat + omega * s_hat
        r = s - omega * t
        rho_prev = rho
        if _LA.norm(r) / b_norm < tol:
            break
        if abs(omega) < 1e-300:
            break
    return (x, it + 1) if return_iters else x


# ---------------------------------------------------------------------------
# Stimulus protocol
# ---------------------------------------------------------------------------
def stim_active(t, onset, dur, period, n_pulses):
    """aniso.py:1290-1301."""
    if t < onset:
        return False
    if not period:
        return t < onset + dur
    k = int((t - onset) // period)
    if n_pulses is not None and k >= n_pulses:
        return False
    return (t - onset) - k * period < dur


def resolve_stim_sites(stim_protocol, root_mask_p, ectopic_region, vox_idx, Np, N_myo,
                       stim_regions=None):
    """Global node ids per site (aniso.py:1204-1287), returned as plain arrays."""
    root_ids = np.nonzero(root_mask_p)[0]
    if ectopic_region is not None:
        inr = ectopic_region[vox_idx[:, 0], vox_idx[:, 1], vox_idx[:, 2]].astype(bool)
        ect = Np + np.nonzero(inr)[0]
    else:
        ect = np.array([], dtype=np.int64)
    reg = {}
    for name, mask in (stim_regions or {}).items():
        inr = np.asarray(mask)[vox_idx[:, 0], vox_idx[:, 1], vox_idx[:, 2]].astype(bool)
        reg[name] = Np + np.nonzero(inr)[0]
    sites = []
    for ev in stim_protocol:
        target = ev.get("target", "root")
        if target == "root":
            ids = root_ids
        elif target == "ectopic":
            ids = ect
        elif target == "region":
            rname = ev.get("region")
            if rname not in reg:
                raise ValueError(f"Stimulus '{ev.get('name', target)}' has target='region' with "
                                 f"region={rname!r}, but stim_regions only defines {sorted(reg)}.")
            ids = reg[rname]
        elif target == "custom":
            ids = np.asarray(ev["node_ids"], dtype=np.int64)
        else:
            raise ValueError(f"Unknown stimulus target '{target}' (expected 'root', 'ectopic', "
                             "'region', or 'custom').")
        sites.append(dict(name=ev.get("name", target), node_ids=np.asarray(ids, dtype=np.int64),
                          amp=float(ev["amp"]), dur=float(ev["dur"]),
                          onset=float(ev.get("onset", 0.0)), period=ev.get("period", None) or None,
                          n_pulses=ev.get("n_pulses", None)))
    return sites


def purkinje_constants(R_P, sigma_P, dx_p, A_M, voxel_size, A_P=None):
    """Derived constants of aniso.py:1506-1538 (A_P = 2/R_P unless given, as in iso.py)."""
    S_P = math.pi * R_P ** 2
    if A_P is None:
        A_P = 2.0 / R_P
    return dict(S_P=S_P, A_P=A_P, w_pp=sigma_P * S_P / dx_p, Ai_P=A_P * S_P * dx_p,
                Ai_M=A_M * voxel_size ** 3)
