"""CPU FP64 restatement of the reference's coupled monodomain time loop
(test infrastructure: imported only by tests, smoke() and bench.py's CPU
baseline; never by the product path).

Reference: purkinje_myocardium_anisotropic_solver.py (``aniso.py``)
  * setup: :1477-1616 (tree discretisation, root zone, derived constants,
    mesh, phi -> phenotype via hetero_solver.py:74-128, PMJ mapping, L, CN)
  * time loop :1749-1791, Lie splitting:
      I_ion, gates   <- TTP06(V^n)                     (hetero_solver.py:141-194)
      V*             = V^n + (dt/Cm)(-I_ion*Cm + I_stim)
      (Cm/dt - theta L) V^{n+1} = (Cm/dt) V* + (1-theta) L V*   (BiCGSTAB)
      V^{n+1}        = clamp(V^{n+1}, -150, 80)

Three integrators are provided:
  * ``run(..., scheme="cn")``: the reference's scheme above (its BiCGSTAB,
    mesh.bicgstab); ``sweeps`` then holds the BiCGSTAB iterations of every
    step.  The GPU's FP64 CN path with its BiCGSTAB solver is compared with
    this one at round-off level (the inner products are summed in another
    order, so the iterates agree to round-off, not bit for bit);
  * ``run(..., scheme="cn_jacobi")``: the same CN step with the linear system
    solved as the HIP kernels solve it -- Jacobi sweeps x <- x + (b - A x)/diag A
    from x0 = V* (with ``warm``, the default: x0 = V* + e, e the previous step's
    increment x - V* extrapolated linearly from the last two steps, the first
    residual then taken as L V* + theta L e - (Cm/dt) e), the first residual
    otherwise L V* (= b - A V*), after each sweep the rows of the junction
    nodes re-solved from the new patch values and the old cable-neighbour
    values, stopping after the sweep whose residual meets
    ||r|| <= cg_tol*||b|| or after ``max_sweeps``; the GPU's FP64 CN path is
    compared with this one at round-off level, and this one with ``cn``
    through the solver tolerance;
  * ``run(..., scheme="explicit")``: the scheme of the HIP kernels -- the
    same reaction step with forward-Euler diffusion on V^n (PMJ entries
    excluded from L), clamp, then the exact two-compartment PMJ exchange.
    GPU results are compared with this one at round-off level and with the
    CN one through activation-time / Vm-trace tolerances.

Activation times follow the reference's frame labels (a snapshot taken after
step i is stamped t = i*dt, aniso.py:1750,1806-1810) and its interpolation
(activation_maps.py:63-89), evaluated on every step.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import mesh as M
from . import ttp06 as O


class Problem:
    """Everything the time loop needs, built from the compute_anisotropic_coupled inputs."""

    def __init__(self, nodes, elements, act_times, S, phi, fibre_dir, *, voxel_size=0.4,
                 dx_p=0.2, sigma_P=0.45, sigma_l=3 * 1.31e-4, sigma_t=1.31e-4, Cm=0.01,
                 A_M=140.0, R_P=0.015, A_P=None, theta=0.5, stim_protocol=None,
                 ectopic_region=None, stim_regions=None, stim_len_mm=2.0, c_pmj=0.05, n_pmj=1,
                 T=400.0, phi_endo_max=0.35, phi_epi_min=0.65, isotropic_sigma=None):
        self.Cm, self.theta = Cm, theta
        nodes_mm = np.asarray(nodes, dtype=np.float32) * voxel_size
        self.comp_nodes, self.comp_edges, self.branch_map = M.discretise_tree(
            nodes_mm, np.asarray(elements), dx_p)
        self.Np = Np = len(self.comp_nodes)
        root = M.root_stim_mask(nodes_mm, act_times, self.comp_nodes, T, stim_len_mm)
        pc = M.purkinje_constants(R_P, sigma_P, dx_p, A_M, voxel_size, A_P)
        self.Ai_P, self.Ai_M, self.w_pp = pc["Ai_P"], pc["Ai_M"], pc["w_pp"]
        if isotropic_sigma is not None:
            myo = M.build_myo_mesh_iso(S, voxel_size, isotropic_sigma)
        else:
            myo = M.build_myo_mesh(S, voxel_size, fibre_dir, sigma_l, sigma_t)
        self.myo = myo
        self.N_myo = N_myo = myo["N_myo"]
        vox = myo["vox_idx"]
        ph = phi[vox[:, 0], vox[:, 1], vox[:, 2]].astype(np.float32)
        ph = np.where(np.isfinite(ph), ph, 0.5)
        ct = np.ones(N_myo, dtype=np.int64)  # mid
        ct[ph < phi_endo_max] = 0
        ct[ph >= phi_epi_min] = 2
        self.celltype = np.concatenate([np.zeros(Np, dtype=np.int64), ct])
        self.pmj_ids = M.pmj_comp_ids(np.asarray(elements), self.branch_map, len(nodes), Np)
        self.pmj_patches = M.map_pmj(self.comp_nodes, self.pmj_ids, myo["nodes_mm"], n_pmj)
        self.c_pmj = c_pmj
        self.L, self.diag_L = M.global_laplacian(
            Np, N_myo, self.comp_edges, self.w_pp, self.Ai_P, myo["edges"], myo["edge_w"],
            self.Ai_M, self.pmj_ids, self.pmj_patches, c_pmj)
        if stim_protocol is None:
            stim_protocol = [dict(name="av_node", target="root", amp=12.0, dur=2.0, onset=0.0)]
        self.sites = M.resolve_stim_sites(stim_protocol, root, ectopic_region, vox, Np, N_myo,
                                          stim_regions)
        self.N = Np + N_myo

    @property
    def L_nopmj(self):
        """L without the PMJ entries (the explicit scheme integrates those separately)."""
        if not hasattr(self, "_L_nopmj"):
            self._L_nopmj, _ = M.global_laplacian(
                self.Np, self.N_myo, self.comp_edges, self.w_pp, self.Ai_P, self.myo["edges"],
                self.myo["edge_w"], self.Ai_M, [], [], 0.0)
        return self._L_nopmj

    def stim_vector(self, t):
        I = np.zeros(self.N)
        for s in self.sites:
            if M.stim_active(t, s["onset"], s["dur"], s["period"], s["n_pulses"]) and s["node_ids"].size:
                I[s["node_ids"]] += s["amp"]
        return I

    def pmj_exchange(self, V, dt, vmin=-150.0, vmax=80.0):
        """Exact two-compartment PMJ exchange over dt (the HIP kernels' scheme)."""
        if self.c_pmj == 0.0 or not self.pmj_ids:
            return V
        Np, Cm = self.Np, self.Cm
        V = V.copy()
        delta = np.zeros(self.N)
        for p, patch in zip(self.pmj_ids, self.pmj_patches):
            idx = Np + np.asarray(patch)
            n = len(idx)
            vm = V[idx]
            mean = vm.mean()
            Cp, Cq = self.Ai_P * Cm, n * self.Ai_M * Cm
            vbar = (Cp * V[p] + Cq * mean) / (Cp + Cq)
            r1 = math.exp(-self.c_pmj * dt * (1 / Cp + 1 / Cq))
            r2 = math.exp(-(self.c_pmj / n) * dt / (self.Ai_M * Cm))
            vp_new = vbar + (V[p] - vbar) * r1
            mean_new = vbar + (mean - vbar) * r1
            np.add.at(delta, idx, mean_new + (vm - mean) * r2 - vm)
            V[p] = min(max(vp_new, vmin), vmax)
        touched = np.unique(np.concatenate([Np + np.asarray(p) for p in self.pmj_patches]))
        V[touched] = np.clip(V[touched] + delta[touched], vmin, vmax)
        return V


def cn_jacobi_solve(A, M_inv, L, Vs, b, tol, max_sweeps, junctions=(), Np=0, e=None,
                    theta=0.5, cm_dt=None):
    """The HIP path's CN solve (see the module docstring); returns (x, sweeps).

    ``junctions``: the cable nodes with PMJ couplings (global ids < Np); after
    each sweep row p is solved for x_p with its myocardial entries at the new
    iterate and its cable entries at the sweep's input.  ``e`` (optional) is the
    warm start: x_0 = V* + e, whose residual b - A x_0 is formed as
    L V* + theta L e - (Cm/dt) e (L V* alone when e is None; cm_dt = Cm/dt)."""
    bb = float(b @ b)
    if e is None:
        x = Vs.copy()
        r = L @ Vs
    else:
        x = Vs + e
        r = (L @ Vs) + theta * (L @ e) - cm_dt * e
    Ad = A.diagonal()
    rows = [A.getrow(p) for p in junctions]
    for k in range(1, max_sweeps + 1):
        if k > 1:
            r = b - A @ x
        xn = x + r * M_inv
        for p, row in zip(junctions, rows):
            idx, val = row.indices, row.data
            off = idx != p
            cab = off & (idx < Np)
            myo = off & (idx >= Np)
            xn[p] = (b[p] - val[cab] @ x[idx[cab]] - val[myo] @ xn[idx[myo]]) / Ad[p]
        x = xn
        if float(r @ r) <= tol * tol * bb:
            return x, k
    return x, max_sweeps


def run(prob: Problem, T, dt, *, scheme="cn", cg_tol=1e-5, cg_max_iter=500, dtype=torch.float64,
        record_every=0, record_nodes=None, at_thr=-40.0, max_steps=None, state_init=None,
        V_init=None, g_ks_scale=None, g_to_scale=None, max_sweeps=10, dot=None, warm=True,
        progress=None):
    """Integrate the problem; returns dict(V, state, at, trace_t, trace_V, steps, wall_s).

    ``at`` is (N,) first activation time (NaN if none), on every node.
    ``trace_V`` holds V of ``record_nodes`` (all nodes if None) every
    ``record_every`` steps (0 = off), stamped like the reference's frames.
    ``g_ks_scale``/``g_to_scale``: optional per-node multipliers of the
    phenotype conductances (the ensemble parameter block).
    ``progress``: optional callable(step, n_steps), called every 1000 steps.
    """
    N, Cm = prob.N, prob.Cm
    n_steps = int(T / dt)
    if max_steps is not None:
        n_steps = min(n_steps, max_steps)
    V = np.full(N, O.V_REST) if V_init is None else np.array(V_init, dtype=np.float64)
    st = O.init_state(N, dtype) if state_init is None else state_init.clone()
    ct = torch.from_numpy(prob.celltype)
    gks_s = None if g_ks_scale is None else torch.as_tensor(np.asarray(g_ks_scale), dtype=dtype)
    gto_s = None if g_to_scale is None else torch.as_tensor(np.asarray(g_to_scale), dtype=dtype)
    sweeps = []
    d = e = None  # cn_jacobi warm start: last increment x - V*, its extrapolation
    if scheme in ("cn", "cn_jacobi"):
        A, M_inv = M.cn_matrix(prob.L, prob.diag_L, Cm / dt, prob.theta)
    at = np.full(N, np.nan)
    tr_t, tr_V = [], []
    t0 = time.perf_counter()
    for i in range(n_steps):
        t = i * dt
        I_stim = prob.stim_vector(t)
        Iion, st = O.step_celltype(torch.from_numpy(V).to(dtype), st, dt, ct, gks_s, gto_s)
        Iion = Iion.double().numpy()
        if scheme == "cn":
            Vs = V + (dt / Cm) * (-Iion * Cm + I_stim)
            b = (Cm / dt) * Vs + (1 - prob.theta) * (prob.L @ Vs)
            Vn, k = M.bicgstab(A, M_inv, b, Vs, cg_tol, cg_max_iter, return_iters=True, dot=dot)
            sweeps.append(k)
            Vn = np.clip(Vn, -150.0, 80.0)
        elif scheme == "cn_jacobi":
            Vs = V + (dt / Cm) * (-Iion * Cm + I_stim)
            b = (Cm / dt) * Vs + (1 - prob.theta) * (prob.L @ Vs)
            jn = sorted(set(int(q) for q in prob.pmj_ids)) if prob.c_pmj != 0.0 else []
            Vn, k = cn_jacobi_solve(A, M_inv, prob.L, Vs, b, cg_tol, max_sweeps, jn, prob.Np,
                                    e if warm else None, prob.theta, Cm / dt)
            sweeps.append(k)
            if warm:  # next start: the increment, extrapolated from the last two steps
                dn = Vn - Vs
                e = dn if d is None else dn + (dn - d)
                d = dn
            Vn = np.clip(Vn, -150.0, 80.0)
        elif scheme == "explicit":
            Vn = V - dt * Iion + (dt / Cm) * (I_stim + prob.L_nopmj @ V)
            Vn = np.clip(Vn, -150.0, 80.0)
            Vn = prob.pmj_exchange(Vn, dt)
        else:
            raise ValueError(scheme)
        cross = (V < at_thr) & (Vn >= at_thr) & np.isnan(at)
        if cross.any():
            at[cross] = (i - 1) * dt + dt * (at_thr - V[cross]) / (Vn[cross] - V[cross])
        V = Vn
        if progress is not None and (i + 1) % 1000 == 0:
            progress(i + 1, n_steps)
        if record_every and i % record_every == 0:
            tr_t.append(t)
            tr_V.append(V.copy() if record_nodes is None else V[record_nodes].copy())
    return dict(V=V, state=st, at=at, trace_t=np.array(tr_t),
                trace_V=np.array(tr_V) if tr_V else None, steps=n_steps,
                wall_s=time.perf_counter() - t0, sweeps=np.array(sweeps, dtype=np.int64))
