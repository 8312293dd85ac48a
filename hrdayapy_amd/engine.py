"""Device-resident monodomain engine: owns the HBM buffers and launches the
HIP kernels of one time step (myocardium fused step, Purkinje cable step,
PMJ exchange, stimulus/step-counter advance), optionally captured in a HIP
graph of K steps.

Global node numbering follows the reference: Purkinje computational nodes
0..Np-1, then myocardial nodes Np..Np+N_myo-1 in ``np.argwhere(S)`` order
(purkinje_myocardium_anisotropic_solver.py:611-613).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as nat
from .layout import BrickLayout
from .ttp06 import STATE_NAMES, INIT, V_REST

KI_BASE, NAI_BASE = INIT["Ki"], INIT["Nai"]


def stim_groups(site_node_ids, n_nodes):
    """Assign each node the set of sites it belongs to, encoded as a group id.

    Returns (group (n_nodes,) int, group_sites (n_groups, 2) uint32 pairs of
    (site base, site bitmask)); at most 32 sites and 31 site combinations.
    """
    if len(site_node_ids) > 32:
        raise ValueError("at most 32 stimulus sites (use ensemble instances for more)")
    mask = np.zeros(n_nodes, dtype=np.uint64)
    for s, ids in enumerate(site_node_ids):
        ids = np.asarray(ids, dtype=np.int64)
        if ids.size:
            mask[ids] |= np.uint64(1 << s)
    uniq, inv = np.unique(mask, return_inverse=True)
    nz = uniq != 0
    gid = np.zeros(len(uniq), dtype=np.int64)
    gid[nz] = np.arange(1, int(nz.sum()) + 1)
    if nz.sum() > 31:
        raise ValueError("more than 31 distinct stimulus-site combinations")
    group_sites = np.zeros((32, 2), dtype=np.uint32)
    group_sites[gid[nz], 1] = uniq[nz].astype(np.uint32)
    return gid[inv], group_sites


def ensemble_stim_groups(inst_of_node, site_inst, site_node_ids, n_inst):
    """Per-instance stimulus groups: each instance has at most 3 sites; the group
    of a node is (inst << 3) | local combination (1..7).  Returns (local group
    per node, group_sites (8*n_inst, 2))."""
    n_nodes = len(inst_of_node)
    local = np.zeros(n_nodes, dtype=np.int64)
    group_sites = np.zeros((8 * n_inst, 2), dtype=np.uint32)
    order = np.argsort(site_inst, kind="stable")
    base = np.zeros(n_inst + 1, dtype=np.int64)
    np.add.at(base, np.asarray(site_inst)[order] + 1, 1)
    base = np.cumsum(base)
    for i in range(n_inst):
        sites = order[base[i]:base[i + 1]]
        if len(sites) > 3:
            raise ValueError("at most 3 stimulus sites per ensemble instance")
        for b, s in enumerate(sites):
            local[np.asarray(site_node_ids[s], dtype=np.int64)] |= 1 << b
        for c in range(1, 8):
            group_sites[8 * i + c] = (base[i], c)
    return local, group_sites, order


class MonodomainEngine:
    """One coupled Purkinje + myocardium monodomain system on one GPU.

    Parameters are in the reference's units (mm, ms, mV, mS, uF, uA).
    ``D6`` is (6, N_myo) D00,D11,D22,D01,D02,D12 in mS/mm; ``celltype_myo``
    phenotype codes 0/1/2; ``cable`` dict(n, edges, w_pp, Ai_P, celltype);
    ``pmj`` dict(ids, patches, c_pmj) with patches in myocardial node ids;
    ``stim_sites`` list of dict(node_ids (global), amp, dur, onset, period, n_pulses).

    ``scheme``: "explicit" (forward-Euler diffusion fused with the reaction,
    exact PMJ exchange; stable below ``tissue.explicit_dt_limit``) or "cn"
    (the reference's Crank-Nicolson step with ``theta``;
    purkinje_myocardium_anisotropic_solver.py:757-849, 1746-1791).  The CN
    system is solved by ``cn_solver`` "bicgstab" (the reference's
    Jacobi-preconditioned BiCGSTAB with its stopping rules at ``cg_tol``, at
    most ``max_sweeps`` iterations per step: the reference's iterates) or
    "jacobi" (Jacobi sweeps to the same relative residual, at most
    ``max_sweeps``: the exact CN solution to well within the tolerance, where
    the reference's BiCGSTAB stops with an error of its own at the fronts).
    With ``cn_warm`` (Jacobi only) the sweeps start from V* plus the last
    steps' increment x - V*, extrapolated linearly, instead of V*: the same
    stopping rule is met after fewer sweeps (a step's increment changes little
    from one step to the next).
    """

    def __init__(self, layout: BrickLayout, D6, celltype_myo, *, h, A_M, Cm, dt,
                 dtype=torch.float32, device="cuda", cable=None, pmj=None, stim_sites=None,
                 at_thr=-40.0, vmin=-150.0, vmax=80.0, V0=V_REST, inst=None, inst_params=None,
                 alloc="torch", scheme="explicit", theta=0.5, cg_tol=1e-5, max_sweeps=10,
                 cn_solver="bicgstab", cn_warm=True):
        nat.load()
        if scheme not in ("explicit", "cn"):
            raise ValueError(f"unknown scheme {scheme!r} (explicit or cn)")
        self.scheme = scheme
        self.lay = layout
        self.dtype = dtype
        self.dev = torch.device(device)
        self.sfx = nat.sfx(dtype)
        self.dt, self.Cm = float(dt), float(Cm)
        self.vmin, self.vmax = float(vmin), float(vmax)
        self.N_myo = layout.n_nodes
        self.Np = int(cable["n"]) if cable else 0
        dev, dt_ = self.dev, dtype
        f64 = dtype == torch.float64
        self.ki_base, self.nai_base = (0.0, 0.0) if f64 else (KI_BASE, NAI_BASE)

        # stimulus groups over the global node set
        sites = stim_sites or []
        if inst is not None:
            if self.Np:
                raise ValueError("ensemble instances run without a Purkinje cable")
            n_inst = int(np.asarray(inst).max()) + 1
            gid, group_sites, order = ensemble_stim_groups(
                np.asarray(inst), [s["inst"] for s in sites], [s["node_ids"] for s in sites],
                n_inst)
            sites = [sites[k] for k in order]
        else:
            gid, group_sites = stim_groups([s["node_ids"] for s in sites], self.Np + self.N_myo)
        sp = np.zeros((max(1, len(sites)), 5))
        for k, s in enumerate(sites):
            per = s.get("period") or 0.0
            npl = s.get("n_pulses")
            sp[k] = (s.get("onset", 0.0), s["dur"], per, -1.0 if npl is None else npl, s["amp"])
        self.n_sites = len(sites)
        self.n_groups = len(group_sites)
        self.site_params = torch.tensor(sp.ravel(), dtype=torch.float64, device=dev)
        self.group_sites = torch.from_numpy(group_sites.ravel().view(np.int32).copy()).to(dev)
        self.group_amp = torch.zeros(self.n_groups, dtype=dt_, device=dev)
        self.step_ctr = torch.zeros(1, dtype=torch.int64, device=dev)

        # myocardium
        L = layout
        lab = L.labels(np.asarray(celltype_myo), gid[self.Np:])
        if pmj and self.Np and pmj["c_pmj"] != 0.0 and len(pmj["ids"]):
            tgt_nodes = np.unique(np.concatenate([np.asarray(p, np.int64) for p in pmj["patches"]]))
            lab[L.slot_of_node[tgt_nodes]] |= np.uint8(0x80)
        # the resident arrays share one allocation (2 MiB-aligned views); the
        # cell state holds the myocardial slots only (no brick padding)
        n = L.n_slots
        groups, state_of_node = L.state_groups()
        self.state_ld = -(-L.n_nodes // 64) * 64
        specs = [("state", (18, self.state_ld), dt_), ("D6", (6, n), dt_), ("V0", n, dt_),
                 ("V1", n, dt_), ("at", n, dt_), ("label", n, torch.uint8),
                 ("groups", (len(groups), 4), torch.int32)]
        if inst is not None:
            specs.append(("inst", n, torch.int16))
        self.arena = nat.DeviceArena(specs, dev, contiguous=(alloc == "contiguous"))
        ar = self.arena
        self.label = ar["label"]
        self.label.copy_(torch.from_numpy(lab))
        self.D6 = ar["D6"]
        self.D6.copy_(torch.from_numpy(L.to_slots(np.asarray(D6, dtype=np.float64))).to(dt_))
        self.nbr = torch.from_numpy(L.nbr).to(dev)
        self.slot_of_node = torch.from_numpy(L.slot_of_node).to(dev)
        self.groups = ar["groups"]
        self.groups.copy_(torch.from_numpy(groups))
        self.state_of_node = torch.from_numpy(state_of_node).to(dev)
        self.V = [ar["V0"], ar["V1"]]
        for v in self.V:
            v.fill_(V0)
        self.state = self._init_state(ar["state"])
        self.at = ar["at"]
        self.at.fill_(float("nan"))
        self.inst = None
        if inst is not None:
            self.inst = ar["inst"]
            self.inst.copy_(torch.from_numpy(
                L.to_slots(np.asarray(inst, dtype=np.uint16)).view(np.int16)))
        self.inst_params = None if inst_params is None else torch.as_tensor(
            np.asarray(inst_params, dtype=np.float64).reshape(-1, 4)).to(dev, dt_).contiguous()
        self.bricks = nat.HpBricks(L.n_bricks, *L.brick, 0, nat.ptr(self.nbr), nat.ptr(self.label),
                                   nat.ptr(self.D6), nat.ptr(self.inst), nat.ptr(self.groups))
        self.p_myo = self._params(1.0 / (A_M * h * h), at_thr, self.at)
        # the myocardial slots in compact-state order: the CN reaction pass runs
        # one lane per cell over them (no idle padding lanes)
        self.cell_slot = torch.from_numpy(
            np.sort(np.asarray(L.slot_of_node)).astype(np.int32)).to(dev)
        self.p_myo.cell_slot = nat.ptr(self.cell_slot)
        self.p_myo.n_cells = int(self.cell_slot.numel())

        # Purkinje cable
        self.cable = None
        if self.Np:
            n = self.Np
            e = np.asarray(cable["edges"], dtype=np.int64).reshape(-1, 2)
            a = np.concatenate([e[:, 0], e[:, 1]])
            b = np.concatenate([e[:, 1], e[:, 0]])
            order = np.lexsort((b, a))
            a, b = a[order], b[order]
            indptr = np.zeros(n + 1, dtype=np.int32)
            np.add.at(indptr, a + 1, 1)
            indptr = np.cumsum(indptr).astype(np.int32)
            ctp = np.asarray(cable.get("celltype", np.zeros(n)), dtype=np.uint8) & 3
            lab_pk = ctp | (gid[:n].astype(np.uint8) << 2)
            self.pk_indptr = torch.from_numpy(indptr).to(dev)
            self.pk_indices = torch.from_numpy(b.astype(np.int32)).to(dev)
            self.pk_label = torch.from_numpy(lab_pk.astype(np.uint8)).to(dev)
            self.Vpk = [torch.full((n,), V0, dtype=dt_, device=dev) for _ in range(2)]
            self.state_pk = self._init_state(torch.empty((18, n), dtype=dt_, device=dev))
            self.cable = nat.HpCable(n, nat.ptr(self.pk_indptr), nat.ptr(self.pk_indices),
                                     nat.ptr(self.pk_label), float(cable["w_pp"]))
            self.p_pk = self._params(1.0 / float(cable["Ai_P"]), at_thr, None)

        # PMJ
        self.pmj = None
        if pmj and self.Np and pmj["c_pmj"] != 0.0 and len(pmj["ids"]):
            patches = [np.asarray(p, dtype=np.int64) for p in pmj["patches"]]
            ptrs = np.zeros(len(patches) + 1, dtype=np.int32)
            ptrs[1:] = np.cumsum([len(p) for p in patches])
            entries = L.slot_of_node[np.concatenate(patches)].astype(np.int32)
            order = np.argsort(entries, kind="stable")
            tgt, first = np.unique(entries[order], return_index=True)
            tptr = np.append(first, len(entries)).astype(np.int32)
            self.pmj_t = [torch.from_numpy(x).to(dev) for x in (
                np.asarray(pmj["ids"], dtype=np.int32), ptrs, entries, tgt.astype(np.int32), tptr,
                order.astype(np.int32))]
            self.pmj_scratch = torch.zeros(len(entries), dtype=dt_, device=dev)
            # per-junction constants of the exact exchange over dt (FP64)
            npatch = np.diff(ptrs).astype(np.float64)
            c, Cp = float(pmj["c_pmj"]), float(pmj["Ai_P"]) * self.Cm
            Cq = npatch * float(pmj["Ai_M"]) * self.Cm
            coef = np.stack([Cp / (Cp + Cq), np.exp(-c * self.dt * (1.0 / Cp + 1.0 / Cq)),
                             np.exp(-(c / npatch) * self.dt / (float(pmj["Ai_M"]) * self.Cm)),
                             1.0 / npatch], axis=1)
            self.pmj_coef = torch.from_numpy(np.ascontiguousarray(coef)).to(dev)
            pk_, ptr_, ent_, tgt_, tptr_, tent_ = self.pmj_t
            self.pmj = nat.HpPmj(len(patches), len(entries), nat.ptr(pk_), nat.ptr(ptr_),
                                 nat.ptr(ent_), len(tgt), nat.ptr(tgt_), nat.ptr(tptr_),
                                 nat.ptr(tent_), float(pmj["c_pmj"]), float(pmj["Ai_P"]),
                                 float(pmj["Ai_M"]), self.Cm, nat.ptr(self.pmj_coef),
                                 int(len(tgt) == len(entries)), 0)
            self.pmj_at = [nat.HpPmjAt(nat.ptr(self.at), nat.ptr(self.V[k]),
                                       nat.ptr(self.step_ctr), self.dt, float(at_thr))
                           for k in range(2)]
        self.cn = None
        if scheme == "cn":
            self._setup_cn(L, pmj, float(h), float(A_M), float(theta), float(cg_tol),
                           int(max_sweeps), cable, cn_solver, bool(cn_warm))
        self.p = 0
        self.overlap_cable = True
        self.fuse_cable = True
        self._side = torch.cuda.Stream(device=dev) if self.Np else None
        self._eval_stim()

    # ------------------------------------------------------------------ setup
    def _setup_cn(self, L, pmj, h, A_M, theta, tol, max_sweeps, cable, solver, warm=True):
        """Work vectors, PMJ rows and device scalars of the Crank-Nicolson step."""
        if not 0.0 <= theta <= 1.0:
            raise ValueError(f"theta must lie in [0, 1] (got {theta})")
        if not tol >= 0.0 or max_sweeps < 1:
            raise ValueError("cg_tol must be >= 0 and max_sweeps >= 1")
        if solver not in ("bicgstab", "jacobi"):
            raise ValueError(f"cn_solver must be 'bicgstab' or 'jacobi' (got {solver!r})")
        self.cn_solver = solver
        bicg = solver == "bicgstab"
        dev, dt_ = self.dev, self.dtype
        n, Np = L.n_slots, self.Np
        z = lambda m: torch.zeros(max(m, 1), dtype=dt_, device=dev)
        self.cn_work_t = ([z(n), None, None, z(Np), None, None] if bicg else
                          [z(n), z(n), z(n), z(Np), z(Np), z(Np)])
        # BiCGSTAB: r, r^, M^-1, t, p (2), v (2) per side
        self.cn_kr_t = [z(n) for _ in range(8)] if bicg else []
        self.cn_kr_pk_t = [z(Np) for _ in range(8)] if bicg and Np else []
        self.cn_scalars = torch.zeros(24, dtype=torch.float64, device=dev)
        # per pass workgroup (<= 520 cable): 2 partial columns (sweeps), 8 (BiCGSTAB)
        part_len = (8 if bicg else 2) * (min(L.n_bricks, 8192) + 1024)
        self.cn_part = torch.zeros(part_len, dtype=torch.float64, device=dev)
        Ai_P = float(cable["Ai_P"]) if Np else 1.0
        Ai_M = A_M * h ** 3
        self.cn_pmj = None
        if pmj and Np and pmj["c_pmj"] != 0.0 and len(pmj["ids"]):
            # every (junction node, patch voxel) pair: conductance c_pmj/|patch|
            # (purkinje_myocardium_anisotropic_solver.py:684-692)
            node, slot, w = [], [], []
            for q, patch in zip(pmj["ids"], pmj["patches"]):
                patch = np.asarray(patch, dtype=np.int64)
                node.append(np.full(len(patch), int(q), np.int64))
                slot.append(L.slot_of_node[patch].astype(np.int64))
                w.append(np.full(len(patch), float(pmj["c_pmj"]) / len(patch)))
            node, slot, w = np.concatenate(node), np.concatenate(slot), np.concatenate(w)
            o = np.lexsort((node, slot))  # myocardial side: by target slot, then node
            tgt, first = np.unique(slot[o], return_index=True)
            tptr = np.append(first, len(o)).astype(np.int32)
            o2 = np.lexsort((slot, node))  # cable side: by node, then slot
            nptr = np.zeros(Np + 1, dtype=np.int64)
            np.add.at(nptr, node[o2] + 1, 1)
            nptr = np.cumsum(nptr).astype(np.int32)
            jn = np.unique(node).astype(np.int32)
            tens = [tgt.astype(np.int32), tptr, node[o].astype(np.int32), w[o] / Ai_M, nptr,
                    slot[o2].astype(np.int32), w[o2] / Ai_P, jn]
            self.cn_pmj_t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in tens]
            t = self.cn_pmj_t
            self.cn_pmj = nat.HpCnPmj(len(tgt), *(nat.ptr(a) for a in t[:7]), len(jn),
                                      nat.ptr(t[7]))
        kr = (C.c_void_p * 8)(*(nat.ptr(a) for a in self.cn_kr_t)) if bicg else (C.c_void_p * 8)()
        kr_pk = ((C.c_void_p * 8)(*(nat.ptr(a) for a in self.cn_kr_pk_t)) if self.cn_kr_pk_t
                 else (C.c_void_p * 8)())
        # Jacobi sweeps start from V* plus the last steps' increment, extrapolated
        # (d, e per slot and per cable node; zero before the first step)
        self.cn_warm = warm and not bicg
        self.cn_warm_t = [z(n), z(n), z(Np), z(Np)] if self.cn_warm else []
        wv = ((C.c_void_p * 4)(*(nat.ptr(a) for a in self.cn_warm_t)) if self.cn_warm
              else (C.c_void_p * 4)())
        self.cn_work = nat.HpCnWork(*(nat.ptr(a) if a is not None else None
                                      for a in self.cn_work_t), kr, kr_pk, wv)
        self.cn = nat.HpCnParams(self.Cm / self.dt, theta, 1.0 / (A_M * h * h), 1.0 / Ai_P, tol,
                                 max_sweeps, 1 if bicg else 0, nat.ptr(self.cn_scalars),
                                 nat.ptr(self.cn_part), part_len)

    def cn_stats(self):
        """Solver statistics of the CN steps so far: steps, mean sweeps, steps that
        stopped at max_sweeps without meeting the tolerance and their worst
        relative residual (one host synchronisation)."""
        if self.cn is None:
            return None
        sc = self.cn_scalars.cpu().numpy()
        steps = int(sc[7])
        return dict(steps=steps, mean_sweeps=float(sc[6] / steps) if steps else 0.0,
                    max_used=int(sc[16]), not_converged=int(sc[4]),
                    worst_rel_res=float(sc[5]), max_sweeps=int(self.cn.max_sweeps),
                    tol=float(self.cn.tol), solver=self.cn_solver)

    def set_cn_max_iters(self, n: int):
        """Bound the CN solve at n sweeps (BiCGSTAB iterations) per step from the
        next step on; graphs captured with the old bound are dropped."""
        if self.cn is None or n < 1:
            raise ValueError("set_cn_max_iters needs a CN engine and n >= 1")
        if int(n) != self.cn.max_sweeps:
            self.cn.max_sweeps = int(n)
            self._b_cur = int(n)

    def _init_state(self, st):
        """Resting TTP06 state into st (18, n) (Ki, Nai as offsets from their bases)."""
        for k, name in enumerate(STATE_NAMES):
            v = INIT[name]
            if name == "Ki":
                v -= self.ki_base
            elif name == "Nai":
                v -= self.nai_base
            st[k].fill_(v)
        return st

    def _params(self, dscale, at_thr, at):
        return nat.HpMonoParams(
            self.dt, 1.0 / self.Cm, dscale, self.ki_base, self.nai_base, self.vmin, self.vmax,
            float(at_thr), 1, 0 if self.inst_params is None else self.inst_params.shape[0],
            nat.ptr(self.inst_params), nat.ptr(self.step_ctr), nat.ptr(self.group_amp),
            nat.ptr(at), None)

    def _eval_stim(self):
        nat.call(f"hp_stim_eval_{self.sfx}", nat.ptr(self.site_params), self.n_sites,
                 nat.ptr(self.group_sites), self.n_groups, nat.ptr(self.step_ctr), self.dt,
                 nat.ptr(self.group_amp), nat.stream_handle())

    # ------------------------------------------------------------------- step
    def _launch_step(self):
        """One time step: the myocardium and the cable in one launch
        (hp_myo_cable_step; with fuse_cable off, the cable kernel on a forked
        stream or after the myocardium kernel);
        one tail kernel then does the PMJ exchange and advances the step
        counter / stimulus amplitudes."""
        cur = torch.cuda.current_stream(self.dev)
        s = cur.cuda_stream
        a, b = self.p, 1 - self.p
        if self.cn is not None:
            cab = self.cable is not None
            nat.call(f"hp_cn_step_{self.sfx}", C.byref(self.bricks), nat.ptr(self.V[a]),
                     nat.ptr(self.V[b]), nat.ptr(self.state), self.state_ld, C.byref(self.p_myo),
                     C.byref(self.cable) if cab else None,
                     nat.ptr(self.Vpk[a]) if cab else None, nat.ptr(self.Vpk[b]) if cab else None,
                     nat.ptr(self.state_pk) if cab else None, self.Np,
                     C.byref(self.p_pk) if cab else None,
                     C.byref(self.cn_pmj) if self.cn_pmj is not None else None,
                     C.byref(self.cn), C.byref(self.cn_work), s)
            nat.call(f"hp_step_tail_{self.sfx}", None, nat.ptr(self.Vpk[b]) if cab else None,
                     nat.ptr(self.V[b]), None, self.dt, self.vmin, self.vmax, None,
                     nat.ptr(self.site_params), self.n_sites, nat.ptr(self.group_sites),
                     self.n_groups, nat.ptr(self.step_ctr), nat.ptr(self.group_amp), s)
            self.p = b
            return
        if self.cable is not None and self.fuse_cable:
            nat.call(f"hp_myo_cable_step_{self.sfx}", C.byref(self.bricks), nat.ptr(self.V[a]),
                     nat.ptr(self.V[b]), nat.ptr(self.state), self.state_ld, C.byref(self.p_myo),
                     C.byref(self.cable), nat.ptr(self.Vpk[a]), nat.ptr(self.Vpk[b]),
                     nat.ptr(self.state_pk), self.Np, C.byref(self.p_pk), s)
            self._launch_tail(a, b, s)
            return
        fork = self.cable is not None and self.overlap_cable
        if fork:
            self._side.wait_stream(cur)
            self._launch_cable(a, b, self._side.cuda_stream)
        nat.call(f"hp_myo_step_{self.sfx}", C.byref(self.bricks), nat.ptr(self.V[a]),
                 nat.ptr(self.V[b]), nat.ptr(self.state), self.state_ld, C.byref(self.p_myo), s)
        if self.cable is not None:
            if fork:
                cur.wait_stream(self._side)
            else:
                self._launch_cable(a, b, s)
        self._launch_tail(a, b, s)

    def _launch_tail(self, a, b, s):
        nat.call(f"hp_step_tail_{self.sfx}", None if self.pmj is None else C.byref(self.pmj),
                 nat.ptr(self.Vpk[b]) if self.Np else None, nat.ptr(self.V[b]),
                 nat.ptr(self.pmj_scratch) if self.pmj is not None else None, self.dt, self.vmin,
                 self.vmax, C.byref(self.pmj_at[a]) if self.pmj is not None else None,
                 nat.ptr(self.site_params), self.n_sites, nat.ptr(self.group_sites),
                 self.n_groups, nat.ptr(self.step_ctr), nat.ptr(self.group_amp), s)
        self.p = b

    def _launch_cable(self, a, b, s):
        nat.call(f"hp_cable_step_{self.sfx}", C.byref(self.cable), nat.ptr(self.Vpk[a]),
                 nat.ptr(self.Vpk[b]), nat.ptr(self.state_pk), self.Np, C.byref(self.p_pk), s)

    def step(self, n=1):
        for _ in range(n):
            self._launch_step()

    def capture(self, k: int):
        """Capture k (even) steps into a HIP graph; returns the graph."""
        if k <= 0 or k % 2:
            raise ValueError("graph chunk must be a positive even number of steps")
        p0 = self.p
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                for _ in range(k):
                    self._launch_step()
        torch.cuda.current_stream(self.dev).wait_stream(side)
        self.p = p0
        return g

    def _graph(self, graph_chunk):
        key = ("graph", graph_chunk, self.p, self.fuse_cable, self.overlap_cable,
               self.cn.max_sweeps if self.cn is not None else 0)
        if not hasattr(self, "_graphs"):
            self._graphs = {}
        if key not in self._graphs:
            self._graphs[key] = self.capture(graph_chunk)
        return self._graphs[key]

    def run(self, n_steps: int, graph_chunk: int = 0, adaptive: bool = False):
        """Advance n_steps; with graph_chunk > 0, replay a captured graph of that many steps.

        A graph holds the buffer pointers of the V/V_pk ping-pong as they stood
        when it was captured, so graphs are kept per parity of the current
        buffer (and per cable launch mode and sweep bound): a replay after an
        odd number of eager steps uses the graph captured from that parity.

        adaptive (CN with Jacobi sweeps, graph_chunk > 0): every step of a
        graph launches as many sweeps as its bound, also after the solve has
        converged (those return at once, but each launch costs a few
        microseconds).  Each chunk is therefore replayed with the bound the
        previous chunk needed; the chunk's state is kept first, and a chunk in
        which a step stopped at that bound without meeting the tolerance is
        replayed from it with the full bound (max_sweeps), so every step still
        meets the stopping rule.  One host synchronisation per chunk."""
        if adaptive and graph_chunk and self.cn is not None and self.cn_solver == "jacobi":
            n_steps = self._run_adaptive(n_steps, graph_chunk)
        elif graph_chunk and n_steps >= graph_chunk:
            g = self._graph(graph_chunk)
            for _ in range(n_steps // graph_chunk):
                g.replay()
            n_steps %= graph_chunk
        self.step(n_steps)

    def _mutable(self):
        """The device buffers a CN step writes that outlive the step."""
        ts = [self.V[self.p], self.state, self.at, self.group_amp, self.step_ctr, self.cn_scalars]
        if self.Np:
            ts += [self.Vpk[self.p], self.state_pk]
        return ts + list(self.cn_warm_t)

    def _run_adaptive(self, n_steps, k):
        self._run_chunks_adaptive(n_steps // k, lambda: self._graph(k).replay())
        return n_steps % k

    def _run_chunks_adaptive(self, n_chunks, replay, extra=()):
        """Replay n_chunks graph chunks, each with the sweep bound the previous
        one needed; `replay` replays one chunk captured at self.cn.max_sweeps,
        `extra` lists further device buffers the chunk writes (kept with the
        engine's own for a replay from the chunk's start)."""
        b_max = self.cn.max_sweeps
        b = min(getattr(self, "_b_cur", b_max), b_max)
        sc = self.cn_scalars
        n0 = float(sc[4].item()) if n_chunks > 0 else 0.0  # steps stopped at a bound so far
        try:
            while n_chunks > 0:
                keep = b < b_max
                if keep:
                    src = self._mutable() + list(extra)
                    if getattr(self, "_shadow", None) is None or len(self._shadow) != len(src) \
                            or any(d.shape != t.shape for d, t in zip(self._shadow, src)):
                        self._shadow = [torch.empty_like(t) for t in src]
                    for d, t in zip(self._shadow, src):
                        d.copy_(t)
                sc[18].zero_()
                self.cn.max_sweeps = b
                replay()
                n1, used = (float(v) for v in sc[[4, 18]].cpu().tolist())
                if keep and n1 > n0:  # a step stopped at the bound: again with the full one
                    for d, t in zip(self._shadow, self._mutable() + list(extra)):
                        t.copy_(d)
                    b = b_max
                    self.retries = getattr(self, "retries", 0) + 1
                    continue
                b = max(1, min(b_max, int(used)))
                n0 = n1
                n_chunks -= 1
        finally:
            self.cn.max_sweeps = b_max
            self._b_cur = b

    # --------------------------------------------------------------- readback
    @property
    def step_index(self) -> int:
        return int(self.step_ctr.item())

    def V_myo_nodes(self, out=None):
        """Myocardial V in reference node order (N_myo,)."""
        out = torch.empty(self.N_myo, dtype=self.dtype, device=self.dev) if out is None else out
        nat.call(f"hp_gather_{self.sfx}", nat.ptr(self.V[self.p]), nat.ptr(self.slot_of_node),
                 nat.ptr(out), self.N_myo, nat.stream_handle())
        return out

    def V_global(self):
        """[V_purkinje | V_myo] in the reference's global order."""
        vm = self.V_myo_nodes()
        if self.Np:
            return torch.cat([self.Vpk[self.p], vm])
        return vm

    def at_nodes(self):
        out = torch.empty(self.N_myo, dtype=self.dtype, device=self.dev)
        nat.call(f"hp_gather_{self.sfx}", nat.ptr(self.at), nat.ptr(self.slot_of_node),
                 nat.ptr(out), self.N_myo, nat.stream_handle())
        return out

    def state_nodes(self):
        """(N_global, 18) state in the reference's column order, absolute Ki/Nai."""
        sm = self.state[:, self.state_of_node.long()]
        if self.Np:
            sm = torch.cat([self.state_pk, sm], dim=1)
        sm = sm.t().contiguous()
        sm[:, 0] += self.ki_base
        sm[:, 1] += self.nai_base
        return sm
