"""ECG computed inside the monodomain time loop.

The reference computes the ECG offline from saved Vm snapshots
(compute_bspm.py:488-492).  Here the heart source is taken directly from the
engine's Vm in HBM every ``ecg_every_ms``, the torso potential is solved by
FP32 Jacobi-PCG inside FP64 iterative refinement (warm-started from the
previous solve), and the electrode potentials are accumulated in FP64 -- no
host transfer besides the electrode values.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from .forward import Electrodes, HeartSource, LeadField, PoissonSolver, leads_from_electrodes


class CoupledECG:
    """method "forward": source + refined torso solve + electrodes every frame;
    method "leadfield": the precomputed lead field (LeadField) applied to
    batches of up to 16 Vm frames copied from the engine."""

    def __init__(self, eng, meta, grid, registration, electrodes_mm, *, sigma_i=0.2625,
                 h_fine_mm=None, tol=1e-8, inner_tol=1e-4, k_nearest=4, method="forward",
                 precond="jacobi", leadfield_tol=1e-10, log=None):
        if eng.dtype != torch.float32:
            raise ValueError("the in-loop ECG reads the FP32 engine's Vm")
        if method not in ("forward", "leadfield"):
            raise ValueError(f"unknown ECG method {method!r}")
        self.eng = eng
        self.method = method
        self.tol, self.inner_tol = tol, inner_tol
        h = float(h_fine_mm if h_fine_mm is not None else meta.get("voxel_size", 0.4))
        self.ps = PoissonSolver(grid, eng.dev, precond=precond)
        self.src = HeartSource(eng.lay, meta["nodes_mm"], grid, registration, sigma_i, h, eng.dev,
                               bricks=eng.bricks)
        self.el = Electrodes(grid, electrodes_mm, k=k_nearest, device=eng.dev)
        self.b = torch.zeros(self.ps.n, dtype=torch.float64, device=eng.dev)
        self.x = torch.zeros(self.ps.n, dtype=torch.float64, device=eng.dev)
        self.stats = []
        self.lf = None
        if method == "leadfield":
            self.lf = LeadField(self.ps, self.src, self.el, tol=leadfield_tol, log=log)
            self.frames = torch.zeros((self.lf.batch, eng.lay.n_slots), dtype=torch.float32,
                                      device=eng.dev)

    def solve(self):
        """Source from the current Vm, refined solve, electrode potentials (FP64, V)."""
        self.b.zero_()
        self.src(self.eng.V[self.eng.p], self.b, self.ps.ws)
        self.ps.solve_refined(self.b, self.x, tol=self.tol, inner_tol=self.inner_tol)
        self.stats.append(dict(self.ps.last_stats))
        return self.el(self.x)

    def run(self, T_ms, ecg_every_ms=1.0, graph_chunk=50):
        """Advance the engine by T_ms, computing the ECG every ecg_every_ms.
        Returns (t_ms (n,), electrode potentials (n, n_el) in mV, wall seconds)."""
        eng = self.eng
        every = int(round(ecg_every_ms / eng.dt))
        n = int(round(T_ms / ecg_every_ms))
        phis = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nb = 0
        for k in range(n):
            eng.run(every, graph_chunk=graph_chunk if every >= graph_chunk else 0)
            if self.lf is None:
                phis.append((self.solve() * 1e3).clone())
            else:
                self.frames[nb].copy_(eng.V[eng.p])
                nb += 1
                if nb == self.lf.batch or k == n - 1:
                    phis.append((self.lf.apply(self.frames, nb) * 1e3).clone())
                    nb = 0
        phi = torch.cat([p.reshape(-1, len(self.el.names)) for p in phis]).cpu().numpy()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        t = eng.dt * every * np.arange(1, n + 1) + (eng.step_index - every * n) * eng.dt
        return t, phi, wall

    def leads(self, phi):
        d = {n: phi[:, i] for i, n in enumerate(self.el.names)}
        L = leads_from_electrodes(d)
        return np.stack([L[k] for k in ("I", "II", "III", "aVR", "aVL", "aVF", "V1", "V2", "V3",
                                        "V4", "V5", "V6")], axis=1)
