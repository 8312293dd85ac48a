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

from .forward import Electrodes, HeartSource, PoissonSolver, leads_from_electrodes


class CoupledECG:
    def __init__(self, eng, meta, grid, registration, electrodes_mm, *, sigma_i=0.2625,
                 h_fine_mm=None, tol=1e-8, inner_tol=1e-4, k_nearest=4):
        if eng.dtype != torch.float32:
            raise ValueError("the in-loop ECG reads the FP32 engine's Vm")
        self.eng = eng
        self.tol, self.inner_tol = tol, inner_tol
        h = float(h_fine_mm if h_fine_mm is not None else meta.get("voxel_size", 0.4))
        self.ps = PoissonSolver(grid, eng.dev)
        self.src = HeartSource(eng.lay, meta["nodes_mm"], grid, registration, sigma_i, h, eng.dev,
                               bricks=eng.bricks)
        self.el = Electrodes(grid, electrodes_mm, k=k_nearest, device=eng.dev)
        self.b = torch.zeros(self.ps.n, dtype=torch.float64, device=eng.dev)
        self.x = torch.zeros(self.ps.n, dtype=torch.float64, device=eng.dev)
        self.stats = []

    def solve(self):
        """Source from the current Vm, refined solve, electrode potentials (FP64, V)."""
        self.src(self.eng.V[self.eng.p], self.b, self.ps.ws)
        self.ps.solve_refined(self.b, self.x, tol=self.tol, inner_tol=self.inner_tol)
        self.stats.append(dict(self.ps.last_stats))
        return self.el(self.x)

    def run(self, T_ms, ecg_every_ms=1.0, graph_chunk=50):
        """Advance the engine by T_ms, solving the ECG every ecg_every_ms.
        Returns (t_ms (n,), electrode potentials (n, n_el) in mV, wall seconds)."""
        eng = self.eng
        every = int(round(ecg_every_ms / eng.dt))
        n = int(round(T_ms / ecg_every_ms))
        ts, phis = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(n):
            eng.run(every, graph_chunk=graph_chunk if every >= graph_chunk else 0)
            phis.append((self.solve() * 1e3).clone())
            ts.append(eng.step_index * eng.dt)
        phi = torch.stack(phis).cpu().numpy()
        torch.cuda.synchronize()
        return np.array(ts), phi, time.perf_counter() - t0

    def leads(self, phi):
        d = {n: phi[:, i] for i, n in enumerate(self.el.names)}
        L = leads_from_electrodes(d)
        return np.stack([L[k] for k in ("I", "II", "III", "aVR", "aVL", "aVF", "V1", "V2", "V3",
                                        "V4", "V5", "V6")], axis=1)
