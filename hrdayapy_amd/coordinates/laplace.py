"""Graph-Laplace solves on the brick layout of a voxel mask (hp_laplace_cg_*).

The rows are the reference transmural solve's (generate_transmural_coordinate.py
:85-183): identity rows on the Dirichlet voxels, the 6-neighbour graph
Laplacian with Neumann faces elsewhere.  CG in FP32 with FP64 inner products
and the reference's stopping rule ||r||/||r0|| < tol (:35-78), its decisions
on the device; the host looks at the done flag every `check_every`
iterations.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _native as nat
from ..layout import INACTIVE, BrickLayout, choose_brick


class LaplaceSolver:
    def __init__(self, S, device="cuda", brick=None):
        nat.load()
        S = np.asarray(S).astype(bool)
        self.dev = dev = torch.device(device)
        self.lay = L = BrickLayout(S, brick or choose_brick(S))
        n = L.n_slots
        d6 = np.zeros((6, n), np.float32)
        d6[:3, L.slot_of_node] = 1.0                    # unit conductivity tensor
        self._d6 = torch.from_numpy(d6).to(dev)
        self._nbr = torch.from_numpy(L.nbr).to(dev)
        self._label = torch.full((n,), INACTIVE, dtype=torch.uint8, device=dev)
        self.bricks = nat.HpBricks(L.n_bricks, *L.brick, 0, nat.ptr(self._nbr),
                                   nat.ptr(self._label), nat.ptr(self._d6), None, None)
        f32 = dict(dtype=torch.float32, device=dev)
        self.x, self.r, self.p0, self.p1, self.ap = (torch.zeros(n, **f32) for _ in range(5))
        self.scalars = torch.zeros(8, dtype=torch.float64, device=dev)
        self.part = torch.zeros(max(L.n_bricks, 1), dtype=torch.float64, device=dev)
        self.last = {}

    def solve(self, dirichlet, x0, tol=1e-6, maxiter=10_000, check_every=64):
        """x (N,) in node order (np.argwhere order of the mask) with x = x0 on
        the Dirichlet nodes (bool mask, node order) and the Laplace equation
        elsewhere; x0 is the start.  Returns the solution as a numpy array."""
        L = self.lay
        lab = np.where(np.asarray(dirichlet, bool), 1, 0).astype(np.uint8)
        self._label.copy_(torch.from_numpy(L.to_slots(lab, fill=INACTIVE)))
        self.x.copy_(torch.from_numpy(L.to_slots(np.asarray(x0, np.float32))))
        cg = nat.HpLapCg(nat.ptr(self.x), nat.ptr(self.r),
                         (C.c_void_p * 2)(nat.ptr(self.p0), nat.ptr(self.p1)), nat.ptr(self.ap),
                         nat.ptr(self.scalars), nat.ptr(self.part), self.part.numel(),
                         float(tol), 0)
        s = nat.stream_handle()
        nat.call("hp_laplace_cg_begin_f32", C.byref(self.bricks), C.byref(cg), s)
        while True:
            sc = self.scalars.cpu().numpy()
            if sc[4] != 0.0 or cg.launched >= maxiter:
                break
            k = int(min(check_every, maxiter - cg.launched))
            nat.call("hp_laplace_cg_iterate_f32", C.byref(self.bricks), C.byref(cg), k,
                     int(maxiter), s)
        self.last = dict(iterations=int(sc[5]), rel_res=float(sc[6]),
                         converged=bool(sc[4] == 1.0))
        return self.x[torch.from_numpy(L.slot_of_node.astype(np.int64)).to(self.dev)].cpu().numpy()
