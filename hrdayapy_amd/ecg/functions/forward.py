"""ECG forward solve on the GPU: heart source, torso Poisson PCG with FP64
iterative refinement, electrodes and the 12-lead ECG.

Reference: hrdayapy/ecg/functions/compute_bspm.py
  * build_projection_map   :126-157  (heart voxel -> coarse torso row)
  * solve_frame source     :381-419  (sigma_i/h^2 K Vm, masked, mean-removed,
                                      scattered, / n_expected, pinned)
  * _pcg_gpu               :164-263  (Jacobi PCG, warm start, rms metric)
The refinement mode (FP32 PCG inside, FP64 residual outside, relative
residual target) has no counterpart in the reference; the 'rms' mode follows
its stopping rule.  Electrodes and limb/precordial leads are not in the
reference either (it outputs raw surface potentials); they use the textbook
Einthoven/Goldberger/Wilson definitions.
"""
from __future__ import annotations

import ctypes as C
import warnings
import time

import numpy as np
import torch

from ... import _native as nat
from ...layout import BrickLayout
from .registration import projection_rows


def _face_sum(Gx, Gy, Gz):
    """Sum of the six face conductances of every cell of a dense box (faces stored
    on the + side of each cell)."""
    S = Gx + Gy + Gz
    S[1:] += Gx[:-1]
    S[:, 1:] += Gy[:, :-1]
    S[:, :, 1:] += Gz[:, :, :-1]
    return S


def coarsen_faces(Gx, Gy, Gz):
    """2x2x2 agglomeration of a dense box of + face conductances: a coarse face is
    1/2 of the sum of the four fine faces it covers (the operator at 2h)."""
    pad = [(0, n % 2) for n in Gx.shape]
    Gx, Gy, Gz = (np.pad(G, pad) for G in (Gx, Gy, Gz))
    Cx = 0.5 * (Gx[1::2, 0::2, 0::2] + Gx[1::2, 1::2, 0::2] + Gx[1::2, 0::2, 1::2]
                + Gx[1::2, 1::2, 1::2])
    Cy = 0.5 * (Gy[0::2, 1::2, 0::2] + Gy[1::2, 1::2, 0::2] + Gy[0::2, 1::2, 1::2]
                + Gy[1::2, 1::2, 1::2])
    Cz = 0.5 * (Gz[0::2, 0::2, 1::2] + Gz[1::2, 0::2, 1::2] + Gz[0::2, 1::2, 1::2]
                + Gz[1::2, 1::2, 1::2])
    return Cx, Cy, Cz


def dense_operator(Gx, Gy, Gz):
    """The (n x n) operator of a dense box (Neumann), FP64, for small boxes."""
    shape = Gx.shape
    n = int(np.prod(shape))
    idx = np.arange(n).reshape(shape)
    A = np.zeros((n, n))
    A[idx.ravel(), idx.ravel()] = _face_sum(Gx, Gy, Gz).ravel()
    pairs = [(Gx[:-1], idx[:-1], idx[1:]), (Gy[:, :-1], idx[:, :-1], idx[:, 1:]),
             (Gz[:, :, :-1], idx[:, :, :-1], idx[:, :, 1:])]
    for g, a, b in pairs:
        g, a, b = g.ravel(), a.ravel(), b.ravel()
        A[a, b] -= g
        A[b, a] -= g
    return A


class Multigrid:
    """V-cycle preconditioner of a TorsoGrid's operator, resident in HBM
    (hp_mg): level 0 is the column-run grid, coarser levels are dense boxes
    down to at most max_coarse cells, solved there with a dense inverse.  The
    coarse cell holding the pinned cell is grounded on every level.  That suits
    right-hand sides of zero sum (the heart source: its mean is removed); a
    right-hand side with a net charge excites the near-constant mode, which the
    widened coarse ground corrects poorly -- LeadField moves such a charge onto
    the pin's neighbours first."""

    def __init__(self, grid, device="cuda", nu=2, omega=0.8, max_coarse=1024):
        dev = torch.device(device)
        L = grid.layout
        shape = tuple(int(x) for x in grid["shape"])
        flat = grid.flat_of_cell
        G = []
        for key in ("gx", "gy", "gz"):
            a = np.zeros(int(np.prod(shape)), dtype=np.float64)
            a[flat] = L[key]
            G.append(a.reshape(shape))
        pin_cell = int(L["pin"])
        pin = np.array(np.unravel_index(int(flat[pin_cell]), shape))
        S0 = _face_sum(*G).ravel()[flat]
        inv0 = np.where(S0 > 0, 1.0 / L["diag"].astype(np.float64), 0.0)
        inv0[pin_cell] = 0.0
        self.keep = [torch.from_numpy(inv0.astype(np.float32)).to(dev),
                     torch.zeros(int(L["n_cells"]), dtype=torch.float32, device=dev)]
        levels, self.shapes, self.active = [], [], []
        while True:
            G = list(coarsen_faces(*G))
            pin = pin // 2
            S = _face_sum(*G)
            inv = np.where(S > 0, 1.0 / np.where(S > 0, S, 1.0), 0.0)
            inv[tuple(pin)] = 0.0   # the cell holding the pinned cell is grounded
            levels.append((G, inv))
            self.shapes.append(G[0].shape)
            self.active.append(int((inv > 0).sum()))
            if G[0].size <= max_coarse:
                break
            if len(levels) == nat.MG_MAX_LEVELS:
                raise ValueError("torso grid needs more multigrid levels than supported")
        # exact inverse on the coarsest level's active cells
        Gc, invc = levels[-1]
        self.coarse_faces = Gc
        A = dense_operator(*Gc)
        act = np.nonzero(invc.ravel() > 0)[0]
        Ainv = np.zeros_like(A)
        if len(act):
            Ainv[np.ix_(act, act)] = np.linalg.pinv(A[np.ix_(act, act)])
        self.coarse_inv = torch.from_numpy(Ainv.astype(np.float32)).to(dev)
        mg = nat.HpMg()
        mg.n_levels, mg.nu, mg.omega = len(levels), int(nu), float(omega)
        mg.inv_diag0, mg.tmp0 = self.keep[0].data_ptr(), self.keep[1].data_ptr()
        mg.coarse_inv = self.coarse_inv.data_ptr()
        for l, (Gl, inv) in enumerate(levels):
            ni, nj, nk = Gl[0].shape
            n = ni * nj * nk
            ts = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32).ravel()).to(dev)
                  for a in (*Gl, inv)]
            ts += [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
            self.keep += ts
            mg.level[l] = nat.HpMgLevel(ni, nj, nk, 0, n, *[t.data_ptr() for t in ts])
        self.mg = mg

    def vcycle(self, torso, r32, z32):
        nat.call("hp_mg_vcycle_f32", C.byref(torso), C.byref(self.mg), r32.data_ptr(),
                 z32.data_ptr(), None, nat.stream_handle())


class PoissonSolver:
    """The torso operator of a TorsoGrid resident in HBM, plus PCG work vectors.

    precond: "jacobi" (the reference's diagonal preconditioner) or "mg" (one
    V-cycle of the geometric multigrid of this module)."""

    def __init__(self, grid, device="cuda", precond="jacobi"):
        nat.load()
        self.grid = grid
        self.dev = dev = torch.device(device)
        L = grid.layout
        self.n = n = int(L["n_cells"])
        self._t = {k: torch.from_numpy(np.ascontiguousarray(L[k])).to(dev)
                   for k in ("col_k0", "col_len", "col_off", "gx", "gy", "gz", "diag")}
        ni, nj, nk = grid["shape"]
        tab = nat.torso_chunks((ni, nj, nk), L["col_k0"], L["col_len"], L["col_off"])
        self._t["chunks"] = torch.from_numpy(tab).to(dev)
        t = self._t
        self.torso = nat.HpTorso(ni, nj, nk, 0, n, t["col_k0"].data_ptr(), t["col_len"].data_ptr(),
                                 t["col_off"].data_ptr(), t["gx"].data_ptr(), t["gy"].data_ptr(),
                                 t["gz"].data_ptr(), t["diag"].data_ptr(), int(L["pin"]),
                                 t["chunks"].data_ptr(), len(tab))
        self.part = torch.zeros(8 * 4096, dtype=torch.float64, device=dev)
        self.counter = torch.zeros(4, dtype=torch.int32, device=dev)
        self.out = torch.zeros(4, dtype=torch.float64, device=dev)
        self.sc = torch.zeros(16, dtype=torch.float64, device=dev)
        self.ws = nat.HpPcgWs(self.part.data_ptr(), self.counter.data_ptr(), self.out.data_ptr(),
                              self.sc.data_ptr())
        f32 = dict(dtype=torch.float32, device=dev)
        self.r32, self.d32, self.p32, self.Ap32 = (torch.zeros(n, **f32) for _ in range(4))
        self.r64 = torch.zeros(n, dtype=torch.float64, device=dev)
        self.last_stats = {}
        # the pinned cell's neighbours and the share of each in the conductance
        # to the pin: A_p^-1 of this vector is constant (1/G) on the unpinned cells
        pc = int(L["pin"])
        flat = grid.flat_of_cell
        pi, pj, pk = np.unravel_index(int(flat[pc]), grid["shape"])
        shape = grid["shape"]
        nbr, gw = [], []
        col_k0, col_len, col_off = L["col_k0"], L["col_len"], L["col_off"]

        def cell_at(i, j, k):
            if not (0 <= i < shape[0] and 0 <= j < shape[1]):
                return -1
            c = i * shape[1] + j
            dk = k - int(col_k0[c])
            return int(col_off[c]) + dk if 0 <= dk < int(col_len[c]) else -1
        for (i, j, k, g_of) in ((pi + 1, pj, pk, lambda n: L["gx"][pc]),
                                (pi - 1, pj, pk, lambda n: L["gx"][n]),
                                (pi, pj + 1, pk, lambda n: L["gy"][pc]),
                                (pi, pj - 1, pk, lambda n: L["gy"][n]),
                                (pi, pj, pk + 1, lambda n: L["gz"][pc]),
                                (pi, pj, pk - 1, lambda n: L["gz"][n])):
            cn = cell_at(i, j, k)
            if cn >= 0 and float(g_of(cn)) > 0:
                nbr.append(cn)
                gw.append(float(g_of(cn)))
        gw = np.array(gw)
        self.pin_G = float(gw.sum())
        self.pin_nbr = torch.tensor(nbr, dtype=torch.int64, device=dev)
        self.pin_w = torch.from_numpy(gw / max(self.pin_G, 1e-300)).to(dev)
        if precond not in ("jacobi", "mg"):
            raise ValueError(f"unknown preconditioner {precond!r}")
        self.precond = precond
        self.mg = None
        if precond == "mg":
            self.mg = Multigrid(grid, dev)
            self.z32 = torch.zeros(n, **f32)

    # ------------------------------------------------------------ primitives
    def spmv32(self, x, y):
        nat.call("hp_poisson_spmv_f32", C.byref(self.torso), x.data_ptr(), y.data_ptr(),
                 nat.stream_handle())

    def residual(self, b, x, r=None):
        """r = b - A x (FP64); returns (||r||, ||b||) as floats (one host sync)."""
        r = self.r64 if r is None else r
        nat.call("hp_poisson_residual_f64", C.byref(self.torso), b.data_ptr(), x.data_ptr(),
                 r.data_ptr(), C.byref(self.ws), nat.stream_handle())
        o = self.out[:2].cpu().numpy()
        return float(np.sqrt(o[0])), float(np.sqrt(o[1]))

    def _pcg(self, tol, max_iter, mode, check_every):
        s = nat.stream_handle()
        mg = self.mg is not None
        if mg and mode != 0:
            raise ValueError("the multigrid PCG stops on the relative residual only")
        if mg:
            check_every = max(1, check_every // 8)
            nat.call("hp_mgpcg_begin_f32", C.byref(self.torso), C.byref(self.mg.mg),
                     self.r32.data_ptr(), self.d32.data_ptr(), self.z32.data_ptr(),
                     self.p32.data_ptr(), C.byref(self.ws), float(tol), int(max_iter), s)
        else:
            n_rows = len(self.grid.cell_of_row) if mode else self.n
            nat.call("hp_pcg_begin_metric_f32", C.byref(self.torso), self.r32.data_ptr(),
                     self.d32.data_ptr(), self.p32.data_ptr(), C.byref(self.ws), float(tol),
                     int(max_iter), int(mode), int(n_rows), s)
        done_iters = 0
        while True:
            n = min(check_every, max_iter - done_iters) if max_iter > done_iters else 0
            if n > 0:
                if mg:
                    nat.call("hp_mgpcg_iterate_f32", C.byref(self.torso), C.byref(self.mg.mg),
                             self.r32.data_ptr(), self.d32.data_ptr(), self.z32.data_ptr(),
                             self.p32.data_ptr(), self.Ap32.data_ptr(), C.byref(self.ws), n, s)
                else:
                    nat.call("hp_pcg_iterate_f32", C.byref(self.torso), self.r32.data_ptr(),
                             self.d32.data_ptr(), self.p32.data_ptr(), self.Ap32.data_ptr(),
                             C.byref(self.ws), n, s)
                done_iters += n
            sc = self.sc[:11].cpu().numpy()
            if sc[7] != 0.0 or n == 0:
                return int(sc[8]), sc

    # ---------------------------------------------------------------- solves
    def solve_refined(self, b, x=None, tol=1e-8, inner_tol=1e-4, max_outer=30, max_inner=20000,
                      check_every=32):
        """Solve A x = b to ||b - A x|| <= tol ||b|| (FP64 residual) by FP32 PCG
        corrections.  b, x: FP64 device vectors in cell order; x is updated in place.
        Each outer step solves A d = r/||r|| to inner_tol relative to that
        (current) residual.  last_stats["converged"] says whether tol was met;
        a RuntimeWarning says so when max_outer ends the loop first."""
        x = torch.zeros(self.n, dtype=torch.float64, device=self.dev) if x is None else x
        inner_total, outer = 0, 0
        rn, bn = self.residual(b, x)
        while bn > 0 and rn > tol * bn and outer < max_outer:
            nat.call("hp_cast_scale_f32", C.byref(self.torso), self.r64.data_ptr(),
                     self.r32.data_ptr(), 1.0 / rn, nat.stream_handle())
            it, _ = self._pcg(inner_tol, max_inner, 0, check_every)
            nat.call("hp_refine_axpy_f64", C.byref(self.torso), x.data_ptr(), self.d32.data_ptr(),
                     rn, nat.stream_handle())
            inner_total += it
            outer += 1
            rn, bn = self.residual(b, x)
        conv = not (bn > 0 and rn > tol * bn)
        self.last_stats = dict(outer=outer, inner=inner_total, rel_res=rn / bn if bn else 0.0,
                               converged=conv)
        if not conv:
            warnings.warn(f"solve_refined: relative residual {rn / bn:.3e} > tol {tol:g} after "
                          f"{outer} outer steps ({inner_total} inner iterations)", RuntimeWarning,
                          stacklevel=2)
        return x

    METRICS = {"rms": 1, "linf": 2, "p99": 3}

    def solve_metric(self, b, x=None, tol_V=1e-5, max_iter=500, metric="rms", check_every=16):
        """The reference's solve: FP32 Jacobi PCG from the warm start x until the
        chosen statistic of z = M^-1 r over the system's rows -- rms, max |z|
        ('linf') or the 99th percentile of |z| ('p99') -- is below tol_V, tested
        after each iteration (compute_bspm.py:164-263).  Returns x (FP64
        storage); last_stats holds the iterations and the statistic in V."""
        if metric not in self.METRICS:
            raise ValueError(f"metric must be one of 'rms', 'p99', 'linf', got {metric!r}")
        x = torch.zeros(self.n, dtype=torch.float64, device=self.dev) if x is None else x
        self.residual(b, x)
        nat.call("hp_cast_scale_f32", C.byref(self.torso), self.r64.data_ptr(),
                 self.r32.data_ptr(), 1.0, nat.stream_handle())
        it, sc = self._pcg(tol_V, max_iter, self.METRICS[metric], check_every)
        nat.call("hp_refine_axpy_f64", C.byref(self.torso), x.data_ptr(), self.d32.data_ptr(), 1.0,
                 nat.stream_handle())
        if metric == "p99":  # the final statistic itself (the test ran without a sort)
            if not hasattr(self, "_row_cells"):
                self._row_cells = torch.from_numpy(self.grid.cell_of_row).to(self.dev)
            z = self.r32[self._row_cells].abs()
            if z.numel() >= (1 << 24):  # torch.quantile's element limit on the device
                z = z.cpu()
            stat = float(torch.quantile(z, 0.99))
        else:
            stat = float(self.sc[13].item())
        self.last_stats = dict(inner=it, metric=metric, err_V=stat)
        if metric == "rms":
            self.last_stats["rms_z"] = stat
        return x

    def solve_rms(self, b, x=None, tol_V=1e-5, max_iter=500, check_every=16):
        """solve_metric with the reference's default 'rms' rule."""
        return self.solve_metric(b, x, tol_V, max_iter, "rms", check_every)


class HeartSource:
    """Per-frame torso right-hand side from the heart's Vm on the brick layout."""

    def __init__(self, layout: BrickLayout, coords_mm, grid, registration, sigma_i, h_fine_mm,
                 device="cuda", bricks=None):
        self.dev = dev = torch.device(device)
        self.lay = layout
        R, t, sc = registration["R"], registration["t"], float(registration["scale"])
        proj = projection_rows(coords_mm, R, t, sc, grid["origin_mm"], grid["dx_mm"],
                               grid["active_id_map"], grid["shape"], dev)
        valid = proj >= 0
        self.n_valid = int(valid.sum())
        if self.n_valid == 0:
            raise ValueError("no heart voxel projects into the torso's active cells")
        slots = layout.slot_of_node[valid].astype(np.int32)
        cells = grid.cell_of_row[proj[valid]]
        self.valid_cells = torch.from_numpy(cells.astype(np.int64)).to(dev)
        order = np.argsort(cells, kind="stable")
        rows_c, first = np.unique(cells[order], return_index=True)
        self.valid_slots = torch.from_numpy(slots).to(dev)
        self.rows = torch.from_numpy(rows_c.astype(np.int32)).to(dev)
        self.ptr = torch.from_numpy(np.append(first, len(order)).astype(np.int32)).to(dev)
        self.fine = torch.from_numpy(slots[order]).to(dev)
        h_m = h_fine_mm * 1e-3
        self.coeff = sigma_i / h_m ** 2
        self.n_expected = (grid["dx_mm"] / h_fine_mm) ** 3
        self.pin = int(grid.layout["pin"])
        # smoothed boundary: each torso row's source is scaled by its psi
        # (compute_bspm.py:413-414); None for the sharp boundary
        psi = getattr(grid, "psi_cell", None)
        self.psi = None if psi is None else torch.as_tensor(psi, dtype=torch.float64).to(dev)
        self.lap = torch.zeros(layout.n_slots, dtype=torch.float64, device=dev)
        if bricks is None:
            self._label = torch.from_numpy(layout.labels(np.zeros(layout.n_nodes, np.uint8))).to(dev)
            self._nbr = torch.from_numpy(layout.nbr).to(dev)
            bricks = nat.HpBricks(layout.n_bricks, *layout.brick, 0, self._nbr.data_ptr(),
                                  self._label.data_ptr(), None, None)
        self.bricks = bricks
        self.n_miss = int((~valid).sum())

    def __call__(self, V_slots_f32, b64, ws):
        """b64[hit cells] <- source of V (mV, FP32, slot order); other cells untouched."""
        s = nat.stream_handle()
        nat.call("hp_heart_lap6_f32", C.byref(self.bricks), V_slots_f32.data_ptr(),
                 self.lap.data_ptr(), 1e-3, s)
        nat.call("hp_bspm_source_f64", self.lap.data_ptr(), self.valid_slots.data_ptr(),
                 self.n_valid, self.rows.data_ptr(), self.ptr.data_ptr(), self.fine.data_ptr(),
                 int(self.rows.numel()), self.coeff, self.n_expected, nat.ptr(self.psi),
                 b64.data_ptr(), self.pin, C.byref(ws), s)


class LeadField:
    """Electrode potentials as a fixed linear functional of Vm.

    The torso potential is x = A^-1 b(Vm) with b linear in Vm (the heart
    source), and an electrode reads E_e^T x, so E_e^T x = (A^-1 E_e)^T b(Vm)
    (reciprocity; the pinned operator restricted to the unpinned cells is
    symmetric, and both E_e and b vanish on the pinned cell).  One refined
    solve per electrode gives y_e = A^-1 E_e; pulling y_e back through the
    source map (gather at each heart voxel's torso cell, remove the mean over
    the valid voxels, apply the symmetric heart Laplacian, scale) gives the
    per-voxel weights W[:, e].  Every ECG frame is then W^T Vm: one pass over
    W for up to 16 frames, instead of a torso solve per frame.
    """

    def __init__(self, ps: "PoissonSolver", src: "HeartSource", el: "Electrodes", tol=1e-10,
                 batch=16, log=None):
        dev = ps.dev
        self.dev, self.batch = dev, int(batch)
        lay = src.lay
        self.ne = ne = len(el.names)
        self.names = el.names
        t0 = time.perf_counter()
        active = np.nonzero(lay.labels(np.zeros(lay.n_nodes, np.uint8)) != 0xFF)[0]
        self.row_slot = torch.from_numpy(active.astype(np.int32)).to(dev)
        n_rows = len(active)
        W = torch.zeros((n_rows, ne), dtype=torch.float32, device=dev)
        rhs = torch.zeros(ps.n, dtype=torch.float64, device=dev)
        y = torch.zeros(ps.n, dtype=torch.float64, device=dev)
        h = torch.zeros(lay.n_slots, dtype=torch.float64, device=dev)
        kh = torch.zeros(lay.n_slots, dtype=torch.float64, device=dev)
        cells = torch.from_numpy(el.cells.astype(np.int64)).to(dev)
        kk = el.cells.shape[1]
        vslots = src.valid_slots.long()
        scale = src.coeff / src.n_expected
        self.solve_stats = []
        for e in range(ne):
            rhs.zero_()
            rhs.index_add_(0, cells[e], torch.full((kk,), 1.0 / kk, dtype=torch.float64,
                                                   device=dev))
            rhs[src.pin] = 0.0
            # move the electrode's net charge onto the pin's neighbours: the
            # solution changes by a constant, which the source map cancels
            # (it removes the mean over the heart voxels)
            q = float(rhs.sum())
            rhs.index_add_(0, ps.pin_nbr, -q * ps.pin_w)
            y.zero_()
            ps.solve_refined(rhs, y, tol=tol, max_outer=8, max_inner=2000)
            self.solve_stats.append(dict(ps.last_stats))
            g = y[src.valid_cells]
            if src.psi is not None:  # the adjoint of the row scale psi
                g = g * src.psi[src.valid_cells]
            h.zero_()
            h[vslots] = g - g.mean()
            nat.call("hp_heart_lap6_f64", C.byref(src.bricks), h.data_ptr(), kh.data_ptr(), 1e-3,
                     nat.stream_handle())
            W[:, e] = (kh[self.row_slot.long()] * scale).float()
            if log and (e + 1) % 8 == 0:
                log(f"[leadfield] {e + 1}/{ne} electrodes, {time.perf_counter() - t0:.1f} s")
        self.W = W
        self.setup_s = time.perf_counter() - t0
        self.part = torch.zeros(4096 * self.batch * ne, dtype=torch.float64, device=dev)
        self.out = torch.zeros((self.batch, ne), dtype=torch.float64, device=dev)

    def apply(self, V_frames, n_frames=None, ldv=None):
        """Electrode potentials (n_frames, ne) in V (FP64, on the device) of
        n_frames Vm frames in slot layout: V_frames (n_frames, n_slots) FP32, mV."""
        nf = int(V_frames.shape[0] if n_frames is None else n_frames)
        if nf > self.batch:
            raise ValueError(f"at most {self.batch} frames per call")
        ldv = int(V_frames.shape[-1] if ldv is None else ldv)
        nat.call("hp_leadfield_apply_f32", self.W.data_ptr(), self.row_slot.data_ptr(),
                 int(self.W.shape[0]), self.ne, V_frames.data_ptr(), ldv, nf,
                 self.part.data_ptr(), int(self.part.numel()), self.out.data_ptr(),
                 nat.stream_handle())
        return self.out[:nf]


LEAD_NAMES = ["I", "II", "III", "aVR", "aVL", "aVF", "V1", "V2", "V3", "V4", "V5", "V6"]
LIMB = ["RA", "LA", "LL"]
PRECORDIAL = ["V1", "V2", "V3", "V4", "V5", "V6"]


def leads_from_electrodes(phi):
    """12 leads from electrode potentials (dict name -> array), textbook definitions."""
    RA, LA, LL = phi["RA"], phi["LA"], phi["LL"]
    wct = (RA + LA + LL) / 3.0
    out = {"I": LA - RA, "II": LL - RA, "III": LL - LA, "aVR": RA - 0.5 * (LA + LL),
           "aVL": LA - 0.5 * (RA + LL), "aVF": LL - 0.5 * (RA + LA)}
    for v in PRECORDIAL:
        out[v] = phi[v] - wct
    return out


class Electrodes:
    """Electrodes on the body surface: each is the mean of the k surface cells
    nearest to its position (FP64 accumulation on the device)."""

    def __init__(self, grid, positions_mm: dict, k=4, device="cuda"):
        from scipy.spatial import cKDTree
        self.names = list(positions_mm)
        xyz = grid["surface_xyz"].astype(np.float64)
        cells = grid.cell_of_row[grid["surface_rows"]]
        kk = min(k, len(xyz))
        _, nn = cKDTree(xyz).query(np.array([positions_mm[n] for n in self.names]), k=kk)
        nn = np.asarray(nn).reshape(len(self.names), kk)
        self.cells = cells[nn]
        dev = torch.device(device)
        self.ptr = torch.arange(0, kk * len(self.names) + 1, kk, dtype=torch.int32, device=dev)
        self.idx = torch.from_numpy(self.cells.ravel().astype(np.int32)).to(dev)
        self.w = torch.full((kk * len(self.names),), 1.0 / kk, dtype=torch.float64, device=dev)
        self.out = torch.zeros(len(self.names), dtype=torch.float64, device=dev)

    def __call__(self, phi64):
        nat.call("hp_leads_accumulate_f64", phi64.data_ptr(), self.ptr.data_ptr(),
                 self.idx.data_ptr(), self.w.data_ptr(), self.out.data_ptr(), len(self.names),
                 nat.stream_handle())
        return self.out


def standard_electrodes(grid, heart_centre_mm, n_bspm=256, seed=0):
    """Electrode positions on a torso: limb electrodes RA/LA/LL, precordial
    V1-V6 around the heart on the anterior surface, and n_bspm seeded sites on
    a regular (azimuth, height) grid over the outer surface."""
    xyz = grid["surface_xyz"].astype(np.float64)
    c = xyz.mean(axis=0)
    ext = xyz.max(axis=0) - xyz.min(axis=0)
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)

    def ray(direction, origin):
        d = np.asarray(direction, float)
        d /= np.linalg.norm(d)
        v = xyz - origin
        along = v @ d
        perp = np.linalg.norm(v - along[:, None] * d, axis=1)
        score = perp - 0.05 * along
        score[along <= 0] = np.inf
        return xyz[int(np.argmin(score))]

    hc = np.asarray(heart_centre_mm, float)
    pos = {}
    # frame: x = left-right, y = anterior(-)/posterior(+), z = foot(-)/head(+)
    top = c + np.array([0, 0, 0.35 * ext[2]])
    bottom = c - np.array([0, 0, 0.35 * ext[2]])
    pos["RA"] = ray([-1, 0, 0.3], top)
    pos["LA"] = ray([1, 0, 0.3], top)
    pos["LL"] = ray([0.3, 0, -1], bottom)
    for i, ang in enumerate(np.deg2rad([-20, 0, 20, 40, 60, 80])):
        pos[f"V{i + 1}"] = ray([np.sin(ang), -np.cos(ang), 0.0], hc)
    if n_bspm <= 0:
        return pos
    rng = np.random.default_rng(seed)
    nz = int(np.sqrt(n_bspm))
    na = n_bspm // nz
    k = 0
    for iz in range(nz):
        z = lo[2] + (hi[2] - lo[2]) * (0.1 + 0.8 * (iz + 0.5) / nz)
        for ia in range(na):
            a = 2 * np.pi * (ia + 0.5 + 0.2 * rng.uniform(-1, 1)) / na
            pos[f"E{k:03d}"] = ray([np.cos(a), np.sin(a), 0.0], np.array([c[0], c[1], z]))
            k += 1
    return pos
