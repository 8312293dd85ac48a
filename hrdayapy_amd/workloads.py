"""Measurement workloads besides the default coupled BiV step: the torso
Poisson solve (ECG) and the batched ensemble.  Each returns a dict that the
benchmark script prints as one JSON line.  GPU only."""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _native as nat
from . import geometry as G

HBM_PEAK_GBS = 8000.0
PCG_BYTES_PER_CELL = 52    # SpMV p 4 + Ap 4 + 3 faces 12 + x,r,p r/w 24 + Ap 4 + M^-1 4
SPMV_BYTES_PER_CELL = 20   # p 4 + Ap 4 + 3 face conductances 12
FUSED_BYTES_PER_VOXEL = 177


def _event_us(fn, reps):
    st = torch.cuda.current_stream()
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps * 1e3


def torso_solve(eng, meta, log=print, shape=(384, 384, 256), dx_mm=1.5, warm_steps=2000,
                n_bspm=256, tol=1e-8):
    """The 384x384x256 @ 1.5 mm torso with the engine's heart registered at the
    blood-pool centre: FP32 Jacobi-PCG inside FP64 refinement to relative
    residual tol, cold and warm-started; per-iteration kernel times."""
    import ctypes as C
    from .ecg.functions.forward import (Electrodes, HeartSource, PoissonSolver,
                                        standard_electrodes)
    from .ecg.functions.torso import TorsoGrid
    t0 = time.perf_counter()
    lab, origin, hc = G.torso(shape, dx_mm)
    grid = TorsoGrid.from_labels(lab, G.TORSO_SIGMA, dx_mm, origin)
    t_grid = time.perf_counter() - t0
    ps = PoissonSolver(grid, eng.dev)
    extent = meta["nodes_mm"].max(axis=0)
    reg = dict(R=np.eye(3), t=hc - extent / 2, scale=1.0)
    src = HeartSource(eng.lay, meta["nodes_mm"], grid, reg, 0.2625, meta["voxel_size"], eng.dev,
                      bricks=eng.bricks)
    el = Electrodes(grid, standard_electrodes(grid, hc, n_bspm), device=eng.dev)
    log(f"[torso] {shape} @ {dx_mm} mm: {ps.n:,} stored cells "
        f"({int((lab != 0).sum()):,} tissue), grid {t_grid:.1f} s; heart voxels in torso "
        f"{src.n_valid:,}")
    eng.run(warm_steps, graph_chunk=50)
    b = torch.zeros(ps.n, dtype=torch.float64, device=eng.dev)
    x = torch.zeros(ps.n, dtype=torch.float64, device=eng.dev)
    src(eng.V[eng.p], b, ps.ws)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ps.solve_refined(b, x, tol=tol)
    torch.cuda.synchronize()
    cold = dict(ps.last_stats, wall_s=time.perf_counter() - t1)
    log(f"[torso] cold solve: {cold}")
    eng.run(50, graph_chunk=50)
    src(eng.V[eng.p], b, ps.ws)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ps.solve_refined(b, x, tol=tol)
    torch.cuda.synchronize()
    warm = dict(ps.last_stats, wall_s=time.perf_counter() - t1)
    log(f"[torso] warm solve (1 ms later): {warm}")
    phi = el(x)
    # kernel times: SpMV and one fused PCG iteration
    s = torch.cuda.current_stream().cuda_stream
    spmv_us = _event_us(lambda: ps.spmv32(ps.p32, ps.Ap32), 50)
    nat.call("hp_pcg_begin_f32", C.byref(ps.torso), ps.r32.data_ptr(), ps.d32.data_ptr(),
             ps.p32.data_ptr(), C.byref(ps.ws), 0.0, 1 << 30, 0, s)
    it_us = _event_us(lambda: nat.call("hp_pcg_iterate_f32", C.byref(ps.torso),
                                       ps.r32.data_ptr(), ps.d32.data_ptr(), ps.p32.data_ptr(),
                                       ps.Ap32.data_ptr(), C.byref(ps.ws), 10, s), 5) / 10
    res_us = _event_us(lambda: ps.residual(b, x), 10)
    n = ps.n
    out = dict(cells=n, tissue_cells=int((lab != 0).sum()), grid_s=round(t_grid, 1),
               cold=cold, warm=warm, spmv_us=round(spmv_us, 2),
               spmv_GBs=SPMV_BYTES_PER_CELL * n / spmv_us / 1e3,
               pcg_iter_us=round(it_us, 2), pcg_iter_GBs=PCG_BYTES_PER_CELL * n / it_us / 1e3,
               residual_f64_us=round(res_us, 2), electrodes=len(el.names),
               phi_range_mV=[float(phi.min()) * 1e3, float(phi.max()) * 1e3])
    out["pcg_roofline_frac"] = out["pcg_iter_GBs"] / HBM_PEAK_GBS
    log(f"[torso] SpMV {spmv_us:.1f} us ({out['spmv_GBs']:.0f} GB/s), PCG iteration "
        f"{it_us:.1f} us ({out['pcg_iter_GBs']:.0f} GB/s)")
    return out


def ensemble_run(n_inst=64, seed=0, steps=1000, warmup=100, graph_chunk=50, brick=None,
                 size_range=(200_000, 2_000_000), log=print):
    """The seeded 64-instance ensemble (slabs and BiV shells of 0.2M-2M voxels,
    sigma x U(0.7,1.3), G_Ks/G_to x U(0.5,1.5), 1-3 pacing sites), FP32,
    one fused launch per step over all instances' bricks."""
    from .ensemble import Ensemble, draw_instances
    t0 = time.perf_counter()
    specs = draw_instances(n_inst, seed, size_range)
    ens = Ensemble(specs, dtype=torch.float32, device="cuda", brick=brick)
    eng = ens.eng
    log(f"[ensemble] {n_inst} instances, {ens.n_nodes:,} voxels in {ens.layout.n_bricks:,} "
        f"bricks {ens.layout.brick}, occupancy {ens.layout.occupancy:.3f}, "
        f"geometry {ens.t_geometry:.1f} s, setup {time.perf_counter() - t0:.1f} s")
    eng.run(warmup, graph_chunk=graph_chunk)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.run(steps, graph_chunk=graph_chunk)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    at = ens.at_instances()
    act = [float(np.isfinite(a).mean()) for a in at]
    sizes = [int(o1 - o0) for o0, o1 in zip(ens.node_offset[:-1], ens.node_offset[1:])]
    return dict(n_inst=n_inst, voxels=ens.n_nodes, sizes_min=min(sizes), sizes_max=max(sizes),
                kinds={k: sum(s.kind == k for s in specs) for k in ("slab", "biv")},
                brick=list(ens.layout.brick), steps=steps, wall_s=wall,
                value=ens.n_nodes * steps / wall, ms_per_step=wall / steps * 1e3,
                activated_fraction_mean=float(np.mean(act)), graph_chunk=graph_chunk)
