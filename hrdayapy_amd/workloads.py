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


def graph_us(fn, reps, warm=2, timed=3):
    """Average device time (us) of fn(i), i = 0..reps-1, captured into one HIP
    graph; after `warm` replays, `timed` replays between HIP events on the
    replaying stream:
    no host launch overhead between the kernels (eager launches through ctypes
    can starve the GPU and time the host instead)."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            for i in range(reps):
                fn(i)
    torch.cuda.current_stream().wait_stream(side)
    for _ in range(warm):
        g.replay()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(timed):
        g.replay()
    e1.record(st)
    e1.synchronize()
    del g
    return e0.elapsed_time(e1) / (reps * timed) * 1e3


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
    t1 = time.perf_counter()
    ps = PoissonSolver(grid, eng.dev, precond="mg")
    t_mg = time.perf_counter() - t1
    extent = meta["nodes_mm"].max(axis=0)
    reg = dict(R=np.eye(3), t=hc - extent / 2, scale=1.0)
    src = HeartSource(eng.lay, meta["nodes_mm"], grid, reg, 0.2625, meta["voxel_size"], eng.dev,
                      bricks=eng.bricks)
    el = Electrodes(grid, standard_electrodes(grid, hc, n_bspm), device=eng.dev)
    log(f"[torso] {shape} @ {dx_mm} mm: {ps.n:,} stored cells "
        f"({int((lab != 0).sum()):,} tissue), grid {t_grid:.1f} s, multigrid {t_mg:.1f} s "
        f"{ps.mg.shapes}; heart voxels in torso {src.n_valid:,}")
    eng.run(warm_steps, graph_chunk=50)
    b = torch.zeros(ps.n, dtype=torch.float64, device=eng.dev)
    V0 = eng.V[eng.p].clone()
    eng.run(50, graph_chunk=50)
    V1 = eng.V[eng.p].clone()
    solves = {}
    x = torch.zeros(ps.n, dtype=torch.float64, device=eng.dev)
    res = {}
    for tag, V in (("cold", V0), ("warm", V1)):
        b.zero_()
        src(V, b, ps.ws)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ps.solve_refined(b, x, tol=tol)
        torch.cuda.synchronize()
        res[tag] = dict(ps.last_stats, wall_s=time.perf_counter() - t1)
        log(f"[torso] multigrid PCG {tag} solve to {tol:g}: {res[tag]}")
    solves["mg"] = res
    phi = el(x)
    mg_keep = ps.mg
    # the diagonal preconditioner (the reference's) for comparison: a bounded
    # number of iterations from the cold start, reporting the residual reached
    ps.mg, ps.precond = None, "jacobi"
    xj = torch.zeros(ps.n, dtype=torch.float64, device=eng.dev)
    b.zero_()
    src(V0, b, ps.ws)
    rn0, _ = ps.residual(b, xj)
    jac = []
    for chunk in range(4):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ps.solve_refined(b, xj, tol=tol, max_outer=1, max_inner=1000)
        torch.cuda.synchronize()
        jac.append(dict(ps.last_stats, wall_s=time.perf_counter() - t1))
        log(f"[torso] Jacobi PCG, +1000 iterations: {jac[-1]}")
    solves["jacobi_4x1000"] = jac
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
    vc_us = _event_us(lambda: mg_keep.vcycle(ps.torso, ps.r32, ps.p32), 5)
    log(f"[torso] V-cycle {vc_us:.0f} us")
    out = dict(cells=n, tissue_cells=int((lab != 0).sum()), grid_s=round(t_grid, 1),
               mg_setup_s=round(t_mg, 1), mg_levels=[list(x) for x in mg_keep.shapes],
               solves=solves, vcycle_us=round(vc_us, 1), spmv_us=round(spmv_us, 2),
               spmv_GBs=SPMV_BYTES_PER_CELL * n / spmv_us / 1e3,
               pcg_iter_us=round(it_us, 2), pcg_iter_GBs=PCG_BYTES_PER_CELL * n / it_us / 1e3,
               residual_f64_us=round(res_us, 2), electrodes=len(el.names),
               phi_range_mV=[float(phi.min()) * 1e3, float(phi.max()) * 1e3])
    out["pcg_roofline_frac"] = out["pcg_iter_GBs"] / HBM_PEAK_GBS
    log(f"[torso] SpMV {spmv_us:.1f} us ({out['spmv_GBs']:.0f} GB/s), PCG iteration "
        f"{it_us:.1f} us ({out['pcg_iter_GBs']:.0f} GB/s)")
    return out


# Static instruction counts of hp::ttp06_strided_kernel<float> in the gfx950
# code object (per cell, straight-line path): 543 VALU instructions, of which
# 100 transcendental (48 v_exp, 47 v_rcp, 4 v_log, 1 v_rsq).
IONIC_VALU_PER_CELL = 543
IONIC_TRANS_PER_CELL = 100
IONIC_BYTES_PER_CELL = 152      # 18 state r+w (144) + V read + I_ion write
# MI355X vector issue: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz one FP32 VALU
# operation per lane and cycle; transcendentals at a quarter of that rate.
VALU_PEAK_PER_S = 256 * 4 * 16 * 2.4e9
TRANS_PEAK_PER_S = VALU_PEAK_PER_S / 4


def ionic_run(n=2_006_132, reps=200, seed=0):
    """The TTP06 cell step alone (hp_ttp06_step_f32) over n cells of seeded
    random phenotype and membrane potential, SoA state (stride_var = n),
    resting initial state: device time per launch from a graph of reps
    launches, and the VALU / transcendental / HBM rates it implies."""
    import ctypes as C
    from .ttp06 import INIT, STATE_NAMES
    KI_BASE, NAI_BASE = INIT["Ki"], INIT["Nai"]
    dev = torch.device("cuda")
    rng = np.random.default_rng(seed)
    V = torch.from_numpy(rng.uniform(-90.0, 30.0, n).astype(np.float32)).to(dev)
    ct = torch.from_numpy(rng.integers(0, 3, n).astype(np.uint8)).to(dev)
    st = torch.empty((18, n), dtype=torch.float32, device=dev)
    for k, name in enumerate(STATE_NAMES):
        st[k].fill_(INIT[name] - (KI_BASE if name == "Ki" else NAI_BASE if name == "Nai" else 0.0))
    I = torch.empty(n, dtype=torch.float32, device=dev)

    def launch(i):
        nat.call("hp_ttp06_step_f32", nat.ptr(V), nat.ptr(st), nat.ptr(I), nat.ptr(ct), 0, None,
                 None, n, n, 1, 0.02, KI_BASE, NAI_BASE, nat.stream_handle())

    us = graph_us(launch, reps)
    cps = n / (us * 1e-6)
    return dict(cells=n, us_per_launch=us, cell_steps_per_s=cps,
                valu_per_s=cps * IONIC_VALU_PER_CELL, valu_frac=cps * IONIC_VALU_PER_CELL / VALU_PEAK_PER_S,
                trans_per_s=cps * IONIC_TRANS_PER_CELL,
                trans_frac=cps * IONIC_TRANS_PER_CELL / TRANS_PEAK_PER_S,
                hbm_GBs=cps * IONIC_BYTES_PER_CELL / 1e9,
                finite=bool(torch.isfinite(I).all()))


STENCIL_BYTES_PER_VOXEL = 33   # V read, L.V written, six FP32 D, one label byte


def stencil_run(eng, reps=200):
    """The 19-point operator alone (hp_stencil19_apply_f32: out = L V on the
    brick layout) over the engine's bricks: device time per launch from a
    graph of reps launches, and the algorithmic HBM rate."""
    import ctypes as C
    out = torch.empty_like(eng.V[0])
    V = eng.V[eng.p]
    dscale = eng.p_myo.dscale

    def launch(i):
        nat.call("hp_stencil19_apply_f32", C.byref(eng.bricks), nat.ptr(V), nat.ptr(out), dscale,
                 nat.stream_handle())

    us = graph_us(launch, reps)
    n = eng.N_myo
    return dict(voxels=n, bricks=eng.lay.n_bricks, brick=list(eng.lay.brick), us_per_launch=us,
                voxels_per_s=n / (us * 1e-6),
                hbm_GBs=n * STENCIL_BYTES_PER_VOXEL / (us * 1e-6) / 1e9,
                finite=bool(torch.isfinite(out).all()))


def ensemble_run(n_inst=64, seed=0, steps=1000, warmup=100, graph_chunk=50, brick=None,
                 size_range=(200_000, 2_000_000), log=print):
    """The seeded 64-instance ensemble (slabs and BiV shells of 0.2M-2M voxels,
    sigma x U(0.7,1.3), G_Ks/G_to x U(0.5,1.5), 1-3 pacing sites), FP32,
    one fused launch per step over all instances' bricks."""
    from .ensemble import Ensemble, draw_instances, instance_geometries
    t0 = time.perf_counter()
    specs = draw_instances(n_inst, seed, size_range)
    tuned = None
    if brick is None:
        # brick shape and steps per launch by measurement on the first 8 instances
        from .autotune import tune
        sample = specs[:8]
        geoms = instance_geometries(sample)
        brick, chunk_t, tuned = tune(
            lambda b: Ensemble(sample, geoms=geoms, dtype=torch.float32, device="cuda",
                               brick=b).eng, chunks=(50, 100, 200), steps=200, log=log)
        if graph_chunk <= 0:
            graph_chunk = chunk_t
        log(f"[ensemble] autotuned brick {brick}, {graph_chunk} steps per graph launch: {tuned}")
    if graph_chunk <= 0:
        graph_chunk = 50
    if warmup % graph_chunk or steps % graph_chunk:
        warmup += -warmup % graph_chunk
        steps += -steps % graph_chunk
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
                activated_fraction_mean=float(np.mean(act)), graph_chunk=graph_chunk,
                autotune=tuned)


class EcgSetup:
    """The ECG of a heart engine inside the 384x384x256 @ 1.5 mm synthetic torso:
    torso operator, multigrid-preconditioned refined solver, heart source,
    12-lead + n_bspm electrodes and the lead field."""

    def __init__(self, eng, meta, shape=(384, 384, 256), dx_mm=1.5, n_bspm=256, tol=1e-10,
                 log=print):
        from .ecg.functions.forward import (Electrodes, HeartSource, LeadField, PoissonSolver,
                                            standard_electrodes)
        from .ecg.functions.torso import TorsoGrid
        t0 = time.perf_counter()
        lab, origin, hc = G.torso(shape, dx_mm)
        self.grid = TorsoGrid.from_labels(lab, G.TORSO_SIGMA, dx_mm, origin)
        self.t_grid = time.perf_counter() - t0
        t1 = time.perf_counter()
        self.ps = PoissonSolver(self.grid, eng.dev, precond="mg")
        self.t_mg = time.perf_counter() - t1
        extent = meta["nodes_mm"].max(axis=0)
        self.reg = dict(R=np.eye(3), t=hc - extent / 2, scale=1.0)
        self.src = HeartSource(eng.lay, meta["nodes_mm"], self.grid, self.reg, 0.2625,
                               meta["voxel_size"], eng.dev, bricks=eng.bricks)
        self.el = Electrodes(self.grid, standard_electrodes(self.grid, hc, n_bspm),
                             device=eng.dev)
        log(f"[ecg] torso {shape} @ {dx_mm} mm: {self.ps.n:,} stored cells, grid "
            f"{self.t_grid:.1f} s, multigrid {self.t_mg:.1f} s ({len(self.ps.mg.shapes)} levels "
            f"down to {self.ps.mg.shapes[-1]}); {len(self.el.names)} electrodes; heart voxels "
            f"in torso {self.src.n_valid:,}")
        self.lf = LeadField(self.ps, self.src, self.el, tol=tol, log=log)
        inner = [s["inner"] for s in self.lf.solve_stats]
        self.info = dict(torso=list(shape), dx_mm=dx_mm, cells=self.ps.n, grid_s=round(self.t_grid, 1),
                         mg_setup_s=round(self.t_mg, 1), mg_levels=len(self.ps.mg.shapes),
                         electrodes=len(self.el.names), leadfield_setup_s=round(self.lf.setup_s, 1),
                         leadfield_tol=tol, leadfield_inner_iters_mean=float(np.mean(inner)),
                         heart_voxels_in_torso=self.src.n_valid)
        log(f"[ecg] lead field: {self.lf.setup_s:.1f} s for {len(inner)} solves "
            f"(mean {np.mean(inner):.1f} MG-PCG iterations)")

    def forward_solve_time(self, eng, n=3, tol=1e-8):
        """Per-frame forward solve (source + MG-PCG in FP64 refinement to tol,
        warm-started), the reference's per-frame method: mean seconds."""
        b = torch.zeros(self.ps.n, dtype=torch.float64, device=eng.dev)
        x = torch.zeros(self.ps.n, dtype=torch.float64, device=eng.dev)
        self.src(eng.V[eng.p], b, self.ps.ws)
        self.ps.solve_refined(b, x, tol=tol)
        ts, st = [], []
        for _ in range(n):
            eng.run(50, graph_chunk=50)
            b.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            self.src(eng.V[eng.p], b, self.ps.ws)
            self.ps.solve_refined(b, x, tol=tol)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            st.append(dict(self.ps.last_stats))
        return float(np.mean(ts)), st


class EcgRecorder:
    """Copies the engine's Vm every call into a frame batch and applies the lead
    field when the batch is full; potentials() returns (n_frames, ne) in mV."""

    def __init__(self, lf, eng):
        self.lf, self.eng = lf, eng
        self.frames = torch.zeros((lf.batch, eng.lay.n_slots), dtype=torch.float32,
                                  device=eng.dev)
        self.nb = 0
        self.out = []

    def record(self):
        self.frames[self.nb].copy_(self.eng.V[self.eng.p])
        self.nb += 1
        if self.nb == self.lf.batch:
            self.flush()

    def flush(self):
        if self.nb:
            self.out.append((self.lf.apply(self.frames, self.nb) * 1e3).clone())
            self.nb = 0

    def potentials(self):
        self.flush()
        return torch.cat(self.out).cpu().numpy() if self.out else np.zeros((0, self.lf.ne))


def twelve_leads(names, phi):
    from .ecg.functions.forward import LEAD_NAMES, leads_from_electrodes
    d = {n: phi[:, i] for i, n in enumerate(names)}
    L = leads_from_electrodes(d)
    return np.stack([L[k] for k in LEAD_NAMES], axis=1)


def lead_agreement(a, b, per_lead=False):
    """min over the 12 leads of the correlation of a with b, and the max relative
    L2 error; with per_lead, also the list of (r, relative L2) of every lead."""
    rs, rel = [], []
    for i in range(a.shape[1]):
        x, y = a[:, i], b[:, i]
        if np.std(x) == 0 or np.std(y) == 0:
            rs.append(None)
            rel.append(None)
            continue
        rs.append(float(np.corrcoef(x, y)[0, 1]))
        rel.append(float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30)))
    r_ok = [r for r in rs if r is not None]
    e_ok = [e for e in rel if e is not None]
    out = (min(r_ok) if r_ok else None), (max(e_ok) if e_ok else None)
    if per_lead:
        return out + ([(None if r is None else round(r, 6), None if e is None else round(e, 4))
                       for r, e in zip(rs, rel)],)
    return out
