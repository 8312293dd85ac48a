#!/usr/bin/env python3
"""Benchmark: voxel-steps/s of the monodomain hot path on one MI355X.

Workload (default): the 2M-voxel synthetic biventricular shell -- two
truncated confocal ellipsoids voxelised at 0.4 mm in a 256^3 grid, the
Laplace transmural coordinate and the rule-based helix fibres (+60 -> -60
deg) computed on the GPU, sigma_l = 0.24, sigma_t = 0.08 S/m, endo/mid/epi
TTP06 by transmural thirds, a seeded Purkinje tree with 800 PMJs coupled at
c_pmj = 2 mS (500 Ohm), FP32, dt = 0.02 ms, plain Rush-Larsen; stimulus at the
Purkinje root and at an LV endocardial apical site (2 ms, 52 pA/pF), ECG (12
leads + 256 BSPM electrodes) every 1 ms in a 384x384x256 @ 1.5 mm torso.

One bench step is 1 ms of simulated time: 50 solver steps of 0.02 ms of the
whole coupled system (myocardium + Purkinje + PMJ + stimulus) and one ECG
frame; --steps K times exactly K of them.  Forms, each run from t = 0 over the
full 600 ms beat (the timed window is part of that run):
  plain     the reference's scheme in FP32: plain Rush-Larsen TTP06 reaction,
            then Crank-Nicolson diffusion of the coupled system solved to the
            reference's relative residual cg_tol = 1e-5 by Jacobi sweeps
            (hp_cn_step, solver 0)
  bicgstab  the same with the reference's own solver, Jacobi-preconditioned
            BiCGSTAB with its stopping rules (solver 1, the drop-in default)
  explicit  the fast form: forward-Euler diffusion fused with the reaction
  bigstep / surrogate  the multirate forms (one 0.2 ms reaction per
            macro-step by the exact model / by the exact model corrected by
            the learned network, plain integration where the upstroke
            indicator fires)
  ref64     the reference scheme with its BiCGSTAB in FP64 (the oracle's
            algorithm, equal to it at round-off level on the test problems):
            the gates' reference
Every form's activation times and 12 leads over the full beat are compared
with ref64 (AT error <= 1 ms, lead correlation >= 0.999).  The headline value
is the surrogate form's throughput if it meets both gates, else the plain
form's.

Prints one JSON line with the throughput, a roofline object for the dominant
kernel of the plain form (the CN Jacobi sweep: 19-point stencil, measured
alone with HIP events on its stream; the BiCGSTAB form's product pass beside
it) and a CPU baseline (the oracle, CN + BiCGSTAB, timed on a bounded number
of steps of the same workload).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "voxel-steps/s, 2M-voxel BiV: surrogate vs plain FP32 at AT err≤1ms, ECG r≥.999"
HBM_PEAK_GBS = 8000.0
# CN Jacobi sweep, per myocardial voxel: x read 4, D tensor 24, label 1, b read 4,
# x written 4 (the first sweep of a step writes b instead of reading it)
SWEEP_BYTES_PER_VOXEL = 37
# BiCGSTAB product pass t = A M^-1 (r - alpha v), per myocardial voxel: r, v,
# M^-1 read 12, D tensor 24, label 1, t written 4
APPLY_BYTES_PER_VOXEL = 41
# CN reaction pass, per voxel: 18 state variables read and written 144, V read 4,
# V* written 4, label 1
REACT_BYTES_PER_VOXEL = 153
STEPS_PER_MS = 50
# the 22-128-128-128-19 network: 2 x (22*128 + 128*128 + 128*128 + 128*19) FLOP per voxel
MLP_FLOP_PER_VOXEL = 76032
BF16_PEAK_TFS = 2500.0
# CN solve per step captured in the step's graph: BiCGSTAB iterations (the
# reference's solver) / Jacobi sweeps; the ones after the solve ended return
# at once
CN_MAX_ITERS = {"bicgstab": 6, "jacobi": 5}
TRAFFIC_FILE = os.path.join(HERE, "profiles", "cn_sweep_biv256_traffic.json")


def measured_traffic():
    """HBM bytes per launch of the dominant kernel (one CN Jacobi sweep) from
    the committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE,
    scaled by the streaming-copy calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def build_geometry(n=256, h=0.4, seed=0):
    from hrdayapy_amd import geometry as G
    t0 = time.perf_counter()
    geo = G.biv_shell(n, h, seed=seed, coords="laplace")
    nodes, el, act = G.purkinje_tree(geo, 800, seed=seed)
    kw = dict(voxel_size=h, dx_p=0.2, sigma_P=0.45, sigma_l=0.24, sigma_t=0.08, Cm=0.01,
              A_M=140.0, R_P=0.015, c_pmj=2.0, n_pmj=1, stim_len_mm=2.0, T=600.0,
              phi_endo_max=1.0 / 3.0, phi_epi_min=2.0 / 3.0,
              ectopic_region=G.apical_stimulus_region(geo, 3.0),
              stim_protocol=[dict(name="av_node", target="root", amp=12.0, dur=2.0, onset=0.0),
                             dict(name="apex", target="ectopic", amp=0.52, dur=2.0, onset=0.0)])
    return geo, (nodes, el, act), kw, time.perf_counter() - t0


def build_engine(geo, tree, kw, dtype=None, brick=None, alloc="torch", scheme="cn",
                 max_sweeps=None, cn_solver="bicgstab"):
    import torch
    from hrdayapy_amd.anisotropic_simulation.functions.purkinje_myocardium_anisotropic_solver import (
        build_coupled_system)
    nodes, el, act = tree
    return build_coupled_system(nodes, el, act, geo["S"], geo["phi"], geo["fibre_dir"], dt=0.02,
                                dtype=dtype or torch.float32, device="cuda", brick=brick,
                                alloc=alloc, scheme=scheme, cn_solver=cn_solver,
                                max_sweeps=max_sweeps or CN_MAX_ITERS[cn_solver], **kw)


def build_workload(n=256, h=0.4, seed=0, dtype=None, brick=None, scheme="cn"):
    geo, tree, kw, t_geo = build_geometry(n, h, seed)
    eng, meta = build_engine(geo, tree, kw, dtype, brick, scheme=scheme)
    return geo, tree, kw, eng, meta, t_geo


FORM_SETUP = {  # form -> (scheme, dtype name, multirate rule, engine options)
    # the reference's step with its own solver, CN + Jacobi-preconditioned
    # BiCGSTAB, in FP64: the gates' reference
    "ref64": ("cn", "float64", None, dict(cn_solver="bicgstab")),
    # plain Rush-Larsen reaction + CN diffusion in FP32, the CN system solved
    # by Jacobi sweeps to the reference's relative residual
    "plain": ("cn", "float32", None, dict(cn_solver="jacobi")),
    # the same in FP32 with the reference's BiCGSTAB (the drop-in's default)
    "bicgstab": ("cn", "float32", None, dict(cn_solver="bicgstab")),
    "explicit": ("explicit", "float32", None, {}),
    # multirate forms: one reaction per 0.2 ms macro-step (exact big step /
    # network), ten substeps of the plain form's CN diffusion driven by the
    # rate (fallback voxels run the full cell model every substep)
    "bigstep": ("cn", "float32", "bigstep", dict(cn_solver="jacobi")),
    "surrogate": ("cn", "float32", "mlp", dict(cn_solver="jacobi")),
    # the big-step rule with cold-started sweeps (x0 = V*, no warm start)
    "bigstep_cold": ("cn", "float32", "bigstep", dict(cn_solver="jacobi", cn_warm=False)),
    # the same rules on explicit substeps (the earlier multirate form)
    "bigstep_explicit": ("explicit", "float32", "bigstep", {}),
    "surrogate_explicit": ("explicit", "float32", "mlp", {}),
    # the CN step with theta = 0 and one sweep: V^{n+1} = V* + (dt/Cm) L V*
    # exactly (forward Euler after the reaction, the reference's Lie order)
    "lie0": ("cn", "float32", None, dict(theta=0.0, max_sweeps=1, cn_solver="jacobi")),
    # Jacobi sweeps, at most two per step
    "cn2": ("cn", "float32", None, dict(max_sweeps=2, cn_solver="jacobi"))}


def run_form_beat(form, geo, tree, kw, brick, ecg, beat_ms, warm_ms, timed_ms, sample_slots,
                  net=None, log=print, graph_macros=5):
    """One form from t = 0 over the whole beat, ECG every 1 ms (lead field) and a
    Vm sample every 1 ms; bench steps [warm_ms, warm_ms + timed_ms) are timed
    (device synchronised at both ends).  Returns the timing, the activation
    times, the electrode potentials and the Vm sample (device)."""
    import torch
    from hrdayapy_amd import surrogate as SG
    from hrdayapy_amd import workloads as W
    scheme, dname, rule, opts = FORM_SETUP[form]
    kwf = dict(kw, **{k: v for k, v in opts.items() if k not in ("max_sweeps", "cn_solver")})
    eng, meta = build_engine(geo, tree, kwf, dtype=getattr(torch, dname), brick=brick,
                             scheme=scheme, max_sweeps=opts.get("max_sweeps"),
                             cn_solver=opts.get("cn_solver", "bicgstab"))
    run = SG.SurrogateRunner(eng, net, mode=rule, graph_macros=graph_macros) if rule else None
    # Jacobi forms: each 1 ms chunk bounded by the sweeps the last one needed
    # (replayed from its kept state with the full bound if a step stopped there)
    jac = eng.cn is not None and eng.cn_solver == "jacobi"
    advance = (lambda: run.run(STEPS_PER_MS // SG.N_SUB, adaptive=jac)) if run else (
        lambda: eng.run(STEPS_PER_MS, graph_chunk=STEPS_PER_MS, adaptive=jac))
    rec = W.EcgRecorder(ecg.lf, eng) if ecg else None
    vm = torch.empty((beat_ms, len(sample_slots)), dtype=torch.float32, device=eng.dev)
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    timed = None
    for k in range(beat_ms):
        if k == warm_ms:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        advance()
        if rec:
            rec.record()
        vm[k] = eng.V[eng.p][sample_slots].float()
        if k == warm_ms + timed_ms - 1:
            torch.cuda.synchronize()
            timed = time.perf_counter() - t0
        if (k + 1) % 100 == 0:
            log(f"[bench] {form}: {k + 1}/{beat_ms} ms")
    phi = rec.potentials() if rec else None
    torch.cuda.synchronize()
    beat_s = time.perf_counter() - t_all
    n_myo = meta["N_myo"]
    out = dict(value=n_myo * STEPS_PER_MS * timed_ms / timed, ms_per_step=timed / timed_ms * 1e3,
               timed_window_ms=[warm_ms, warm_ms + timed_ms],
               beat_s=round(beat_s, 3), beat_voxel_steps_per_s=n_myo * STEPS_PER_MS * beat_ms / beat_s,
               setup_s=round(meta["setup_s"], 1))
    if eng.cn is not None:
        out["cn_solver"] = eng.cn_stats()
        if jac:
            out["cn_solver"].update(adaptive_bound=True, warm_start=bool(eng.cn_warm),
                                    chunks_replayed=int(getattr(eng, "retries", 0)))
    if run is not None:
        out["fallback_fraction"] = round(run.fallback_fraction(), 4)
        out["reaction_us"] = round(W.graph_us(lambda i: run.launch_reaction(), 20), 2)
        out["substep_us"] = round(W.graph_us(lambda i: run.launch_substep(), 20), 2)
    at = eng.at_nodes().double().cpu().numpy()
    return out, at, phi, vm, eng, meta


def gates(at, phi, vm, ref, names):
    """Activation-time, Vm-sample and 12-lead agreement of a form with ref64."""
    from hrdayapy_amd import workloads as W
    at_r, phi_r, vm_r = ref
    both = np.isfinite(at) & np.isfinite(at_r)
    g = dict(at_err_max_ms=float(np.max(np.abs(at[both] - at_r[both]))) if both.any() else None,
             at_err_mean_ms=float(np.mean(np.abs(at[both] - at_r[both]))) if both.any() else None,
             activated=int(np.isfinite(at).sum()), activated_ref=int(np.isfinite(at_r).sum()),
             vm_rms_mV=float(((vm - vm_r) ** 2).mean().sqrt()),
             vm_max_abs_mV=float((vm - vm_r).abs().max()))
    if phi is not None:
        r, rel, per = W.lead_agreement(W.twelve_leads(names, phi), W.twelve_leads(names, phi_r),
                                       per_lead=True)
        from hrdayapy_amd.ecg.functions.forward import LEAD_NAMES
        g.update(ecg_r_min=r, ecg_rel_l2_max=rel, ecg_per_lead_r_rel=dict(zip(LEAD_NAMES, per)))
    g["pass"] = bool(g["at_err_max_ms"] is not None and g["at_err_max_ms"] <= 1.0
                     and g["activated"] == g["activated_ref"]
                     and (phi is None or (g.get("ecg_r_min") or 0.0) >= 0.999))
    return g


def stream_copy_bandwidth(mib, reps=20):
    """Achieved read+write bandwidth of the plain streaming-copy kernel."""
    import torch
    from hrdayapy_amd import _native as nat
    from hrdayapy_amd import workloads as W
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    us = W.graph_us(lambda i: nat.call("hp_stream_copy_f32", nat.ptr(x), nat.ptr(y), n,
                                       nat.stream_handle()), reps)
    del x, y
    torch.cuda.empty_cache()
    return dict(mib=mib, us=round(us, 2), GBs=2 * n * 4 / (us * 1e-6) / 1e9)


def time_cn_kernels(eng, eng_b, reps=100, steps=True):
    """Device times (us), each replayed from a HIP graph between HIP events on
    the replaying stream: one Jacobi sweep of the myocardial rows
    (hp_cn_sweep_f32, the plain form's dominant kernel, the 19-point stencil)
    and the plain form's whole step (eng, solver 0); one BiCGSTAB product pass
    t = A M^-1 (r - alpha v) (hp_cn_apply_f32) and the whole step of the
    BiCGSTAB form (eng_b, solver 1)."""
    import torch
    from hrdayapy_amd import _native as nat
    from hrdayapy_amd import workloads as W
    Vs, bvec, x = eng.cn_work_t[:3]

    def sweep(i):
        src, dst = (Vs, x) if i % 2 == 0 else (x, Vs)
        nat.call("hp_cn_sweep_f32", C.byref(eng.bricks), nat.ptr(src), nat.ptr(dst),
                 nat.ptr(bvec), C.byref(eng.cn), nat.stream_handle())
    out = dict(sweep_us=W.graph_us(sweep, reps))
    if steps:
        out["step_us"] = W.graph_us(lambda i: eng._launch_step(), reps)
    sc = torch.zeros(16, dtype=torch.float64, device=eng_b.dev)  # done = 0, alpha = 0
    cp = nat.HpCnParams.from_buffer_copy(eng_b.cn)
    cp.scalars = nat.ptr(sc)
    out["apply_us"] = W.graph_us(lambda i: nat.call("hp_cn_apply_f32", C.byref(eng_b.bricks),
                                                    C.byref(cp), C.byref(eng_b.cn_work),
                                                    nat.stream_handle()), reps)
    if steps:
        out["bicgstab_step_us"] = W.graph_us(lambda i: eng_b._launch_step(), reps)
    return out


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, aff


def cpu_baseline(args_geo, kw, n_steps, dtype_name="float32"):
    """The oracle (CPU restatement of the reference: Lie split, TTP06, CN + BiCGSTAB)
    timed on this host over n_steps of the same workload, on every core this
    process may use (bounded by the OMP_NUM_THREADS share the host sets)."""
    import torch
    from oracle import monodomain as OMD
    geo, (nodes, el, act) = args_geo
    model, aff = cpu_info()
    share = int(os.environ.get("OMP_NUM_THREADS", aff) or aff)
    cores = max(1, min(aff, share))
    torch.set_num_threads(cores)
    prob = OMD.Problem(nodes, el, act, geo["S"], geo["phi"], geo["fibre_dir"], **kw)
    dt_ = torch.float32 if dtype_name == "float32" else torch.float64
    OMD.run(prob, 0.02 * 2, 0.02, scheme="cn", dtype=dt_)  # warm caches
    r = OMD.run(prob, 0.02 * (n_steps + 0.5), 0.02, scheme="cn", dtype=dt_, max_steps=n_steps)
    vs = prob.N_myo * r["steps"] / r["wall_s"]
    return dict(value=vs, unit="voxel-steps/s", cores=cores, kind="port",
                cpu_model=model, affinity_cpus=aff,
                sample=f"{r['steps']} CN+BiCGSTAB steps of the same {prob.N_myo:,}-voxel BiV "
                       f"(+{prob.Np:,} Purkinje nodes), TTP06 in {dtype_name}, sparse algebra "
                       f"FP64 (scipy), {r['wall_s']:.1f} s on {cores} threads")


def config1_workload(a):
    """Configuration 1 in full: the 64x64x16 slab at 0.25 mm, isotropic 0.13 S/m,
    endocardial TTP06, FP64, dt 0.02 ms for 500 ms, 52 pA/pF for 2 ms on the 4x4x4
    corner, inert Purkinje stub: the oracle (CN + BiCGSTAB) on the host cores
    and the GPU CN path in FP64 and FP32, with their activation times compared."""
    import torch
    from hrdayapy_amd import geometry as G
    from hrdayapy_amd.anisotropic_simulation.functions.purkinje_myocardium_anisotropic_solver import (
        build_coupled_system)
    from oracle import monodomain as OMD
    log = lambda m: print(m, flush=True)
    S, _, fib = G.slab()
    nodes, el, act = G.stub_tree(S.shape)
    region = np.zeros(S.shape, dtype=bool)
    region[:4, :4, :4] = True
    T = a.config1_ms
    kw = dict(voxel_size=0.25, dx_p=0.25, sigma_P=0.0, sigma_l=0.13, sigma_t=0.13, Cm=0.01,
              A_M=140.0, c_pmj=0.0, T=T, ectopic_region=region, phi_endo_max=0.35,
              phi_epi_min=0.65,
              stim_protocol=[dict(name="corner", target="ectopic", amp=0.52, dur=2.0, onset=0.0)])
    phi = np.zeros(S.shape, np.float32)
    out = {}
    for dname in ("float64", "float32"):
        eng, meta = build_coupled_system(nodes, el, act, S, phi, fib, dt=0.02,
                                         dtype=getattr(torch, dname), device="cuda", scheme="cn",
                                         **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.run(int(round(T / 0.02)), graph_chunk=50)
        torch.cuda.synchronize()
        w = time.perf_counter() - t0
        out[dname] = dict(wall_s=w, voxel_steps_per_s=meta["N_myo"] * T / 0.02 / w,
                          at=eng.at_nodes().double().cpu().numpy(), cn=eng.cn_stats())
        log(f"[config1] GPU CN {dname}: {w:.2f} s ({out[dname]['voxel_steps_per_s']:.3e} "
            f"voxel-steps/s) {out[dname]['cn']}")
    model, aff = cpu_info()
    cores = max(1, min(aff, int(os.environ.get("OMP_NUM_THREADS", aff) or aff)))
    torch.set_num_threads(cores)
    prob = OMD.Problem(nodes, el, act, S, phi, fib, **kw)
    steps = int(round(T / 0.02)) if a.config1_cpu_steps <= 0 else a.config1_cpu_steps
    log(f"[config1] oracle on {cores} threads: {steps} steps ({model})")
    t_cpu = time.perf_counter()
    r = OMD.run(prob, 0.02 * (steps + 0.5), 0.02, scheme="cn", max_steps=steps,
                progress=lambda i, n: log(f"[config1] oracle step {i}/{n} "
                                          f"({time.perf_counter() - t_cpu:.0f} s)"))
    cpu_vs = prob.N_myo * r["steps"] / r["wall_s"]
    res = {"metric": "voxel-steps/s, configuration 1 (64x64x16 slab, FP64, 500 ms)",
           "value": out["float64"]["voxel_steps_per_s"], "unit": "voxel-steps/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"64x64x16 slab @0.25 mm, {prob.N_myo} voxels, {T} ms"},
           "gpu": {k: dict(wall_s=v["wall_s"], voxel_steps_per_s=v["voxel_steps_per_s"],
                           cn=v["cn"]) for k, v in out.items()},
           "cpu_baseline": {"value": cpu_vs, "unit": "voxel-steps/s", "cores": cores,
                            "kind": "port", "cpu_model": model, "steps": r["steps"],
                            "wall_s": r["wall_s"]}}
    if r["steps"] == int(round(T / 0.02)):
        ra = r["at"][prob.Np:]
        for dname in ("float64", "float32"):
            at = out[dname]["at"]
            both = np.isfinite(at) & np.isfinite(ra)
            res["gpu"][dname]["at_err_max_vs_oracle_ms"] = float(np.abs(at[both] - ra[both]).max())
            res["gpu"][dname]["activated"] = [int(np.isfinite(at).sum()), int(np.isfinite(ra).sum())]
    print(json.dumps(res))


def other_workload(a):
    """Torso Poisson, ensemble, ionic-kernel or stencil measurement: one JSON line of its own."""
    from hrdayapy_amd import workloads as W
    log = lambda m: print(m, flush=True)
    if a.workload == "torso":
        geo, tree, kw, _ = build_geometry(n=a.grid, h=0.4 * 256 / a.grid)
        eng, meta = build_engine(geo, tree, kw, brick=(4, 4, 4))
        r = W.torso_solve(eng, meta, log=log)
        out = {"metric": "torso Poisson PCG iteration bandwidth", "value": r["pcg_iter_GBs"],
               "unit": "GB/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32+f64",
               "data": "synthetic",
               "config": {"workload": "torso 384x384x256 @1.5 mm, 2M-voxel BiV source, FP32 "
                                      "Jacobi-PCG in FP64 refinement to rel. residual 1e-8"},
               "result": r,
               "roofline": {"bound": "hbm", "achieved": r["pcg_iter_GBs"], "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": r["pcg_roofline_frac"], "traffic": None}}
    elif a.workload == "stencil":
        geo, tree, kw, _ = build_geometry(n=a.grid, h=0.4 * 256 / a.grid)
        brick = (4, 4, 4) if a.brick == "auto" else tuple(int(x) for x in a.brick.split("x"))
        eng, meta = build_engine(geo, tree, kw, brick=brick)
        eng.run(500, graph_chunk=a.graph_chunk)  # a wave under way
        r = W.stencil_run(eng, reps=a.kernel_reps)
        log(f"[stencil] {r}")
        gbs = r["hbm_GBs"]
        out = {"metric": "19-point stencil sweep bandwidth (L.V alone)", "value": gbs, "unit": "GB/s",
               "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
               "config": {"workload": f"synthetic BiV 256^3 @0.4 mm, {r['voxels']:,} voxels in "
                                      f"{r['bricks']:,} bricks {tuple(r['brick'])}"},
               "result": r,
               "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                            "bytes_per_voxel": W.STENCIL_BYTES_PER_VOXEL}}
    elif a.workload == "mlp":
        # the learned correction's inference kernel with every voxel of the BiV
        # through the network (rate set to a number: no fallback, no voxel at rest)
        import torch
        from hrdayapy_amd import _native as nat
        from hrdayapy_amd import surrogate as SG
        geo, tree, kw, _ = build_geometry(n=a.grid, h=0.4 * 256 / a.grid)
        brick = (8, 4, 4) if a.brick == "auto" else tuple(int(x) for x in a.brick.split("x"))
        eng, meta = build_engine(geo, tree, kw, brick=brick, cn_solver="jacobi")
        net = SG.SurrogateNet("cuda", seed=0)
        net.correct = True
        # output layer zeroed: the correction is 0, so repeated calls leave the
        # state and the rate as they are (same work per call)
        net.W[3].zero_()
        net.b[3].zero_()
        net.refresh()
        run = SG.SurrogateRunner(eng, net, mode="mlp")
        state0 = eng.state.clone()
        res = {}
        for variant in (2, 1, 0):
            sp = nat.HpSurrogateParams.from_buffer_copy(run.sp)
            sp.variant = variant
            run.rate.fill_(1.0)

            def launch(i, sp=sp):
                nat.call("hp_surrogate_step", C.byref(eng.bricks), nat.ptr(eng.V[eng.p]),
                         nat.ptr(run.V_prev), nat.ptr(eng.state), eng.state_ld,
                         nat.ptr(run.rate), C.byref(net.hp_w), C.byref(net.hp_norm),
                         C.byref(sp), 0, nat.stream_handle())
            res[variant] = W.graph_us(launch, a.kernel_reps)
            eng.state.copy_(state0)
            log(f"[mlp] variant {variant}: {res[variant]:.1f} us per call")
        n_myo = meta["N_myo"]
        flops = MLP_FLOP_PER_VOXEL * n_myo
        tfs = flops / (res[2] * 1e-6) / 1e12
        out = {"metric": "learned-correction inference, voxels/s (every voxel through the network)",
               "value": n_myo / (res[2] * 1e-6), "unit": "voxels/s", "n_gpus": 1,
               "higher_is_better": True, "dtype": "bf16 MFMA, fp32 accumulation",
               "data": "synthetic (random weights, resting state)",
               "config": {"workload": f"synthetic BiV 256^3 @0.4 mm, {n_myo:,} voxels, "
                                      "22-128-128-128-19 SiLU MLP"},
               "result": {"us_per_call": {str(k): round(v, 2) for k, v in res.items()}},
               "roofline": {"bound": "mfma", "achieved": tfs, "peak": BF16_PEAK_TFS,
                            "unit": "TFLOP/s", "frac": tfs / BF16_PEAK_TFS, "traffic": None,
                            "flop_per_voxel": MLP_FLOP_PER_VOXEL}}
    elif a.workload == "ionic":
        r = W.ionic_run(reps=a.kernel_reps)
        log(f"[ionic] {r}")
        out = {"metric": "TTP06 cell-steps/s, ionic kernel alone", "value": r["cell_steps_per_s"],
               "unit": "cell-steps/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
               "data": "synthetic",
               "config": {"workload": f"{r['cells']:,} cells, seeded phenotype and Vm, resting "
                                      "state, dt 0.02 ms, SoA state"},
               "result": r,
               "roofline": {"bound": "valu", "achieved": r["valu_per_s"] / 1e12,
                            "peak": W.VALU_PEAK_PER_S / 1e12, "unit": "T VALU lane-ops/s",
                            "frac": r["valu_frac"], "traffic": None,
                            "transcendental_frac": r["trans_frac"]}}
    else:
        brick = None if a.brick == "auto" else tuple(int(x) for x in a.brick.split("x"))
        r = W.ensemble_run(a.ensemble_size, 0, a.steps, a.warmup,
                           0 if a.graph_chunk_auto else a.graph_chunk, brick, log=log)
        out = {"metric": "voxel-steps/s, batched ensemble", "value": r["value"],
               "unit": "voxel-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
               "ms_per_step": r["ms_per_step"], "higher_is_better": True, "dtype": "f32",
               "data": "synthetic",
               "config": {"workload": f"{a.ensemble_size} instances (slabs and BiV shells, "
                                      "0.2M-2M voxels), FP32, dt 0.02 ms"},
               "result": r}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50,
                    help="bench steps timed (1 ms of simulated time each: 50 solver steps of "
                         "0.02 ms and one ECG frame)")
    ap.add_argument("--warmup", type=int, default=5, help="bench steps run before the timed ones")
    ap.add_argument("--beat-ms", type=int, default=600,
                    help="length of every form's run (the gates compare the whole run)")
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--brick", type=str, default="auto", help="e.g. 8x4x4, or auto")
    ap.add_argument("--forms", type=str, default="plain,bicgstab,explicit,bigstep,surrogate")
    ap.add_argument("--train-samples", type=int, default=50_000_000)
    ap.add_argument("--train-epochs", type=int, default=14)
    ap.add_argument("--train-biv", type=int, default=160,
                    help="grid of the extra BiV training run (another seed than the bench's "
                         "BiV; 0 = slab runs only)")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=100)
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the CN Jacobi sweep and BiCGSTAB product pass alone, then exit")
    ap.add_argument("--calib-copy-mb", type=int, default=1024,
                    help="size of the streaming-copy calibration (0 = off)")
    ap.add_argument("--workload", type=str, default="biv",
                    choices=["biv", "torso", "ensemble", "ionic", "stencil", "config1", "mlp"],
                    help="biv: the coupled 2M-voxel BiV with ECG (default); torso: the "
                         "384x384x256 Poisson solve; ensemble: 64 instances in one brick list; "
                         "config1: the 64x64x16 FP64 slab in full, GPU and CPU oracle")
    ap.add_argument("--ensemble-size", type=int, default=64)
    ap.add_argument("--graph-chunk", type=int, default=50)
    ap.add_argument("--graph-macros", type=int, default=5,
                    help="multirate forms: 0.2 ms macro-steps per captured graph (and per "
                         "adaptive sweep-bound chunk); must divide 5")
    ap.add_argument("--config1-ms", type=float, default=500.0)
    ap.add_argument("--config1-cpu-steps", type=int, default=0, help="0 = the whole run")
    ap.add_argument("--ecg", type=str, default="leadfield", choices=["leadfield", "none"])
    ap.add_argument("--bspm", type=int, default=256, help="BSPM electrodes besides the 12-lead set")
    a = ap.parse_args()
    a.graph_chunk_auto = False
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implemented path)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.workload == "config1":
        return config1_workload(a)
    if a.workload != "biv":
        return other_workload(a)
    from hrdayapy_amd import workloads as W
    forms = [f for f in a.forms.split(",") if f]
    log = lambda m: print(m, flush=True)
    if a.warmup + a.steps > a.beat_ms:
        raise SystemExit("--warmup + --steps must fit in --beat-ms")
    t_start = time.perf_counter()
    log(f"[bench] building geometry grid={a.grid}")
    geo, tree, kw, t_geo = build_geometry(n=a.grid, h=0.4 * 256 / a.grid)
    tune = None
    if a.brick == "auto":
        from hrdayapy_amd.autotune import tune as autotune
        brick, _, tune = autotune(lambda b: build_engine(geo, tree, kw, brick=b,
                                                         cn_solver="jacobi")[0],
                                  shapes=((4, 4, 4), (8, 4, 4), (8, 8, 4)), chunks=(50,),
                                  steps=100, log=log)
        log(f"[bench] autotuned brick {brick} for the plain (CN) form: {tune}")
    else:
        brick = tuple(int(x) for x in a.brick.split("x"))
    eng0, meta = build_engine(geo, tree, kw, brick=brick, cn_solver="jacobi")
    n_myo = meta["N_myo"]
    log(f"[bench] N_myo={n_myo:,} Np={meta['Np']:,} bricks={eng0.lay.n_bricks:,} "
        f"{eng0.lay.brick} occ={eng0.lay.occupancy:.3f} PMJ={len(meta['pmj_ids'])} "
        f"geo {t_geo:.1f}s setup {meta['setup_s']:.1f}s")
    ecg = W.EcgSetup(eng0, meta, n_bspm=a.bspm, log=log) if a.ecg == "leadfield" else None
    # dominant kernel of the plain form and its whole step (device time); the
    # BiCGSTAB form's product pass and step
    eng_b, _ = build_engine(geo, tree, kw, brick=brick, cn_solver="bicgstab")
    if a.kernels_only:  # the two solver kernels alone (for a profiler run)
        kt = time_cn_kernels(eng0, eng_b, a.kernel_reps, steps=False)
        kt["copy"] = stream_copy_bandwidth(a.calib_copy_mb) if a.calib_copy_mb else None
        log(f"[bench] kernels only: {json.dumps(kt)}")
        return
    kt = time_cn_kernels(eng0, eng_b, a.kernel_reps)
    del eng_b
    sweep_us, step_us = kt["sweep_us"], kt["step_us"]
    sweep_bytes = SWEEP_BYTES_PER_VOXEL * n_myo
    achieved = sweep_bytes / (sweep_us * 1e-6) / 1e9
    apply_gbs = APPLY_BYTES_PER_VOXEL * n_myo / (kt["apply_us"] * 1e-6) / 1e9
    log(f"[bench] CN Jacobi sweep {sweep_us:.1f} us/launch -> {achieved:.0f} GB/s algorithmic "
        f"({sweep_bytes / 1e6:.1f} MB); plain CN step replayed {step_us:.1f} us; BiCGSTAB "
        f"product pass {kt['apply_us']:.1f} us ({apply_gbs:.0f} GB/s), BiCGSTAB step "
        f"{kt['bicgstab_step_us']:.1f} us")
    ecg_info = None
    if ecg:
        rec0 = W.EcgRecorder(ecg.lf, eng0)
        for _ in range(ecg.lf.batch):
            rec0.record()
        lf_us = W.graph_us(lambda i: ecg.lf.apply(rec0.frames, ecg.lf.batch), 5)
        fw_s, fw_stats = ecg.forward_solve_time(eng0)
        ecg_info = dict(ecg.info, method="leadfield", every_ms=1.0,
                        leadfield_apply_us_per_frame=round(lf_us / ecg.lf.batch, 1),
                        forward_solve_ms_per_frame=round(fw_s * 1e3, 2),
                        forward_solve_stats=fw_stats)
        log(f"[bench] ECG: lead field {lf_us / ecg.lf.batch:.1f} us/frame; per-frame forward "
            f"solve (MG-PCG, FP64 refinement to 1e-8) {fw_s * 1e3:.1f} ms")
        del rec0
    copy = stream_copy_bandwidth(a.calib_copy_mb) if a.calib_copy_mb else None
    if copy:
        log(f"[bench] streaming copy {a.calib_copy_mb} MiB: {copy['us']:.1f} us -> "
            f"{copy['GBs']:.0f} GB/s (read+write)")
    rng = np.random.default_rng(0)
    sample_nodes = np.sort(rng.choice(n_myo, size=min(n_myo, 32768), replace=False))
    sample_slots = torch.from_numpy(eng0.lay.slot_of_node[sample_nodes].astype(np.int64)).to(
        eng0.dev)
    del eng0
    torch.cuda.empty_cache()
    net, net_info = None, None
    if any(f.startswith("surrogate") for f in forms):
        from hrdayapy_amd import surrogate as SG
        t1 = time.perf_counter()
        X, D = SG.training_runs("cuda", a.train_samples,
                                biv=(a.train_biv, 1) if a.train_biv else None)
        t_data = time.perf_counter() - t1
        net, tr, st = SG.train_surrogate(X, D, "cuda", epochs=a.train_epochs,
                                         log=lambda m: log("[bench] train " + m))
        net_info = dict(samples=int(X.shape[0]), data_s=round(t_data, 1),
                        train_s=round(st["train_s"], 1), role="correction of the big step",
                        heldout_dv_mae_mV=st["heldout_dv_mae"],
                        heldout_dv_max_mV=st["heldout_dv_max"],
                        heldout_bigstep_dv_mae_mV=st.get("heldout_bigstep_dv_mae"),
                        heldout_bigstep_dv_max_mV=st.get("heldout_bigstep_dv_max"))
        del X, D, tr
        torch.cuda.empty_cache()
        log(f"[bench] surrogate trained: {net_info}")
    names = ecg.lf.names if ecg else None
    results, ref = {}, None
    for form in ["ref64"] + forms:
        r, at, phi, vm, eng, m = run_form_beat(form, geo, tree, kw, brick, ecg, a.beat_ms,
                                               a.warmup, a.steps, sample_slots,
                                               net=net if form.startswith("surrogate") else None,
                                               log=log, graph_macros=a.graph_macros)
        if form == "ref64":
            ref = (at, phi, vm)
            r["activated"] = int(np.isfinite(at).sum())
        else:
            r["vs_ref64"] = gates(at, phi, vm, ref, names)
        if form.startswith("surrogate"):
            r["network"] = net_info
        results[form] = r
        log(f"[bench] {form}: {json.dumps(r)}")
        del eng, vm
        torch.cuda.empty_cache()
    plain_v = results["plain"]["value"] if "plain" in results else None
    for f in ("explicit", "bigstep", "surrogate", "bigstep_cold", "bigstep_explicit",
              "surrogate_explicit",
              "cn2"):
        if f in results and plain_v:
            results[f]["speedup_vs_plain"] = results[f]["value"] / plain_v
            results[f]["beat_speedup_vs_plain"] = (results["plain"]["beat_s"]
                                                   / results[f]["beat_s"])
    cpu = None
    if not a.no_cpu_baseline:
        log(f"[bench] CPU baseline ({a.cpu_steps} steps)")
        cpu = cpu_baseline((geo, tree), kw, a.cpu_steps)
    sr = results.get("surrogate")
    head = "surrogate" if (sr and sr["vs_ref64"]["pass"]) else (
        "plain" if "plain" in results else forms[0])
    hr = results[head]
    e2e = dict(geometry_s=round(t_geo, 1), engine_setup_s=hr["setup_s"],
               leadfield_build_s=ecg_info["leadfield_setup_s"] if ecg_info else 0.0,
               torso_grid_s=ecg_info["grid_s"] if ecg_info else 0.0,
               mg_setup_s=ecg_info["mg_setup_s"] if ecg_info else 0.0,
               beat_s=hr["beat_s"])
    e2e["total_s"] = round(sum(e2e.values()), 1)
    wl = (f"synthetic BiV shell 256^3 @0.4 mm ({n_myo:,} voxels), FP32, dt 0.02 ms, Purkinje + "
          f"PMJ" + (f", ECG every 1 ms in a 384x384x256 @1.5 mm torso "
                    f"({ecg.info['electrodes']} electrodes, lead field)" if ecg else ", no ECG")
          + "; bench step = 1 ms simulated (50 solver steps + 1 ECG frame)")
    out = {
        "metric": METRIC, "value": hr["value"], "unit": "voxel-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": hr["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl, "form": head, "n_myo_voxels": n_myo,
                   "n_purkinje_nodes": meta["Np"], "n_pmj": len(meta["pmj_ids"]),
                   "brick": list(brick), "brick_autotune_us": tune,
                   "brick_occupancy": round(meta["layout"].occupancy, 4),
                   "solver_steps_per_bench_step": STEPS_PER_MS, "beat_ms": a.beat_ms},
        "forms": results,
        "ecg": ecg_info,
        "end_to_end": e2e,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "cn_sweep_kernel<float,8,4,4> (plain form: one Jacobi sweep "
                               "of the CN solve, the 19-point stencil; ~3 per solver step)",
                     "kernel_us": sweep_us, "step_us": step_us, "bytes_per_launch": sweep_bytes,
                     "bytes_per_voxel": SWEEP_BYTES_PER_VOXEL,
                     "bicgstab_apply": {"kernel": "cn_bi_pass_kernel<float,8,4,4,S> (the "
                                        "BiCGSTAB product pass t = A M^-1 s; two stencil "
                                        "passes per iteration)",
                                        "kernel_us": kt["apply_us"], "achieved": apply_gbs,
                                        "bytes_per_voxel": APPLY_BYTES_PER_VOXEL,
                                        "step_us": kt["bicgstab_step_us"]}},
        "cpu_baseline": cpu,
        "stream_copy": copy,
        "bench_wall_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
