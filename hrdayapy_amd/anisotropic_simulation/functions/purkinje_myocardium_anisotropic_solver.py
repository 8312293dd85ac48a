"""Coupled Purkinje + myocardium anisotropic monodomain solver on MI355X.

Drop-in for hrdayapy.anisotropic_simulation.functions.
purkinje_myocardium_anisotropic_solver (reference file
purkinje_myocardium_anisotropic_solver.py:1308-1881): same arguments, same
returned dict keys and the same ``vm_snapshots.npz`` schema (:1923-1993).

What differs is the machinery: the state lives in HBM in a brick layout
(hrdayapy_amd.layout) and the steps are HIP kernels replayed from captured
HIP graphs.  Two schemes:

* ``scheme="cn"`` (default): the reference's step -- Lie splitting, then
  Crank-Nicolson diffusion with ``theta`` on the whole coupled system (PMJ
  couplings included), solved by the reference's Jacobi-preconditioned
  BiCGSTAB with its stopping rules at ``cg_tol`` (:786-849, :1746-1791),
  matrix-free on the device (hp_cn_step, solver 1); the iterations launched
  per step are bounded by cg_max_iter and fitted to the largest count any step
  has needed (plus two) as the run goes, and a warning names the steps that
  stopped at the bound.  ``cn_solver="jacobi"`` solves the same system by
  Jacobi sweeps to the same relative residual (at most min(cg_max_iter, 12)):
  the exact CN solution to well within the tolerance, which differs from the
  reference's by the error its BiCGSTAB leaves at the fronts.
* ``scheme="explicit"``: the fast form -- one fused kernel per step (TTP06 +
  forward-Euler 19-point diffusion on V^n + stimulus + clamp + activation),
  the cable in the same launch, the PMJ exchange integrated exactly; ``dt``
  above the stencil's stability limit is split into substeps.  It does not
  use theta/cg_tol/cg_max_iter (a warning says so when they are changed) and
  differs from the CN scheme by the schemes' truncation error.
"""
from __future__ import annotations

import math
import time
import warnings
from pathlib import Path

import numpy as np
import torch

from ... import tissue
from ...engine import MonodomainEngine
from ...layout import BrickLayout, choose_brick


# BiCGSTAB iterations / Jacobi sweeps of the CN solve per step captured in the
# step's graph at the start (each step launches them all; the ones after the
# solve ended return at once).  After every output chunk the bound is fitted to
# the largest count any step has needed plus CN_ITER_HEADROOM, and doubled
# when a step reached it without meeting the tolerance; never above
# cg_max_iter.
CN_MAX_ITERS = {"bicgstab": 16, "jacobi": 12}
CN_ITER_HEADROOM = 2


def _fit_cn_bound(eng, cap, seen):
    """Refit the engine's per-step iteration bound from its solver statistics
    (one host synchronisation); seen holds the non-converged count so far."""
    st = eng.cn_stats()
    if st["not_converged"] > seen[0]:
        seen[0] = st["not_converged"]
        eng.set_cn_max_iters(min(cap, 2 * st["max_sweeps"]))
    else:
        eng.set_cn_max_iters(min(cap, max(st["max_used"] + CN_ITER_HEADROOM, 2)))


def build_coupled_system(nodes, elements, activation_times, S, phi, fibre_dir=None, *,
                         voxel_size=0.4, dx_p=0.2, sigma_P=0.45, sigma_l=3 * 1.31e-4,
                         sigma_t=1.31e-4, sigma_M=None, Cm=0.01, A_M=140.0, R_P=0.015, A_P=None,
                         dt=0.02, T=400.0, stim_protocol=None, ectopic_region=None,
                         stim_regions=None, stim_len_mm=2.0, c_pmj=0.05, n_pmj=1,
                         phi_endo_max=0.35, phi_epi_min=0.65, dtype=torch.float32,
                         device="cuda", brick=None, verbose=False, sigma_scale=None,
                         inst=None, inst_params=None, auto_substep=False, alloc="torch",
                         scheme="explicit", theta=0.5, cg_tol=1e-5, max_sweeps=10,
                         cn_solver="bicgstab", cn_warm=True):
    """Set up the tissue on the host and upload it into a MonodomainEngine.

    ``sigma_M`` selects the isotropic 6-point operator (simulation.compute_coupled);
    otherwise sigma_l/sigma_t and fibre_dir give the 19-point anisotropic one.
    ``scheme="explicit"``: ``dt`` above the forward-Euler limit of the stencil
    or the cable raises ValueError; with ``auto_substep`` the engine instead
    steps dt/n_sub, n_sub = meta["n_sub"] engine steps per dt.
    ``scheme="cn"``: the reference's Crank-Nicolson step (any dt).
    Returns (engine, meta) where meta holds the reference-order bookkeeping.
    """
    t0 = time.perf_counter()
    S = np.asarray(S).astype(bool)
    nodes_mm = np.asarray(nodes, dtype=np.float32) * voxel_size
    comp_nodes, comp_edges, branch_map = tissue.discretise_tree(nodes_mm, elements, dx_p)
    Np = len(comp_nodes)
    root = tissue.root_zone(nodes_mm, activation_times, comp_nodes, T, stim_len_mm)
    pc = tissue.purkinje_constants(R_P, sigma_P, dx_p, A_M, voxel_size, A_P)
    lay = BrickLayout(S, brick or choose_brick(S))
    vox = lay.vox_idx
    if sigma_M is not None:
        D6 = tissue.isotropic_D6(lay.n_nodes, sigma_M)
    else:
        D6 = tissue.conductivity_D6(fibre_dir, vox, sigma_l, sigma_t, sigma_scale)
    ct = tissue.phenotype_from_phi(phi, vox, phi_endo_max, phi_epi_min)
    myo_nodes_mm = vox[:, [1, 0, 2]].astype(np.float32) * voxel_size
    pmj_ids = tissue.pmj_sites(elements, branch_map, len(nodes), Np)
    patches = tissue.pmj_patches(comp_nodes, pmj_ids, myo_nodes_mm, n_pmj)
    if stim_protocol is None:
        stim_protocol = [dict(name="av_node", target="root", amp=12.0, dur=2.0, onset=0.0)]
    sites = tissue.stim_sites(stim_protocol, root, ectopic_region, vox, Np, stim_regions)
    dt_lim = tissue.explicit_dt_limit(D6, voxel_size, A_M, Cm, comp_edges, pc["w_pp"], pc["Ai_P"])
    n_sub = 1
    if scheme == "explicit" and dt > dt_lim:
        if not auto_substep:
            raise ValueError(f"dt={dt} ms exceeds the explicit diffusion limit "
                             f"{dt_lim:.4g} ms of this tissue (pass a smaller dt)")
        n_sub = int(math.ceil(dt / dt_lim))
    eng = MonodomainEngine(
        lay, D6, ct, h=voxel_size, A_M=A_M, Cm=Cm, dt=dt / n_sub, dtype=dtype, device=device,
        cable=dict(n=Np, edges=comp_edges, w_pp=pc["w_pp"], Ai_P=pc["Ai_P"]),
        pmj=dict(ids=pmj_ids, patches=patches, c_pmj=c_pmj, Ai_P=pc["Ai_P"], Ai_M=pc["Ai_M"]),
        stim_sites=sites, inst=inst, inst_params=inst_params, alloc=alloc, scheme=scheme,
        theta=theta, cg_tol=cg_tol, max_sweeps=max_sweeps, cn_solver=cn_solver, cn_warm=cn_warm)
    meta = dict(Np=Np, N_myo=lay.n_nodes, n_sub=n_sub, dt_limit=dt_lim, scheme=scheme,
                comp_nodes=comp_nodes, comp_edges=comp_edges,
                branch_map=branch_map, vox_idx=vox, nodes_mm=myo_nodes_mm, celltype=ct,
                pmj_ids=pmj_ids, pmj_patches=patches, sites=sites, layout=lay,
                constants=pc, voxel_size=voxel_size, setup_s=time.perf_counter() - t0)
    if verbose:
        print(f"[hrdaya] Purkinje {Np:,} nodes, myocardium {lay.n_nodes:,} voxels in "
              f"{lay.n_bricks:,} bricks of {lay.brick} (occupancy {lay.occupancy:.2f}), "
              f"{len(pmj_ids)} PMJs, setup {meta['setup_s']:.1f} s")
    return eng, meta


def _surface_node_ids(S, Z, vox):
    Nx, Ny, Nz = S.shape
    surf = np.zeros(len(vox), dtype=bool)
    for off in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        c = vox + np.array(off)
        inb = ((c[:, 0] >= 0) & (c[:, 0] < Nx) & (c[:, 1] >= 0) & (c[:, 1] < Ny)
               & (c[:, 2] >= 0) & (c[:, 2] < Nz))
        cc = np.clip(c, 0, np.array(S.shape) - 1)
        surf |= ~(inb & (np.asarray(Z)[cc[:, 0], cc[:, 1], cc[:, 2]] == 1))
    return np.nonzero(surf)[0].astype(np.int32)


def purkinje_myocardium_anisotropic_solver(
        nodes, elements, activation_times, S, Z, phi, fibre_dir, *, voxel_size=0.4, dx_p=0.2,
        sigma_P=0.225, sigma_l=3.0 * 6.5625e-5, sigma_t=6.5625e-5, Cm=0.01, A_P=None, A_M=140.0,
        R_P=0.015, dt=0.1, T=400.0, theta=0.5, stim_protocol=None, ectopic_region=None,
        stim_regions=None, stim_len_mm=2.0, c_pmj=0.05, n_pmj=1, cg_tol=1e-5, cg_max_iter=100,
        n_frames=100, frame_save_dt=None, anim_tmp_dir=None, device=None, dtype=torch.float32,
        out_npz=None, phi_endo_max=0.35, phi_epi_min=0.65, vm_save_dt=None, out_vm_npz=None,
        vm_tmp_dir=None, graph_chunk=50, scheme="cn", cn_solver="bicgstab", _sigma_M=None):
    """Run the coupled simulation; see the module docstring for what differs from the reference."""
    if scheme not in ("cn", "explicit"):
        raise ValueError(f"scheme must be 'cn' or 'explicit' (got {scheme!r})")
    if scheme == "explicit" and (theta != 0.5 or cg_tol != 1e-5):
        warnings.warn("scheme='explicit' integrates diffusion by forward Euler: theta and cg_tol "
                      "apply to scheme='cn' only and are not used", stacklevel=2)
    if scheme == "cn" and (not 0.0 <= theta <= 1.0 or cg_tol < 0 or int(cg_max_iter) < 1
                           or cn_solver not in CN_MAX_ITERS):
        raise ValueError("scheme='cn' needs 0 <= theta <= 1, cg_tol >= 0, cg_max_iter >= 1 and "
                         "cn_solver 'bicgstab' or 'jacobi'")
    if A_P is not None and _sigma_M is None:
        raise TypeError(
            "A_P is no longer a settable parameter -- it's mechanically determined by R_P "
            "(A_P = 2.0/R_P for a cylindrical fibre) and specifying it independently risks it "
            "silently disagreeing with R_P. Remove the A_P= argument; only set R_P.")
    dev = torch.device(device or "cuda")
    if dev.type != "cuda":
        raise RuntimeError("the hrdaya solver runs on the GPU only (HIP kernels); got device "
                           f"{dev}")
    eng, meta = build_coupled_system(
        nodes, elements, activation_times, S, phi, fibre_dir, voxel_size=voxel_size, dx_p=dx_p,
        sigma_P=sigma_P, sigma_l=sigma_l, sigma_t=sigma_t, sigma_M=_sigma_M, Cm=Cm, A_M=A_M,
        R_P=R_P, A_P=A_P if _sigma_M is not None else None, dt=dt, T=T,
        stim_protocol=stim_protocol, ectopic_region=ectopic_region, stim_regions=stim_regions,
        stim_len_mm=stim_len_mm, c_pmj=c_pmj, n_pmj=n_pmj, phi_endo_max=phi_endo_max,
        phi_epi_min=phi_epi_min, dtype=dtype, device=dev, verbose=True, auto_substep=True,
        scheme=scheme, theta=theta, cg_tol=cg_tol, cn_solver=cn_solver,
        max_sweeps=min(int(cg_max_iter), CN_MAX_ITERS.get(cn_solver, 1)))
    Np, N_myo, n_sub = meta["Np"], meta["N_myo"], meta["n_sub"]
    if n_sub > 1:
        print(f"[hrdaya] dt={dt} ms is above the explicit diffusion limit "
              f"{meta['dt_limit']:.4g} ms: {n_sub} substeps of {dt / n_sub:.4g} ms per step")
    vox = meta["vox_idx"]
    n_steps = int(T / dt)
    save_every = (max(1, int(round(frame_save_dt / dt))) if frame_save_dt is not None
                  else max(1, n_steps // n_frames))
    surf_ids = _surface_node_ids(np.asarray(S), Z, vox)
    run_vm = vm_save_dt is not None and out_vm_npz is not None
    vm_every = max(1, int(round(vm_save_dt / dt))) if run_vm else 0
    bmaps = [np.asarray(b, dtype=np.int64) for b in meta["branch_map"]]
    # Output frames stay in HBM while the simulation runs (one device gather per
    # frame, no host synchronisation) and travel to the host once at the end.
    n_anim = (n_steps - 1) // save_every + 1 if n_steps else 0
    n_vm = (n_steps - 1) // vm_every + 1 if run_vm and n_steps else 0
    f32 = dict(dtype=torch.float32, device=dev)
    surf_t = torch.from_numpy(surf_ids.astype(np.int64)).to(dev)
    bcat = np.concatenate(bmaps) if bmaps else np.zeros(0, np.int64)
    bcat_t = torch.from_numpy(bcat.astype(np.int64)).to(dev)
    surf_buf = torch.empty((n_anim, len(surf_ids)), **f32)
    branch_buf = torch.empty((n_anim, len(bcat)), **f32)
    vm_buf = torch.empty((n_vm, N_myo), **f32)
    times, vm_times = [], []
    # the loop advances to the next output step with a captured graph chunk
    cadence = [c for c in (save_every, vm_every) if c]
    t0 = time.perf_counter()
    step = 0
    cn_seen = [0]
    while step < n_steps:
        # output after step `step` is due when step % every == 0 (frame stamped step*dt)
        nxt = min([((step // c) + (1 if step % c else 0)) * c for c in cadence] + [n_steps - 1])
        nxt = max(nxt, step)
        k = (nxt - step + 1) * n_sub
        # Jacobi sweeps: each chunk bounded by what the last one needed, replayed
        # with the full bound if a step stopped at it (MonodomainEngine.run)
        jac = scheme == "cn" and cn_solver == "jacobi"
        eng.run(k, graph_chunk=graph_chunk if k >= graph_chunk else 0, adaptive=jac)
        if scheme == "cn" and not jac:
            _fit_cn_bound(eng, int(cg_max_iter), cn_seen)
        t = nxt * dt
        want_anim, want_vm = nxt % save_every == 0, run_vm and nxt % vm_every == 0
        if want_anim or want_vm:
            vm_now = eng.V_myo_nodes()
            if want_anim:
                surf_buf[len(times)] = vm_now[surf_t]
                if len(bcat):
                    branch_buf[len(times)] = eng.Vpk[eng.p][bcat_t]
                times.append(t)
            if want_vm:
                vm_buf[len(vm_times)] = vm_now
                vm_times.append(t)
        step = nxt + 1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"[hrdaya] {n_steps * n_sub:,} steps x {N_myo:,} voxels in {wall:.2f} s "
          f"({n_steps * n_sub * N_myo / max(wall, 1e-9):.3e} voxel-steps/s)")
    cn = eng.cn_stats()
    if cn is not None:
        what = "BiCGSTAB iterations" if cn["solver"] == "bicgstab" else "Jacobi sweeps"
        print(f"[hrdaya] CN: {cn['mean_sweeps']:.2f} {what} per step on average, "
              f"{cn['max_used']} at most (tolerance {cn['tol']:g}, bound {cn['max_sweeps']})")
        if cn["not_converged"]:
            warnings.warn(f"CN solve: {cn['not_converged']} of {cn['steps']} steps stopped at "
                          f"the bound on {what} with relative residual up to "
                          f"{cn['worst_rel_res']:.2e} > cg_tol={cn['tol']:g}", RuntimeWarning,
                          stacklevel=2)
    if run_vm and vm_times:
        _save_vm_snapshots(out_vm_npz, np.arange(N_myo, dtype=np.int32) + Np, meta["nodes_mm"],
                           vox, vm_buf[:len(vm_times)].cpu().numpy(),
                           np.array(vm_times, np.float32), dt, vm_every * dt)
    frames_surf = surf_buf[:len(times)].cpu().numpy()
    bf = branch_buf[:len(times)].cpu().numpy()
    offs = np.cumsum([0] + [len(b) for b in bmaps])
    branch_frames = [bf[:, offs[i]:offs[i + 1]] for i in range(len(bmaps))]
    results = dict(
        comp_nodes=meta["comp_nodes"], comp_edges=meta["comp_edges"],
        branch_map=meta["branch_map"],
        branch_frames_p=[np.ascontiguousarray(f, dtype=np.float32) for f in branch_frames],
        surf_verts=np.zeros((0, 3), np.float32), surf_faces=np.zeros((0, 3), np.int32),
        vert_to_surf_node=np.zeros(0, np.int32),
        frames_surf=np.asarray(frames_surf, dtype=np.float32), times=times,
        surface_node_ids=surf_ids, activation_times=eng.at_nodes().float().cpu().numpy(),
        wall_s=wall, cn_stats=cn)
    if out_npz is not None:
        _save_npz(results, out_npz)
    return results


def _save_npz(results, path):
    """Animation NPZ, schema of the reference's _save_npz (:1888-1916)."""
    path = Path(path).with_suffix(".npz")
    d = {"time": np.array(results["times"], dtype=np.float32),
         "comp_nodes": results["comp_nodes"], "comp_edges": results["comp_edges"],
         "surf_verts": results["surf_verts"], "surf_faces": results["surf_faces"],
         "vert_to_surf_node": results["vert_to_surf_node"],
         "surf_frames": np.asarray(results["frames_surf"], dtype=np.float32)}
    for b, bm in enumerate(results["branch_map"]):
        d[f"bmap_{b}"] = np.array(bm, dtype=np.int32)
        d[f"branch_{b}"] = np.asarray(results["branch_frames_p"][b], dtype=np.float32)
    np.savez_compressed(str(path), **d)


def _save_vm_snapshots(path, node_ids, coords_mm, vox_idx, Vm, times, dt_solver, vm_save_dt):
    """vm_snapshots.npz, schema of the reference's _save_vm_snapshots (:1923-1993)."""
    path = Path(path).with_suffix(".npz")
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(str(path), node_ids=node_ids.astype(np.int32),
                        coords_mm=coords_mm.astype(np.float32), vox_idx=vox_idx.astype(np.int32),
                        Vm=np.asarray(Vm, dtype=np.float32), time_vm=times.astype(np.float32),
                        dt_solver=np.float32(dt_solver), vm_save_dt=np.float32(vm_save_dt))
