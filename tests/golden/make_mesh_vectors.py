"""Write golden vectors of the reference's mesh, coupling, solver, stimulus and
ECG-source functions by running the reference's own modules.

Run in the development container only (the reference tree is not present on
GPU machines); the output ``mesh_golden.npz`` is committed next to this
script, and the tests read it (tests/test_oracle_golden.py).

What is imported, by file path under synthetic parent packages so that the
modules' relative imports resolve to the real files and no package
``__init__`` runs (the package chain needs pynrrd/pyvista, absent here):
  * hrdayapy/anisotropic_simulation/functions/
    purkinje_myocardium_anisotropic_solver.py (with the ionic modules its
    top-level imports name: ionic_models/gpu_ttp06.py, hetero_solver.py,
    ttp06_phenotype_constants.py);
  * hrdayapy/ecg/functions/compute_bspm.py and build_torso_grid.py.
Only module-level functions are called; the time loop (which needs
scikit-image for its surface mesh) is not run.

Functions pinned, each on seeded synthetic inputs that are stored too:
  _discretise_tree, _find_terminals, build_myo_mesh_anisotropic, map_pmj,
  _build_global_laplacian_gpu (PMJ rows included), _build_cn_matrix,
  _bicgstab, _resolve_stim_sites, _stim_active; _assemble_stiffness,
  _find_surface_ids, build_heart_laplacian, build_projection_map and
  make_gpu_solver's solve_frame (source + _pcg_gpu, three stopping metrics,
  two frames with the warm start), on the CPU device.

The PMJ site ids passed to map_pmj follow the solver's inline rule
(purkinje_myocardium_anisotropic_solver.py:1566-1573: the last computational
node of every element whose second original node is a terminal); the rule
itself is pinned separately on a hand-worked tree in the tests.
"""
from __future__ import annotations

import importlib.util
import sys
import types
from pathlib import Path

import numpy as np
import torch

REF = Path("/root/reference/hrdayapy")
HERE = Path(__file__).resolve().parent
OUT = HERE / "mesh_golden.npz"
sys.path.insert(0, str(HERE.parent.parent))


def _package(name, path):
    pkg = types.ModuleType(name)
    pkg.__path__ = [str(path)]
    sys.modules[name] = pkg


def _module(pkg, name, path):
    spec = importlib.util.spec_from_file_location(f"{pkg}.{name}", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def _load():
    fa = REF / "anisotropic_simulation" / "functions"
    _package("_ref_af", fa)
    _package("_ref_af.ionic_models", fa / "ionic_models")
    for n in ("ttp06_phenotype_constants", "gpu_ttp06", "hetero_solver"):
        _module("_ref_af.ionic_models", n, fa / "ionic_models" / f"{n}.py")
    aniso = _module("_ref_af", "purkinje_myocardium_anisotropic_solver",
                    fa / "purkinje_myocardium_anisotropic_solver.py")
    fe = REF / "ecg" / "functions"
    _package("_ref_ef", fe)
    bspm = _module("_ref_ef", "compute_bspm", fe / "compute_bspm.py")
    torso = _module("_ref_ef", "build_torso_grid", fe / "build_torso_grid.py")
    return aniso, bspm, torso


def _slab():
    """10x8x6 slab (padded by one voxel) with a hole, rotating fibres."""
    from tests._problems import slab
    S, phi, f = slab(10, 8, 6, seed=3)
    S[4:6, 3:5, 2:4] = 0
    return S, phi, f


def _tree(S):
    """Root above the slab, a few branches, one of them subdivided unevenly."""
    from tests._problems import small_tree
    nodes, el, act = small_tree(S, 0.25, seed=4, n_branch=4)
    # an extra side branch off an interior node (a junction of degree 3)
    nodes = np.vstack([nodes, nodes[1] + np.array([2.5, -1.0, -1.5])])
    el = np.vstack([el, [1, len(nodes) - 1]])
    act = np.append(act, 1.7)
    return nodes, el, act


def main():
    aniso, bspm, torso = _load()
    out = {}
    rng = np.random.default_rng(20260301)
    h = 0.25
    S, phi, fib = _slab()
    nodes, el, act = _tree(S)
    out.update(slab_S=S, slab_phi=phi, slab_fibre=fib, tree_nodes=nodes, tree_el=el,
               tree_act=act)

    # --- tree discretisation, terminals
    nodes_mm = nodes.astype(np.float32) * h
    comp_nodes, comp_edges, branch_map = aniso._discretise_tree(nodes_mm, el, 0.2)
    out.update(dt_comp_nodes=comp_nodes, dt_comp_edges=comp_edges,
               dt_branch_ptr=np.cumsum([0] + [len(b) for b in branch_map]),
               dt_branch_cat=np.concatenate([np.asarray(b) for b in branch_map]))
    terms = aniso._find_terminals(el, len(nodes))
    out["terminals"] = np.asarray(terms)

    # --- myocardial mesh
    sl, st = 0.24, 0.08
    myo = aniso.build_myo_mesh_anisotropic(S, S, h, fib, sl, st)
    out.update(myo_edges=np.asarray(myo["edges"]), myo_edge_w=np.asarray(myo["edge_w"]),
               myo_vox_idx=np.asarray(myo["vox_idx"]),
               myo_nodes_mm=np.asarray(myo["nodes_mm"]), myo_N=int(myo["N_myo"]))

    # --- PMJ patches (site ids by the solver's inline rule)
    tset = set(terms)
    pmj_ids = sorted({bm[-1] for i, bm in enumerate(branch_map) if el[i][1] in tset})
    out["pmj_ids"] = np.asarray(pmj_ids)
    for n_pmj in (1, 3):
        patches = aniso.map_pmj(comp_nodes, pmj_ids, myo, n_pmj=n_pmj)
        out[f"pmj_patches_{n_pmj}"] = np.asarray(patches)

    # --- global Laplacian, CN matrix, BiCGSTAB
    R_P, sigma_P, dx_p, A_M = 0.015, 0.45, 0.2, 140.0
    S_P = np.pi * R_P ** 2
    w_pp, Ai_P, Ai_M = sigma_P * S_P / dx_p, (2.0 / R_P) * S_P * dx_p, A_M * h ** 3
    Np, N_myo = len(comp_nodes), int(myo["N_myo"])
    patches = aniso.map_pmj(comp_nodes, pmj_ids, myo, n_pmj=3)
    c_pmj = 2.0
    L, diag_L = aniso._build_global_laplacian_gpu(
        Np, N_myo, comp_edges, w_pp, Ai_P, myo["edges"], myo["edge_w"], Ai_M, pmj_ids,
        patches, c_pmj, torch.device("cpu"), torch.float64)
    X = rng.standard_normal((Np + N_myo, 3))
    out.update(lap_consts=np.array([w_pp, Ai_P, Ai_M, c_pmj]), lap_X=X,
               lap_LX=torch.sparse.mm(L, torch.from_numpy(X)).numpy(),
               lap_diag=diag_L.numpy())
    Cm, dt, theta = 0.01, 0.02, 0.5
    A_csr, M_inv = aniso._build_cn_matrix(L, diag_L, Cm / dt, theta)
    out["cn_AX"] = torch.stack([torch.mv(A_csr, torch.from_numpy(X[:, k]))
                                for k in range(3)], 1).numpy()
    out["cn_Minv"] = M_inv.numpy()
    x0 = -80.0 + 5.0 * rng.standard_normal(Np + N_myo)
    x0[:Np] += 40.0
    b = (Cm / dt) * x0 + (1 - theta) * torch.sparse.mm(
        L, torch.from_numpy(x0)[:, None]).squeeze(1).numpy()
    for tol in (1e-5, 1e-10):
        x = aniso._bicgstab(A_csr, M_inv, torch.from_numpy(b), torch.from_numpy(x0), tol, 500)
        out[f"bicg_x_{tol:g}"] = x.numpy()
    out.update(bicg_b=b, bicg_x0=x0)

    # --- stimulus sites and schedule
    root = np.zeros(Np, dtype=bool)
    root[[0, 1, 5]] = True
    ect = np.zeros(S.shape, dtype=bool)
    ect[1:4, 1:4, 1:3] = True
    reg = {"apex": np.zeros(S.shape, dtype=bool)}
    reg["apex"][8:11, 6:9, 4:7] = True
    proto = [dict(name="av", target="root", amp=12.0, dur=2.0, onset=0.0),
             dict(name="e", target="ectopic", amp=0.52, dur=1.0, onset=3.0, period=6.0,
                  n_pulses=3),
             dict(name="r", target="region", region="apex", amp=0.3, dur=0.5, onset=1.0),
             dict(name="c", target="custom", node_ids=[Np + 3, Np + 7], amp=0.2, dur=1.0)]
    sites = aniso._resolve_stim_sites(proto, root, ect, np.asarray(myo["vox_idx"]), Np, N_myo,
                                      torch.device("cpu"), torch.float64, reg)
    out["stim_amp"] = np.stack([s["amp_tensor"].numpy() for s in sites])
    out["stim_root"] = root
    out["stim_ect"] = ect
    out["stim_reg"] = reg["apex"]
    ts = np.round(np.arange(0.0, 40.0, 0.02), 10)
    ts = np.concatenate([ts, [3.0, 9.0, 15.0, 21.0, 4.0, 10.0, 2.9999999]])
    act_tab = np.array([[aniso._stim_active(float(t), *a) for t in ts] for a in
                        ((3.0, 1.0, 6.0, 3), (0.1, 0.3, 0.9, 7), (0.0, 2.0, None, None),
                         (0.5, 0.25, 0.75, None))])
    out.update(stim_t=ts, stim_active=act_tab)

    # --- torso: stiffness, surface, pin; heart source and frame solves
    n = (16, 14, 12)
    ax = [np.linspace(-1, 1, m) for m in n]
    Xg, Yg, Zg = np.meshgrid(*ax, indexing="ij")
    lab = np.zeros(n, dtype=np.int32)
    lab[Xg ** 2 / 0.9 + Yg ** 2 / 0.8 + Zg ** 2 / 0.95 <= 1.0] = 1
    lab[(Xg + 0.35) ** 2 / 0.1 + Yg ** 2 / 0.2 + Zg ** 2 / 0.25 <= 1] = 3
    lab[(Xg - 0.25) ** 2 / 0.06 + (Yg + 0.1) ** 2 / 0.06 + Zg ** 2 / 0.08 <= 1] = 4
    smap = {1: 0.2, 3: 0.05, 4: 0.7}
    sig = np.zeros(n)
    for k, v in smap.items():
        sig[lab == k] = v
    active = lab > 0
    active_ids = np.flatnonzero(active.ravel())
    id_map = np.full(active.size, -1, dtype=np.int64)
    id_map[active_ids] = np.arange(len(active_ids))
    dxc = 1.5
    A = torso._assemble_stiffness(sig, dxc * 1e-3, active, active_ids, id_map)
    Xt = rng.standard_normal((A.shape[0], 2))
    out.update(torso_lab=lab, torso_sig=sig, torso_AX=A @ Xt, torso_X=Xt,
               torso_diag=A.diagonal())
    sflat, srows = torso._find_surface_ids(lab, active, id_map)
    interior = np.where(active.ravel())[0]
    pin_row = int(id_map[int(interior[len(interior) // 2])])
    out.update(torso_surf_flat=sflat, torso_surf_rows=srows, torso_pin=pin_row)
    # heart: a small block of fine voxels inside the torso box
    hf = 0.5
    hv = np.argwhere(np.ones((9, 8, 7), dtype=bool)) + np.array([3, 2, 1])
    keep = ~((hv[:, 0] == 5) & (hv[:, 1] >= 4))  # a notch, so rows differ in degree
    hv = hv[keep].astype(np.int32)
    K = bspm.build_heart_laplacian(hv, len(hv), hf)
    Vt = rng.standard_normal((len(hv), 2))
    out.update(heart_vox=hv, heart_KV=K @ Vt, heart_V=Vt)
    coords = hv[:, [1, 0, 2]].astype(np.float64) * hf
    ang = 0.3
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    t = np.array([6.0, 4.5, 3.0])
    scale = 1.1
    origin = np.array([-0.75, -0.75, -0.75])
    proj, valid = bspm.build_projection_map(coords, R, t, scale, origin, dxc, id_map, n)
    out.update(proj_R=R, proj_t=t, proj_scale=scale, proj_origin=origin, proj_rows=proj,
               proj_valid=valid)
    K_t = bspm._scipy_csr_to_torch(K.tocsr(), torch.device("cpu"), torch.float32)
    Vm = np.stack([-85.0 + 110.0 * (hv[:, 0] > 5 + 2 * k) for k in range(2)]).astype(np.float32)
    Vm += rng.standard_normal(Vm.shape).astype(np.float32)
    out["frame_Vm"] = Vm
    for metric in ("rms", "p99", "linf"):
        solve, iters = bspm.make_gpu_solver(
            A, pin_row, proj, valid, K_t, 0.2625, hf * 1e-3, dxc, srows, A.shape[0],
            torch.device("cpu"), cg_tol_mV=0.01, cg_max=500, cg_metric=metric)
        res = []
        for k in range(2):
            phi_s, n_it, res_mV, full = solve(Vm[k], capture_full=True)
            res.append((phi_s, n_it, res_mV, full))
        out[f"frame_{metric}_phi_surf"] = np.stack([r[0] for r in res])
        out[f"frame_{metric}_iters"] = np.array([r[1] for r in res])
        out[f"frame_{metric}_res_mV"] = np.array([r[2] for r in res])
        out[f"frame_{metric}_phi_full"] = np.stack([r[3] for r in res])
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1024:.0f} KiB, {len(out)} arrays)")


if __name__ == "__main__":
    main()
