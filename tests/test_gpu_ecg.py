"""GPU parity of the torso Poisson path (operator, refined PCG, reference-rule
PCG, heart source, electrodes) against the CPU oracle and Frank's sphere."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from hrdayapy_amd.ecg.functions.forward import (Electrodes, HeartSource, PoissonSolver,
                                                leads_from_electrodes)
from hrdayapy_amd.ecg.functions.torso import TorsoGrid
from hrdayapy_amd.layout import BrickLayout
from oracle import ecg as E
from tests.test_oracle_ecg import _torso

pytestmark = pytest.mark.gpu


def _setup(n=(30, 26, 22)):
    lab, smap = _torso(n=n)
    lab[12:14, 10:12, 5:8] = 0
    g = TorsoGrid.from_labels(lab, smap, 1.5)
    sig = np.zeros(lab.shape)
    for k, v in smap.items():
        sig[lab == k] = v
    A, ids, idm = E.assemble_stiffness(sig, 1.5e-3, lab != 0)
    Ap = E.pin(A, g["pin_row"])
    return lab, g, Ap


def _to_cells(g, v_rows, dev):
    out = np.zeros(g.layout["n_cells"])
    out[g.cell_of_row] = v_rows
    return torch.from_numpy(out).to(dev)


@pytest.mark.parametrize("n", [(30, 26, 22), (72, 64, 160)])
def test_operator_matches_oracle(cuda, n):
    """FP64 residual and FP32 product against the oracle stiffness; the larger
    grid has runs of several 64-cell chunks and a full grid split between
    the XCDs."""
    print(f"[case] torso operator {n}", flush=True)
    lab, g, Ap = _setup(n)
    ps = PoissonSolver(g, cuda)
    if n[2] > 64:
        assert ps.torso.n_chunks >= 8 * 4 * 8 and int(g.layout["col_len"].max()) > 64
    x = np.random.default_rng(0).standard_normal(Ap.shape[0])
    x[g["pin_row"]] = 0.0
    b = np.zeros_like(x)
    xc, bc = _to_cells(g, x, cuda), _to_cells(g, b, cuda)
    r = torch.empty_like(xc)
    rn, bn = ps.residual(bc, xc, r)
    ax = -r.cpu().numpy()[g.cell_of_row]
    ref = Ap @ x
    assert np.max(np.abs(ax - ref)) < 1e-6 * np.abs(ref).max()
    # the block-partial reduction of ||r||^2 (last block to arrive)
    assert bn == 0.0 and abs(rn - np.linalg.norm(ref)) < 1e-6 * np.linalg.norm(ref)
    x32 = xc.float()
    y32 = torch.empty_like(x32)
    ps.spmv32(x32, y32)
    y = y32.double().cpu().numpy()[g.cell_of_row]
    assert np.max(np.abs(y - ref)) < 1e-5 * np.abs(ref).max()


def test_refined_solve_matches_direct(cuda):
    print("[case] refined PCG vs direct", flush=True)
    lab, g, Ap = _setup()
    rng = np.random.default_rng(1)
    b = rng.standard_normal(Ap.shape[0]) * 1e3
    b -= b.mean()
    b[g["pin_row"]] = 0.0
    xd = spla.spsolve(Ap.tocsc(), b)
    ps = PoissonSolver(g, cuda)
    x = ps.solve_refined(_to_cells(g, b, cuda), tol=1e-10)
    st = ps.last_stats
    print(f"  outer {st['outer']} inner {st['inner']} rel_res {st['rel_res']:.2e}")
    assert st["rel_res"] <= 1e-10
    xs = x.cpu().numpy()[g.cell_of_row]
    assert np.max(np.abs(xs - xd)) < 1e-7 * np.abs(xd).max()
    # warm start from the answer: no iteration needed
    ps.solve_refined(_to_cells(g, b, cuda), x=x, tol=1e-8)
    assert ps.last_stats["outer"] == 0


def test_rms_rule_matches_reference_pcg(cuda):
    print("[case] rms-metric PCG vs oracle", flush=True)
    lab, g, Ap = _setup()
    rng = np.random.default_rng(2)
    b = rng.standard_normal(Ap.shape[0]) * 10.0
    b -= b.mean()
    b[g["pin_row"]] = 0.0
    xo, it_o, err_o = E.pcg(Ap, b, 1.0 / Ap.diagonal(), tol_abs_mV=0.01, max_iter=500)
    ps = PoissonSolver(g, cuda)
    x = ps.solve_rms(_to_cells(g, b, cuda), tol_V=1e-5, max_iter=500)
    it = ps.last_stats["inner"]
    print(f"  iterations gpu {it} oracle {it_o}")
    assert abs(it - it_o) <= 2
    xs = x.cpu().numpy()[g.cell_of_row]
    assert np.sqrt(np.mean((xs - xo) ** 2)) < 2e-5  # both within the 0.01 mV rule


def test_frank_sphere_reference_config(cuda):
    """experiment_4's configuration (R = 100 mm, dx = 2 mm, d = 8 mm): RDM < 0.05."""
    print("[case] Frank sphere R=100 dx=2", flush=True)
    R, dx, sigma, p, d = 100.0, 2.0, 0.2, 1e-6, 8.0
    lab, origin = E.sphere_labels(R, dx)
    g = TorsoGrid.from_labels(lab, {1: sigma}, dx, origin)
    ps = PoissonSolver(g, cuda)
    axis = np.array([0.0, 0.0, 1.0])
    I = p / (d * 1e-3)
    b = np.zeros(len(g["active_ids"]))
    for sgn in (1, -1):
        ijk = np.round((sgn * axis * d / 2 - origin) / dx).astype(int)
        row = g["active_id_map"][(ijk[0] * lab.shape[1] + ijk[1]) * lab.shape[2] + ijk[2]]
        b[row] += sgn * I / (dx * 1e-3) ** 3
    x = ps.solve_refined(_to_cells(g, b, cuda), tol=1e-8)
    phi = x.cpu().numpy()[g.cell_of_row]
    sx = g["surface_xyz"].astype(float)
    sr = g["surface_rows"]
    ref = int(np.argmin(np.linalg.norm(sx - np.array([R, 0, 0]), axis=1)))
    num = phi[sr] - phi[sr[ref]]
    ana = E.frank_sphere(p, sigma, R, sx, np.zeros(3), axis)
    ana -= E.frank_sphere(p, sigma, R, sx[ref:ref + 1], np.zeros(3), axis)[0]
    rdm = np.linalg.norm(num - ana) / np.linalg.norm(ana)
    print(f"  RDM {rdm:.4f}, outer {ps.last_stats['outer']}, inner {ps.last_stats['inner']}")
    assert rdm < 0.05


def test_heart_source_matches_oracle(cuda):
    print("[case] heart source", flush=True)
    lab, g, Ap = _setup(n=(40, 36, 30))
    S = np.zeros((40, 36, 24), dtype=bool)
    S[8:28, 6:26, 4:18] = True
    S[15:18, 10:14, 8:11] = False
    lay = BrickLayout(S, (8, 8, 4))
    h_fine = 0.5
    coords = lay.vox_idx[:, [1, 0, 2]].astype(np.float32) * h_fine
    reg = dict(R=np.eye(3), t=np.array([12.0, 14.0, 9.0]), scale=1.0)
    Vm = np.random.default_rng(3).uniform(-90, 30, lay.n_nodes)
    src = HeartSource(lay, coords, g, reg, 0.2625, h_fine, cuda)
    b = torch.zeros(g.layout["n_cells"], dtype=torch.float64, device=cuda)
    Vs = torch.from_numpy(lay.to_slots(Vm.astype(np.float32))).to(cuda)
    src(Vs, b, PoissonSolver(g, cuda).ws)
    got = b.cpu().numpy()[g.cell_of_row]
    K = E.heart_laplacian(lay.vox_idx)
    proj = E.projection(coords.astype(np.float64), reg["R"], reg["t"], 1.0, g["origin_mm"],
                        g["dx_mm"], g["active_id_map"], g["shape"])
    assert (proj >= 0).sum() == src.n_valid
    ref = E.source(Vm.astype(np.float32).astype(np.float64), K, proj, 0.2625, h_fine * 1e-3,
                   g["dx_mm"], len(g["active_ids"]), g["pin_row"])
    assert np.max(np.abs(got - ref)) < 1e-9 * np.abs(ref).max()


def test_electrodes_and_leads(cuda):
    print("[case] electrodes", flush=True)
    lab, g, Ap = _setup()
    pos = {n: g["surface_xyz"][i * 7].astype(float) for i, n in
           enumerate(["RA", "LA", "LL", "V1", "V2", "V3", "V4", "V5", "V6"])}
    el = Electrodes(g, pos, k=3, device=cuda)
    phi = np.random.default_rng(4).standard_normal(g.layout["n_cells"])
    out = el(torch.from_numpy(phi).to(cuda)).cpu().numpy()
    ref = phi[el.cells].mean(axis=1)
    np.testing.assert_allclose(out, ref, rtol=1e-14, atol=1e-15)
    L = leads_from_electrodes(dict(zip(el.names, out)))
    assert abs(L["I"] + L["III"] - L["II"]) < 1e-12          # Einthoven's law
    assert abs(L["aVR"] + L["aVL"] + L["aVF"]) < 1e-12       # Goldberger sum


def test_multigrid_vcycle_is_spd(cuda):
    """The V-cycle is a symmetric positive definite operator B: u.Bv = v.Bu and
    v.Bv > 0 for random residual vectors (zero on the pinned cell and holes)."""
    print("[case] multigrid V-cycle symmetry", flush=True)
    lab, g, Ap = _setup()
    ps = PoissonSolver(g, cuda, precond="mg")
    rng = np.random.default_rng(5)
    n = Ap.shape[0]
    vs = []
    for _ in range(3):
        v = rng.standard_normal(n)
        v[g["pin_row"]] = 0.0
        vs.append(_to_cells(g, v, cuda).float())
    Bv = []
    for v in vs:
        z = torch.zeros_like(v)
        ps.mg.vcycle(ps.torso, v, z)
        Bv.append(z.double())
    for a in range(3):
        assert float(vs[a].double() @ Bv[a]) > 0
        for b in range(a + 1, 3):
            x, y = float(vs[a].double() @ Bv[b]), float(vs[b].double() @ Bv[a])
            print(f"  u.Bv {x:.6e}  v.Bu {y:.6e}")
            assert abs(x - y) <= 1e-5 * max(abs(x), abs(y))


def test_multigrid_pcg_matches_direct(cuda):
    print("[case] multigrid PCG vs direct", flush=True)
    lab, g, Ap = _setup()
    rng = np.random.default_rng(1)
    b = rng.standard_normal(Ap.shape[0]) * 1e3
    b -= b.mean()
    b[g["pin_row"]] = 0.0
    xd = spla.spsolve(Ap.tocsc(), b)
    its = {}
    for pc in ("jacobi", "mg"):
        ps = PoissonSolver(g, cuda, precond=pc)
        x = ps.solve_refined(_to_cells(g, b, cuda), tol=1e-10)
        st = ps.last_stats
        its[pc] = st["inner"]
        print(f"  {pc}: outer {st['outer']} inner {st['inner']} rel_res {st['rel_res']:.2e}")
        assert st["rel_res"] <= 1e-10
        xs = x.cpu().numpy()[g.cell_of_row]
        assert np.max(np.abs(xs - xd)) < 1e-7 * np.abs(xd).max()
    assert its["mg"] * 4 < its["jacobi"]
