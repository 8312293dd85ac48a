"""Pins the CPU oracle of the stencil, Laplacian, CN/BiCGSTAB and stimulus (oracle/mesh.py)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import mesh as M
from tests._problems import slab


def _K(myo, N):
    e = myo["edges"]
    w = myo["edge_w"]
    K = sp.coo_matrix((np.r_[w, w], (np.r_[e[:, 0], e[:, 1]], np.r_[e[:, 1], e[:, 0]])),
                      shape=(N, N)).tocsr()
    return K - sp.diags(np.asarray(K.sum(axis=1)).ravel())


def test_row_sums_zero_and_symmetric():
    S, phi, f = slab(10, 8, 6)
    myo = M.build_myo_mesh(S, 0.25, f, 0.24, 0.08)
    K = _K(myo, myo["N_myo"])
    assert abs(K.sum(axis=1)).max() < 1e-15
    assert abs(K - K.T).max() < 1e-18
    L, dL = M.global_laplacian(0, myo["N_myo"], np.zeros((0, 2)), 0, 1, myo["edges"],
                               myo["edge_w"], 140 * 0.25 ** 3, [], [], 0.0)
    assert abs(L.sum(axis=1)).max() < 1e-12
    np.testing.assert_allclose(L.diagonal(), dL, rtol=1e-14)


def test_isotropic_reduces_to_seven_point():
    S, phi, f = slab(8, 7, 5)
    a = M.build_myo_mesh(S, 0.4, f, 0.13, 0.13)
    b = M.build_myo_mesh_iso(S, 0.4, 0.13)
    N = a["N_myo"]
    Ka, Kb = _K(a, N), _K(b, N)
    assert abs(Ka - Kb).max() < 1e-15


def test_uniform_tensor_exact_on_quadratics():
    # interior rows of L applied to a quadratic u = x^T Q x / 2 give tr(D Q)/A_M
    h, A_M = 0.3, 140.0
    S = np.ones((9, 9, 9), dtype=np.uint8)
    fdir = np.array([0.6, 0.48, 0.64])
    f = np.broadcast_to(fdir / np.linalg.norm(fdir), (9, 9, 9, 3)).astype(np.float64)
    sl, st = 0.24, 0.08
    myo = M.build_myo_mesh(S, h, f, sl, st)
    N = myo["N_myo"]
    L, _ = M.global_laplacian(0, N, np.zeros((0, 2)), 0, 1, myo["edges"], myo["edge_w"],
                              A_M * h ** 3, [], [], 0.0)
    Q = np.array([[1.0, 0.3, -0.2], [0.3, -0.5, 0.7], [-0.2, 0.7, 0.9]])
    x = myo["vox_idx"] * h
    u = 0.5 * np.einsum("ni,ij,nj->n", x, Q, x) + x @ np.array([0.1, -0.2, 0.3])
    ff = fdir / np.linalg.norm(fdir)
    D = st * np.eye(3) + (sl - st) * np.outer(ff, ff)
    interior = np.all((myo["vox_idx"] > 0) & (myo["vox_idx"] < 8), axis=1)
    np.testing.assert_allclose((L @ u)[interior], np.trace(D @ Q) / A_M, rtol=1e-10)


def test_harmonic_face_mean_gives_series_flux():
    # two voxels along x with different axial conductivity (fibre along x in the
    # first, along y in the second): the face weight must be the conductance of
    # two half-cells in series, h^2 / ((h/2)/s_a + (h/2)/s_b)
    S = np.ones((2, 1, 1), dtype=np.uint8)
    f = np.zeros((2, 1, 1, 3))
    f[0, 0, 0, 0] = 1.0
    f[1, 0, 0, 1] = 1.0
    h, sl, st = 0.5, 0.24, 0.03
    myo = M.build_myo_mesh(S, h, f, sl, st)
    face = [k for k, (a, b) in enumerate(myo["edges"]) if {a, b} == {0, 1}]
    assert len(face) == 1
    np.testing.assert_allclose(myo["edge_w"][face[0]], h * h / ((h / 2) / sl + (h / 2) / st),
                               rtol=1e-14)


def test_bicgstab_matches_direct_solve():
    S, phi, f = slab(10, 8, 6)
    myo = M.build_myo_mesh(S, 0.25, f, 0.24, 0.08)
    N = myo["N_myo"]
    L, dL = M.global_laplacian(0, N, np.zeros((0, 2)), 0, 1, myo["edges"], myo["edge_w"],
                               140 * 0.25 ** 3, [], [], 0.0)
    A, Mi = M.cn_matrix(L, dL, 0.01 / 0.02, 0.5)
    b = np.random.default_rng(0).standard_normal(N)
    x = M.bicgstab(A, Mi, b, np.zeros(N), tol=1e-12, maxiter=500)
    np.testing.assert_allclose(x, spla.spsolve(A.tocsc(), b), rtol=1e-9, atol=1e-12)


def test_stim_active_schedule():
    assert not M.stim_active(0.5, 1.0, 2.0, None, None)
    assert M.stim_active(1.0, 1.0, 2.0, None, None)
    assert not M.stim_active(3.0, 1.0, 2.0, None, None)
    assert M.stim_active(101.5, 1.0, 2.0, 100.0, None)
    assert not M.stim_active(101.5, 1.0, 2.0, 100.0, 1)
    assert not M.stim_active(104.0, 1.0, 2.0, 100.0, None)


def test_discretise_tree_segments():
    nodes = np.array([[0, 0, 0], [1.0, 0, 0], [1.0, 0.45, 0]], dtype=np.float32)
    cn, ce, bm = M.discretise_tree(nodes, np.array([[0, 1], [1, 2]]), 0.2)
    assert [len(b) for b in bm] == [6, 3]
    assert len(cn) == 3 + 4 + 1
    seg = np.linalg.norm(cn[ce[:, 0]] - cn[ce[:, 1]], axis=1)
    np.testing.assert_allclose(seg[:5], 0.2, rtol=1e-6)


def test_cn_jacobi_solve_against_direct_and_bicgstab():
    """The oracle's restatement of the kernels' CN solve (Jacobi sweeps, the
    junction rows re-solved after each sweep) against a direct solve of the
    same CN system on a tree + slab with stiff junctions, and its stopping
    rule against the reference's BiCGSTAB at the same tolerance."""
    import scipy.sparse.linalg as spla
    from oracle import monodomain as OMD
    from tests.test_gpu_monodomain import _case
    args, kw = _case("tree")
    nodes, el, act, S, phi, f = args
    prob = OMD.Problem(nodes, el, act, S, phi, f, **kw)
    A, Minv = M.cn_matrix(prob.L, prob.diag_L, prob.Cm / 0.02, 0.5)
    rng = np.random.default_rng(5)
    Vs = -80.0 + 10.0 * rng.standard_normal(prob.N)
    Vs[:prob.Np] += 60.0  # junctions far from their patches: the stiff case
    b = (prob.Cm / 0.02) * Vs + 0.5 * (prob.L @ Vs)
    xd = spla.spsolve(A.tocsc(), b)
    jn = sorted(set(int(q) for q in prob.pmj_ids))
    x, k = OMD.cn_jacobi_solve(A, Minv, prob.L, Vs, b, 1e-12, 200, jn, prob.Np)
    assert k < 200 and np.max(np.abs(x - xd)) < 1e-9 * np.abs(xd).max()
    # tol 1e-5: the returned iterate is one sweep past the first that met the
    # rule; its own residual meets it too, as BiCGSTAB's answer does
    x5, k5 = OMD.cn_jacobi_solve(A, Minv, prob.L, Vs, b, 1e-5, 200, jn, prob.Np)
    xb = M.bicgstab(A, Minv, b, Vs, 1e-5, 500)
    bn = np.linalg.norm(b)
    assert np.linalg.norm(b - A @ x5) <= 1e-5 * bn
    assert np.linalg.norm(b - A @ xb) <= 1.0001e-5 * bn
    # the same solution to within the tolerance (A >= Cm/dt I in the D-norm)
    assert np.max(np.abs(x5 - xd)) < 0.05 and np.max(np.abs(xb - xd)) < 0.05
    # without the junction re-solve the same rule takes more sweeps
    _, kj = OMD.cn_jacobi_solve(A, Minv, prob.L, Vs, b, 1e-5, 200, [], prob.Np)
    assert k5 < kj


def test_cn_jacobi_warm_start():
    """The warm start x0 = V* + e: with e = 0 it is the cold solve bit for bit;
    its first residual L V* + theta L e - (Cm/dt) e equals b - A x0 (a sign or
    factor slip in that form fails here); with a good guess the rule is met in
    fewer sweeps and the answer still meets it."""
    from oracle import monodomain as OMD
    from tests.test_gpu_monodomain import _case
    args, kw = _case("tree")
    nodes, el, act, S, phi, f = args
    prob = OMD.Problem(nodes, el, act, S, phi, f, **kw)
    cm_dt = prob.Cm / 0.02
    A, Minv = M.cn_matrix(prob.L, prob.diag_L, cm_dt, 0.5)
    rng = np.random.default_rng(6)
    Vs = -80.0 + 10.0 * rng.standard_normal(prob.N)
    b = cm_dt * Vs + 0.5 * (prob.L @ Vs)
    jn = sorted(set(int(q) for q in prob.pmj_ids))
    x0, k0 = OMD.cn_jacobi_solve(A, Minv, prob.L, Vs, b, 1e-5, 50, jn, prob.Np)
    xz, kz = OMD.cn_jacobi_solve(A, Minv, prob.L, Vs, b, 1e-5, 50, jn, prob.Np,
                                 np.zeros(prob.N), 0.5, cm_dt)
    assert kz == k0 and np.array_equal(xz, x0)
    e = rng.standard_normal(prob.N)
    r_form = prob.L @ Vs + 0.5 * (prob.L @ e) - cm_dt * e
    r_true = b - A @ (Vs + e)
    assert np.abs(r_form - r_true).max() <= 1e-12 * np.abs(r_true).max()
    x_ex, _ = OMD.cn_jacobi_solve(A, Minv, prob.L, Vs, b, 1e-13, 200, jn, prob.Np)
    guess = (x_ex - Vs) + 1e-4 * rng.standard_normal(prob.N)
    xw, kw_ = OMD.cn_jacobi_solve(A, Minv, prob.L, Vs, b, 1e-5, 50, jn, prob.Np, guess, 0.5,
                                  cm_dt)
    assert kw_ < k0
    assert np.linalg.norm(b - A @ xw) <= 1e-5 * np.linalg.norm(b)


def test_cn_jacobi_trajectory_matches_reference_scheme():
    """Over 8 ms of a tree-driven activation the Jacobi-solved CN trajectory
    stays within a few hundredths of a mV of the BiCGSTAB-solved one."""
    from oracle import monodomain as OMD
    from tests.test_gpu_monodomain import _case
    args, kw = _case("tree")
    nodes, el, act, S, phi, f = args
    prob = OMD.Problem(nodes, el, act, S, phi, f, **kw)
    a = OMD.run(prob, 8.0, 0.02, scheme="cn", record_every=25)
    b = OMD.run(prob, 8.0, 0.02, scheme="cn_jacobi", record_every=25, max_sweeps=12)
    c = OMD.run(prob, 8.0, 0.02, scheme="cn_jacobi", record_every=25, max_sweeps=12, warm=False)
    for tr in (b, c):  # warm-started (the default) and cold-started sweeps
        d = a["trace_V"] - tr["trace_V"]
        assert np.sqrt(np.mean(d ** 2)) < 0.05 and np.abs(d).max() < 1.0
        assert tr["sweeps"].max() < 12
    assert b["sweeps"].sum() < c["sweeps"].sum()  # the warm start saves sweeps
    assert a["trace_V"][:, :prob.Np].max() > 0.0  # the tree fired


@pytest.mark.parametrize("dt,bound,max_iter_diff", [(0.02, 2e-5, 0), (0.1, 2e-3, 6)])
def test_bicgstab_trajectory_moves_little_with_the_summation_order(dt, bound, max_iter_diff):
    """The bars of the GPU's FP64 BiCGSTAB parity tests: over 20 ms of the
    tree-driven case, the reference scheme with its inner products summed in
    reverse order stays within `bound` of itself (measured 6.4e-6 mV at
    dt = 0.02 with the same iteration count in every step; 5.6e-4 mV at
    dt = 0.1, where the solve takes ~6 iterations, with 3 fewer in total)."""
    from oracle import monodomain as OMD
    from tests.test_gpu_monodomain import _case
    args, kw = _case("tree")
    prob = OMD.Problem(*args, **kw)
    a = OMD.run(prob, 20.0, dt, scheme="cn")
    b = OMD.run(prob, 20.0, dt, scheme="cn",
                dot=lambda u, w: float(np.dot(u[::-1], w[::-1])))
    d = np.abs(a["V"] - b["V"]).max()
    assert 0.0 < d < bound
    assert abs(int(a["sweeps"].sum()) - int(b["sweeps"].sum())) <= max_iter_diff
    assert a["sweeps"].max() >= 3  # a case where the solve takes several iterations
