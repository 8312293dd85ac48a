"""Host-side construction of the torso multigrid hierarchy: coarse operators,
pinned cell, dense coarsest inverse."""
import numpy as np

from hrdayapy_amd.ecg.functions.forward import (Multigrid, _face_sum, coarsen_faces,
                                                dense_operator)
from hrdayapy_amd.ecg.functions.torso import TorsoGrid
from tests.test_oracle_ecg import _torso


def test_uniform_box_coarsens_to_the_2h_operator():
    """Uniform conductance g on a full 8x8x8 box: the coarse face conductance is
    2g, i.e. (sigma/(2h)^2) x 8 -- the 2h operator applied to the restricted
    (summed) residual of the 8 children."""
    g = 3.0
    G = [np.full((8, 8, 8), g) for _ in range(3)]
    G[0][-1], G[1][:, -1], G[2][:, :, -1] = 0.0, 0.0, 0.0  # no faces out of the box
    C = coarsen_faces(*G)
    assert C[0].shape == (4, 4, 4)
    assert np.allclose(C[0][:-1], 2 * g) and np.allclose(C[0][-1], 0.0)
    assert np.allclose(C[1][:, :-1], 2 * g) and np.allclose(C[2][:, :, :-1], 2 * g)
    S = _face_sum(*C)
    assert np.isclose(S[1, 1, 1], 12 * g) and np.isclose(S[0, 0, 0], 6 * g)


def test_odd_box_pads_without_new_faces():
    G = [np.ones((5, 3, 7)) for _ in range(3)]
    G[0][-1], G[1][:, -1], G[2][:, :, -1] = 0.0, 0.0, 0.0
    C = coarsen_faces(*G)
    assert C[0].shape == (3, 2, 4)
    # the last coarse layers only hold padded children on the + side: no + faces
    assert np.all(C[0][-1] == 0) and np.all(C[1][:, -1] == 0) and np.all(C[2][:, :, -1] == 0)


def test_dense_operator_symmetric_zero_rowsum():
    rng = np.random.default_rng(0)
    G = [rng.uniform(0.1, 1.0, (3, 4, 2)) for _ in range(3)]
    G[0][-1], G[1][:, -1], G[2][:, :, -1] = 0.0, 0.0, 0.0
    A = dense_operator(*G)
    assert np.allclose(A, A.T)
    assert np.allclose(A.sum(axis=1), 0.0)


def test_hierarchy_of_a_torso():
    lab, smap = _torso(n=(30, 26, 22))
    g = TorsoGrid.from_labels(lab, smap, 1.5)
    mg = Multigrid(g, "cpu")
    assert mg.shapes[0] == (15, 13, 11)
    assert np.prod(mg.shapes[-1]) <= 1024
    assert mg.mg.n_levels == len(mg.shapes)
    # coarsest inverse: inverse of the operator on the active (non-pinned) cells
    inv = mg.coarse_inv.numpy().astype(np.float64)
    lvl = mg.mg.level[mg.mg.n_levels - 1]
    n = lvl.n
    assert inv.shape == (n, n)
    act = np.nonzero(np.abs(inv).sum(axis=1) > 0)[0]
    assert len(act) == mg.active[-1]
    A = dense_operator(*mg.coarse_faces)
    prod = inv[np.ix_(act, act)] @ A[np.ix_(act, act)]
    assert np.abs(prod - np.eye(len(act))).max() < 1e-4
    # the grounded coarse cell (parent of the pinned cell) is excluded, which
    # makes the operator on the remaining cells nonsingular
    assert mg.active[-1] < int((_face_sum(*mg.coarse_faces) > 0).sum())
    assert np.linalg.matrix_rank(A[np.ix_(act, act)]) == len(act)
    # level-0 inverse diagonal: zero exactly on the pinned cell and the holes
    inv0 = mg.keep[0].numpy()
    assert inv0[g.layout["pin"]] == 0.0
    holes = ~g["active_mask"].ravel()[g.flat_of_cell]
    assert np.all(inv0[holes] == 0.0)


def test_solver_setup_on_host_memory():
    """PoissonSolver with the multigrid builds on host tensors (no kernel is
    launched at set-up): vectors sized to the grid, and the pin's neighbours
    carry weights that sum to one over the conductance to the pin."""
    import torch
    from hrdayapy_amd.ecg.functions.forward import PoissonSolver
    lab, smap = _torso(n=(30, 26, 22))
    g = TorsoGrid.from_labels(lab, smap, 1.5)
    ps = PoissonSolver(g, "cpu", precond="mg")
    assert ps.z32.numel() == ps.n == g.layout["n_cells"]
    assert abs(float(ps.pin_w.sum()) - 1.0) < 1e-12
    L = g.layout
    faces = L["gx"][ps.pin_nbr.numpy()]  # neighbours exist and conduct
    assert len(ps.pin_nbr) >= 1 and ps.pin_G > 0 and np.all(np.isfinite(faces))
    assert int(ps.torso.n_col_list) == int((L["col_len"] > 0).sum())
