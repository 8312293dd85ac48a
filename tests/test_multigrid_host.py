"""Host-side construction of the torso multigrid hierarchy: coarse operators,
pinned cell, dense coarsest inverse."""
import numpy as np
import pytest

from hrdayapy_amd.ecg.functions.forward import (Multigrid, _face_sum, coarsen_faces,
                                                dense_operator)
from tests._torso_host import host_grid
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
    g = host_grid(lab, smap, 1.5)
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
    g = host_grid(lab, smap, 1.5)
    ps = PoissonSolver(g, "cpu", precond="mg")
    assert ps.z32.numel() == ps.n == g.layout["n_cells"]
    assert abs(float(ps.pin_w.sum()) - 1.0) < 1e-12
    L = g.layout
    faces = L["gx"][ps.pin_nbr.numpy()]  # neighbours exist and conduct
    assert len(ps.pin_nbr) >= 1 and ps.pin_G > 0 and np.all(np.isfinite(faces))
    assert int(ps.torso.n_chunks) == int(((L["col_len"] + 63) // 64).sum())


def test_torso_chunk_table():
    """hp_torso_chunks: the chunks cover every stored cell exactly once, in
    stored order, and each record's neighbour base/span give the stored index
    of the +x/-x/+y/-y cell at the same k (or nothing) as the column arrays do."""
    from hrdayapy_amd import _native as nat
    lab, smap = _torso(n=(30, 26, 150))   # runs longer than 64 cells
    g = host_grid(lab, smap, 1.5)
    L = g.layout
    ni, nj, nk = g["shape"]
    k0, ln, off = (np.asarray(L[k]) for k in ("col_k0", "col_len", "col_off"))
    tab = nat.torso_chunks((ni, nj, nk), k0, ln, off)
    assert tab.shape[1] == 16 and np.all(tab[:, 14:] == 0)
    assert np.all((tab[:, 1] >= 1) & (tab[:, 1] <= 64)) and ln.max() > 64
    cells = np.concatenate([r[4] + r[0] + np.arange(r[1]) for r in tab])
    assert np.array_equal(cells, np.arange(int(L["n_cells"])))

    def stored(i, j, k):
        if not (0 <= i < ni and 0 <= j < nj):
            return -1
        c = i * nj + j
        return int(off[c] + k - k0[c]) if ln[c] > 0 and k0[c] <= k < k0[c] + ln[c] else -1

    rng = np.random.default_rng(3)
    for r in tab[rng.choice(len(tab), 400, replace=False)]:
        i, j = int(r[2]), int(r[3])
        for k in range(r[0] - 1, r[0] + r[1] + 1):
            for q, (di, dj) in enumerate(((0, 0), (1, 0), (-1, 0), (0, 1), (0, -1))):
                lo, hi = r[9 + q] & 0xFFFF, r[9 + q] >> 16
                got = int(r[4 + q] + k) if lo <= k < hi else -1
                assert got == stored(i + di, j + dj, k), (i, j, k, q)
    import ctypes
    a = [np.ascontiguousarray(v, dtype=np.int32) for v in (k0, ln, off)]
    small, cnt = np.zeros((2, 16), np.int32), ctypes.c_int64(0)
    with pytest.raises(RuntimeError, match="capacity"):
        nat.call("hp_torso_chunks", ni, nj, nk, *(v.ctypes.data for v in a), small.ctypes.data, 1,
                 ctypes.byref(cnt))


def test_standard_electrodes_with_and_without_bspm():
    """The 12-lead set alone (n_bspm = 0) and with a BSPM grid: limb and
    precordial electrodes on the body surface, the BSPM sites added after them."""
    from hrdayapy_amd import geometry as G
    from hrdayapy_amd.ecg.functions.forward import standard_electrodes
    lab, origin, hc = G.torso((64, 64, 48), 6.0)
    grid = host_grid(lab, G.TORSO_SIGMA, 6.0, origin)
    twelve = ["RA", "LA", "LL"] + [f"V{i}" for i in range(1, 7)]
    pos = standard_electrodes(grid, hc, n_bspm=0)
    assert list(pos) == twelve
    surf = {tuple(p) for p in grid["surface_xyz"].astype(float)}
    assert all(tuple(np.asarray(p, float)) in surf for p in pos.values())
    pos16 = standard_electrodes(grid, hc, n_bspm=16)
    assert list(pos16)[:9] == twelve and len(pos16) == 25
