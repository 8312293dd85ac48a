"""Pins the CPU oracle of the ECG forward solve (oracle/ecg.py) (CPU only)."""
import numpy as np
import scipy.sparse.linalg as spla

from oracle import ecg as E


def _torso(seed=0, n=(18, 16, 14)):
    rng = np.random.default_rng(seed)
    ni, nj, nk = n
    ax = [np.linspace(-1, 1, m) for m in n]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    lab = np.zeros(n, dtype=np.int32)
    lab[X ** 2 + Y ** 2 + Z ** 2 <= 0.95] = 1
    lab[(X + 0.35) ** 2 / 0.1 + Y ** 2 / 0.3 + Z ** 2 / 0.3 <= 1] = 3
    lab[(X - 0.2) ** 2 + Y ** 2 + Z ** 2 <= 0.08] = 4
    lab[0, 0, 0] = 0
    return lab, {1: 0.2, 3: 0.05, 4: 0.7}


def test_stiffness_symmetric_zero_rowsum():
    lab, smap = _torso()
    sig = np.zeros(lab.shape)
    for k, v in smap.items():
        sig[lab == k] = v
    A, ids, idm = E.assemble_stiffness(sig, 1.5e-3, lab != 0)
    assert abs(A - A.T).max() < 1e-6 * abs(A).max()
    assert np.abs(np.asarray(A.sum(axis=1)).ravel()).max() < 1e-9 * abs(A).max()


def test_pcg_matches_direct_solve():
    lab, smap = _torso()
    sig = np.zeros(lab.shape)
    for k, v in smap.items():
        sig[lab == k] = v
    A, ids, idm = E.assemble_stiffness(sig, 1.5e-3, lab != 0)
    pr = E.pin_row_of(lab != 0, idm)
    Ap = E.pin(A, pr)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A.shape[0])
    b -= b.mean()
    b[pr] = 0.0
    xd = spla.spsolve(Ap.tocsc(), b)
    x, it, err = E.pcg(Ap, b, 1.0 / Ap.diagonal(), tol_abs_mV=1e-13, max_iter=5000)
    assert it < 5000
    np.testing.assert_allclose(x, xd, rtol=1e-6, atol=1e-9 * np.abs(xd).max())


def test_frank_sphere_dipole():
    # Frank's closed form on the pinned FDM system; the reference accepts
    # RDM < 0.05 at dx/R = 0.02, this coarser grid (dx/R = 0.04) at 0.1
    R, dx, sigma, p = 50.0, 2.0, 0.2, 1e-6
    lab, origin = E.sphere_labels(R, dx)
    sig = lab * sigma
    A, ids, idm = E.assemble_stiffness(sig.astype(float), dx * 1e-3, lab != 0)
    pr = E.pin_row_of(lab != 0, idm)
    Ap = E.pin(A, pr)
    axis = np.array([0.0, 0.0, 1.0])
    d = 8.0
    h = dx * 1e-3
    I = p / (d * 1e-3)

    def row(pt):
        ijk = np.round((pt - origin) / dx).astype(int)
        return idm[(ijk[0] * lab.shape[1] + ijk[1]) * lab.shape[2] + ijk[2]]

    b = np.zeros(A.shape[0])
    b[row(axis * d / 2)] += I / h ** 3
    b[row(-axis * d / 2)] -= I / h ** 3
    phi, it, _ = E.pcg(Ap, b, 1.0 / Ap.diagonal(), tol_abs_mV=1e-10, max_iter=20000)
    assert it < 20000
    sf = E.surface_flat(lab, lab != 0)
    sxyz = origin + np.stack(np.unravel_index(sf, lab.shape), 1) * dx
    srow = idm[sf]
    ref = int(np.argmin(np.linalg.norm(sxyz - np.array([R, 0, 0]), axis=1)))
    num = phi[srow] - phi[srow[ref]]
    ana = E.frank_sphere(p, sigma, R, sxyz, np.zeros(3), axis)
    ana = ana - E.frank_sphere(p, sigma, R, sxyz[ref:ref + 1], np.zeros(3), axis)[0]
    rdm = np.linalg.norm(num - ana) / np.linalg.norm(ana)
    assert rdm < 0.1, rdm


def test_source_uniform_zero_and_linear():
    vox = np.argwhere(np.ones((6, 5, 4), bool))
    K = E.heart_laplacian(vox)
    proj = np.arange(len(vox)) % 7
    proj[::11] = -1  # unprojected voxels break the cancellation; the mean removal restores it
    b0 = E.source(np.full(len(vox), -80.0), K, proj, 0.2625, 0.4e-3, 2.0, 8, 7)
    assert np.abs(b0).max() < 1e-9
    rng = np.random.default_rng(0)
    v1, v2 = rng.standard_normal(len(vox)), rng.standard_normal(len(vox))
    s = lambda v: E.source(v, K, proj, 0.2625, 0.4e-3, 2.0, 8, 7)
    np.testing.assert_allclose(s(2 * v1 + v2), 2 * s(v1) + s(v2), atol=1e-9)
    assert abs(s(v1).sum()) < 1e-9 * np.abs(s(v1)).max()  # sum(b) = 0 (Neumann compatibility)


