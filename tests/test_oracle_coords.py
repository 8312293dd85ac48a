"""Pins the oracle's transmural coordinate and rule-based fibres
(oracle/coordinates.py) against the reference's own outputs
(tests/golden/coord_golden.npz, written by tests/golden/make_coord_vectors.py)
and against the invariants of the construction."""
import os

import numpy as np
import pytest

from oracle import coordinates as OC

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "coord_golden.npz"))


def test_transmural_matches_reference(gold):
    S, lab = gold["S"].astype(bool), gold["surface_label"]
    phi, it = OC.transmural_coordinate(S, lab, tol=1e-6)
    ref = gold["phi"]
    assert np.array_equal(np.isnan(phi), np.isnan(ref))
    d = np.nanmax(np.abs(phi - ref))
    print(f"oracle phi vs reference: max |d| {d:.2e} after {it} CG iterations")
    assert d < 2e-5
    assert np.all(phi[lab == -1] == 0.0) and np.all(phi[lab == 1] == 1.0)


def test_transmural_system_is_the_graph_laplacian(gold):
    """Interior rows sum to zero and are symmetric among interior voxels;
    Dirichlet rows are identity rows."""
    S, lab = gold["S"].astype(bool), gold["surface_label"]
    A, b, vox, lv = OC.transmural_system(S, lab)
    inner = lv == 0
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert np.allclose(rs[inner], 0.0)
    Ai = A[inner][:, inner]
    assert abs(Ai - Ai.T).max() == 0.0
    Ad = A[~inner]
    assert np.array_equal(Ad.indices, np.nonzero(~inner)[0]) and np.all(Ad.data == 1.0)
    assert np.array_equal(b, (lv == 1).astype(np.float32))


@pytest.mark.parametrize("key,spacing,angles", [("fibre_iso", (2.0, 2.0, 2.0), (60.0, -60.0)),
                                                 ("fibre_aniso", (1.0, 1.5, 2.0), (50.0, -70.0))])
def test_fibres_match_reference(gold, key, spacing, angles):
    S = gold["S"].astype(bool)
    f = OC.fibre_direction(S, gold["phi"], gold["psi"], spacing, *angles)
    ref = gold[key]
    d = np.abs(f - ref).max()
    print(f"{key}: max |d| {d:.2e}")
    assert d < 2e-6


def test_fibre_invariants(gold):
    """|f| = 1 on the wall, f orthogonal to grad(phi) and at the helix angle
    of the law alpha(phi) against the circumferential direction."""
    from scipy.ndimage import binary_erosion
    S = gold["S"].astype(bool)
    phi, psi = gold["phi"], gold["psi"]
    f = OC.fibre_direction(S, phi, psi, (2.0, 2.0, 2.0))
    Fp = OC.fill_nan_nearest(phi, S)
    g = np.stack(np.gradient(Fp, 2.0, 2.0, 2.0), axis=-1)
    t = g / np.maximum(np.linalg.norm(g, axis=-1, keepdims=True), 1e-8)
    Fs = OC.fill_nan_nearest(psi, S)
    gs = np.stack(np.gradient(Fs, 2.0, 2.0, 2.0), axis=-1)
    el = gs - (gs * t).sum(-1, keepdims=True) * t
    eln = np.linalg.norm(el, axis=-1)
    el /= np.maximum(eln, 1e-8)[..., None]
    # where both gradients define the frame (not at the apex, where grad(psi)
    # is parallel to grad(phi), nor where grad(phi) vanishes)
    inner = (binary_erosion(S, iterations=1) & (np.linalg.norm(g, axis=-1) > 1e-3)
             & (eln > 1e-3))
    assert inner.sum() > 0.5 * S.sum()
    assert np.allclose(np.linalg.norm(f[inner], axis=1), 1.0, atol=1e-5)
    assert np.abs((f * t).sum(-1)[inner]).max() < 1e-5
    # helix angle: f = cos(a) e_c + sin(a) e_l, so a = atan2(f.e_l, f.e_c)
    ec = np.cross(t, el)
    a = np.degrees(np.arctan2((f * el).sum(-1), (f * ec).sum(-1)))
    law = 60.0 - 120.0 * Fp
    assert np.abs(a - law)[inner].max() < 1e-3
