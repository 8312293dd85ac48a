"""The oracle's torso-grid build (sharp and smoothed boundary) and its
registration against the reference's own outputs (golden vectors written by
tests/golden/make_torso_vectors.py), plus the properties that hold for any
correct restatement."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import ecg as OE
from oracle import registration as OR

GOLD = os.path.join(os.path.dirname(__file__), "golden", "torso_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


def _A(g, tag):
    n = len(g[f"{tag}_active_ids"])
    return sp.csr_matrix((g[f"{tag}_A_data"], g[f"{tag}_A_indices"], g[f"{tag}_A_indptr"]),
                         shape=(n, n))


def _sigma(g, coarse):
    s = np.zeros(coarse.shape)
    for k, v in zip(g["sigma_keys"], g["sigma_vals"]):
        s[coarse == k] = v
    return s


def test_resample_and_smoothed_domain_match_reference(gold):
    coarse = OE.resample_labels(gold["fine_labels"], gold["fine_spacing"], float(gold["dx_coarse"]))
    assert np.array_equal(coarse, gold["coarse_labels"])
    psi = OE.smoothed_domain(gold["fine_labels"], gold["fine_spacing"], float(gold["dx_coarse"]),
                             coarse.shape)
    assert np.array_equal(psi, gold["psi"])
    # psi is a domain function: 0 or in [cutoff, 1], ~1 deep inside, 0 far outside
    a = psi[psi > 0]
    assert a.min() >= 1e-3 and a.max() <= 1.0
    assert psi[0, 0, 0] == 0.0


@pytest.mark.parametrize("tag", ["sharp", "sbm"])
def test_assembly_matches_reference(gold, tag):
    coarse = gold["coarse_labels"]
    sigma = _sigma(gold, coarse)
    tissue = coarse != 0
    h_m = float(gold["dx_coarse"]) * 1e-3
    if tag == "sbm":
        psi = gold["psi"]
        active = psi > 0
        sig = OE.extend_conductivity(sigma, tissue, active)
        assert np.array_equal(sig, gold["sbm_sigma"])
        A, ids, _ = OE.assemble_stiffness(sig, h_m, active, psi_vol=psi)
    else:
        A, ids, _ = OE.assemble_stiffness(sigma, h_m, tissue)
    assert np.array_equal(ids, gold[f"{tag}_active_ids"])
    R = _A(gold, tag)
    A = A.tocsr()
    A.sort_indices()
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    np.testing.assert_allclose(A.data, R.data, rtol=1e-14, atol=0)
    # symmetric, zero row sums (Neumann): holds for both boundaries
    assert abs(A - A.T).max() <= 1e-9 * abs(A).max()
    assert np.abs(np.asarray(A.sum(axis=1)).ravel()).max() <= 1e-9 * abs(A).max()


def test_step_function_psi_reproduces_sharp_assembly(gold):
    """psi = 1 on tissue and 0 elsewhere gives the sharp-boundary matrix exactly."""
    coarse = gold["coarse_labels"]
    sigma = _sigma(gold, coarse)
    tissue = coarse != 0
    h_m = float(gold["dx_coarse"]) * 1e-3
    A0, _, _ = OE.assemble_stiffness(sigma, h_m, tissue)
    A1, _, _ = OE.assemble_stiffness(sigma, h_m, tissue, psi_vol=tissue.astype(float))
    assert abs(A0 - A1).max() == 0.0


def test_surface_matches_reference(gold):
    coarse = gold["coarse_labels"]
    sf = OE.surface_flat(coarse, coarse != 0)
    assert np.array_equal(sf, gold["surface_flat"])


def test_registration_matches_reference(gold):
    src = gold["reg_src"]
    tgt = gold["reg_tgt"].astype(np.float64)
    R, t, scale, dice, all_d = OR.register(src, tgt, 3.0)
    assert np.array_equal(all_d, gold["reg_all_dice"])
    assert np.array_equal(R, gold["reg_R"])
    assert np.array_equal(t, gold["reg_t"]) and scale == float(gold["reg_scale"])
    assert dice == float(gold["reg_dice"])
    # the known signed permutation is the one recovered
    assert np.array_equal(R, gold["reg_R_true"])


def test_voxel_dice_closed_cases():
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 30, size=(500, 3))
    assert OR.voxel_dice(a, a, 3.0) == 1.0
    assert OR.voxel_dice(a, a + 100.0, 3.0) == 0.0
    # two points in one voxel and one in another: |A| = 1, |B| = 2, |A and B| = 1
    assert OR.voxel_dice(np.array([[0.5, 0.5, 0.5]]),
                         np.array([[0.2, 0.2, 0.2], [4.0, 0.2, 0.2]]), 3.0) == pytest.approx(2 / 3)
    assert len(OR.signed_permutations()) == 48
    dets = [round(np.linalg.det(R)) for R in OR.signed_permutations()]
    assert dets.count(1) == 24 and dets.count(-1) == 24
