"""Closed-form pins of the explicit form's junction exchange
(oracle.monodomain.Problem.pmj_exchange, the scheme of hp_pmj_couple /
hp_step_tail): over dt the junction conductances alone integrate the linear
ODE system
    Ai_P Cm dVp/dt = sum_m (c/n) (Vm - Vp),   Ai_M Cm dVm/dt = (c/n) (Vp - Vm),
whose exact solution is exp(dt M) V0 (scipy's matrix exponential, an
independent evaluation); charge Ai_P Vp + Ai_M sum Vm is conserved; shared
patch voxels receive the sum of their junctions' increments."""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import monodomain as OMD


def _problem(n_patch, c=2.0, Ai_P=0.0188, Ai_M=2.1875, Cm=0.01, n_myo=12, patches=None,
             ids=None):
    p = OMD.Problem.__new__(OMD.Problem)
    p.Np = 4
    p.N = p.Np + n_myo
    p.Cm, p.c_pmj, p.Ai_P, p.Ai_M = Cm, c, Ai_P, Ai_M
    p.pmj_ids = ids if ids is not None else [3]
    p.pmj_patches = patches if patches is not None else [list(range(n_patch))]
    return p


def _ode_matrix(p):
    """The linear system of the junction conductances over the global vector."""
    Mx = np.zeros((p.N, p.N))
    for q, patch in zip(p.pmj_ids, p.pmj_patches):
        w = p.c_pmj / len(patch)
        for m in patch:
            i = p.Np + m
            Mx[q, q] -= w / (p.Ai_P * p.Cm)
            Mx[q, i] += w / (p.Ai_P * p.Cm)
            Mx[i, i] -= w / (p.Ai_M * p.Cm)
            Mx[i, q] += w / (p.Ai_M * p.Cm)
    return Mx


@pytest.mark.parametrize("n_patch", [1, 3, 7])
@pytest.mark.parametrize("dt", [0.02, 0.1])
def test_exchange_is_the_exact_solution(n_patch, dt):
    p = _problem(n_patch)
    rng = np.random.default_rng(n_patch)
    V0 = -80.0 + 30.0 * rng.standard_normal(p.N)
    V0[3] = 20.0
    got = p.pmj_exchange(V0, dt, vmin=-1e9, vmax=1e9)
    ref = expm(dt * _ode_matrix(p)) @ V0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9 * np.abs(V0).max())
    # charge conservation
    q0 = p.Ai_P * V0[3] + p.Ai_M * V0[p.Np:p.Np + n_patch].sum()
    q1 = p.Ai_P * got[3] + p.Ai_M * got[p.Np:p.Np + n_patch].sum()
    assert abs(q1 - q0) < 1e-12 * abs(q0)
    # untouched nodes stay put
    np.testing.assert_array_equal(got[p.Np + n_patch:], V0[p.Np + n_patch:])
    np.testing.assert_array_equal(got[:3], V0[:3])


def test_shared_voxels_sum_their_junctions():
    """Two junctions whose patches share voxel 2: the exchange splits the
    coupling by junction (each integrated exactly against its own patch mean)
    and the shared voxel takes the sum of the two increments; to first order
    in dt that is the full linear system's step."""
    p = _problem(0, patches=[[0, 1, 2], [2, 3, 4]], ids=[2, 3])
    rng = np.random.default_rng(9)
    V0 = -80.0 + 30.0 * rng.standard_normal(p.N)
    for dt in (1e-4, 1e-5):
        got = p.pmj_exchange(V0, dt, vmin=-1e9, vmax=1e9)
        ref = expm(dt * _ode_matrix(p)) @ V0
        lin = (ref - V0).max()
        assert np.abs(got - ref).max() < 0.05 * abs(lin) + 1e-12
    got = p.pmj_exchange(V0, 0.02, vmin=-1e9, vmax=1e9)
    q0 = p.Ai_P * V0[2:4].sum() + p.Ai_M * V0[p.Np:p.Np + 5].sum()
    q1 = p.Ai_P * got[2:4].sum() + p.Ai_M * got[p.Np:p.Np + 5].sum()
    assert abs(q1 - q0) < 1e-12 * abs(q0)


def test_exchange_clamps_like_the_step():
    p = _problem(1)
    V0 = np.full(p.N, -80.0)
    V0[3] = 500.0  # far outside the clamp range
    got = p.pmj_exchange(V0, 0.02)
    assert got.max() <= 80.0 and got.min() >= -150.0
