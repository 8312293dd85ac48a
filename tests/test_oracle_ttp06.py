"""Pins the CPU oracle of the TTP06 step (oracle/ttp06.py)."""
import numpy as np
import pytest
import torch

from oracle import ttp06 as O

def _gold():
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    return np.load(os.path.join(here, "golden", "ttp06_golden.npz"))


@pytest.mark.parametrize("ph", ["endo", "mid", "epi"])
def test_oracle_matches_reference_vectors(ph):
    g = _gold()
    p = O.PHENOTYPES[ph]
    I, S = O.currents_and_gates(torch.from_numpy(g["V"]), torch.from_numpy(g["state"]),
                                float(g["dt"]), p.g_Ks, p.g_to, p.g_Kr, p.g_CaL, p.endo_s)
    np.testing.assert_allclose(I.numpy(), g[f"I_{ph}"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(S.numpy(), g[f"S_{ph}"], rtol=1e-13, atol=0)


def test_celltype_dispatch_matches_phenotype_classes():
    g = _gold()
    n = len(g["V"])
    ct = torch.from_numpy(np.arange(n) % 3)
    I, S = O.step_celltype(torch.from_numpy(g["V"]), torch.from_numpy(g["state"]), 0.02, ct)
    for code, ph in enumerate(["endo", "mid", "epi"]):
        sel = (np.arange(n) % 3) == code
        np.testing.assert_allclose(I.numpy()[sel], g[f"I_{ph}"][sel], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(S.numpy()[sel], g[f"S_{ph}"][sel], rtol=1e-13)


def test_rush_larsen_closed_form_frozen_voltage():
    # with V frozen, n Rush-Larsen steps of dt equal one step of n*dt for the
    # gates whose x_inf/tau depend on V only (m, h, j, Xr1, Xr2, Xs, d, f, f2, s, r)
    V = torch.tensor([-20.0, 10.0, -60.0], dtype=torch.float64)
    st = O.init_state(3)
    s1 = st.clone()
    for _ in range(10):
        _, s1 = O.currents_and_gates(V, s1, 0.01, 0.392, 0.073)
    _, s2 = O.currents_and_gates(V, st.clone(), 0.1, 0.392, 0.073)
    gates = [O.COL[k] for k in ("m", "h", "j", "Xr1", "Xr2", "Xs", "d", "f", "f2", "s", "r")]
    np.testing.assert_allclose(s1[:, gates].numpy(), s2[:, gates].numpy(), rtol=1e-12, atol=1e-15)


def test_gates_bounded_and_rest_near_equilibrium():
    g = _gold()
    _, S = O.currents_and_gates(torch.from_numpy(g["V"]), torch.from_numpy(g["state"]), 0.02,
                                0.392, 0.294, endo_s=False)
    gates = S[:, 3:15].numpy()
    assert gates.min() >= 0.0 and gates.max() <= 1.0
    V = torch.tensor([O.V_REST], dtype=torch.float64)
    I, _ = O.currents_and_gates(V, O.init_state(1), 0.02, 0.392, 0.073)
    assert abs(float(I[0])) < 0.5  # |dV/dt| < 0.5 mV/ms at the CellML rest state


def test_trace_matches_reference_single_cell_ap():
    g = _gold()
    tr = g["trace_epi"]
    # literature sanity of the reference's own first-beat AP (TTP06 epi, 1 Hz protocol)
    assert -87.5 < tr[-1] < -84.0
    assert 25.0 < tr.max() < 45.0
    m = O.ap_metrics(np.arange(len(tr)) * 0.2, tr)
    assert 240.0 < m["APD90_ms"] < 320.0


def test_pace_single_cell_matches_reference_trace():
    """The oracle's single-cell pacing loop (experiment 1's protocol) against
    the reference's own trace, first 40 ms (upstroke, notch, early plateau)."""
    g = _gold()
    t, V = O.pace_single_cell("epi", dt=0.02, bcl=40.0)
    ref = g["trace_epi"][:len(V[::10])]
    np.testing.assert_allclose(V[::10], ref, rtol=0, atol=1e-9)
    assert ref.max() > 20.0
