"""The kernels' cell algebra run on the CPU (hp_ttp06_step_host_*) against the
golden vectors of the reference's TTP06 and the oracle.

The FP32 kernels use an algebraic form with merged reciprocals (Nernst
potentials as differences of logs, Rush-Larsen exponents from the rates);
these tests check that form without a GPU, so that a wrong constant or a
wrongly merged fraction fails here before any kernel runs.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

from hrdayapy_amd import _native as nat
from oracle import ttp06 as O

HERE = os.path.dirname(os.path.abspath(__file__))
KB, NB = 138.4, 10.355


def _gold():
    return np.load(os.path.join(HERE, "golden", "ttp06_golden.npz"))


def _host_step(V, st, dt, ct, dtype, kb=0.0, nb=0.0):
    """V (n,), st (n,18) absolute state; returns (I, new state (n,18) absolute)."""
    npdt = np.float32 if dtype == "f32" else np.float64
    V = np.ascontiguousarray(V, dtype=npdt)
    S = np.array(st, dtype=np.float64)
    S[:, 0] -= kb
    S[:, 1] -= nb
    S = np.ascontiguousarray(S.astype(npdt))
    I = np.empty(len(V), dtype=npdt)
    ctt = np.ascontiguousarray(ct, dtype=np.uint8)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    nat.call(f"hp_ttp06_step_host_{dtype}", p(V), p(S), p(I), p(ctt), 0, None, None, len(V), 1,
             18, 0.02 if dt is None else dt, kb, nb)
    S = S.astype(np.float64)
    S[:, 0] += kb
    S[:, 1] += nb
    return I.astype(np.float64), S


def test_host_f64_matches_golden():
    g = _gold()
    n = len(g["V"])
    for code, ph in enumerate(["endo", "mid", "epi"]):
        I, S = _host_step(g["V"], g["state"], 0.02, np.full(n, code), "f64")
        np.testing.assert_allclose(I, g[f"I_{ph}"], rtol=1e-10, atol=1e-11)
        np.testing.assert_allclose(S, g[f"S_{ph}"], rtol=1e-10, atol=1e-14)


def test_host_f32_form_matches_oracle():
    """FP32 merged-reciprocal form vs the FP64 oracle: every current term
    enters I_ion, so a wrong Nernst constant or fraction shows up here."""
    g = _gold()
    n = len(g["V"])
    for code in range(3):
        ct = np.full(n, code)
        Io, So = O.step_celltype(torch.from_numpy(g["V"]), torch.from_numpy(g["state"]), 0.02,
                                 torch.from_numpy(ct))
        I, S = _host_step(g["V"], g["state"], 0.02, ct, "f32", KB, NB)
        sing = np.abs(g["V"] - 15.0) < 1e-6  # I_CaL removable singularity, checked below
        err = np.abs(I - Io.numpy()) / np.maximum(np.abs(Io.numpy()), 1.0)
        assert np.max(err[~sing]) < 5e-5, (code, np.argmax(np.where(sing, 0, err)))
        np.testing.assert_allclose(S[~sing], So.numpy()[~sing], rtol=2e-4, atol=2e-7)
        # at V = 15 the FP32 form takes the limit x/expm1(x) -> 1: compare with the
        # reference form evaluated 1e-4 mV away, where its guard is inactive
        Ip, _ = O.step_celltype(torch.from_numpy(g["V"][sing] + 1e-4),
                                torch.from_numpy(g["state"][sing]), 0.02,
                                torch.from_numpy(ct[sing]))
        np.testing.assert_allclose(I[sing], Ip.numpy(), rtol=2e-4, atol=1e-3)


def test_host_f32_form_detects_a_planted_error():
    """The comparison above has the power to see a small change of one
    conductance: scaling the oracle's g_Ks by 1.01 must break the 5e-5 bar."""
    g = _gold()
    n = len(g["V"])
    ct = np.full(n, 2)
    V = torch.from_numpy(g["V"])
    st = torch.from_numpy(g["state"])
    Io, _ = O.currents_and_gates(V, st, 0.02, g_Ks=0.392 * 1.01, g_to=0.294, g_Kr=0.153,
                                 g_CaL=3.98e-5, endo_s=False)
    I, _ = _host_step(g["V"], g["state"], 0.02, ct, "f32", KB, NB)
    err = np.abs(I - Io.numpy()) / np.maximum(np.abs(Io.numpy()), 1.0)
    assert np.max(err) > 5e-5


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-6), ("f32", 0.5)])
def test_host_single_cell_ap_trace(dtype, tol):
    """Experiment-1 protocol (52 uA/uF for 1 ms, dt 0.02 ms) vs the reference's
    epicardial trace, sampled every 0.2 ms."""
    g = _gold()
    ref = g["trace_epi"]
    npdt = np.float32 if dtype == "f32" else np.float64
    kb, nb = (KB, NB) if dtype == "f32" else (0.0, 0.0)
    st = np.array([O.INIT[k] for k in O.STATE_NAMES], dtype=np.float64)
    st[0] -= kb
    st[1] -= nb
    S = np.ascontiguousarray(st.astype(npdt))
    V = np.full(1, O.V_REST, dtype=npdt)
    I = np.empty(1, dtype=npdt)
    ct = np.full(1, 2, dtype=np.uint8)
    out = np.empty(len(ref))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    fn = getattr(nat.load(), f"hp_ttp06_step_host_{dtype}")
    dt = npdt(0.02)
    for k in range(len(ref) * 10):
        rc = fn(p(V), p(S), p(I), p(ct), 0, None, None, 1, 1, 1, 0.02, kb, nb)
        assert rc == 0
        V[0] = V[0] + dt * (npdt(52.0 if k < 50 else 0.0) - I[0])
        if k % 10 == 0:
            out[k // 10] = V[0]
    rms = float(np.sqrt(np.mean((out - ref) ** 2)))
    print(f"host {dtype}: Vm RMS vs reference trace = {rms:.3e} mV")
    assert rms < tol
