"""The learned correction's training targets on the CPU: what ten exact
0.02 ms substeps add to one 0.2 ms big step, in the network's output space,
and the update rule the inference kernel applies to it (V + d; state + d, or
state * exp(d) for the log-scaled variables).  The cell steps run through the
library's host entry points (hp_ttp06_step_host_*), the same algebra as the
device kernels."""
import ctypes

import numpy as np
import torch

from hrdayapy_amd import _native as nat
from hrdayapy_amd import surrogate as SG


def _host_step(V, S, ct, dt, dtype):
    """One cell step; V (n,), S (n,18) absolute; returns (I, new absolute S)."""
    npdt = np.float32 if dtype == "f32" else np.float64
    kb, nb = (SG.KI_BASE, SG.NAI_BASE) if dtype == "f32" else (0.0, 0.0)
    Vc = np.ascontiguousarray(V, dtype=npdt)
    St = np.array(S, dtype=np.float64)
    St[:, 0] -= kb
    St[:, 1] -= nb
    St = np.ascontiguousarray(St.astype(npdt))
    I = np.empty(len(Vc), dtype=npdt)
    ctt = np.ascontiguousarray(ct, dtype=np.uint8)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    nat.call(f"hp_ttp06_step_host_{dtype}", p(Vc), p(St), p(I), p(ctt), 0, None, None, len(Vc), 1,
             18, dt, kb, nb)
    St = St.astype(np.float64)
    St[:, 0] += kb
    St[:, 1] += nb
    return I.astype(np.float64), St


def _states():
    """States on paced action potentials (two cells per phenotype, stimulated
    at 0 and 3 ms, 400 ms at 0.02 ms), sampled every 0.2 ms where the
    multirate rule would not fall back (|dV/dt| <= 1 mV/ms over the last
    macro-step, V outside (-70, -40) mV)."""
    n = 6
    ct = np.repeat(np.arange(3), 2)
    onset = np.tile([0.0, 3.0], 3)
    V = np.full(n, -86.2)
    S = np.tile(np.array([SG.INIT[k] for k in SG.STATE_NAMES]), (n, 1))
    dt = SG.T_MACRO / SG.N_SUB
    Vs, Ss, Cs, V_prev = [], [], [], None
    for k in range(int(round(400.0 / dt))):
        if k % SG.N_SUB == 0:
            if V_prev is not None:
                keep = (np.abs(V - V_prev) / SG.T_MACRO <= 1.0) & ~((V > -70) & (V < -40))
                Vs.append(V[keep])
                Ss.append(S[keep])
                Cs.append(ct[keep])
            V_prev = V.copy()
        I, S = _host_step(V, S, ct, dt, "f64")
        t = k * dt
        V = V + dt * (np.where((t >= onset) & (t < onset + 1.0), 52.0, 0.0) - I)
    return np.concatenate(Vs), np.concatenate(Ss), np.concatenate(Cs)


def test_correction_target_and_update_rule_recover_the_exact_substeps():
    V, S, ct = _states()
    # ten exact substeps (FP64) and the FP32 big step
    Ve, Se = V.copy(), S.copy()
    for _ in range(SG.N_SUB):
        I, Se = _host_step(Ve, Se, ct, SG.T_MACRO / SG.N_SUB, "f64")
        Ve = Ve - SG.T_MACRO / SG.N_SUB * I
    Ib, Sb = _host_step(V, S, ct, SG.T_MACRO, "f32")
    Vb = V - SG.T_MACRO * Ib
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    D = SG.output_increments(t(Vb), t(Sb), t(Ve), t(Se)).double().numpy()
    # the kernel's update applied to the big step's result gives the substeps
    lg = SG.IN_LOG[1:19].astype(bool)
    V1 = Vb + D[:, 0]
    S1 = np.where(lg[None, :], Sb * np.exp(D[:, 1:]), Sb + D[:, 1:])
    np.testing.assert_allclose(V1, Ve, rtol=0, atol=1e-4)
    np.testing.assert_allclose(S1, Se, rtol=1e-5, atol=1e-9)
    # the correction is much smaller than the whole increment it completes
    inc = np.abs(Ve - V)
    assert len(V) > 5000
    assert np.mean(np.abs(D[:, 0])) < 0.05 * np.mean(inc)


def test_a_planted_base_mistake_fails():
    """The target taken relative to the state at the start of the macro-step
    (the whole increment) instead of the big step's result does not recover
    the substeps through the correction update."""
    V, S, ct = _states()
    Ve, Se = V.copy(), S.copy()
    for _ in range(SG.N_SUB):
        I, Se = _host_step(Ve, Se, ct, SG.T_MACRO / SG.N_SUB, "f64")
        Ve = Ve - SG.T_MACRO / SG.N_SUB * I
    Ib, Sb = _host_step(V, S, ct, SG.T_MACRO, "f32")
    Vb = V - SG.T_MACRO * Ib
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    D = SG.output_increments(t(V), t(S), t(Ve), t(Se)).double().numpy()
    assert np.max(np.abs(Vb + D[:, 0] - Ve)) > 1e-3
