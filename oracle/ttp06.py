"""CPU FP64 restatement of the reference TTP06 reaction step (test infrastructure).

This module is part of the ORACLE: it is imported only by tests, by
``__graft_entry__.smoke()`` and by ``bench.py``'s CPU-baseline leg, always as
the checker and never as the product path.

Reference restated: ``hrdayapy/simulation/functions/ionic_models/gpu_ttp06.py``
  * constants                        gpu_ttp06.py:95-147
  * initial state / column order     gpu_ttp06.py:150-176
  * _currents_and_gates              gpu_ttp06.py:198-412
  * phenotype conductances           ttp06_phenotype_constants.py:54-61
The rate equations are the published ten Tusscher & Panfilov (2006) model, so
every expression has to be the same as the reference's; the code is written
from the model equations in this project's own form (torch, any dtype).

Parity pins (tests/test_oracle_ttp06.py):
  * golden per-step vectors written by importing the reference module in the
    development container (tests/golden/make_ttp06_vectors.py) -- per-step,
    formula-level parity for all three phenotypes;
  * Rush-Larsen closed form for frozen V (gates relax exactly to x_inf);
  * gates stay in [0, 1]; rest state is near equilibrium;
  * single-cell AP metrics against the TTP06 paper (RMP, overshoot, APD);
  * phenotype APD ordering mid > endo > epi.
"""
from __future__ import annotations

import math
from typing import NamedTuple

import torch

# Physical constants (gpu_ttp06.py:95-100)
R_GAS = 8314.472
T_K = 310.0
F_C = 96485.3415
RTF = R_GAS * T_K / F_C
FRT = F_C / (R_GAS * T_K)

# Extracellular concentrations (gpu_ttp06.py:102-105)
K_O, NA_O, CA_O = 5.4, 140.0, 2.0

# Shared conductances / pump parameters (gpu_ttp06.py:107-124)
G_NA, G_K1, G_BNA, G_BCA, G_PCA, G_PK = 14.838, 5.405, 0.00029, 0.000592, 0.1238, 0.0146
P_NAK, K_MK, K_MNA = 2.724, 1.0, 40.0
K_NACA, K_SAT, ALPHA_NACA, GAMMA_NACA, KM_CA, KM_NAI = 1000.0, 0.1, 2.5, 0.35, 1.38, 87.5
P_KNA, K_PCA = 0.03, 0.0005

# SR / Ca handling (gpu_ttp06.py:126-147)
K1P, K2P, K3, K4 = 0.15, 0.045, 0.06, 0.005
EC, MAX_SR, MIN_SR = 1.5, 2.5, 1.0
V_REL, V_XFER, K_UP, V_LEAK, VMAX_UP = 0.102, 0.0038, 0.00025, 0.00036, 0.006375
BUF_C, K_BUF_C, BUF_SR, K_BUF_SR, BUF_SS, K_BUF_SS = 0.2, 0.001, 10.0, 0.3, 0.4, 0.00025
V_C, V_SR, V_SS = 0.016404, 0.001094, 0.00005468
CM_MODEL = 1.0

# Column order of the (N, 18) state (gpu_ttp06.py:172-176)
STATE_NAMES = ["Ki", "Nai", "Cai", "Xr1", "Xr2", "Xs", "m", "h", "j", "d", "f",
               "f2", "fCass", "s", "r", "CaSR", "Rprime", "Cass"]
N_STATES = len(STATE_NAMES)
COL = {n: i for i, n in enumerate(STATE_NAMES)}

# Initial conditions (gpu_ttp06.py:150-169)
INIT = {"Ki": 138.4, "Nai": 10.355, "Cai": 0.00013, "Xr1": 0.00448, "Xr2": 0.476,
        "Xs": 0.0087, "m": 0.00155, "h": 0.7573, "j": 0.7225, "d": 3.164e-5,
        "f": 0.8009, "f2": 0.9778, "fCass": 0.9953, "s": 0.3212, "r": 2.235e-8,
        "CaSR": 3.715, "Rprime": 0.9068, "Cass": 0.00036}
V_REST = -86.709  # hetero_solver.py:82


class Phenotype(NamedTuple):
    """ttp06_phenotype_constants.py:41-61."""
    g_Ks: float
    g_to: float
    g_Kr: float
    g_CaL: float
    endo_s: bool


PHENOTYPES = {
    "endo": Phenotype(0.392, 0.073, 0.153, 3.98e-5, True),
    "mid": Phenotype(0.098, 0.294, 0.153, 3.98e-5, True),
    "epi": Phenotype(0.392, 0.294, 0.153, 3.98e-5, False),
}
CELLTYPE_CODE = {"endo": 0, "mid": 1, "epi": 2}


def init_state(n: int, dtype=torch.float64, device="cpu") -> torch.Tensor:
    """(n, 18) resting state, gpu_ttp06.py:179-184."""
    row = torch.tensor([INIT[k] for k in STATE_NAMES], dtype=dtype, device=device)
    return row.expand(n, N_STATES).clone()


def currents_and_gates(V, state, dt, g_Ks, g_to, g_Kr=0.153, g_CaL=3.98e-5, endo_s=True):
    """One TTP06 step; returns (I_ion [mV/ms], new_state). gpu_ttp06.py:198-412.

    ``g_Ks``/``g_to`` may be scalars or per-node tensors (the ensemble form
    scales them per instance); ``endo_s`` may be a bool or a per-node bool
    tensor (True selects the endo/M-cell s-gate, gpu_ttp06.py:305-313).
    """
    ex, lg, sq = torch.exp, torch.log, torch.sqrt
    c = [state[:, i] for i in range(N_STATES)]
    Ki, Nai, Cai, Xr1, Xr2, Xs, m, h, j, d, f, f2, fCass, s, r, CaSR, Rp, Cass = c

    # Nernst potentials (:242-247)
    E_Na = RTF * lg(NA_O / Nai)
    E_K = RTF * lg(K_O / Ki)
    E_Ks = RTF * lg((K_O + P_KNA * NA_O) / (Ki + P_KNA * Nai))
    E_Ca = 0.5 * RTF * lg(CA_O / Cai)

    # I_Na gates (:250-276)
    m_inf = (1.0 / (1.0 + ex((-56.86 - V) / 9.03))) ** 2
    a_m = 1.0 / (1.0 + ex((-60.0 - V) / 5.0))
    b_m = 0.1 / (1.0 + ex((V + 35.0) / 5.0)) + 0.1 / (1.0 + ex((V - 50.0) / 200.0))
    tau_m = a_m * b_m
    h_inf = (1.0 / (1.0 + ex((V + 71.55) / 7.43))) ** 2
    low = V < -40.0
    zero = torch.zeros_like(V)
    a_h = torch.where(low, 0.057 * ex(-(V + 80.0) / 6.8), zero)
    b_h = torch.where(low, 2.7 * ex(0.079 * V) + 3.1e5 * ex(0.3485 * V),
                      0.77 / (0.13 * (1.0 + ex((V + 10.66) / -11.1))))
    tau_h = 1.0 / (a_h + b_h + 1e-30)
    j_inf = h_inf
    a_j = torch.where(low, ((-25428.0 * ex(0.2444 * V) - 6.948e-6 * ex(-0.04391 * V)) * (V + 37.78))
                      / (1.0 + ex(0.311 * (V + 79.23))), zero)
    b_j = torch.where(low, 0.02424 * ex(-0.01052 * V) / (1.0 + ex(-0.1378 * (V + 40.14))),
                      0.6 * ex(0.057 * V) / (1.0 + ex(-0.1 * (V + 32.0))))
    tau_j = 1.0 / (a_j + b_j + 1e-30)
    i_Na = G_NA * m ** 3 * h * j * (V - E_Na)

    # I_Kr (:279-287)
    xr1_inf = 1.0 / (1.0 + ex((-26.0 - V) / 7.0))
    tau_xr1 = 450.0 / (1.0 + ex((-45.0 - V) / 10.0)) * 6.0 / (1.0 + ex((V + 30.0) / 11.5))
    xr2_inf = 1.0 / (1.0 + ex((V + 88.0) / 24.0))
    tau_xr2 = 3.0 / (1.0 + ex((-60.0 - V) / 20.0)) * 1.12 / (1.0 + ex((V - 60.0) / 20.0))
    sqrt_ko = math.sqrt(K_O / 5.4)
    i_Kr = g_Kr * sqrt_ko * Xr1 * Xr2 * (V - E_K)

    # I_Ks (:290-294)
    xs_inf = 1.0 / (1.0 + ex((-5.0 - V) / 14.0))
    a_xs = 1400.0 / sq(1.0 + ex((5.0 - V) / 6.0))
    b_xs = 1.0 / (1.0 + ex((V - 35.0) / 15.0))
    tau_xs = a_xs * b_xs + 80.0
    i_Ks = g_Ks * Xs ** 2 * (V - E_Ks)

    # I_K1 (:297-302)
    vek = V - E_K
    a_k1 = 0.1 / (1.0 + ex(0.06 * (vek - 200.0)))
    b_k1 = (3.0 * ex(2e-4 * (vek + 100.0)) + ex(0.1 * (vek - 10.0))) / (1.0 + ex(-0.5 * vek))
    xk1_inf = a_k1 / (a_k1 + b_k1)
    i_K1 = G_K1 * sqrt_ko * xk1_inf * vek

    # I_to with phenotype s-gate (:305-317)
    s_inf_endo = 1.0 / (1.0 + ex((V + 28.0) / 5.0))
    tau_s_endo = 1000.0 * ex(-(V + 67.0) ** 2 / 1000.0) + 8.0
    s_inf_epi = 1.0 / (1.0 + ex((V + 20.0) / 5.0))
    tau_s_epi = 85.0 * ex(-(V + 45.0) ** 2 / 320.0) + 5.0 / (1.0 + ex((V - 20.0) / 5.0)) + 3.0
    if isinstance(endo_s, torch.Tensor):
        s_inf = torch.where(endo_s, s_inf_endo, s_inf_epi)
        tau_s = torch.where(endo_s, tau_s_endo, tau_s_epi)
    elif endo_s:
        s_inf, tau_s = s_inf_endo, tau_s_endo
    else:
        s_inf, tau_s = s_inf_epi, tau_s_epi
    r_inf = 1.0 / (1.0 + ex((20.0 - V) / 6.0))
    tau_r = 9.5 * ex(-(V + 40.0) ** 2 / 1800.0) + 0.8
    i_to = g_to * r * s * vek

    # I_CaL (:320-340)
    d_inf = 1.0 / (1.0 + ex((-8.0 - V) / 7.5))
    tau_d = ((1.4 / (1.0 + ex((-35.0 - V) / 13.0)) + 0.25) * 1.4 / (1.0 + ex((V + 5.0) / 5.0))
             + 1.0 / (1.0 + ex((50.0 - V) / 20.0)))
    f_inf = 1.0 / (1.0 + ex((V + 20.0) / 7.0))
    tau_f = (1102.5 * ex(-((V + 27.0) ** 2) / 225.0) + 200.0 / (1.0 + ex((13.0 - V) / 10.0))
             + 180.0 / (1.0 + ex((V + 30.0) / 10.0)) + 20.0)
    f2_inf = 0.67 / (1.0 + ex((V + 35.0) / 7.0)) + 0.33
    tau_f2 = (562.0 * ex(-((V + 27.0) ** 2) / 240.0) + 31.0 / (1.0 + ex((25.0 - V) / 10.0))
              + 80.0 / (1.0 + ex((V + 30.0) / 10.0)))
    cs2 = (Cass / 0.05) ** 2
    fCass_inf = 0.6 / (1.0 + cs2) + 0.4
    tau_fCass = 80.0 / (1.0 + cs2) + 2.0
    e2v = ex(2.0 * (V - 15.0) * FRT)
    den = e2v - 1.0
    den = torch.where(den.abs() < 1e-10, torch.full_like(den, 1e-10), den)
    i_CaL = (g_CaL * d * f * f2 * fCass * 4.0 * (V - 15.0) * F_C ** 2 / (R_GAS * T_K)
             * (0.25 * Cass * e2v - CA_O) / den)

    # background, pumps, exchanger (:343-357)
    i_bNa = G_BNA * (V - E_Na)
    i_bCa = G_BCA * (V - E_Ca)
    i_pCa = G_PCA * Cai / (Cai + K_PCA)
    i_pK = G_PK * vek / (1.0 + ex((25.0 - V) / 5.98))
    i_NaK = (P_NAK * K_O / (K_O + K_MK) * Nai / (Nai + K_MNA)
             / (1.0 + 0.1245 * ex(-0.1 * V * FRT) + 0.0353 * ex(-V * FRT)))
    e_g = ex(GAMMA_NACA * V * FRT)
    e_gm1 = ex((GAMMA_NACA - 1.0) * V * FRT)
    i_NaCa = (K_NACA * (e_g * Nai ** 3 * CA_O - e_gm1 * NA_O ** 3 * Cai * ALPHA_NACA)
              / ((KM_NAI ** 3 + NA_O ** 3) * (KM_CA + CA_O) * (1.0 + K_SAT * e_gm1)))

    # total current from the OLD gates (:360-361)
    I_ion = (i_K1 + i_to + i_Kr + i_Ks + i_CaL + i_NaK + i_Na + i_bNa + i_NaCa + i_bCa
             + i_pCa + i_pK)

    # SR / subspace (:364-375)
    kcasr = MAX_SR - (MAX_SR - MIN_SR) / (1.0 + (EC / CaSR) ** 2)
    k1 = K1P / kcasr
    k2 = K2P * kcasr
    O = k1 * Cass ** 2 * Rp / (K3 + k1 * Cass ** 2 + 1e-30)
    i_rel = V_REL * O * (CaSR - Cass)
    i_up = VMAX_UP / (1.0 + K_UP ** 2 / (Cai ** 2 + 1e-30))
    i_leak = V_LEAK * (CaSR - Cai)
    i_xfer = V_XFER * (Cass - Cai)
    bufc = 1.0 / (1.0 + BUF_C * K_BUF_C / (Cai + K_BUF_C) ** 2)
    bufsr = 1.0 / (1.0 + BUF_SR * K_BUF_SR / (CaSR + K_BUF_SR) ** 2)
    bufss = 1.0 / (1.0 + BUF_SS * K_BUF_SS / (Cass + K_BUF_SS) ** 2)

    # concentration derivatives (:378-384); dKi excludes I_stim (as the reference)
    dNai = -(i_Na + i_bNa + 3.0 * i_NaK + 3.0 * i_NaCa) * CM_MODEL / (V_C * F_C)
    dKi = -(i_K1 + i_to + i_Kr + i_Ks + i_pK - 2.0 * i_NaK) * CM_MODEL / (V_C * F_C)
    dCai = bufc * ((i_leak - i_up) * V_SR / V_C + i_xfer
                   - (i_bCa + i_pCa - 2.0 * i_NaCa) * CM_MODEL / (2.0 * V_C * F_C))
    dCaSR = bufsr * (i_up - i_rel - i_leak)
    dCass = bufss * (-i_CaL * CM_MODEL / (2.0 * V_SS * F_C) + i_rel * V_SR / V_SS
                     - i_xfer * V_C / V_SS)

    def rl(x, xi, tau):  # Rush-Larsen, gpu_ttp06.py:146-148
        return xi - (xi - x) * ex(-dt / tau)

    out = torch.empty_like(state)
    out[:, COL["Ki"]] = (Ki + dt * dKi).clamp(min=1e-3)
    out[:, COL["Nai"]] = (Nai + dt * dNai).clamp(min=1e-3)
    out[:, COL["Cai"]] = (Cai + dt * dCai).clamp(min=1e-7)
    out[:, COL["Xr1"]] = rl(Xr1, xr1_inf, tau_xr1)
    out[:, COL["Xr2"]] = rl(Xr2, xr2_inf, tau_xr2)
    out[:, COL["Xs"]] = rl(Xs, xs_inf, tau_xs)
    out[:, COL["m"]] = rl(m, m_inf, tau_m)
    out[:, COL["h"]] = rl(h, h_inf, tau_h)
    out[:, COL["j"]] = rl(j, j_inf, tau_j)
    out[:, COL["d"]] = rl(d, d_inf, tau_d)
    out[:, COL["f"]] = rl(f, f_inf, tau_f)
    out[:, COL["f2"]] = rl(f2, f2_inf, tau_f2)
    out[:, COL["fCass"]] = rl(fCass, fCass_inf, tau_fCass)
    out[:, COL["s"]] = rl(s, s_inf, tau_s)
    out[:, COL["r"]] = rl(r, r_inf, tau_r)
    out[:, COL["CaSR"]] = (CaSR + dt * dCaSR).clamp(min=1e-7)
    out[:, COL["Rprime"]] = Rp + dt * (-k2 * Cass * Rp + K4 * (1.0 - Rp))
    out[:, COL["Cass"]] = (Cass + dt * dCass).clamp(min=1e-7)
    return I_ion, out


def step_celltype(V, state, dt, celltype, g_ks_scale=None, g_to_scale=None):
    """Step nodes of mixed phenotype; ``celltype`` is 0=endo, 1=mid, 2=epi per node.

    Mirrors HeteroIonicSolver.step (hetero_solver.py:141-194): each node uses its
    phenotype's conductances; per-node scale factors model the ensemble form.
    """
    ct = celltype.long()
    gks_tab = torch.tensor([PHENOTYPES[k].g_Ks for k in ("endo", "mid", "epi")], dtype=V.dtype)
    gto_tab = torch.tensor([PHENOTYPES[k].g_to for k in ("endo", "mid", "epi")], dtype=V.dtype)
    g_Ks = gks_tab[ct]
    g_to = gto_tab[ct]
    if g_ks_scale is not None:
        g_Ks = g_Ks * g_ks_scale
    if g_to_scale is not None:
        g_to = g_to * g_to_scale
    endo_s = ct != 2
    return currents_and_gates(V, state, dt, g_Ks, g_to, 0.153, 3.98e-5, endo_s)


def pace_single_cell(phenotype="epi", dt=0.02, bcl=1000.0, stim_amp=52.0, stim_dur=1.0,
                     n_beats=1, dtype=torch.float64):
    """Single-cell pacing, experiments/experiment_1/run_experiment1_single.py:128-173.

    Each step: I_ion from the model, then V += dt * (I_stim - I_ion) with the
    stimulus in pA/pF for the first round(stim_dur/dt) steps of a beat.
    Returns (t, V) of the last beat, V recorded after each update.
    """
    import numpy as np
    p = PHENOTYPES[phenotype]
    V = torch.tensor([V_REST], dtype=dtype)
    st = init_state(1, dtype)
    n_steps = int(round(bcl / dt))
    stim_steps = int(round(stim_dur / dt))
    Vs = np.empty(n_steps)
    for _beat in range(n_beats):
        for k in range(n_steps):
            i_st = stim_amp if k < stim_steps else 0.0
            I, st = currents_and_gates(V, st, dt, p.g_Ks, p.g_to, p.g_Kr, p.g_CaL, p.endo_s)
            V = V + dt * (i_st - I)
            Vs[k] = float(V[0])
    return np.arange(n_steps) * dt, Vs


def ap_metrics(t, V, plateau_offset_ms=100.0):
    """APD90 / upstroke metrics, run_experiment1_single.py:74-101."""
    import numpy as np
    dVdt = np.diff(V) / np.diff(t)
    i_up = int(np.argmax(dVdt))
    t_up = t[i_up]
    rmp = float(V[0])
    peak = float(V.max())
    i_peak = int(V.argmax())
    target = rmp + 0.1 * (peak - rmp)
    below = np.nonzero(V[i_peak:] <= target)[0]
    apd90 = float(t[i_peak + below[0]] - t_up) if below.size else float(t[-1] - t_up)
    ip = int(np.searchsorted(t, t_up + plateau_offset_ms))
    plateau = float(V[ip]) if ip < len(V) else float("nan")
    return dict(RMP_mV=rmp, overshoot_mV=peak, plateau_mV=plateau, APD90_ms=apd90,
                dVdt_max_Vs=float(dVdt[i_up]))
