// One TTP06 (ten Tusscher & Panfilov 2006) step for one cell, in registers.
//
// Model and step semantics follow the reference's _currents_and_gates
// (hrdayapy/simulation/functions/ionic_models/gpu_ttp06.py:198-412):
//   * I_ion is evaluated from the OLD gates and concentrations (:360-361);
//   * 12 gates advance by Rush-Larsen with V^n (:387-399);
//   * R' and the five concentrations advance by forward Euler, with the
//     clamps Nai,Ki >= 1e-3 and Cai,CaSR,Cass >= 1e-7 (:402-410).
// The rate expressions are the published model equations and are therefore
// the same as the reference's.
//
// Two algebraic forms of the same equations:
//   kRef = true  (FP64 kernels): the reference's own arrangement, so that the
//                FP64 GPU path meets the FP64 oracle at round-off level;
//   kRef = false (FP32 kernels): reciprocals merged where the model allows it
//                -- Nernst potentials as differences of logs, Rush-Larsen
//                exponents -dt/tau computed directly from the rate
//                (h, j: -dt*(alpha+beta); Xr1, Xr2, m: products of the
//                sigmoid denominators), buffer factors q^2/(q^2+c), one
//                reciprocal for x_K1,inf, R' opening, I_up, I_NaK -- which
//                removes about a third of the v_rcp_f32 instructions -- and
//                exponentials of a common slope in V shared: exp(a(V+b)) =
//                e^{ab} exp(aV) for the slopes 0.1, 0.2 (= 0.1 squared), 0.05,
//                1/7, -1/6 and F/RT, which replaces 24 exponentials and an
//                expm1 by 11 hardware exp/rcp.  The FP32 path uses the hardware
//                exp/log/rcp (1-2 ulp).
//
// Storage of Ki and Nai: the state holds Ki - ki_base and Nai - nai_base.
// With base 0 this is the reference's state.  The FP32 engine passes the
// resting values (138.4, 10.355) so that the per-step increments (~1e-5 mM,
// about one FP32 ulp of 138) are accumulated at full relative precision.
//
// FP32 I_CaL: the factor (V-15)/(exp(2(V-15)F/RT)-1) is evaluated as
// x/(e^x - 1) with its series at |x| < 0.01, instead of the reference's guard
// |den| < 1e-10 (below FP32 resolution).  FP64 keeps the reference form.
#pragma once
#include "common.h"

namespace hp {

enum StateCol {
  S_KI = 0, S_NAI, S_CAI, S_XR1, S_XR2, S_XS, S_M, S_H, S_J, S_D, S_F, S_F2,
  S_FCASS, S_S, S_R, S_CASR, S_RPRIME, S_CASS, N_STATE
};

// Phenotype conductances, ttp06_phenotype_constants.py:54-61 (endo, mid, epi).
__device__ __constant__ static const double kGKs[3] = {0.392, 0.098, 0.392};
__device__ __constant__ static const double kGto[3] = {0.073, 0.294, 0.294};

// Host copies of the two tables, for the host entry points.
static const double kGKsHost[3] = {0.392, 0.098, 0.392};
static const double kGtoHost[3] = {0.073, 0.294, 0.294};

template <typename T, bool kRef>
HP_HD T ttp06_cell(T V, T (&s)[N_STATE], T dt, T g_Ks, T g_to,
                                         bool endo_s, T ki_base, T nai_base) {
  typedef M<T> m_;
#define C(x) T(x)
  const T R = C(8314.472), TK = C(310.0), F = C(96485.3415);
  const T RTF = C(8314.472 * 310.0 / 96485.3415);
  const T FRT = C(96485.3415 / (8314.472 * 310.0));
  const T Ko = C(5.4), Nao = C(140.0), Cao = C(2.0);
  const T g_Kr = C(0.153), g_CaL = C(3.98e-5);

  const T Ki = ki_base + s[S_KI];
  const T Nai = nai_base + s[S_NAI];
  const T Cai = s[S_CAI], Xr1 = s[S_XR1], Xr2 = s[S_XR2], Xs = s[S_XS];
  const T mg = s[S_M], hg = s[S_H], jg = s[S_J], d = s[S_D], f = s[S_F], f2 = s[S_F2];
  const T fCass = s[S_FCASS], sg = s[S_S], rg = s[S_R], CaSR = s[S_CASR], Rp = s[S_RPRIME];
  const T Cass = s[S_CASS];
  const T one = C(1.0);
  // The FP32 form evaluates every exponential as exp2 of an argument already
  // scaled by log2(e): exp(a*V + b) = exp2(V*(a log2 e) + b log2 e), one FMA
  // and one v_exp_f32.  L2E(c) is the constant c*log2(e).
#define L2E(c) C((c) * 1.4426950408889634)
  // Rush-Larsen with the exponent supplied: x_inf - (x_inf - x) exp(-dt/tau);
  // the exponent is -dt/tau (reference form) or -dt log2(e)/tau (FP32 form)
#define RLE(x, xi, ndt_over_tau) \
  ((xi) - ((xi) - (x)) * (kRef ? m_::ex(ndt_over_tau) : m_::ex2(ndt_over_tau)))
  const T ndt0 = kRef ? -dt : -dt * L2E(1.0);
  // exp(ref) in the reference arrangement, `fast` (a shared exponential times
  // a constant) in the FP32 form
#define XP(ref, fast) (kRef ? m_::ex(ref) : (fast))
  // exp(ref) = exp(a*u + b) in the reference arrangement / FP32 form
#define LX(ref, u, a, b) (kRef ? m_::ex(ref) : m_::ex2((u) * L2E(a) + L2E(b)))
  // shared exponentials of the FP32 form (V is clamped to [-150, 80], so none
  // of these leaves the FP32 range)
  T Q = one, iQ = one, P = one, iP = one, R5 = one, iR5 = one, S7 = one, iS7 = one, W6 = one,
    E = one, iE = one;
  if (!kRef) {
    Q = m_::ex2(V * L2E(0.1));       // e^{0.1 V}
    iQ = m_::rcp(Q);
    P = Q * Q;                       // e^{0.2 V}
    iP = iQ * iQ;
    R5 = m_::ex2(V * L2E(0.05));     // e^{0.05 V}
    iR5 = m_::rcp(R5);
    S7 = m_::ex2(V * L2E(1.0 / 7.0));  // e^{V/7}
    iS7 = m_::rcp(S7);
    W6 = m_::ex2(V * L2E(-1.0 / 6.0));  // e^{-V/6}
    E = m_::ex2(V * L2E(96485.3415 / (8314.472 * 310.0)));  // e^{V F/RT}
    iE = m_::rcp(E);
  }

  // Nernst potentials (gpu_ttp06.py:242-247)
  T E_Na, E_K, E_Ks, E_Ca;
  if (kRef) {
    E_Na = RTF * m_::lg(m_::dv(Nao, Nai));
    E_K = RTF * m_::lg(m_::dv(Ko, Ki));
    E_Ks = RTF * m_::lg(m_::dv(Ko + C(0.03) * Nao, Ki + C(0.03) * Nai));
    E_Ca = C(0.5) * RTF * m_::lg(m_::dv(Cao, Cai));
  } else {
    E_Na = RTF * (C(4.941642422609304) - m_::lg(Nai));          // log(140)
    E_K = RTF * (C(1.6863989535702288) - m_::lg(Ki));           // log(5.4)
    E_Ks = RTF * (C(2.2617630984737906) - m_::lg(Ki + C(0.03) * Nai));  // log(5.4+0.03*140)
    E_Ca = C(0.5) * RTF * (C(0.6931471805599453) - m_::lg(Cai));  // log(2)
  }

  // I_Na (:250-276)
  const T e_m1 = XP((C(-60.0) - V) * C(0.2), C(6.14421235332821e-06) * iP);   // e^-12
  const T e_m2 = XP((V + C(35.0)) * C(0.2), C(1096.6331584284585) * P);       // e^7
  const T e_m3 = LX((V - C(50.0)) * C(1.0 / 200.0), V, 1.0 / 200.0, -50.0 / 200.0);
  T t1 = m_::rcp(one + LX((C(-56.86) - V) * C(1.0 / 9.03), V, -1.0 / 9.03, -56.86 / 9.03));
  const T m_inf = t1 * t1;
  T ndt_m;  // -dt / tau_m, tau_m = a_m * b_m
  if (kRef) {
    const T a_m = m_::rcp(one + e_m1);
    const T b_m = C(0.1) * m_::rcp(one + e_m2) + C(0.1) * m_::rcp(one + e_m3);
    ndt_m = ndt0 * m_::rcp(a_m * b_m);
  } else {
    // 1/tau_m = (1+e1)(1+e2)(1+e3) / (0.1 (2 + e2 + e3))
    ndt_m = ndt0 * (one + e_m1) * (one + e_m2) * (one + e_m3) *
            m_::rcp(C(0.1) * (C(2.0) + e_m2 + e_m3));
  }
  t1 = m_::rcp(one + LX((V + C(71.55)) * C(1.0 / 7.43), V, 1.0 / 7.43, 71.55 / 7.43));
  const T h_inf = t1 * t1;
  T a_h, b_h, a_j, b_j;
  if (V < C(-40.0)) {
    a_h = C(0.057) * LX(-(V + C(80.0)) * C(1.0 / 6.8), V, -1.0 / 6.8, -80.0 / 6.8);
    b_h = C(2.7) * LX(C(0.079) * V, V, 0.079, 0.0) + C(3.1e5) * LX(C(0.3485) * V, V, 0.3485, 0.0);
    a_j = m_::dv((C(-25428.0) * LX(C(0.2444) * V, V, 0.2444, 0.0) -
                  C(6.948e-6) * LX(C(-0.04391) * V, V, -0.04391, 0.0)) *
                     (V + C(37.78)),
                 one + LX(C(0.311) * (V + C(79.23)), V, 0.311, 0.311 * 79.23));
    b_j = m_::dv(C(0.02424) * LX(C(-0.01052) * V, V, -0.01052, 0.0),
                 one + LX(C(-0.1378) * (V + C(40.14)), V, -0.1378, -0.1378 * 40.14));
  } else {
    a_h = C(0.0);
    b_h = m_::dv(C(0.77), C(0.13) * (one + LX((V + C(10.66)) * C(-1.0 / 11.1), V, -1.0 / 11.1,
                                              -10.66 / 11.1)));
    a_j = C(0.0);
    b_j = m_::dv(C(0.6) * LX(C(0.057) * V, V, 0.057, 0.0),
                 one + XP(C(-0.1) * (V + C(32.0)), C(0.04076220397836621) * iQ));  // e^-3.2
  }
  T ndt_h, ndt_j;
  if (kRef) {
    ndt_h = ndt0 * m_::rcp(m_::rcp(a_h + b_h + C(1e-30)));
    ndt_j = ndt0 * m_::rcp(m_::rcp(a_j + b_j + C(1e-30)));
  } else {
    ndt_h = ndt0 * (a_h + b_h + C(1e-30));
    ndt_j = ndt0 * (a_j + b_j + C(1e-30));
  }
  const T i_Na = C(14.838) * mg * mg * mg * hg * jg * (V - E_Na);

  // I_Kr (:279-287); sqrt(Ko/5.4) = 1
  const T xr1_inf = m_::rcp(one + XP((C(-26.0) - V) * C(1.0 / 7.0),
                                     C(0.0243728440732796) * iS7));  // e^(-26/7)
  const T e_r1 = XP((C(-45.0) - V) * C(0.1), C(0.011108996538242306) * iQ);  // e^-4.5
  const T e_r2 = LX((V + C(30.0)) * C(1.0 / 11.5), V, 1.0 / 11.5, 30.0 / 11.5);
  const T xr2_inf = m_::rcp(one + LX((V + C(88.0)) * C(1.0 / 24.0), V, 1.0 / 24.0, 88.0 / 24.0));
  const T e_r3 = XP((C(-60.0) - V) * C(0.05), C(0.049787068367863944) * iR5);  // e^-3
  const T e_r4 = XP((V - C(60.0)) * C(0.05), C(0.049787068367863944) * R5);
  T ndt_xr1, ndt_xr2;
  if (kRef) {
    const T tau_xr1 = C(450.0) * m_::rcp(one + e_r1) * C(6.0) * m_::rcp(one + e_r2);
    const T tau_xr2 = C(3.0) * m_::rcp(one + e_r3) * C(1.12) * m_::rcp(one + e_r4);
    ndt_xr1 = ndt0 * m_::rcp(tau_xr1);
    ndt_xr2 = ndt0 * m_::rcp(tau_xr2);
  } else {
    ndt_xr1 = ndt0 * (one + e_r1) * (one + e_r2) * C(1.0 / 2700.0);
    ndt_xr2 = ndt0 * (one + e_r3) * (one + e_r4) * C(1.0 / 3.36);
  }
  const T vek = V - E_K;
  const T i_Kr = g_Kr * Xr1 * Xr2 * vek;

  // I_Ks (:290-294)
  const T xs_inf = m_::rcp(one + LX((C(-5.0) - V) * C(1.0 / 14.0), V, -1.0 / 14.0, -5.0 / 14.0));
  const T e_xs = XP((C(5.0) - V) * C(1.0 / 6.0), C(2.300975890892825) * W6);  // e^(5/6)
  T a_xs;
  if (kRef) a_xs = m_::dv(C(1400.0), m_::sq(one + e_xs));
  else a_xs = C(1400.0) * m_::rsq(one + e_xs);
  const T b_xs = m_::rcp(one + LX((V - C(35.0)) * C(1.0 / 15.0), V, 1.0 / 15.0, -35.0 / 15.0));
  const T tau_xs = a_xs * b_xs + C(80.0);
  const T i_Ks = g_Ks * Xs * Xs * (V - E_Ks);

  // I_K1 (:297-302)
  const T e_k1 = LX(C(0.06) * (vek - C(200.0)), vek, 0.06, -12.0);
  const T nb_k1 = C(3.0) * LX(C(2e-4) * (vek + C(100.0)), vek, 2e-4, 0.02) +
                  LX(C(0.1) * (vek - C(10.0)), vek, 0.1, -1.0);
  const T e_k2 = LX(C(-0.5) * vek, vek, -0.5, 0.0);
  T xk1_inf;
  if (kRef) {
    const T a_k1 = C(0.1) * m_::rcp(one + e_k1);
    const T b_k1 = m_::dv(nb_k1, one + e_k2);
    xk1_inf = m_::dv(a_k1, a_k1 + b_k1);
  } else {
    // a/(a+b) = 0.1(1+e2) / (0.1(1+e2) + nb (1+e1))
    const T num = C(0.1) * (one + e_k2);
    xk1_inf = num * m_::rcp(num + nb_k1 * (one + e_k1));
  }
  const T i_K1 = C(5.405) * xk1_inf * vek;

  // I_to and the phenotype s-gate (:305-317)
  T s_inf, tau_s;
  if (endo_s) {
    s_inf = m_::rcp(one + XP((V + C(28.0)) * C(0.2), C(270.42640742615254) * P));  // e^5.6
    const T u = V + C(67.0);
    tau_s = C(1000.0) * LX(-u * u * C(1.0 / 1000.0), u * u, -1.0 / 1000.0, 0.0) + C(8.0);
  } else {
    s_inf = m_::rcp(one + XP((V + C(20.0)) * C(0.2), C(54.598150033144236) * P));  // e^4
    const T u = V + C(45.0);
    tau_s = C(85.0) * LX(-u * u * C(1.0 / 320.0), u * u, -1.0 / 320.0, 0.0) +
            C(5.0) * m_::rcp(one + XP((V - C(20.0)) * C(0.2), C(0.01831563888873418) * P)) +
            C(3.0);  // e^-4
  }
  const T r_inf = m_::rcp(one + XP((C(20.0) - V) * C(1.0 / 6.0),
                                  C(28.03162489452614) * W6));  // e^(20/6)
  const T u40 = V + C(40.0);
  const T tau_r = C(9.5) * LX(-u40 * u40 * C(1.0 / 1800.0), u40 * u40, -1.0 / 1800.0, 0.0) + C(0.8);
  const T i_to = g_to * rg * sg * vek;

  // I_CaL (:320-340)
  const T d_inf = m_::rcp(one + LX((C(-8.0) - V) * C(1.0 / 7.5), V, -1.0 / 7.5, -8.0 / 7.5));
  const T tau_d = (C(1.4) * m_::rcp(one + LX((C(-35.0) - V) * C(1.0 / 13.0), V, -1.0 / 13.0,
                                             -35.0 / 13.0)) +
                   C(0.25)) * C(1.4) *
                      m_::rcp(one + XP((V + C(5.0)) * C(0.2), C(2.718281828459045) * P)) +
                  m_::rcp(one + XP((C(50.0) - V) * C(0.05), C(12.182493960703473) * iR5));
  const T f_inf = m_::rcp(one + XP((V + C(20.0)) * C(1.0 / 7.0),
                                   C(17.41170806332765) * S7));  // e^(20/7)
  const T u27 = V + C(27.0);
  const T r30 = m_::rcp(one + XP((V + C(30.0)) * C(0.1), C(20.085536923187668) * Q));  // e^3
  const T tau_f = C(1102.5) * LX(-(u27 * u27) * C(1.0 / 225.0), u27 * u27, -1.0 / 225.0, 0.0) +
                  C(200.0) * m_::rcp(one + XP((C(13.0) - V) * C(0.1),
                                              C(3.6692966676192444) * iQ)) +  // e^1.3
                  C(180.0) * r30 +
                  C(20.0);
  const T f2_inf = C(0.67) * m_::rcp(one + XP((V + C(35.0)) * C(1.0 / 7.0),
                                             C(148.4131591025766) * S7)) +  // e^5
                   C(0.33);
  const T tau_f2 = C(562.0) * LX(-(u27 * u27) * C(1.0 / 240.0), u27 * u27, -1.0 / 240.0, 0.0) +
                   C(31.0) * m_::rcp(one + XP((C(25.0) - V) * C(0.1),
                                              C(12.182493960703473) * iQ)) +  // e^2.5
                   C(80.0) * r30;
  const T cs = Cass * C(20.0);  // Cass / 0.05
  const T cs2 = cs * cs;
  const T rcs = m_::rcp(one + cs2);
  const T fCass_inf = C(0.6) * rcs + C(0.4);
  const T tau_fCass = C(80.0) * rcs + C(2.0);
  const T xcal = C(2.0) * (V - C(15.0)) * FRT;
  // e^{2(V-15)F/RT} = e^{2VF/RT} e^{-30F/RT}
  const T e2v = XP(xcal, E * E * C(0.3252969814969026));
  T i_CaL;
  if (kRef) {
    T den = e2v - one;
    if (m_::ab(den) < C(1e-10)) den = C(1e-10);
    i_CaL = g_CaL * d * f * f2 * fCass * C(4.0) * (V - C(15.0)) * F * F / (R * TK) *
            (C(0.25) * Cass * e2v - Cao) / den;
  } else {
    // x / (exp(x) - 1): three terms of its series for |x| < 0.01 (truncation
    // < 2e-11), else the quotient (cancellation error < 1e-5 relative there)
    T xe;
    if (m_::ab(xcal) < C(0.01)) xe = one - C(0.5) * xcal + xcal * xcal * C(1.0 / 12.0);
    else xe = xcal * m_::rcp(e2v - one);
    i_CaL = g_CaL * C(2.0 * 96485.3415) * d * f * f2 * fCass * (C(0.25) * Cass * e2v - Cao) * xe;
  }

  // background, pump, exchanger (:343-357)
  const T i_bNa = C(0.00029) * (V - E_Na);
  const T i_bCa = C(0.000592) * (V - E_Ca);
  const T i_pCa = m_::dv(C(0.1238) * Cai, Cai + C(0.0005));
  const T i_pK = m_::dv(C(0.0146) * vek,
                        one + LX((C(25.0) - V) * C(1.0 / 5.98), V, -1.0 / 5.98, 25.0 / 5.98));
  const T nak_den = one +
                    C(0.1245) * LX(C(-0.1) * V * FRT, V, -0.1 * 96485.3415 / (8314.472 * 310.0), 0.0) +
                    C(0.0353) * XP(-V * FRT, iE);
  T i_NaK;
  if (kRef) i_NaK = C(2.724 * 5.4 / (5.4 + 1.0)) * m_::dv(Nai, Nai + C(40.0)) * m_::rcp(nak_den);
  else i_NaK = C(2.724 * 5.4 / (5.4 + 1.0)) * Nai * m_::rcp((Nai + C(40.0)) * nak_den);
  const T e_g = LX(C(0.35) * V * FRT, V, 0.35 * 96485.3415 / (8314.472 * 310.0), 0.0);
  const T e_gm1 = XP(C(0.35 - 1.0) * V * FRT, e_g * iE);
  const T i_NaCa = m_::dv(C(1000.0) * (e_g * Nai * Nai * Nai * Cao - e_gm1 * C(140.0 * 140.0 * 140.0) * Cai * C(2.5)),
                          C((87.5 * 87.5 * 87.5 + 140.0 * 140.0 * 140.0) * (1.38 + 2.0)) *
                              (one + C(0.1) * e_gm1));

  // total current from the old state (:360-361)
  const T I_ion = i_K1 + i_to + i_Kr + i_Ks + i_CaL + i_NaK + i_Na + i_bNa + i_NaCa + i_bCa + i_pCa + i_pK;

  // SR / subspace Ca (:364-375)
  const T c2 = Cass * Cass;
  T kcasr, O;
  if (kRef) {
    const T ec = m_::dv(C(1.5), CaSR);
    kcasr = C(2.5) - C(1.5) * m_::rcp(one + ec * ec);
    const T k1 = m_::dv(C(0.15), kcasr);
    O = m_::dv(k1 * c2 * Rp, C(0.06) + k1 * c2 + C(1e-30));
  } else {
    const T sr2 = CaSR * CaSR;
    kcasr = C(2.5) - C(1.5) * sr2 * m_::rcp(sr2 + C(2.25));
    // k1 = 0.15/kcasr folded in: O = 0.15 c2 Rp / (0.06 kcasr + 0.15 c2)
    O = C(0.15) * c2 * Rp * m_::rcp(C(0.06) * kcasr + C(0.15) * c2 + C(1e-30));
  }
  const T k2 = C(0.045) * kcasr;
  const T i_rel = C(0.102) * O * (CaSR - Cass);
  T i_up, bufc, bufsr, bufss;
  const T qc = Cai + C(0.001), qsr = CaSR + C(0.3), qss = Cass + C(0.00025);
  if (kRef) {
    i_up = m_::dv(C(0.006375), one + m_::dv(C(0.00025 * 0.00025), Cai * Cai + C(1e-30)));
    bufc = m_::rcp(one + m_::dv(C(0.2 * 0.001), qc * qc));
    bufsr = m_::rcp(one + m_::dv(C(10.0 * 0.3), qsr * qsr));
    bufss = m_::rcp(one + m_::dv(C(0.4 * 0.00025), qss * qss));
  } else {
    const T ca2 = Cai * Cai + C(1e-30);
    i_up = C(0.006375) * ca2 * m_::rcp(ca2 + C(0.00025 * 0.00025));
    bufc = qc * qc * m_::rcp(qc * qc + C(0.2 * 0.001));
    bufsr = qsr * qsr * m_::rcp(qsr * qsr + C(10.0 * 0.3));
    bufss = qss * qss * m_::rcp(qss * qss + C(0.4 * 0.00025));
  }
  const T i_leak = C(0.00036) * (CaSR - Cai);
  const T i_xfer = C(0.0038) * (Cass - Cai);

  // concentration derivatives (:378-384)
  const T inv_vcf = C(1.0 / (0.016404 * 96485.3415));
  const T dNai = -(i_Na + i_bNa + C(3.0) * i_NaK + C(3.0) * i_NaCa) * inv_vcf;
  const T dKi = -(i_K1 + i_to + i_Kr + i_Ks + i_pK - C(2.0) * i_NaK) * inv_vcf;
  const T dCai = bufc * ((i_leak - i_up) * C(0.001094 / 0.016404) + i_xfer -
                         (i_bCa + i_pCa - C(2.0) * i_NaCa) * C(1.0 / (2.0 * 0.016404 * 96485.3415)));
  const T dCaSR = bufsr * (i_up - i_rel - i_leak);
  const T dCass = bufss * (-i_CaL * C(1.0 / (2.0 * 0.00005468 * 96485.3415)) +
                           i_rel * C(0.001094 / 0.00005468) - i_xfer * C(0.016404 / 0.00005468));

  // Rush-Larsen gates (:387-399)
#define NDT(tau) (ndt0 * m_::rcp(tau))
  s[S_M] = RLE(mg, m_inf, ndt_m);
  s[S_H] = RLE(hg, h_inf, ndt_h);
  s[S_J] = RLE(jg, h_inf, ndt_j);
  s[S_XR1] = RLE(Xr1, xr1_inf, ndt_xr1);
  s[S_XR2] = RLE(Xr2, xr2_inf, ndt_xr2);
  s[S_XS] = RLE(Xs, xs_inf, NDT(tau_xs));
  s[S_D] = RLE(d, d_inf, NDT(tau_d));
  s[S_F] = RLE(f, f_inf, NDT(tau_f));
  s[S_F2] = RLE(f2, f2_inf, NDT(tau_f2));
  s[S_FCASS] = RLE(fCass, fCass_inf, NDT(tau_fCass));
  s[S_S] = RLE(sg, s_inf, NDT(tau_s));
  s[S_R] = RLE(rg, r_inf, NDT(tau_r));
#undef NDT
#undef RLE
#undef XP
#undef LX
#undef L2E
  // R' forward Euler (:402-403)
  s[S_RPRIME] = Rp + dt * (-k2 * Cass * Rp + C(0.005) * (one - Rp));
  // concentrations, forward Euler with clamps (:406-410)
  s[S_NAI] = m_::mx(s[S_NAI] + dt * dNai, C(1e-3) - nai_base);
  s[S_KI] = m_::mx(s[S_KI] + dt * dKi, C(1e-3) - ki_base);
  s[S_CAI] = m_::mx(Cai + dt * dCai, C(1e-7));
  s[S_CASR] = m_::mx(CaSR + dt * dCaSR, C(1e-7));
  s[S_CASS] = m_::mx(Cass + dt * dCass, C(1e-7));
#undef C
  return I_ion;
}

}  // namespace hp
