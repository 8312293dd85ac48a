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
// Storage of Ki and Nai: the state holds Ki - ki_base and Nai - nai_base.
// With base 0 this is the reference's state.  The FP32 engine passes the
// resting values (138.4, 10.355) so that the per-step increments (~1e-5 mM,
// about one FP32 ulp of 138) are accumulated at full relative precision.
//
// FP32 I_CaL: the factor (V-15)/(exp(2(V-15)F/RT)-1) is evaluated as
// x/expm1(x) with a series at |x| < 1e-6, instead of the reference's guard
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

template <typename T, bool kExactCaL>
__device__ __forceinline__ T ttp06_cell(T V, T (&s)[N_STATE], T dt, T g_Ks, T g_to,
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

  // Nernst potentials (gpu_ttp06.py:242-247)
  const T E_Na = RTF * m_::lg(m_::dv(Nao, Nai));
  const T E_K = RTF * m_::lg(m_::dv(Ko, Ki));
  const T E_Ks = RTF * m_::lg(m_::dv(Ko + C(0.03) * Nao, Ki + C(0.03) * Nai));
  const T E_Ca = C(0.5) * RTF * m_::lg(m_::dv(Cao, Cai));

  // I_Na (:250-276)
  T t1 = m_::rcp(one + m_::ex((C(-56.86) - V) * C(1.0 / 9.03)));
  const T m_inf = t1 * t1;
  const T a_m = m_::rcp(one + m_::ex((C(-60.0) - V) * C(0.2)));
  const T b_m = C(0.1) * m_::rcp(one + m_::ex((V + C(35.0)) * C(0.2))) +
                C(0.1) * m_::rcp(one + m_::ex((V - C(50.0)) * C(1.0 / 200.0)));
  const T tau_m = a_m * b_m;
  t1 = m_::rcp(one + m_::ex((V + C(71.55)) * C(1.0 / 7.43)));
  const T h_inf = t1 * t1;
  T a_h, b_h, a_j, b_j;
  if (V < C(-40.0)) {
    a_h = C(0.057) * m_::ex(-(V + C(80.0)) * C(1.0 / 6.8));
    b_h = C(2.7) * m_::ex(C(0.079) * V) + C(3.1e5) * m_::ex(C(0.3485) * V);
    a_j = m_::dv((C(-25428.0) * m_::ex(C(0.2444) * V) - C(6.948e-6) * m_::ex(C(-0.04391) * V)) *
                     (V + C(37.78)),
                 one + m_::ex(C(0.311) * (V + C(79.23))));
    b_j = m_::dv(C(0.02424) * m_::ex(C(-0.01052) * V), one + m_::ex(C(-0.1378) * (V + C(40.14))));
  } else {
    a_h = C(0.0);
    b_h = m_::dv(C(0.77), C(0.13) * (one + m_::ex((V + C(10.66)) * C(-1.0 / 11.1))));
    a_j = C(0.0);
    b_j = m_::dv(C(0.6) * m_::ex(C(0.057) * V), one + m_::ex(C(-0.1) * (V + C(32.0))));
  }
  const T tau_h = m_::rcp(a_h + b_h + C(1e-30));
  const T tau_j = m_::rcp(a_j + b_j + C(1e-30));
  const T i_Na = C(14.838) * mg * mg * mg * hg * jg * (V - E_Na);

  // I_Kr (:279-287); sqrt(Ko/5.4) = 1
  const T xr1_inf = m_::rcp(one + m_::ex((C(-26.0) - V) * C(1.0 / 7.0)));
  const T tau_xr1 = C(450.0) * m_::rcp(one + m_::ex((C(-45.0) - V) * C(0.1))) * C(6.0) *
                    m_::rcp(one + m_::ex((V + C(30.0)) * C(1.0 / 11.5)));
  const T xr2_inf = m_::rcp(one + m_::ex((V + C(88.0)) * C(1.0 / 24.0)));
  const T tau_xr2 = C(3.0) * m_::rcp(one + m_::ex((C(-60.0) - V) * C(0.05))) * C(1.12) *
                    m_::rcp(one + m_::ex((V - C(60.0)) * C(0.05)));
  const T vek = V - E_K;
  const T i_Kr = g_Kr * Xr1 * Xr2 * vek;

  // I_Ks (:290-294)
  const T xs_inf = m_::rcp(one + m_::ex((C(-5.0) - V) * C(1.0 / 14.0)));
  const T a_xs = m_::dv(C(1400.0), m_::sq(one + m_::ex((C(5.0) - V) * C(1.0 / 6.0))));
  const T b_xs = m_::rcp(one + m_::ex((V - C(35.0)) * C(1.0 / 15.0)));
  const T tau_xs = a_xs * b_xs + C(80.0);
  const T i_Ks = g_Ks * Xs * Xs * (V - E_Ks);

  // I_K1 (:297-302)
  const T a_k1 = C(0.1) * m_::rcp(one + m_::ex(C(0.06) * (vek - C(200.0))));
  const T b_k1 = m_::dv(C(3.0) * m_::ex(C(2e-4) * (vek + C(100.0))) + m_::ex(C(0.1) * (vek - C(10.0))),
                        one + m_::ex(C(-0.5) * vek));
  const T i_K1 = C(5.405) * m_::dv(a_k1, a_k1 + b_k1) * vek;

  // I_to and the phenotype s-gate (:305-317)
  T s_inf, tau_s;
  if (endo_s) {
    s_inf = m_::rcp(one + m_::ex((V + C(28.0)) * C(0.2)));
    const T u = V + C(67.0);
    tau_s = C(1000.0) * m_::ex(-u * u * C(1.0 / 1000.0)) + C(8.0);
  } else {
    s_inf = m_::rcp(one + m_::ex((V + C(20.0)) * C(0.2)));
    const T u = V + C(45.0);
    tau_s = C(85.0) * m_::ex(-u * u * C(1.0 / 320.0)) +
            C(5.0) * m_::rcp(one + m_::ex((V - C(20.0)) * C(0.2))) + C(3.0);
  }
  const T r_inf = m_::rcp(one + m_::ex((C(20.0) - V) * C(1.0 / 6.0)));
  const T u40 = V + C(40.0);
  const T tau_r = C(9.5) * m_::ex(-u40 * u40 * C(1.0 / 1800.0)) + C(0.8);
  const T i_to = g_to * rg * sg * vek;

  // I_CaL (:320-340)
  const T d_inf = m_::rcp(one + m_::ex((C(-8.0) - V) * C(1.0 / 7.5)));
  const T tau_d = (C(1.4) * m_::rcp(one + m_::ex((C(-35.0) - V) * C(1.0 / 13.0))) + C(0.25)) * C(1.4) *
                      m_::rcp(one + m_::ex((V + C(5.0)) * C(0.2))) +
                  m_::rcp(one + m_::ex((C(50.0) - V) * C(0.05)));
  const T f_inf = m_::rcp(one + m_::ex((V + C(20.0)) * C(1.0 / 7.0)));
  const T u27 = V + C(27.0);
  const T tau_f = C(1102.5) * m_::ex(-(u27 * u27) * C(1.0 / 225.0)) +
                  C(200.0) * m_::rcp(one + m_::ex((C(13.0) - V) * C(0.1))) +
                  C(180.0) * m_::rcp(one + m_::ex((V + C(30.0)) * C(0.1))) + C(20.0);
  const T f2_inf = C(0.67) * m_::rcp(one + m_::ex((V + C(35.0)) * C(1.0 / 7.0))) + C(0.33);
  const T tau_f2 = C(562.0) * m_::ex(-(u27 * u27) * C(1.0 / 240.0)) +
                   C(31.0) * m_::rcp(one + m_::ex((C(25.0) - V) * C(0.1))) +
                   C(80.0) * m_::rcp(one + m_::ex((V + C(30.0)) * C(0.1)));
  const T cs = Cass * C(20.0);  // Cass / 0.05
  const T cs2 = cs * cs;
  const T fCass_inf = C(0.6) * m_::rcp(one + cs2) + C(0.4);
  const T tau_fCass = C(80.0) * m_::rcp(one + cs2) + C(2.0);
  const T xcal = C(2.0) * (V - C(15.0)) * FRT;
  const T e2v = m_::ex(xcal);
  T i_CaL;
  if (kExactCaL) {
    T den = e2v - one;
    if (m_::ab(den) < C(1e-10)) den = C(1e-10);
    i_CaL = g_CaL * d * f * f2 * fCass * C(4.0) * (V - C(15.0)) * F * F / (R * TK) *
            (C(0.25) * Cass * e2v - Cao) / den;
  } else {
    T xe;  // x / (exp(x) - 1)
    if (m_::ab(xcal) < C(1e-6)) xe = one - C(0.5) * xcal;
    else xe = m_::dv(xcal, (T)expm1(xcal));
    i_CaL = g_CaL * C(2.0 * 96485.3415) * d * f * f2 * fCass * (C(0.25) * Cass * e2v - Cao) * xe;
  }

  // background, pump, exchanger (:343-357)
  const T i_bNa = C(0.00029) * (V - E_Na);
  const T i_bCa = C(0.000592) * (V - E_Ca);
  const T i_pCa = m_::dv(C(0.1238) * Cai, Cai + C(0.0005));
  const T i_pK = m_::dv(C(0.0146) * vek, one + m_::ex((C(25.0) - V) * C(1.0 / 5.98)));
  const T i_NaK = C(2.724 * 5.4 / (5.4 + 1.0)) * m_::dv(Nai, Nai + C(40.0)) *
                  m_::rcp(one + C(0.1245) * m_::ex(C(-0.1) * V * FRT) + C(0.0353) * m_::ex(-V * FRT));
  const T e_g = m_::ex(C(0.35) * V * FRT);
  const T e_gm1 = m_::ex(C(0.35 - 1.0) * V * FRT);
  const T i_NaCa = m_::dv(C(1000.0) * (e_g * Nai * Nai * Nai * Cao - e_gm1 * C(140.0 * 140.0 * 140.0) * Cai * C(2.5)),
                          C((87.5 * 87.5 * 87.5 + 140.0 * 140.0 * 140.0) * (1.38 + 2.0)) *
                              (one + C(0.1) * e_gm1));

  // total current from the old state (:360-361)
  const T I_ion = i_K1 + i_to + i_Kr + i_Ks + i_CaL + i_NaK + i_Na + i_bNa + i_NaCa + i_bCa + i_pCa + i_pK;

  // SR / subspace Ca (:364-375)
  const T ec = m_::dv(C(1.5), CaSR);
  const T kcasr = C(2.5) - C(1.5) * m_::rcp(one + ec * ec);
  const T k1 = m_::dv(C(0.15), kcasr);
  const T k2 = C(0.045) * kcasr;
  const T c2 = Cass * Cass;
  const T O = m_::dv(k1 * c2 * Rp, C(0.06) + k1 * c2 + C(1e-30));
  const T i_rel = C(0.102) * O * (CaSR - Cass);
  const T i_up = m_::dv(C(0.006375), one + m_::dv(C(0.00025 * 0.00025), Cai * Cai + C(1e-30)));
  const T i_leak = C(0.00036) * (CaSR - Cai);
  const T i_xfer = C(0.0038) * (Cass - Cai);
  T q = Cai + C(0.001);
  const T bufc = m_::rcp(one + m_::dv(C(0.2 * 0.001), q * q));
  q = CaSR + C(0.3);
  const T bufsr = m_::rcp(one + m_::dv(C(10.0 * 0.3), q * q));
  q = Cass + C(0.00025);
  const T bufss = m_::rcp(one + m_::dv(C(0.4 * 0.00025), q * q));

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
#define RL(x, xi, tau) ((xi) - ((xi) - (x)) * m_::ex(-dt * m_::rcp(tau)))
  s[S_M] = RL(mg, m_inf, tau_m);
  s[S_H] = RL(hg, h_inf, tau_h);
  s[S_J] = RL(jg, h_inf, tau_j);
  s[S_XR1] = RL(Xr1, xr1_inf, tau_xr1);
  s[S_XR2] = RL(Xr2, xr2_inf, tau_xr2);
  s[S_XS] = RL(Xs, xs_inf, tau_xs);
  s[S_D] = RL(d, d_inf, tau_d);
  s[S_F] = RL(f, f_inf, tau_f);
  s[S_F2] = RL(f2, f2_inf, tau_f2);
  s[S_FCASS] = RL(fCass, fCass_inf, tau_fCass);
  s[S_S] = RL(sg, s_inf, tau_s);
  s[S_R] = RL(rg, r_inf, tau_r);
#undef RL
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
