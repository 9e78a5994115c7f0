// Device physics of the explicit finite-volume step (FP64), gfx950.
//
// Every function restates one reference routine (file:line relative to the
// ECOGEN source tree) with the same evaluation order, clamp constants and
// std::min/std::max semantics, so that the only arithmetic differences from
// the reference are those of the power in the isentrope (pow_pos, a few
// ulp; the reference's C library pow is close to correctly rounded).
// Derived quantities the reference caches per cell (phase energies, sound
// speeds, mixture density/pressure/energy, frozen sound speed) are recomputed
// from the stored primitive state (alpha_k, rho_k, p_k, u, E) when needed.
#pragma once
#include <hip/hip_runtime.h>
#include "eco_math.h"
#include "eco_params.h"

namespace eco {

#define ECO_D __device__ __forceinline__

ECO_D double smin(double a, double b) { return (b < a) ? b : a; }
ECO_D double smax(double a, double b) { return (a < b) ? b : a; }

// ----------------------------------------------------------------------------
// EOS: src/Eos/EosSG.cpp:93-254, src/Eos/EosIG.cpp:90-251
// ----------------------------------------------------------------------------
ECO_D double eos_e(const EosParams &q, double rho, double p, double eps) {
  // og = RN(1/(gamma-1)) (host division): the first quotient through qdiv
  if (q.type == EOS_SG) return qdiv(p + q.gpinf, q.gm1, q.og) / smax(rho, eps) + q.eref;
  return qdiv(p, q.gm1, q.og) / smax(rho, eps) + q.eref;
}
ECO_D double eos_p(const EosParams &q, double rho, double e) {
  if (q.type == EOS_SG) return q.gm1 * rho * (e - q.eref) - q.gpinf;
  return q.gm1 * rho * (e - q.eref);
}
ECO_D double eos_c(const EosParams &q, double rho, double p, double eps) {
  if (q.type == EOS_SG) return sqrt(q.gamma * (p + q.pinf) / smax(rho, eps));
  return sqrt(q.gamma * p / smax(rho, eps));
}
// tab: the power's tables in fast memory (pow_pos), or null for the constant ones
// eos_e and eos_c of one phase whose (floored) density rho both divide by:
// the two quotients through one shared reciprocal yrho = RN(1/rho) (qdiv,
// correctly rounded, so bit for bit the two IEEE divisions)
ECO_D double eos_e_y(const EosParams &q, double rho, double yrho, double p) {
  if (q.type == EOS_SG) return qdiv(qdiv(p + q.gpinf, q.gm1, q.og), rho, yrho) + q.eref;
  return qdiv(qdiv(p, q.gm1, q.og), rho, yrho) + q.eref;
}
ECO_D double eos_c_y(const EosParams &q, double rho, double yrho, double p) {
  if (q.type == EOS_SG) return sqrt(qdiv(q.gamma * (p + q.pinf), rho, yrho));
  return sqrt(qdiv(q.gamma * p, rho, yrho));
}
ECO_D double eos_isentrope(const EosParams &q, double p0, double rho0, double rho1, double eps,
                           const double *tab = nullptr) {
  if (q.type == EOS_SG) return (p0 + q.pinf) * pow_pos(rho1 / smax(rho0, eps), q.gamma, tab) - q.pinf;
  return p0 * pow_pos(rho1 / smax(rho0, eps), q.gamma, tab);
}
ECO_D double eos_rho_pfinal(const EosParams &q, double p0, double rho0, double pf, double &drho, double eps) {
  if (q.type == EOS_SG) {
    double num = q.gamma * (pf + q.pinf);
    double denom = num + p0 - pf;
    double rf = rho0 * num / smax(denom, eps);
    drho = rho0 * q.gamma * (p0 + q.pinf) / smax(denom * denom, eps);
    return rf;
  }
  double num = q.gamma * pf;
  double denom = num + p0 - pf;
  double rf = rho0 * num / smax(denom, eps);
  drho = rho0 * q.gamma * p0 / smax(denom * denom, eps);
  return rf;
}
// verifyAndModifyPressure (EosSG.cpp:243-247, EosIG.cpp:241-244)
ECO_D double eos_fix_p(const EosParams &q, double p) {
  if (q.type == EOS_SG) return (p <= q.pfloor) ? q.pfloor : p;
  return (p < 1.e-15) ? 1.e-15 : p;
}

// ----------------------------------------------------------------------------
// Limiters: src/Ordre2/Limiter*.cpp:49-59
// ----------------------------------------------------------------------------
// Branch-free forms (every operand computed, the result selected), so that
// the limited slopes of all variables of a face schedule as one basic block.
// The selected value is the reference's bit for bit.
// a / b: IEEE division in FP64 (the reference's limiters); in FP32 (the
// FP32-storage reconstruction, held to 1e-5 of FP64, not to the bit) the
// hardware reciprocal (1 ulp) and a product -- two instructions where an IEEE
// FP32 division takes about ten
template <typename T>
ECO_D T ldiv(T a, T b) {
  if constexpr (sizeof(T) == sizeof(float)) return a * __builtin_amdgcn_rcpf(b);
  else return a / b;
}

// 1/x in FP64 for the FP32-storage path, whose FP64 arithmetic only has to
// stay far below FP32's rounding: the hardware reciprocal refined by two
// Newton steps (within about an ulp of 1/x for normal x; five
// instructions where an IEEE FP64 division takes about eleven).
ECO_D double rcp_fast(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.);
  y = fma(y, e, y);
  e = fma(-x, y, 1.);
  return fma(y, e, y);
}
// 1/b and a/b: IEEE division, or (FAST, the FP32-storage path) rcp_fast
template <bool FAST>
ECO_D double drcp(double b) {
  if constexpr (FAST) return rcp_fast(b);
  else return 1. / b;
}
template <bool FAST>
ECO_D double ddiv(double a, double b) {
  if constexpr (FAST) return a * rcp_fast(b);
  else return a / b;
}

// log1p and expm1 in FP32 for the isentropic power factor of the FP32-storage
// HLLC.  For the small arguments of nearly every face (|x| <= 1/32, |z| <=
// 1/4) their Taylor series to degree 7 by Horner's rule: the first omitted
// term is below 2^-38 (log1p) and 2^-28 (expm1) relative to the result, under
// FP32's rounding.  Larger arguments take the library's log1pf / expm1f.
ECO_D float log1p_f(float x) {
  if (!(fabsf(x) <= 0x1p-5f)) return log1pf(x);
  float p = 1.f / 7.f;
  p = p * x + (-1.f / 6.f);
  p = p * x + 0.2f;
  p = p * x + (-0.25f);
  p = p * x + (1.f / 3.f);
  p = p * x + (-0.5f);
  p = p * x + 1.f;
  return p * x;
}
ECO_D float expm1_f(float z) {
  if (!(fabsf(z) <= 0.25f)) return expm1f(z);
  float p = 1.f / 5040.f;
  p = p * z + (1.f / 720.f);
  p = p * z + (1.f / 120.f);
  p = p * z + (1.f / 24.f);
  p = p * z + (1.f / 6.f);
  p = p * z + 0.5f;
  p = p * z + 1.f;
  return p * z;
}

template <int LIM, typename T>
ECO_D T limiter_kt(T s1, T s2) {
  auto mn = [](T a, T b) { return (b < a) ? b : a; };
  auto mx = [](T a, T b) { return (a < b) ? b : a; };
  const T prod = s1 * s2;
  if constexpr (LIM == LIM_MINMOD || LIM == LIM_THINC) {
    const T m = mn(fabs(s1), fabs(s2));
    return (prod > T(1e-9)) ? ((s1 < 0) ? -m : m) : T(0);
  } else if constexpr (LIM == LIM_VANLEER) {
    const T sum = s1 + s2;
    const T q = ldiv(T(2) * prod, sum);
    return ((fabs(sum) > T(1e-6)) && (prod > T(1e-6))) ? q : T(0);
  } else if constexpr (LIM == LIM_VANALBADA) {
    const T sum = s1 + s2;
    const T q = ldiv(prod * sum, s1 * s1 + s2 * s2);
    return ((fabs(s1) > T(1e-6)) && (fabs(s2) > T(1e-6)) && (fabs(sum) > T(1e-6)) && (prod > T(1e-6))) ? q : T(0);
  } else if constexpr (LIM == LIM_MC) {
    T m = mn(T(2) * fabs(s1), T(2) * fabs(s2));
    m = mn(m, T(0.5) * (fabs(s1) + fabs(s2)));
    const T slope = (prod > T(1e-9)) ? m : T(0);
    return (s1 < 0) ? -slope : slope;
  } else if constexpr (LIM == LIM_SUPERBEE) {
    const T m = mx(mn(T(2) * fabs(s1), fabs(s2)), mn(fabs(s1), T(2) * fabs(s2)));
    const T slope = (prod > T(1e-6)) ? m : T(0);
    return (s1 < 0) ? -slope : slope;
  } else {
    return T(0);
  }
}

template <int LIM>
ECO_D double limiter_k(double s1, double s2) {
  return limiter_kt<LIM, double>(s1, s2);
}

// runtime limiter choice (LimiterMinmod/VanLeer/VanAlbada/Mc/Superbee
// ::limiteSlope, src/Ordre2/Limiter*.cpp:49-59; THINC limits with minmod),
// with the reference's branches (the division only where it is used)
ECO_D double limiter(int lim, double s1, double s2) {
  double slope = 0., prod = s1 * s2;
  switch (lim) {
    case LIM_MINMOD:
    case LIM_THINC:
      if (prod > 1e-9) { slope = smin(fabs(s1), fabs(s2)); if (s1 < 0) slope = -slope; }
      return slope;
    case LIM_VANLEER: {
      double sum = s1 + s2;
      if ((fabs(sum) > 1e-6) && (prod > 1e-6)) slope = 2. * prod / sum;
      return slope;
    }
    case LIM_VANALBADA: {
      double sum = s1 + s2;
      if ((fabs(s1) > 1e-6) && (fabs(s2) > 1e-6) && (fabs(sum) > 1e-6) && (prod > 1e-6))
        slope = prod * sum / (s1 * s1 + s2 * s2);
      return slope;
    }
    case LIM_MC:
      if (prod > 1e-9) {
        slope = smin(2. * fabs(s1), 2. * fabs(s2));
        slope = smin(slope, 0.5 * (fabs(s1) + fabs(s2)));
      }
      if (s1 < 0.) slope = -slope;
      return slope;
    case LIM_SUPERBEE:
      if (prod > 1e-6) slope = smax(smin(2. * fabs(s1), fabs(s2)), smin(fabs(s1), 2. * fabs(s2)));
      if (s1 < 0.) slope = -slope;
      return slope;
    default:
      return 0.;
  }
}

// ----------------------------------------------------------------------------
// Cell state in registers
// ----------------------------------------------------------------------------
// Multiphase models: per phase alpha, rho, p (+ derived e, c, Y); mixture
// velocity u[3] and total energy E, mixture rho, p, e, frozen sound speed cF.
// Euler: phase 0 carries rho, p, e, c; u[3]; E is the phase total energy.
// ThermalEq: alpha_k and the mixture pressure pm and temperature T are the
// primitives; rho_k = rho_k(pm, T), p_k = pm, and cF is the mixture sound
// speed sqrt(sum Y_k c_k^2) (MixThermalEq).
template <int NPH>
struct CellState {
  double a[NPH], r[NPH], p[NPH];
  double e[NPH], c[NPH], Y[NPH];
  double u[3];
  double E;
  double rho, pm, em, cF;
  double T;
};

template <int NPH>
struct FluxVec {
  double a[NPH], m[NPH], en[NPH];
  double q[3];
  double E;
  double sM;
};

ECO_D double sqn(const double *u) { return u[0] * u[0] + u[1] * u[1] + u[2] * u[2]; }

// PhaseKapila::verifyAndCorrectPhase (PhaseKapila.cpp:294-307),
// PhaseMultiP::verifyAndCorrectPhase (PhaseMultiP.cpp:288-294),
// PhaseThermalEq::verifyAndCorrectPhase,
// PhaseEuler::verifyAndCorrectPhase (PhaseEuler.cpp:349-353)
template <int MODEL, int NPH>
ECO_D void verify_correct(const SchemeParams &S, CellState<NPH> &s, int k) {
  if constexpr (MODEL == MODEL_KAPILA) {
    if (S.eps > 1.e-20) {
      if (s.a[k] < 0.) s.a[k] = 0.;
      if (s.a[k] > 1.) s.a[k] = 1.;
      if (s.r[k] <= 1.e-15) s.r[k] = 1.e-15;
    } else {
      if (s.a[k] < 1e-15) s.a[k] = 1e-14;
      if (s.a[k] > 1. - 1e-15) s.a[k] = 1. - 1e-14;
      if (s.r[k] < 1.e-15) s.r[k] = 1.e-14;
    }
  } else if constexpr (MODEL == MODEL_MULTIP) {
    if (s.a[k] < 1e-10) s.a[k] = 1e-9;
    if (s.a[k] > 1. - 1e-10) s.a[k] = 1. - 1e-9;
    if (s.r[k] < 1.e-10) s.r[k] = 1.e-9;
  } else if constexpr (te_like(MODEL)) {
    // PhaseThermalEq::verifyAndCorrectPhase (PhaseThermalEq.cpp:240-246),
    // PhaseEulerHomogeneous::verifyAndCorrectPhase (PhaseEulerHomogeneous.cpp:189-195)
    if (s.a[k] < 1e-10) s.a[k] = 1e-10;
    if (s.a[k] > 1. - 1e-10) s.a[k] = 1. - 1e-10;
    if (s.r[k] < 1.e-10) s.r[k] = 1.e-10;
  } else {
    if (s.r[k] < 1.e-10) s.r[k] = 1.e-10;
  }
  s.p[k] = eos_fix_p(S.eos[k], s.p[k]);
}

// ThermalEq: the stiffened-gas offset of a phase (Eos::getPInf is 0 for
// EosIG, Eos.h:149) and its density at (p, T) (computeDensity, EosSG.cpp:
// 114-117, EosIG.cpp:111-114)
ECO_D double te_pinf(const EosParams &q) { return (q.type == EOS_SG) ? q.pinf : 0.; }
ECO_D double eos_rho_pt(const EosParams &q, double p, double T, double eps) {
  return ((q.type == EOS_SG) ? p + q.pinf : p) / smax((q.gamma - 1.) * q.cv * T, eps);
}

// ThermalEq phases at the mixture pressure and temperature
// (ModThermalEq::fulfillState, ModThermalEq.cpp:78-89, with
// PhaseThermalEq::extendedCalculusPhase, PhaseThermalEq.cpp:73-78), then
// MixThermalEq::computeMixtureVariables (MixThermalEq.cpp:305-323)
template <int NPH>
ECO_D void te_phases(const SchemeParams &S, CellState<NPH> &s) {
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    s.p[k] = s.pm;
    s.r[k] = eos_rho_pt(S.eos[k], s.pm, s.T, S.eps);
    s.e[k] = eos_e(S.eos[k], s.r[k], s.p[k], S.eps);
    s.c[k] = eos_c(S.eos[k], s.r[k], s.p[k], S.eps);
  }
}
template <int NPH>
ECO_D void te_mixture(CellState<NPH> &s) {
  s.rho = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) s.rho += s.a[k] * s.r[k];
  const double ird = 1. / s.rho;
#pragma unroll
  for (int k = 0; k < NPH; k++) s.Y[k] = qdiv(s.a[k] * s.r[k], s.rho, ird);
  s.em = 0.; s.cF = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) { s.em += s.Y[k] * s.e[k]; s.cF += s.Y[k] * s.c[k] * s.c[k]; }
  s.cF = sqrt(s.cF);
}

// ----------------------------------------------------------------------------
// Liquid-vapour saturation (EulerHomogeneous, RelaxationPTMu)
// ----------------------------------------------------------------------------
// Mixture::computeTsat (Mixture.cpp:63-104): the saturation temperature of
// the liquid and vapour phases S.liq, S.vap at pressure p, Newton from
// 0.1 B / C on f = A + B/T + C log T - log(p + pinf_vap) + D log(p + pinf_liq)
// (the logarithms as the C library's, log_lib); dTsat (may be null)
// receives dTsat/dp.  A soft error of the reference at 50 iterations is not
// reported; the iterate then is the reference's too.
ECO_D double sat_tsat(const SchemeParams &S, double p, double *dTsat) {
  const double pInfL = te_pinf(S.eos[S.liq]), pInfV = te_pinf(S.eos[S.vap]);
  const double A = S.tsat[0], B = S.tsat[1], C = S.tsat[2], D = S.tsat[3];
  const double lV = log_lib(p + pInfV), lL = log_lib(p + pInfL);
  double T = 0.1 * B / C, f = 0., df = 1.;
  int it = 0;
  do {
    T -= f / df;
    if (++it > 50) break;
    f = A + B / T + C * log_lib(T) - lV + D * lL;
    df = C / T - B / (T * T);
  } while (fabs(f) > 1e-10);
  if (dTsat) {
    const double dfdp = -1. / (p + pInfV) + D / (p + pInfL);
    *dTsat = -dfdp / df;
  }
  return T;
}

// EosSG/EosIG::computeDensitySaturation (EosSG.cpp:191-200, EosIG.cpp:188-197)
// with its derivative in p (drho may be null)
ECO_D double eos_rho_sat(const EosParams &q, double p, double T, double dT, double *drho, double eps) {
  const double g = q.gamma, cv = q.cv;
  const double pp = (q.type == EOS_SG) ? p + q.pinf : p;
  if (drho) {
    *drho = (g - 1.) * cv * T - pp * (g - 1.) * cv * dT;
    *drho /= smax((((g - 1.) * cv * T) * ((g - 1.) * cv * T)), eps);
  }
  return pp / smax(((g - 1.) * cv * T), eps);
}

// EosSG/EosIG::computeDensityEnergySaturation (EosSG.cpp:204-210, EosIG.cpp:201-207)
ECO_D double eos_rhoe_sat(const EosParams &q, double p, double rho, double drho, double *drhoe) {
  if (drhoe) *drhoe = 1. / (q.gamma - 1.) + drho * q.eref;
  if (q.type == EOS_SG) return (p + q.gamma * q.pinf) / (q.gamma - 1.) + rho * q.eref;
  return p / (q.gamma - 1.) + rho * q.eref;
}

// MixEulerHomogeneous::computePressure (MixEulerHomogeneous.cpp:93-125): the
// pressure at which the saturated mixture of density masse holds the internal
// energy e, Newton from p_guess (the cell's last pressure)
ECO_D double eh_pressure(const SchemeParams &S, double masse, double e, double p_guess) {
  const EosParams &qL = S.eos[S.liq], &qV = S.eos[S.vap];
  double p = p_guess, f = 0., df = 1.;
  int it = 0;
  do {
    p -= f / df;
    if (++it > 50) break;
    double dT, drV, drL, dreV, dreL;
    const double T = sat_tsat(S, p, &dT);
    const double rV = eos_rho_sat(qV, p, T, dT, &drV, S.eps);
    const double rL = eos_rho_sat(qL, p, T, dT, &drL, S.eps);
    const double aV = (masse - rL) / (rV - rL);
    const double daV = (-drL * (rV - rL) - (masse - rL) * (drV - drL)) / ((rV - rL) * (rV - rL));
    const double reV = eos_rhoe_sat(qV, p, rV, drV, &dreV);
    const double reL = eos_rhoe_sat(qL, p, rL, drL, &dreL);
    f = masse * e - aV * (reV - reL) - reL;
    df = -daV * (reV - reL) - aV * (dreV - dreL) - dreL;
  } while (fabs(f / (masse * e)) > 1e-10);
  return p;
}

// Model::fulfillState: extendedCalculusPhase (energy, sound speed) and
// Mix*::computeMixtureVariables (MixKapila.cpp:147-172, MixMultiP.cpp:142-168).
// Temperatures and Wood's sound speed are not used by the step and skipped.
template <int MODEL, int NPH>
ECO_D void fulfill(const SchemeParams &S, CellState<NPH> &s) {
  const double eps = S.eps;
  // EulerHomogeneous: the stored temperature is the saturation temperature of
  // the stored pressure (buildPrim and the initial state set both), so its
  // phases' saturated densities (p + pinf)/((gamma-1) cv Tsat) are those at
  // (p, T) of ThermalEq (ModEulerHomogeneous::fulfillState, ModEulerHomogeneous.cpp:42-56)
  if constexpr (te_like(MODEL)) {
    te_phases<NPH>(S, s);
    te_mixture<NPH>(s);
  } else if constexpr (MODEL == MODEL_EULER) {
    const double r = smax(s.r[0], eps), yr = 1. / r;
    s.e[0] = eos_e_y(S.eos[0], r, yr, s.p[0]);
    s.c[0] = eos_c_y(S.eos[0], r, yr, s.p[0]);
    s.E = s.e[0] + 0.5 * sqn(s.u);
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double r = smax(s.r[k], eps), yr = 1. / r;
      s.e[k] = eos_e_y(S.eos[k], r, yr, s.p[k]);
      s.c[k] = eos_c_y(S.eos[k], r, yr, s.p[k]);
    }
    s.rho = 0.; s.pm = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) { s.rho += s.a[k] * s.r[k]; s.pm += s.a[k] * s.p[k]; }
    const double rd = (MODEL == MODEL_KAPILA) ? smax(s.rho, eps) : s.rho;
    const double ird = 1. / rd;
#pragma unroll
    for (int k = 0; k < NPH; k++) s.Y[k] = qdiv(s.a[k] * s.r[k], rd, ird);
    s.em = 0.; s.cF = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) { s.em += s.Y[k] * s.e[k]; s.cF += s.Y[k] * s.c[k] * s.c[k]; }
    s.cF = sqrt(s.cF);
  }
}

// FluxKapila::buildCons (FluxKapila.cpp:123-135), FluxMultiP::buildCons
// (FluxMultiP.cpp:122-140), FluxEuler::buildCons (FluxEuler.cpp:95-104).
// Needs s.rho (mixture density) for the multiphase models.
template <int MODEL, int NPH>
ECO_D void build_cons(const SchemeParams &S, const CellState<NPH> &s, FluxVec<NPH> &f) {
  if constexpr (MODEL == MODEL_EULERHOM) {
    // FluxEulerHomogeneous::buildCons (FluxEulerHomogeneous.cpp:70-89): the
    // mixture mass, momentum and total energy, the internal energy summed from
    // the phases' energies at their densities and pressures
    double ein = 0.;
    f.m[0] = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double ar = s.a[k] * s.r[k];
      f.m[0] += ar;
      ein += ar * eos_e(S.eos[k], s.r[k], s.p[k], S.eps);
    }
#pragma unroll
    for (int k = 1; k < NPH; k++) f.m[k] = 0.;
#pragma unroll
    for (int d = 0; d < 3; d++) f.q[d] = f.m[0] * s.u[d];
    f.E = ein + 0.5 * f.m[0] * sqn(s.u);
  } else if constexpr (MODEL == MODEL_THERMALEQ) {
    // FluxThermalEq::buildCons (FluxThermalEq.cpp:113-126)
#pragma unroll
    for (int k = 0; k < NPH; k++) f.m[k] = s.a[k] * s.r[k];
#pragma unroll
    for (int d = 0; d < 3; d++) f.q[d] = s.rho * s.u[d];
    f.E = s.rho * s.E;
  } else if constexpr (MODEL == MODEL_EULER) {
    f.m[0] = s.r[0];
#pragma unroll
    for (int d = 0; d < 3; d++) f.q[d] = s.u[d] * f.m[0];
    f.E = f.m[0] * s.E;
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      f.a[k] = s.a[k];
      f.m[k] = s.a[k] * s.r[k];
      double ein = eos_e(S.eos[k], s.r[k], s.p[k], S.eps);
      f.en[k] = s.a[k] * s.r[k] * ein;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) f.q[d] = s.u[d] * s.rho;
    f.E = s.rho * s.E;
  }
}

ECO_D void zero_small(double *u) {
  if (fabs(u[0]) < 1.e-8) u[0] = 0.;
  if (fabs(u[1]) < 1.e-8) u[1] = 0.;
  if (fabs(u[2]) < 1.e-8) u[2] = 0.;
}

// Flux*::buildPrim: FluxKapila.cpp:139-185, FluxMultiP.cpp:144-183,
// FluxEuler.cpp:107-131.  Leaves the derived quantities of fulfillState set.
// FAST (FP32-storage path, Kapila and MultiP): quotients through rcp_fast.
template <int MODEL, int NPH, bool FAST = false>
ECO_D void build_prim(const SchemeParams &S, FluxVec<NPH> &f, CellState<NPH> &s) {
  const double eps = S.eps;
  if constexpr (MODEL == MODEL_EULERHOM) {
    // FluxEulerHomogeneous::buildPrim (FluxEulerHomogeneous.cpp:93-128): the
    // saturated mixture's pressure (Newton from s.pm, the cell's last
    // pressure), its saturation temperature, the saturated phase densities and
    // the vapour fraction of the mixture density
    if constexpr (NPH == 2) {
      const double masse = f.m[0], irm = 1. / masse;
#pragma unroll
      for (int d = 0; d < 3; d++) s.u[d] = qdiv(f.q[d], masse, irm);
      zero_small(s.u);
      const double ein = qdiv(f.E, masse, irm) - 0.5 * sqn(s.u);
      const double p = eh_pressure(S, masse, ein, s.pm);
      const double T = sat_tsat(S, p, nullptr);
      const double r0 = eos_rho_sat(S.eos[0], p, T, T, nullptr, eps), r1 = eos_rho_sat(S.eos[1], p, T, T, nullptr, eps);
      const double rl = S.liq ? r1 : r0, rv = S.liq ? r0 : r1;
      const double av = (masse - rl) / (rv - rl), al = 1. - av;
      s.r[0] = r0; s.r[1] = r1;
      s.a[0] = S.liq ? av : al;
      s.a[1] = S.liq ? al : av;
      s.pm = p;
      s.T = T;
#pragma unroll
      for (int k = 0; k < NPH; k++) {
        s.p[k] = p;
        s.e[k] = eos_e(S.eos[k], s.r[k], p, eps);
        s.c[k] = eos_c(S.eos[k], s.r[k], p, eps);
      }
      te_mixture<NPH>(s);
      s.E = qdiv(f.E, masse, irm);
    }
  } else if constexpr (MODEL == MODEL_THERMALEQ) {
    // FluxThermalEq::buildPrim (FluxThermalEq.cpp:130-162) with the
    // closed-form equilibrium pressure and temperature of two phases
    // (MixThermalEq::computePressure / computeTemperature, MixThermalEq.cpp:
    // 121-178, Le Martelot 2013)
    double rhoMel = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) rhoMel += f.m[k];
    const double irm = 1. / rhoMel;
#pragma unroll
    for (int d = 0; d < 3; d++) s.u[d] = qdiv(f.q[d], rhoMel, irm);
    zero_small(s.u);
    const double ein = qdiv(f.E, rhoMel, irm) - 0.5 * (s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2]);
    const EosParams &q1 = S.eos[0], &q2 = S.eos[1];
    const double pInf1 = te_pinf(q1), pInf2 = te_pinf(q2);
    const double Y1 = qdiv(f.m[0], rhoMel, irm), Y2 = qdiv(f.m[1], rhoMel, irm);
    const double qm = Y1 * q1.eref + Y2 * q2.eref;
    const double cvMel = Y1 * q1.cv + Y2 * q2.cv;
    const double A1 = Y1 * (q1.gamma - 1.) * q1.cv / cvMel * (rhoMel * (ein - qm) - pInf1);
    const double A2 = Y2 * (q2.gamma - 1.) * q2.cv / cvMel * (rhoMel * (ein - qm) - pInf2);
    const double p = 0.5 * (A1 + A2 - (pInf1 + pInf2)) +
                     sqrt(0.25 * (A2 - A1 - (pInf2 - pInf1)) * (A2 - A1 - (pInf2 - pInf1)) + A1 * A2);
    double T = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double Yk = qdiv(f.m[k], rhoMel, irm);
      T += Yk * (S.eos[k].gamma - 1.) * S.eos[k].cv / (p + te_pinf(S.eos[k]));
    }
    s.T = 1. / (T * rhoMel);
    s.pm = p;
    te_phases<NPH>(S, s);
#pragma unroll
    for (int k = 0; k < NPH; k++) s.a[k] = f.m[k] / s.r[k];
    te_mixture<NPH>(s);
    s.E = qdiv(f.E, rhoMel, irm);
  } else if constexpr (MODEL == MODEL_EULER) {
    s.r[0] = f.m[0];
    const double irm = 1. / f.m[0];
#pragma unroll
    for (int d = 0; d < 3; d++) s.u[d] = qdiv(f.q[d], f.m[0], irm);
    zero_small(s.u);
    s.E = qdiv(f.E, f.m[0], irm);
    s.e[0] = s.E - 0.5 * sqn(s.u);
    s.p[0] = eos_p(S.eos[0], f.m[0], s.e[0]);
    s.c[0] = eos_c(S.eos[0], f.m[0], s.p[0], eps);
  } else {
    double un = 0.;
    if constexpr (MODEL == MODEL_KAPILA) {
      if (eps > 1.e-20) {
#pragma unroll
        for (int k = 0; k < NPH; k++) {
          if (f.a[k] < 0.) f.a[k] = 0.;
          if (f.a[k] > 1.) f.a[k] = 1.;
          un += f.a[k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < NPH; k++) {
          if (f.a[k] <= 1.e-15) f.a[k] = 1e-15;
          if (f.a[k] >= 1. - 1.e-15) f.a[k] = 1.0 - 1e-15;
          un += f.a[k];
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < NPH; k++) {
        if (f.a[k] <= 1.e-10) f.a[k] = 1e-9;
        if (f.a[k] >= 1. - 1.e-10) f.a[k] = 1.0 - 1e-9;
        un += f.a[k];
      }
    }
    const double iun = drcp<FAST>(un);
#pragma unroll
    for (int k = 0; k < NPH; k++) f.a[k] = qdiv(f.a[k], un, iun);
    double rhoMel = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      rhoMel = rhoMel + f.m[k];
      s.a[k] = f.a[k];
      double ein;
      if constexpr (MODEL == MODEL_KAPILA) {
        s.r[k] = ddiv<FAST>(f.m[k], smax(f.a[k], eps));
        ein = ddiv<FAST>(f.en[k], smax(f.m[k], eps));
      } else {
        s.r[k] = ddiv<FAST>(f.m[k], f.a[k]);
        ein = ddiv<FAST>(f.en[k], f.m[k]);
      }
      s.p[k] = eos_p(S.eos[k], s.r[k], ein);
      verify_correct<MODEL, NPH>(S, s, k);
    }
    const double irm = drcp<FAST>(rhoMel);
#pragma unroll
    for (int d = 0; d < 3; d++) s.u[d] = qdiv(f.q[d], rhoMel, irm);
    zero_small(s.u);
    fulfill<MODEL, NPH>(S, s);
    s.E = qdiv(f.E, rhoMel, irm);
  }
}

// FluxMultiP::schemeCorrection (FluxMultiP.cpp:237-260); buf holds U^n.
template <int MODEL, int NPH>
ECO_D void scheme_correction(FluxVec<NPH> &f, const FluxVec<NPH> &buf) {
  if constexpr (MODEL == MODEL_MULTIP) {
    double rhoN = 0., rhoNp1 = 0., sum = 0., dek[NPH];
#pragma unroll
    for (int k = 0; k < NPH; k++) { rhoN += buf.m[k]; rhoNp1 += f.m[k]; dek[k] = f.en[k] - buf.en[k]; }
#pragma unroll
    for (int k = 0; k < NPH; k++) sum += dek[k];
    double dE = f.E - buf.E;
    double dRU2 = sqn(f.q) / rhoNp1 - sqn(buf.q) / rhoN;
    if (fabs(sum) > 1.e-10 * f.E) {
#pragma unroll
      for (int k = 0; k < NPH; k++) f.en[k] = buf.en[k] + dek[k] / sum * (dE - 0.5 * dRU2);
    }
  }
}

// U <- U*mult + U^n (Flux::multiply then addFlux(1.), Cell.cpp:214-227).
template <int MODEL, int NPH>
ECO_D void scale_add(FluxVec<NPH> &U, double mult, const FluxVec<NPH> &buf) {
  if constexpr (MODEL == MODEL_EULER) {
    U.m[0] *= mult;
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) U.m[k] *= mult;
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { U.a[k] *= mult; U.m[k] *= mult; U.en[k] *= mult; }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) U.q[d] *= mult;
  U.E *= mult;
  if constexpr (MODEL == MODEL_EULER) {
    U.m[0] += 1. * buf.m[0];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) U.m[k] += 1. * buf.m[k];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { U.a[k] += 1. * buf.a[k]; U.m[k] += 1. * buf.m[k]; U.en[k] += 1. * buf.en[k]; }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) U.q[d] += buf.q[d] * 1.;
  U.E += 1. * buf.E;
}

// RelaxationP::stiffRelaxation (RelaxationP.cpp:49-111).  Returns false when
// the Newton iteration did not converge in 100 iterations (state untouched,
// as in the reference).
template <int MODEL, int NPH>
ECO_D bool relax_p(const SchemeParams &S, CellState<NPH> &s) {
  if (S.eps > 1.e-20) {
#pragma unroll
    for (int k = 0; k < NPH; k++) if (s.a[k] > (1. - 1.e-5)) return true;
  }
  double ak[NPH], pk[NPH], rk[NPH], rkS[NPH], akS[NPH];
  double pStar = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    verify_correct<MODEL, NPH>(S, s, k);
    ak[k] = s.a[k]; pk[k] = s.p[k]; rk[k] = s.r[k];
    pStar += ak[k] * pk[k];
  }
  int it = 0;
  double f = 0., df = 1.;
  do {
    it++;
    pStar -= f / df;
#pragma unroll
    for (int k = 0; k < NPH; k++) pStar = eos_fix_p(S.eos[k], pStar);
    f = -1.; df = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      double drho;
      rkS[k] = eos_rho_pfinal(S.eos[k], pk[k], rk[k], pStar, drho, S.eps);
      akS[k] = ak[k] * rk[k] / rkS[k];
      double dalpha = ak[k] * rk[k] * drho / (rkS[k] * rkS[k]);
      f += akS[k];
      df -= dalpha;
    }
  } while (fabs(f) > 1e-10 && it < 100);
  if (it < 100) {
#pragma unroll
    for (int k = 0; k < NPH; k++) { s.a[k] = akS[k]; s.r[k] = rkS[k]; s.p[k] = pStar; }
    s.pm = pStar;
    return true;
  }
  return false;
}

// RelaxationPT::stiffRelaxation (RelaxationPT.cpp:49-84) with
// analyticalPressure (88-118) and analyticalTemperature (122-141), two
// phases: from the mixture energy, density and mass fractions of the last
// computeMixtureVariables (s.em, s.rho, s.Y: buildPrim's, before any
// relaxation), the equilibrium pressure and temperature in closed form; the
// phases take them at constant partial densities alpha_k rho_k.
template <int MODEL, int NPH>
ECO_D void relax_pt(const SchemeParams &S, CellState<NPH> &s) {
  if constexpr (NPH == 2 && (MODEL == MODEL_KAPILA || MODEL == MODEL_MULTIP)) {
    if (S.eps > 1.e-20) {
#pragma unroll
      for (int k = 0; k < NPH; k++) if (s.a[k] > (1. - 1.e-5)) return;
    }
    const EosParams &q1 = S.eos[0], &q2 = S.eos[1];
    const double e0 = s.em, rho0 = s.rho;
    const double pInf1 = te_pinf(q1), pInf2 = te_pinf(q2), Y1 = s.Y[0], Y2 = s.Y[1];
    const double qm = Y1 * q1.eref + Y2 * q2.eref;
    const double cvMel = Y1 * q1.cv + Y2 * q2.cv;
    const double A1 = Y1 * (q1.gamma - 1.) * q1.cv / cvMel * (rho0 * (e0 - qm) - pInf1);
    const double A2 = Y2 * (q2.gamma - 1.) * q2.cv / cvMel * (rho0 * (e0 - qm) - pInf2);
    const double pStar = 0.5 * (A1 + A2 - (pInf1 + pInf2)) +
                         sqrt(0.25 * (A2 - A1 - (pInf2 - pInf1)) * (A2 - A1 - (pInf2 - pInf1)) + A1 * A2);
    double T = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) T += s.Y[k] * (S.eos[k].gamma - 1.) * S.eos[k].cv / (pStar + te_pinf(S.eos[k]));
    T = 1. / (T * rho0);
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double rkS = eos_rho_pt(S.eos[k], pStar, T, S.eps);
      s.a[k] = s.a[k] * s.r[k] / rkS;
      s.r[k] = rkS;
      s.p[k] = pStar;
    }
    s.pm = pStar;
  }
}

// RelaxationPTMu::stiffRelaxation (RelaxationPTMu.cpp:51-157), two phases
// (phase 0 the liquid, phase 1 the vapour): from sum alpha_k p_k of s (whose
// phases are verified and corrected), Newton on the pressure at which the
// saturated mixture of density rho holds the internal energy rho*e.  The
// density and energy are the cell's own mixture's and the result goes to the
// cell's own phases (RelaxationPTMu.cpp:47-48, 138-149, no Prim type): out
// is s itself after a time-n update, the time-n state in the predictor.
template <int MODEL, int NPH>
ECO_D void relax_ptmu(const SchemeParams &S, CellState<NPH> &s, CellState<NPH> &out, double rho, double e) {
  if constexpr (NPH == 2 && MODEL == MODEL_KAPILA) {
    double pStar = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      pStar += s.a[k] * s.p[k];
      verify_correct<MODEL, NPH>(S, s, k);
    }
    const double rhoe = rho * e;
    const EosParams &qL = S.eos[0], &qV = S.eos[1];
    double rL = 0., rV = 0., aL = 0., aV = 0.;
    double f = 0., df = 1.;
    int it = 0;
    do {
      pStar -= f / df;
      if (++it > 50) break;
#pragma unroll
      for (int k = 0; k < NPH; k++) pStar = eos_fix_p(S.eos[k], pStar);
      double dT, drL, drV, dreL, dreV;
      const double T = sat_tsat(S, pStar, &dT);
      rL = eos_rho_sat(qL, pStar, T, dT, &drL, S.eps);
      rV = eos_rho_sat(qV, pStar, T, dT, &drV, S.eps);
      if (rL <= rho) { rL = rho + 1e-6; drL = 0.; }
      if (rV >= rho) { rV = rho - 1e-6; drV = 0.; }
      aL = (rho - rV) / (rL - rV);
      const double daL = (-drV * (rL - rV) - (rho - rV) * (drL - drV)) / ((rL - rV) * (rL - rV));
      if (aL <= 0.) aL = 1e-8;
      if (aL >= 1.) aL = 1. - 1e-8;
      aV = 1. - aL;
      const double daV = -daL;
      f = rhoe; df = 0.;
      const double reL = eos_rhoe_sat(qL, pStar, rL, drL, &dreL);
      const double reV = eos_rhoe_sat(qV, pStar, rV, drV, &dreV);
      f -= (aL * reL + aV * reV);
      df -= (daL * reL + aL * dreL + daV * reV + aV * dreV);
      f /= rhoe;
      df /= rhoe;
    } while (fabs(f) > 1e-10);
    out.a[0] = aL; out.r[0] = rL; out.p[0] = pStar;
    out.a[1] = aV; out.r[1] = rV; out.p[1] = pStar;
  }
}

// FluxKapila::correctionEnergy (FluxKapila.cpp:239-268), after
// Mixture::totalEnergyToInternalEnergy (MixKapila.cpp:188-195).  s.rho is the
// mixture density cached by buildPrim (before relaxation).
template <int MODEL, int NPH, bool FAST = false>
ECO_D void correction_energy(const SchemeParams &S, CellState<NPH> &s) {
  if constexpr (MODEL == MODEL_KAPILA) {
    s.em = s.E - 0.5 * sqn(s.u);
    double rhoe = s.rho * s.em;
    double p = rhoe, denom = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      p -= s.a[k] * (S.eos[k].gpg + s.r[k] * S.eos[k].eref);
      denom += s.a[k] * S.eos[k].og;
    }
    p = ddiv<FAST>(p, denom);
#pragma unroll
    for (int k = 0; k < NPH; k++) { s.p[k] = p; verify_correct<MODEL, NPH>(S, s, k); }
  }
}

// ----------------------------------------------------------------------------
// HLLC in the face frame
// ----------------------------------------------------------------------------
// ModKapila::solveRiemannIntern (ModKapila.cpp:107-218) == ModMultiP (ModMultiP.cpp:101-214).
// em_* is the mixture internal energy the reference reads (ΣY_k e_k after
// fulfillState, or E - |u|^2/2 for MultiP first-order interior faces).
// Returns max(|sL|,|sR|) among speeds above the 1e-3 guard (0 if none).
// Eos{IG,SG}::computeDensityIsentropic (EosIG.cpp:146-151, EosSG.cpp:149-154)
// and computeDensityHugoniot (EosIG.cpp:155-162, EosSG.cpp:158-165)
ECO_D double eos_rho_isentropic(const EosParams &q, double p0, double rho0, double pf, double eps) {
  if (q.type == EOS_SG) return rho0 * pow_pos((pf + q.pinf) / smax((p0 + q.pinf), eps), 1. / q.gamma);
  return rho0 * pow_pos(pf / smax(p0, eps), 1. / q.gamma);
}
ECO_D double eos_rho_hugoniot(const EosParams &q, double p0, double rho0, double pf, double eps) {
  double num, denom;
  if (q.type == EOS_SG) {
    num = (q.gamma + 1.) * (pf + q.pinf) + (q.gamma - 1.) * (p0 + q.pinf);
    denom = (q.gamma - 1.) * (pf + q.pinf) + (q.gamma + 1.) * (p0 + q.pinf);
  } else {
    num = (q.gamma + 1.) * pf + (q.gamma - 1.) * p0;
    denom = (q.gamma - 1.) * pf + (q.gamma + 1.) * p0;
  }
  return rho0 * num / smax(denom, eps);
}

// The outflow's half Riemann problem (imposed pressure p0) in the face frame:
// ModKapila::solveRiemannOutflow (ModKapila.cpp:479-549, isentropic left
// wave), ModMultiP::solveRiemannOutflow (ModMultiP.cpp:468-538, Hugoniot left
// wave), ModEuler::solveRiemannOutflow (ModEuler.cpp:427-475).  Where the
// MultiP form leaves mL unset (|v* - 1/rhoL| <= 1e-10) it is rhoL cL, as in
// the Kapila form.  Returns |sL| when above the 1e-3 guard, else 0.
// MixThermalEq::computeTemperatureIsentrope / computeVolumeIsentrope
// (MixThermalEq.cpp:200-226, 256-277): the mixture isentrope at fixed mass
// fractions Y from (p0, T0) to p, as a specific volume
template <int NPH>
ECO_D double te_v_isentrope(const SchemeParams &S, const double *Y, double p0, double T0, double p) {
  double T = T0, cM = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) cM += Y[k] * S.eos[k].gamma * S.eos[k].cv;
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    const EosParams &q = S.eos[k];
    const double puissance = (q.gamma - 1.) * Y[k] * q.cv / cM;
    T *= pow_pos((p + te_pinf(q)) / ((p0 + te_pinf(q))), puissance);
  }
  double v = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    const EosParams &q = S.eos[k];
    v += Y[k] * (((q.gamma - 1.) * q.cv * T) / (p + te_pinf(q)));
  }
  return v;
}

template <int MODEL, int NPH>
ECO_D double outflow_flux(const SchemeParams &S, const CellState<NPH> &L, double p0, FluxVec<NPH> &F) {
  const double eps = S.eps;
#pragma unroll
  for (int k = 0; k < NPH; k++) { F.a[k] = 0.; F.m[k] = 0.; F.en[k] = 0.; }
  double sL;
  if constexpr (MODEL == MODEL_THERMALEQ) {
    // ModThermalEq::solveRiemannOutflow (ModThermalEq.cpp:372-424)
    const double uL = L.u[0], cL = L.cF, pL = L.pm, rhoL = L.rho, TL = L.T, vL = L.u[1], wL = L.u[2];
    double Y[NPH];
#pragma unroll
    for (int k = 0; k < NPH; k++) Y[k] = L.a[k] * L.r[k] / rhoL;
    const double zL = rhoL * cL;
    double pStar = p0;
    double rhoStar = 1. / te_v_isentrope<NPH>(S, Y, pL, TL, p0);
    const double vmv0 = 1. / rhoStar - 1. / rhoL;
    double mL = zL;
    if (fabs(vmv0) > 1e-10) mL = sqrt((pL - p0) / vmv0);
    sL = uL - mL / rhoL;
    double uStar = uL + mL * vmv0;
    if (sL >= 0.) { uStar = uL; pStar = pL; rhoStar = rhoL; }
    const double EStar = (L.em + 0.5 * sqn(L.u)) + (uStar - uL) * (uStar - pL / mL);
#pragma unroll
    for (int k = 0; k < NPH; k++) F.m[k] = rhoStar * Y[k] * uStar;
    F.q[0] = uStar * uStar * rhoStar + pStar;
    F.q[1] = uStar * vL * rhoStar;
    F.q[2] = uStar * wL * rhoStar;
    F.E = (EStar * rhoStar + pStar) * uStar;
    F.sM = uStar;
    (void)eps;
  } else if constexpr (MODEL == MODEL_EULER) {
    const double uL = L.u[0], vL = L.u[1], wL = L.u[2], pL = L.p[0], rhoL = L.r[0], cL = L.c[0];
    const double zL = rhoL * cL;
    sL = uL - cL;
    double pStar = p0;
    double rhoStar = eos_rho_hugoniot(S.eos[0], pL, rhoL, pStar, 0.);
    double uStar = (pL + zL * uL - pStar) / zL;
    if (sL >= 0.) { uStar = uL; pStar = pL; rhoStar = rhoL; }
    else if (uStar < 0) { rhoStar = rhoL; }
    const double eStar = eos_e(S.eos[0], rhoStar, pStar, 0.);
    F.m[0] = rhoStar * uStar;
    F.q[0] = rhoStar * uStar * uStar + pStar; F.q[1] = rhoStar * uStar * vL; F.q[2] = rhoStar * uStar * wL;
    F.E = (rhoStar * (eStar + 0.5 * (uStar * uStar + vL * vL + wL * wL)) + pStar) * uStar;
    F.sM = uStar;
  } else {
    constexpr bool KAP = (MODEL == MODEL_KAPILA);
    const double uL = L.u[0], cL = L.cF, pL = L.pm, rhoL = L.rho;
    const double uyL = L.u[1], uzL = L.u[2], EL = L.em + 0.5 * sqn(L.u);
    const double zL = rhoL * cL;
    double pStar = p0, EStar, vStar = 0., uStar, rk[NPH];
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      if constexpr (KAP) {
        rk[k] = eos_rho_isentropic(S.eos[k], pL, L.r[k], pStar, eps);
        vStar += L.a[k] * L.r[k] / rhoL / smax(rk[k], eps);
      } else {
        rk[k] = eos_rho_hugoniot(S.eos[k], pL, L.r[k], pStar, eps);
        vStar += L.a[k] * L.r[k] / rhoL / rk[k];
      }
    }
    const double vSmvL = vStar - 1. / rhoL;
    double mL = zL;
    if (fabs(vSmvL) > 1e-10) mL = sqrt((pL - pStar) / vSmvL);
    sL = uL - mL / rhoL;
    const double sM = uL + mL * vSmvL;
    if (sL >= 0. || sM < 0) {  // supersonic outflow, or inflow: the left state
      uStar = uL; pStar = pL;
#pragma unroll
      for (int k = 0; k < NPH; k++) rk[k] = L.r[k];
      vStar = 1. / rhoL;
      EStar = EL;
    } else {
      uStar = sM;
      EStar = EL + (uStar - uL) * (uStar - pL / mL);
    }
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double YkL = L.a[k] * L.r[k] / rhoL;
      const double ekStar = eos_e(S.eos[k], rk[k], pStar, eps);
      F.a[k] = KAP ? YkL / smax(rk[k], eps) / vStar * uStar : YkL / rk[k] / vStar * uStar;
      F.m[k] = F.a[k] * rk[k];
      F.en[k] = F.m[k] * ekStar;
    }
    F.q[0] = uStar * uStar / vStar + pStar; F.q[1] = uStar * uyL / vStar; F.q[2] = uStar * uzL / vStar;
    F.E = (EStar / vStar + pStar) * uStar;
    F.sM = uStar;
  }
  return (fabs(sL) > 1.e-3) ? fabs(sL) : 0.;
}

// Eos{IG,SG}::vfpfh, dvdpch, dvdhcp (EosIG.cpp:220-239, EosSG.cpp:223-242):
// specific volume from pressure and enthalpy, and its two partial derivatives.
ECO_D double eos_vfpfh(const EosParams &q, double p, double h, double eps) {
  const double pp = (q.type == EOS_SG) ? q.gamma * (p + q.pinf) : q.gamma * p;
  return q.gm1 * (h - q.eref) / smax(pp, eps);
}
ECO_D double eos_dvdpch(const EosParams &q, double p, double h, double eps) {
  const double pp = (q.type == EOS_SG) ? (p + q.pinf) * (p + q.pinf) : p * p;
  return (1. - q.gamma) / q.gamma * (h - q.eref) / smax(pp, eps);
}
ECO_D double eos_dvdhcp(const EosParams &q, double p, double eps) {
  const double pp = (q.type == EOS_SG) ? (p + q.pinf) : p;
  return q.gm1 / q.gamma / smax(pp, eps);
}

// The injection's half Riemann problem in the face frame (BoundCondInj): a
// fluid of phase total enthalpies H_k0 and mass fractions Y_k0 enters at the
// mass flow m0 per unit area (negative: into the domain, the reference's sign
// after reading, BoundCondInj.cpp:62); the left wave is acoustic.
// ModKapila::solveRiemannInflow (ModKapila.cpp:253-322) == ModMultiP
// (ModMultiP.cpp:249-318); ModEuler::solveRiemannInflow (ModEuler.cpp:215-303)
// is the one-phase form with H0 in H[0].  Newton on the pressure from pL until
// |f| <= 1e-10; where the reference stops after 50 iterations, the 50th
// iterate is kept and `nonconv` set.  Returns |sL| above the 1e-3 guard, else 0.
template <int MODEL, int NPH>
ECO_D double inflow_flux(const SchemeParams &S, const CellState<NPH> &L, double m0, const double *Y,
                         const double *H, FluxVec<NPH> &F, bool &nonconv) {
  const double eps = S.eps;
#pragma unroll
  for (int k = 0; k < NPH; k++) { F.a[k] = 0.; F.m[k] = 0.; F.en[k] = 0.; }
  double uL, vL, wL, pL, rhoL, cL;
  if constexpr (MODEL == MODEL_EULER) {
    uL = L.u[0]; vL = L.u[1]; wL = L.u[2]; pL = L.p[0]; rhoL = L.r[0]; cL = L.c[0];
  } else {
    uL = L.u[0]; vL = L.u[1]; wL = L.u[2]; pL = L.pm; rhoL = L.rho; cL = L.cF;
  }
  const double sL = uL - cL;
  const double zL = rhoL * cL;
  const double e = (MODEL == MODEL_EULER) ? 0. : eps;
  double p = pL, f = 0., df = 1., u = 0., v = 0., vk[NPH];
  int it = 0;
  do {
    p -= f / df;
    if (++it > 50) { nonconv = true; break; }
#pragma unroll
    for (int k = 0; k < NPH; k++) p = eos_fix_p(S.eos[k], p);
    u = uL + (pL - p) / zL;
    if (u >= -1e-6) u = -1e-6;
    const double du = -1. / zL;
    v = u / m0;
    f = v;
    df = du / m0;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double hk = H[k] - 0.5 * u * u;
      vk[k] = eos_vfpfh(S.eos[k], p, hk, e);
      const double dvk = eos_dvdpch(S.eos[k], p, hk, e) - eos_dvdhcp(S.eos[k], p, e) * u * du;
      if constexpr (MODEL == MODEL_EULER) { f -= vk[k]; df -= dvk; }
      else { f -= Y[k] * vk[k]; df -= Y[k] * dvk; }
    }
  } while (fabs(f) > 1e-10);
  if constexpr (MODEL == MODEL_EULER) {
    const double rhoStar = m0 / u;
    const double eStar = eos_e(S.eos[0], rhoStar, p, 0.);
    F.m[0] = rhoStar * u;
    F.q[0] = rhoStar * u * u + p; F.q[1] = rhoStar * u * vL; F.q[2] = rhoStar * u * wL;
    F.E = (rhoStar * (eStar + 0.5 * (u * u + vL * vL + wL * wL)) + p) * u;
  } else {
    double Estar = 0.5 * (u * u + vL * vL + wL * wL);
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double rhok = 1. / vk[k];
      const double ek = eos_e(S.eos[k], rhok, p, eps);
      Estar += Y[k] * ek;
      F.a[k] = Y[k] * vk[k] / v * u;
      F.m[k] = F.a[k] * rhok;
      F.en[k] = F.a[k] * rhok * ek;
    }
    F.q[0] = u * u / v + p; F.q[1] = u * vL / v; F.q[2] = u * wL / v;
    F.E = (Estar / v + p) * u;
  }
  F.sM = u;
  return (fabs(sL) > 1.e-3) ? fabs(sL) : 0.;
}

// computeDensityIsentropic / computeDensityHugoniot with drho/dp
// (EosIG.cpp:146-162, EosSG.cpp:149-165) and computeEnthalpyIsentropic with
// dh/dp (EosIG.cpp:177-184, EosSG.cpp:180-187).
ECO_D double eos_rho_isentropic_d(const EosParams &q, double p0, double rho0, double pf, double eps, double &drho) {
  const double rf = eos_rho_isentropic(q, p0, rho0, pf, eps);
  drho = rf / smax(q.gamma * (pf + q.pinf), eps);
  return rf;
}
ECO_D double eos_rho_hugoniot_d(const EosParams &q, double p0, double rho0, double pf, double eps, double &drho) {
  const double denom = (q.gamma - 1.) * (pf + q.pinf) + (q.gamma + 1.) * (p0 + q.pinf);
  drho = rho0 * 4. * q.gamma * (p0 + q.pinf) / smax(denom * denom, eps);
  return eos_rho_hugoniot(q, p0, rho0, pf, eps);
}
ECO_D double eos_h_isentropic(const EosParams &q, double p0, double rho0, double pf, double eps, double &dh) {
  double drho;
  const double rf = eos_rho_isentropic_d(q, p0, rho0, pf, eps, drho);
  const double pp = pf + q.pinf;
  dh = q.gamma / q.gm1 * (rf - pp * drho) / smax(rf * rf, eps);
  return q.gamma * pp / q.gm1 / smax(rf, eps) + q.eref;
}

// The tank's half Riemann problem in the face frame (BoundCondTank): from the
// left wave at p* = p0, supersonic outflow keeps the left state and subsonic
// outflow the left star state at p0; otherwise fluid flows in from the tank
// (at rest, total enthalpy H0, isentropic) and a Newton solve on p* matches
// the tank-side velocity to the left star velocity.
// ModKapila::solveRiemannTank (ModKapila.cpp:326-475: Hugoniot left wave,
// |f| <= 1e-2), ModMultiP::solveRiemannTank (ModMultiP.cpp:322-465: |f| <=
// 1e-3, no epsilon guards, the first estimate's sM as the contact speed),
// ModEuler::solveRiemannTank (ModEuler.cpp:305-423: isentropic left wave,
// |f| <= 1e-3).  Every iterate's left wave enters the dt bound; where the
// reference stops after 50 iterations, the last iterate is kept and `nonconv`
// set.  The tank's rho_k0, Y_k0 and H0 come from the host.  Returns the
// largest |sL| above the 1e-3 guard, else 0.
template <int MODEL, int NPH>
ECO_D double tank_flux(const SchemeParams &S, const CellState<NPH> &L, double p0, const double *rhok0,
                       const double *Yk0, double H0, FluxVec<NPH> &F, bool &nonconv) {
#pragma unroll
  for (int k = 0; k < NPH; k++) { F.a[k] = 0.; F.m[k] = 0.; F.en[k] = 0.; }
  double smx = 0.;
  auto wave = [&](double sL) { if (fabs(sL) > 1.e-3) smx = smax(smx, fabs(sL)); };
  if constexpr (MODEL == MODEL_EULER) {
    const EosParams &q = S.eos[0];
    const double uL = L.u[0], vL = L.u[1], wL = L.u[2], pL = L.p[0], rhoL = L.r[0], cL = L.c[0];
    const double zL = rhoL * cL;
    double uStar, rhoStar, pStar = p0, vStar, wStar, mL, drho;
    double v = 1. / eos_rho_isentropic(q, pL, rhoL, pStar, 0.);
    double vmv0 = v - 1. / rhoL;
    mL = (fabs(vmv0) > 1e-10) ? sqrt((pL - p0) / vmv0) : zL;
    double sL = uL - mL / rhoL;
    wave(sL);
    const double u = uL + mL * vmv0;
    if (sL >= 0.) { uStar = uL; pStar = pL; rhoStar = rhoL; vStar = vL; wStar = wL; }
    else if (u >= -1e-3) { uStar = u; pStar = p0; rhoStar = 1. / v; vStar = vL; wStar = wL; }
    else {
      int it = 0;
      double p = 0.5 * p0, f = 0., df = 1., dh;
      double uStarR = 0., uStarL = 0.;
      rhoStar = 0.;
      do {
        p -= f / df;
        if (++it > 50) { nonconv = true; break; }
        p = eos_fix_p(q, p);
        if (p > p0) p = p0 - 1e-6;
        rhoStar = eos_rho_isentropic(q, p0, rhok0[0], p, 0.);
        const double h = eos_h_isentropic(q, p0, rhok0[0], p, 0., dh);
        uStarR = -sqrt(2. * (H0 - h));
        const double duStarR = -dh / uStarR;
        v = 1.0 / eos_rho_isentropic_d(q, pL, rhoL, p, 0., drho);
        const double dv = -v * v * drho;
        vmv0 = v - 1. / rhoL;
        double dmL = 0.;
        if (fabs(vmv0) > 1e-10) {
          mL = sqrt((pL - p) / vmv0);
          dmL = 0.5 * (-vmv0 + (p - pL) * dv) / (vmv0 * vmv0) / mL;
        } else {
          mL = zL;
        }
        sL = uL - mL / rhoL;
        wave(sL);
        uStarL = uL + mL * vmv0;
        const double duStarL = dmL * vmv0 + mL * dv;
        f = uStarR - uStarL;
        df = duStarR - duStarL;
      } while (fabs(f) > 1e-3);
      pStar = p;
      uStar = 0.5 * (uStarL + uStarR);
      vStar = 0.; wStar = 0.;
    }
    const double eStar = eos_e(q, rhoStar, pStar, 0.);
    F.m[0] = rhoStar * uStar;
    F.q[0] = rhoStar * uStar * uStar + pStar; F.q[1] = rhoStar * uStar * vStar; F.q[2] = rhoStar * uStar * wStar;
    F.E = (rhoStar * (eStar + 0.5 * (uStar * uStar + vStar * vStar + wStar * wStar)) + pStar) * uStar;
    F.sM = uStar;
  } else {
    constexpr bool KAP = (MODEL == MODEL_KAPILA);
    const double eps = S.eps;
    auto gd = [&](double x) { return KAP ? smax(x, eps) : x; };  // the Kapila form's epsilon guards
    const double uL = L.u[0], cL = L.cF, pL = L.pm, rhoL = L.rho;
    const double uyL = L.u[1], uzL = L.u[2];
    const double zL = rhoL * cL;
    double rkS[NPH], YkS[NPH], pStar = p0, vStar = 0., uStar, rhoStar, uyStar, uzStar, mL;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      rkS[k] = eos_rho_hugoniot(S.eos[k], L.p[k], L.r[k], pStar, eps);
      vStar += L.a[k] * L.r[k] / rhoL / gd(rkS[k]);
    }
    double vmv0 = vStar - 1. / rhoL;
    mL = (fabs(vmv0) > 1e-10) ? sqrt((pL - pStar) / vmv0) : zL;
    double sL = uL - mL / rhoL;
    wave(sL);
    const double sM = uL + mL * vmv0;
    if (sL >= 0.) {
      uStar = uL; pStar = pL;
#pragma unroll
      for (int k = 0; k < NPH; k++) { rkS[k] = L.r[k]; YkS[k] = L.a[k] * L.r[k] / rhoL; }
      rhoStar = rhoL; uyStar = uyL; uzStar = uzL;
    } else if (sM >= -1e-3) {
      uStar = sM; pStar = p0;
#pragma unroll
      for (int k = 0; k < NPH; k++) YkS[k] = L.a[k] * L.r[k] / rhoL;
      rhoStar = 1. / vStar; uyStar = uyL; uzStar = uzL;
    } else {
      int it = 0;
      double p = 0.5 * p0, f = 0., df = 1.;
      double uStarR = 0., uStarL = 0.;
      do {
        p -= f / df;
        if (++it > 50) { nonconv = true; break; }
#pragma unroll
        for (int k = 0; k < NPH; k++) p = eos_fix_p(S.eos[k], p);
        if (p > p0) p = p0 - 1e-6;
        uStarR = H0;
        double duStarR = 0.;
#pragma unroll
        for (int k = 0; k < NPH; k++) {
          double dhk;
          rkS[k] = eos_rho_isentropic(S.eos[k], p0, rhok0[k], p, eps);
          const double hk = eos_h_isentropic(S.eos[k], p0, rhok0[k], p, eps, dhk);
          uStarR -= Yk0[k] * hk;
          duStarR -= Yk0[k] * dhk;
        }
        uStarR = -sqrt(2. * uStarR);
        duStarR = duStarR / uStarR;
        double vStarL = 0., dvStarL = 0.;
#pragma unroll
        for (int k = 0; k < NPH; k++) {
          double drhok;
          const double rhok = eos_rho_hugoniot_d(S.eos[k], L.p[k], L.r[k], p, eps, drhok);
          const double YkL = L.a[k] * L.r[k] / rhoL;
          vStarL += YkL / gd(rhok);
          dvStarL -= YkL / gd(rhok * rhok) * drhok;
        }
        vmv0 = vStarL - 1. / rhoL;
        double dmL = 0.;
        if (fabs(vmv0) > 1e-10) {
          mL = sqrt((pL - p) / vmv0);
          dmL = 0.5 * (-vmv0 + (p - pL) * dvStarL) / (vmv0 * vmv0) / mL;
        } else {
          mL = zL;
        }
        sL = uL - mL / rhoL;
        wave(sL);
        uStarL = uL + mL * vmv0;
        const double duStarL = dmL * vmv0 + mL * dvStarL;
        f = uStarR - uStarL;
        df = duStarR - duStarL;
      } while (fabs(f) > (KAP ? 1e-2 : 1e-3));
      pStar = p;
      uStar = 0.5 * (uStarL + uStarR);
      rhoStar = 0.;
#pragma unroll
      for (int k = 0; k < NPH; k++) { YkS[k] = Yk0[k]; rhoStar += YkS[k] / gd(rkS[k]); }
      rhoStar = 1. / rhoStar;
      uyStar = 0.; uzStar = 0.;
    }
    double EStar = 0.5 * (uStar * uStar + uyStar * uyStar + uzStar * uzStar);
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double ek = eos_e(S.eos[k], rkS[k], pStar, eps);
      EStar += YkS[k] * ek;
      F.a[k] = YkS[k] * rhoStar / gd(rkS[k]) * uStar;
      F.m[k] = F.a[k] * rkS[k];
      F.en[k] = F.m[k] * ek;
    }
    F.q[0] = rhoStar * uStar * uStar + pStar; F.q[1] = rhoStar * uStar * uyStar; F.q[2] = rhoStar * uStar * uzStar;
    F.E = (rhoStar * EStar + pStar) * uStar;
    F.sM = KAP ? uStar : sM;
  }
  return smx;
}

// The wall's half Riemann problem in the face frame: ModKapila::solveRiemannWall
// (ModKapila.cpp:224-249), ModMultiP::solveRiemannWall (ModMultiP.cpp:220-245),
// ModEuler::solveRiemannWall (ModEuler.cpp:181-210).  Only the normal momentum
// flux p* is nonzero and sM = 0; the products keep each model's operand order.
// Returns |sL| when above the 1e-3 guard of the dt estimate, else 0.
template <int MODEL, int NPH>
ECO_D double wall_flux(const CellState<NPH> &L, FluxVec<NPH> &F) {
#pragma unroll
  for (int k = 0; k < NPH; k++) { F.a[k] = 0.; F.m[k] = 0.; F.en[k] = 0.; }
  F.q[1] = 0.; F.q[2] = 0.; F.E = 0.; F.sM = 0.;
  double sL;
  if constexpr (MODEL == MODEL_EULER) {
    const double uL = L.u[0], pL = L.p[0], rhoL = L.r[0], cL = L.c[0];
    sL = smin(uL - cL, -uL - cL);
    F.q[0] = rhoL * uL * (uL - sL) + pL;
  } else {
    const double uL = L.u[0], pL = L.pm, rhoL = L.rho, cL = L.cF;
    sL = smin(uL - cL, -uL - cL);
    F.q[0] = rhoL * (uL - sL) * uL + pL;
  }
  return (fabs(sL) > 1.e-3) ? fabs(sL) : 0.;
}

// ModThermalEq::solveRiemannIntern (ModThermalEq.cpp:94-210): the mixture
// HLLC carrying the phases' partial densities.
template <int NPH>
ECO_D double hllc_te(const CellState<NPH> &L, const CellState<NPH> &R, FluxVec<NPH> &F) {
  double uL = L.u[0], cL = L.cF, pL = L.pm, rhoL = L.rho;
  double uR = R.u[0], cR = R.cF, pR = R.pm, rhoR = R.rho;
  double sL = smin(uL - cL, uR - cR);
  double sR = smax(uR + cR, uL + cL);
  double smx = 0.;
  if (fabs(sL) > 1.e-3) smx = fabs(sL);
  if (fabs(sR) > 1.e-3) smx = smax(smx, fabs(sR));
  double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  if (sL >= 0. || sR <= 0.) {
    const bool left = (sL >= 0.);
    const CellState<NPH> &X = left ? L : R;
    double uS = left ? uL : uR, pS = left ? pL : pR, rhoS = left ? rhoL : rhoR;
#pragma unroll
    for (int k = 0; k < NPH; k++) F.m[k] = X.a[k] * X.r[k] * uS;
    double Et = X.em + 0.5 * sqn(X.u);
    F.q[0] = rhoS * uS * uS + pS;
    F.q[1] = rhoS * X.u[1] * uS;
    F.q[2] = rhoS * X.u[2] * uS;
    F.E = (rhoS * Et + pS) * uS;
  } else {
    const bool left = (sM >= 0.);
    const CellState<NPH> &X = left ? L : R;
    double uS = left ? uL : uR, pS = left ? pL : pR, mS = left ? mL : mR, sS = left ? sL : sR;
    double Et = X.em + 0.5 * sqn(X.u);
    const double dS = sS - sM, idS = 1. / dS;
    double rhoStar = qdiv(mS, dS, idS);
    double EStar = Et + (sM - uS) * (sM + pS / mS);
    double pStar = mS * (sM - uS) + pS;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const double mk = X.a[k] * X.r[k] * (sS - uS);
      F.m[k] = qdiv(mk, dS, idS) * sM;
    }
    F.q[0] = rhoStar * sM * sM + pStar;
    F.q[1] = rhoStar * X.u[1] * sM;
    F.q[2] = rhoStar * X.u[2] * sM;
    F.E = (rhoStar * EStar + pStar) * sM;
  }
  F.sM = sM;
  return smx;
}

template <int MODEL, int NPH>
ECO_D double hllc(const SchemeParams &S, const CellState<NPH> &L, const CellState<NPH> &R, FluxVec<NPH> &F,
                  const double *tab = nullptr) {
  const double eps = S.eps;
  if constexpr (MODEL == MODEL_EULERHOM) {
    // ModEulerHomogeneous::solveRiemannIntern (ModEulerHomogeneous.cpp:60-136):
    // HLLC on the mixture (its energy plus |u|^2/2 and the sound speed
    // sqrt(sum Y_k c_k^2)), strict supersonic branches; the mass flux in
    // phase 0's slot
    (void)eps;
    const double uL = L.u[0], vL = L.u[1], wL = L.u[2], cL = L.cF, pL = L.pm, rhoL = L.rho;
    const double uR = R.u[0], vR = R.u[1], wR = R.u[2], cR = R.cF, pR = R.pm, rhoR = R.rho;
    const double EL = L.em + 0.5 * sqn(L.u), ER = R.em + 0.5 * sqn(R.u);
    const double sL = smin(uL - cL, uR - cR);
    const double sR = smax(uR + cR, uL + cL);
    double smx = 0.;
    if (fabs(sL) > 1.e-3) smx = fabs(sL);
    if (fabs(sR) > 1.e-3) smx = smax(smx, fabs(sR));
    const double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
    double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
    if (fabs(sM) < 1.e-8) sM = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) { F.a[k] = 0.; F.m[k] = 0.; F.en[k] = 0.; }
    if (sL > 0.) {
      F.m[0] = rhoL * uL;
      F.q[0] = rhoL * uL * uL + pL; F.q[1] = rhoL * vL * uL; F.q[2] = rhoL * wL * uL;
      F.E = (rhoL * EL + pL) * uL;
    } else if (sR < 0.) {
      F.m[0] = rhoR * uR;
      F.q[0] = rhoR * uR * uR + pR; F.q[1] = rhoR * vR * uR; F.q[2] = rhoR * wR * uR;
      F.E = (rhoR * ER + pR) * uR;
    } else {
      const bool left = (sM >= 0.);
      const double mS = left ? mL : mR, uS = left ? uL : uR, pS = left ? pL : pR, sS = left ? sL : sR;
      const double ES = left ? EL : ER, vS = left ? vL : vR, wS = left ? wL : wR;
      const double pStar = mS * (sM - uS) + pS;
      const double rhoStar = mS / (sS - sM);
      const double Estar = ES + (sM - uS) * (sM + pS / mS);
      F.m[0] = rhoStar * sM;
      F.q[0] = rhoStar * sM * sM + pStar; F.q[1] = rhoStar * sM * vS; F.q[2] = rhoStar * sM * wS;
      F.E = (rhoStar * Estar + pStar) * sM;
    }
    F.sM = sM;
    return smx;
  } else if constexpr (MODEL == MODEL_THERMALEQ) {
    (void)eps;
    return hllc_te<NPH>(L, R, F);
  } else if constexpr (MODEL == MODEL_EULER) {
    // ModEuler::solveRiemannIntern (ModEuler.cpp:90-175)
    double uL = L.u[0], vL = L.u[1], wL = L.u[2], pL = L.p[0], rhoL = L.r[0], cL = L.c[0], EL = L.E;
    double uR = R.u[0], vR = R.u[1], wR = R.u[2], pR = R.p[0], rhoR = R.r[0], cR = R.c[0], ER = R.E;
    double sL = smin(uL - cL, uR - cR);
    double sR = smax(uR + cR, uL + cL);
    double smx = 0.;
    if (fabs(sL) > 1.e-3) smx = fabs(sL);
    if (fabs(sR) > 1.e-3) smx = smax(smx, fabs(sR));
    double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
    double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
    if (fabs(sM) < 1.e-8) sM = 0.;
    if (sL > 0.) {
      F.m[0] = rhoL * uL;
      F.q[0] = rhoL * uL * uL + pL; F.q[1] = rhoL * vL * uL; F.q[2] = rhoL * wL * uL;
      F.E = (rhoL * EL + pL) * uL;
    } else if (sR < 0.) {
      F.m[0] = rhoR * uR;
      F.q[0] = rhoR * uR * uR + pR; F.q[1] = rhoR * vR * uR; F.q[2] = rhoR * wR * uR;
      F.E = (rhoR * ER + pR) * uR;
    } else {
      const bool left = (sM >= 0.);
      double mS = left ? mL : mR, uS = left ? uL : uR, pS = left ? pL : pR, sS = left ? sL : sR;
      double ES = left ? EL : ER, vS = left ? vL : vR, wS = left ? wL : wR;
      double pStar = mS * (sM - uS) + pS;
      double rhoStar = mS / (sS - sM);
      double Estar = ES + (sM - uS) * (sM + pS / mS);
      F.m[0] = rhoStar * sM;
      F.q[0] = rhoStar * sM * sM + pStar; F.q[1] = rhoStar * sM * vS; F.q[2] = rhoStar * sM * wS;
      F.E = (rhoStar * Estar + pStar) * sM;
    }
    F.sM = sM;
    return smx;
  } else {
    double uL = L.u[0], cL = L.cF, pL = L.pm, rhoL = L.rho;
    double uR = R.u[0], cR = R.cF, pR = R.pm, rhoR = R.rho;
    double sL = smin(uL - cL, uR - cR);
    double sR = smax(uR + cR, uL + cL);
    double smx = 0.;
    if (fabs(sL) > 1.e-3) smx = fabs(sL);
    if (fabs(sR) > 1.e-3) smx = smax(smx, fabs(sR));
    double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
    double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
    if (fabs(sM) < 1.e-8) sM = 0.;
    if (sL >= 0. || sR <= 0.) {
      const bool left = (sL >= 0.);
      const CellState<NPH> &X = left ? L : R;
      double uS = left ? uL : uR, pS = left ? pL : pR, rhoS = left ? rhoL : rhoR;
#pragma unroll
      for (int k = 0; k < NPH; k++) {
        F.a[k] = X.a[k] * sM;
        F.m[k] = X.a[k] * X.r[k] * uS;
        F.en[k] = X.a[k] * X.r[k] * X.e[k] * uS;
      }
      double Et = X.em + 0.5 * sqn(X.u);
      F.q[0] = rhoS * uS * uS + pS;
      F.q[1] = rhoS * X.u[1] * uS;
      F.q[2] = rhoS * X.u[2] * uS;
      F.E = (rhoS * Et + pS) * uS;
    } else {
      const bool left = (sM >= 0.);
      const CellState<NPH> &X = left ? L : R;
      double uS = left ? uL : uR, pS = left ? pL : pR, mS = left ? mL : mR, sS = left ? sL : sR;
      double Et = X.em + 0.5 * sqn(X.u);
      const double dS = sS - sM, idS = 1. / dS;
      double rhoStar = qdiv(mS, dS, idS);
      double EStar = Et + (sM - uS) * (sM + pS / mS);
      double pStar = mS * (sM - uS) + pS;
#pragma unroll
      for (int k = 0; k < NPH; k++) {
        double mk = X.r[k] * (sS - uS);
        double rks = qdiv(mk, dS, idS);
        double pks = eos_isentrope(S.eos[k], X.p[k], X.r[k], rks, eps, tab);
        double eks = eos_e(S.eos[k], rks, pks, eps);
        F.a[k] = X.a[k] * sM;
        F.m[k] = X.a[k] * rks * sM;
        F.en[k] = X.a[k] * rks * eks * sM;
      }
      F.q[0] = rhoStar * sM * sM + pStar;
      F.q[1] = rhoStar * X.u[1] * sM;
      F.q[2] = rhoStar * X.u[2] * sM;
      F.E = (rhoStar * EStar + pStar) * sM;
    }
    F.sM = sM;
    return smx;
  }
}

// ----------------------------------------------------------------------------
// Mixed-precision HLLC of the FP32-storage path (multiphase models).
// The wave-speed estimates sL, sR come in from the FP32 reconstruction; every
// flux component is formed in FP64 from the (FP32-rounded) face states, so
// that the volume-fraction, mass and energy fluxes stay mutually consistent
// to FP64 precision (the stiffened-gas energy offset gamma*pinf/(gamma-1)
// multiplies them).  The isentropic star pressure
//   p_k* = (p_k + pinf)(rho*/rho)^gamma - pinf = p_k + (p_k + pinf) expm1(gamma log1p(delta)),
// delta = (sM - uS)/(sS - sM), evaluates the power factor in FP32 without the
// (p + pinf) - pinf cancellation.
// ----------------------------------------------------------------------------
template <int MODEL, int NPH>
ECO_D void hllc_mixed(const SchemeParams &S, const CellState<NPH> &L, const CellState<NPH> &R, double sL, double sR,
                      FluxVec<NPH> &F) {
  const double eps = S.eps;
  double uL = L.u[0], pL = L.pm, rhoL = L.rho;
  double uR = R.u[0], pR = R.pm, rhoR = R.rho;
  double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) * rcp_fast(mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  if (sL >= 0. || sR <= 0.) {
    const bool left = (sL >= 0.);
    const CellState<NPH> &X = left ? L : R;
    double uS = left ? uL : uR, pS = left ? pL : pR, rhoS = left ? rhoL : rhoR;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      F.a[k] = X.a[k] * sM;
      F.m[k] = X.a[k] * X.r[k] * uS;
      F.en[k] = X.a[k] * X.r[k] * X.e[k] * uS;
    }
    double Et = X.em + 0.5 * sqn(X.u);
    F.q[0] = rhoS * uS * uS + pS;
    F.q[1] = rhoS * X.u[1] * uS;
    F.q[2] = rhoS * X.u[2] * uS;
    F.E = (rhoS * Et + pS) * uS;
  } else {
    const bool left = (sM >= 0.);
    const CellState<NPH> &X = left ? L : R;
    double uS = left ? uL : uR, pS = left ? pL : pR, mS = left ? mL : mR, sS = left ? sL : sR;
    double Et = X.em + 0.5 * sqn(X.u);
    const double iss = rcp_fast(sS - sM);
    double rhoStar = mS * iss;
    double EStar = Et + (sM - uS) * (sM + pS * rcp_fast(mS));
    double pStar = mS * (sM - uS) + pS;
    const double ratio = (sS - uS) * iss;
    const float delta = (float)((sM - uS) * iss);
    const float l1p = log1p_f(delta);
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const EosParams &q = S.eos[k];
      double rks = X.r[k] * ratio;
      double g = (double)expm1_f((float)q.gamma * l1p);
      double pks = X.p[k] + (X.p[k] + q.pinf) * g;
      double eks = (pks + q.gpinf) * q.og * rcp_fast(smax(rks, eps)) + q.eref;
      F.a[k] = X.a[k] * sM;
      F.m[k] = X.a[k] * rks * sM;
      F.en[k] = X.a[k] * rks * eks * sM;
    }
    F.q[0] = rhoStar * sM * sM + pStar;
    F.q[1] = rhoStar * X.u[1] * sM;
    F.q[2] = rhoStar * X.u[2] * sM;
    F.E = (rhoStar * EStar + pStar) * sM;
  }
  F.sM = sM;
}

// FP32 limiter (same thresholds and branch structure as limiter()).
ECO_D float limiter_f(int lim, float s1, float s2) {
  float slope = 0.f, prod = s1 * s2;
  switch (lim) {
    case LIM_MINMOD:
    case LIM_THINC:
      if (prod > 1e-9f) { slope = fminf(fabsf(s1), fabsf(s2)); if (s1 < 0.f) slope = -slope; }
      return slope;
    case LIM_VANLEER: {
      float sum = s1 + s2;
      if ((fabsf(sum) > 1e-6f) && (prod > 1e-6f)) slope = ldiv(2.f * prod, sum);
      return slope;
    }
    case LIM_VANALBADA: {
      float sum = s1 + s2;
      if ((fabsf(s1) > 1e-6f) && (fabsf(s2) > 1e-6f) && (fabsf(sum) > 1e-6f) && (prod > 1e-6f))
        slope = ldiv(prod * sum, s1 * s1 + s2 * s2);
      return slope;
    }
    case LIM_MC:
      if (prod > 1e-9f) {
        slope = fminf(2.f * fabsf(s1), 2.f * fabsf(s2));
        slope = fminf(slope, 0.5f * (fabsf(s1) + fabsf(s2)));
      }
      if (s1 < 0.f) slope = -slope;
      return slope;
    case LIM_SUPERBEE:
      if (prod > 1e-6f) slope = fmaxf(fminf(2.f * fabsf(s1), fabsf(s2)), fminf(fabsf(s1), 2.f * fabsf(s2)));
      if (s1 < 0.f) slope = -slope;
      return slope;
    default:
      return 0.f;
  }
}

// ----------------------------------------------------------------------------
// Face frames (src/Meshes/MeshCartesian.cpp:300-551, src/Maths/Coord.cpp:132-150)
// side: 0 = interior face / upper boundary of direction dir, 1 = lower boundary.
// ----------------------------------------------------------------------------
ECO_D void to_face_frame(int dir, int lower, double *u) {
  double a = u[0], b = u[1], c = u[2];
  if (dir == 0) { if (lower) { u[0] = -a; u[1] = -b; u[2] = c; } }
  else if (dir == 1) { if (lower) { u[0] = -b; u[1] = a; u[2] = c; } else { u[0] = b; u[1] = -a; u[2] = c; } }
  else { if (lower) { u[0] = -c; u[1] = -a; u[2] = b; } else { u[0] = c; u[1] = a; u[2] = b; } }
}
ECO_D void from_face_frame(int dir, int lower, double *f) {
  double a = f[0], b = f[1], c = f[2];
  if (dir == 0) { if (lower) { f[0] = -a; f[1] = -b; f[2] = c; } }
  else if (dir == 1) { if (lower) { f[0] = b; f[1] = -a; f[2] = c; } else { f[0] = -b; f[1] = a; f[2] = c; } }
  else { if (lower) { f[0] = -b; f[1] = c; f[2] = -a; } else { f[0] = b; f[1] = c; f[2] = a; } }
}

}  // namespace eco
