/*
 * CPU oracle for the explicit finite-volume step -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 restatement of the reference ECOGEN step on Cartesian meshes.  Each
 * function below names the reference file:line it follows (paths relative
 * to the ECOGEN source tree).  The restatement keeps the reference's cached
 * per-cell quantities (phase energies, sound speeds, mixture density and
 * internal energy, ...) and reproduces its evaluation order, min/max
 * semantics and clamp constants, so that it matches the reference built with
 * g++ -O3 (no FMA) operation by operation.  Build with -ffp-contract=off.
 *
 * Used by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
 * bench.py, as the checker only.
 */
#include "eco_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXPH ORC_MAXPH
#define MAXTR ORC_MAXTR

/* std::min / std::max / std::abs semantics (<algorithm>, <cmath>). */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* Per-cell state mirroring Phase{Kapila,MultiP,Euler} and Mix{Kapila,MultiP} */
/* ------------------------------------------------------------------------ */
typedef struct {
  double alpha, rho, p, e, c, T, Y, Et; /* Et: phase total energy (MultiP, Euler) */
  double u[3];                          /* Euler phase velocity */
} Phase;

typedef struct {
  double rho, p, u[3], e, E, cF, cW;
  double T; /* ThermalEq: MixThermalEq::m_temperature (cF holds its sound speed) */
} Mix;

typedef struct {
  Phase ph[MAXPH];
  Mix mx;
  double tr[MAXTR]; /* Cell::m_vecTransports (Transport::m_value) */
} State;

/* Flux{Kapila,MultiP,Euler}: alpha/masse/energ per phase, qdm, energMixture, sM. */
typedef struct {
  double alpha[MAXPH], masse[MAXPH], energ[MAXPH];
  double qdm[3];
  double energMix;
  double sM;
  double tr[MAXTR]; /* per face: Transport flux (fluxBufferTransport); per cell: m_consTransports */
} Flux;

typedef struct {
  const orc_cfg *cfg;
  double epsilon; /* src/Eos/Eos.cpp:42,114-122 */
} Ctx;

/* ------------------------------------------------------------------------ */
/* EOS: src/Eos/EosSG.cpp:93-254, src/Eos/EosIG.cpp:90-251                   */
/* ------------------------------------------------------------------------ */
static double eos_T(const orc_eos *q, double rho, double p, double eps) {
  if (q->type == ORC_EOS_SG) return (p + q->pinf) / (q->gamma - 1.) / smax(rho, eps) / q->cv;
  return p / (q->gamma - 1.) / smax(rho, eps) / q->cv;
}
static double eos_e(const orc_eos *q, double rho, double p, double eps) {
  if (q->type == ORC_EOS_SG) return (p + q->gamma * q->pinf) / (q->gamma - 1.) / smax(rho, eps) + q->eref;
  return p / (q->gamma - 1.) / smax(rho, eps) + q->eref;
}
static double eos_p(const orc_eos *q, double rho, double e) {
  if (q->type == ORC_EOS_SG) return (q->gamma - 1.) * rho * (e - q->eref) - q->gamma * q->pinf;
  return (q->gamma - 1.) * rho * (e - q->eref);
}
static double eos_c(const orc_eos *q, double rho, double p, double eps) {
  if (q->type == ORC_EOS_SG) return sqrt(q->gamma * (p + q->pinf) / smax(rho, eps));
  return sqrt(q->gamma * p / smax(rho, eps));
}
static double eos_isentrope(const orc_eos *q, double p0, double rho0, double rho1, double eps) {
  if (q->type == ORC_EOS_SG) return (p0 + q->pinf) * pow(rho1 / smax(rho0, eps), q->gamma) - q->pinf;
  return p0 * pow(rho1 / smax(rho0, eps), q->gamma);
}
static double eos_rho_pfinal(const orc_eos *q, double p0, double rho0, double pf, double *drho, double eps) {
  if (q->type == ORC_EOS_SG) {
    double num = (q->gamma) * (pf + q->pinf);
    double denom = num + p0 - pf;
    double rf = rho0 * num / smax(denom, eps);
    if (drho) *drho = rho0 * q->gamma * (p0 + q->pinf) / smax((denom * denom), eps);
    return rf;
  }
  double num = (q->gamma) * pf;
  double denom = num + p0 - pf;
  double rf = rho0 * num / smax(denom, eps);
  if (drho) *drho = rho0 * q->gamma * p0 / smax((denom * denom), eps);
  return rf;
}
/* computeDensityIsentropic: EosSG.cpp:149-154, EosIG.cpp:146-151 */
static double eos_rho_isentropic(const orc_eos *q, double p0, double rho0, double pf, double eps) {
  if (q->type == ORC_EOS_SG) return rho0 * pow((pf + q->pinf) / smax((p0 + q->pinf), eps), 1. / q->gamma);
  return rho0 * pow(pf / smax(p0, eps), 1. / q->gamma);
}
/* computeDensityHugoniot: EosSG.cpp:158-165, EosIG.cpp:155-162 */
static double eos_rho_hugoniot(const orc_eos *q, double p0, double rho0, double pf, double eps) {
  double num, denom;
  if (q->type == ORC_EOS_SG) {
    num = (q->gamma + 1.) * (pf + q->pinf) + (q->gamma - 1.) * (p0 + q->pinf);
    denom = (q->gamma - 1.) * (pf + q->pinf) + (q->gamma + 1.) * (p0 + q->pinf);
  } else {
    num = (q->gamma + 1.) * pf + (q->gamma - 1.) * p0;
    denom = (q->gamma - 1.) * pf + (q->gamma + 1.) * p0;
  }
  return rho0 * num / smax(denom, eps);
}
/* verifyAndModifyPressure: EosSG.cpp:243-247, EosIG.cpp:241-244 */
static void eos_fix_p(const orc_eos *q, double *p) {
  if (q->type == ORC_EOS_SG) {
    if (*p <= -(1. - 1.e-15) * q->pinf + 1.e-15) *p = -(1. - 1.e-15) * q->pinf + 1.e-15;
  } else {
    if (*p < 1.e-15) *p = 1.e-15;
  }
}
/* computeDensity(p, T): EosSG.cpp:114-117, EosIG.cpp:111-114 */
static double eos_rho_pt(const orc_eos *q, double p, double T, double eps);
/* sendSpecialMixtureEos: EosSG.cpp:217-222, EosIG.cpp:213-218 */
static void eos_special(const orc_eos *q, double *gpg, double *eref, double *og) {
  *gpg = (q->type == ORC_EOS_SG) ? q->gamma * q->pinf / (q->gamma - 1.) : 0.;
  *eref = q->eref;
  *og = 1. / (q->gamma - 1.);
}

/* ------------------------------------------------------------------------ */
/* Limiters: src/Ordre2/Limiter*.cpp:49-59                                   */
/* ------------------------------------------------------------------------ */
static double limiter(int lim, double s1, double s2) {
  double slope = 0., prod = s1 * s2;
  switch (lim) {
    case ORC_LIM_MINMOD:
    case ORC_LIM_THINC: /* LimiterTHINC.cpp:49-58 limits as minmod */
      if (prod > 1e-9) { slope = smin(fabs(s1), fabs(s2)); if (s1 < 0) slope = -slope; }
      return slope;
    case ORC_LIM_VANLEER: {
      double sum = s1 + s2;
      if ((fabs(sum) > 1e-6) && (prod > 1e-6)) slope = 2. * prod / sum;
      return slope;
    }
    case ORC_LIM_VANALBADA: {
      double sum = s1 + s2;
      if ((fabs(s1) > 1e-6) && (fabs(s2) > 1e-6) && (fabs(sum) > 1e-6) && (prod > 1e-6))
        slope = prod * sum / (s1 * s1 + s2 * s2);
      return slope;
    }
    case ORC_LIM_MC:
      if (prod > 1e-9) {
        slope = smin(2. * fabs(s1), 2. * fabs(s2));
        slope = smin(slope, 0.5 * (fabs(s1) + fabs(s2)));
      }
      if (s1 < 0.) slope = -slope;
      return slope;
    case ORC_LIM_SUPERBEE:
      if (prod > 1e-6) slope = smax(smin(2. * fabs(s1), fabs(s2)), smin(fabs(s1), 2. * fabs(s2)));
      if (s1 < 0.) slope = -slope;
      return slope;
    default:
      return 0.;
  }
}

/* ------------------------------------------------------------------------ */
/* Phase / mixture state completion                                          */
/* ------------------------------------------------------------------------ */
static inline double sqnorm(const double *u) { return u[0] * u[0] + u[1] * u[1] + u[2] * u[2]; }

/* PhaseKapila::verifyAndCorrectPhase (PhaseKapila.cpp:294-307),
   PhaseThermalEq::verifyAndCorrectPhase,
   PhaseMultiP::verifyAndCorrectPhase (PhaseMultiP.cpp:288-294),
   PhaseEuler::verifyAndCorrectPhase (PhaseEuler.cpp:349-353) */
static void verify_correct(const Ctx *c, int k, Phase *ph) {
  const orc_cfg *cfg = c->cfg;
  if (cfg->model == ORC_KAPILA) {
    if (c->epsilon > 1.e-20) {
      if (ph->alpha < 0.) ph->alpha = 0.;
      if (ph->alpha > 1.) ph->alpha = 1.;
      if (ph->rho <= 1.e-15) ph->rho = 1.e-15;
    } else {
      if (ph->alpha < 1e-15) ph->alpha = 1e-14;
      if (ph->alpha > 1. - 1e-15) ph->alpha = 1. - 1e-14;
      if (ph->rho < 1.e-15) ph->rho = 1.e-14;
    }
  } else if (cfg->model == ORC_MULTIP) {
    if (ph->alpha < 1e-10) ph->alpha = 1e-9;
    if (ph->alpha > 1. - 1e-10) ph->alpha = 1. - 1e-9;
    if (ph->rho < 1.e-10) ph->rho = 1.e-9;
  } else if (cfg->model == ORC_THERMALEQ || cfg->model == ORC_EULERHOM) {
    /* PhaseThermalEq::verifyAndCorrectPhase (PhaseThermalEq.cpp:240-246),
       PhaseEulerHomogeneous::verifyAndCorrectPhase (PhaseEulerHomogeneous.cpp:189-195) */
    if (ph->alpha < 1e-10) ph->alpha = 1e-10;
    if (ph->alpha > 1. - 1e-10) ph->alpha = 1. - 1e-10;
    if (ph->rho < 1.e-10) ph->rho = 1.e-10;
  } else {
    if (ph->rho < 1.e-10) ph->rho = 1.e-10;
  }
  eos_fix_p(&cfg->eos[k], &ph->p);
}

/* extendedCalculusPhase: PhaseKapila.cpp:97-102, PhaseMultiP.cpp:97-103, PhaseEuler.cpp:112-118 */
static void extended_phase(const Ctx *c, int k, Phase *ph, const double *u) {
  const orc_eos *q = &c->cfg->eos[k];
  double eps = c->epsilon;
  if (c->cfg->model == ORC_EULER) {
    ph->e = eos_e(q, ph->rho, ph->p, eps);
    ph->c = eos_c(q, ph->rho, ph->p, eps);
    ph->Et = ph->e + 0.5 * sqnorm(ph->u);
    ph->T = eos_T(q, ph->rho, ph->p, eps);
    return;
  }
  ph->T = eos_T(q, ph->rho, ph->p, eps);
  ph->e = eos_e(q, ph->rho, ph->p, eps);
  ph->c = eos_c(q, ph->rho, ph->p, eps);
  if (c->cfg->model == ORC_MULTIP) ph->Et = ph->e + 0.5 * sqnorm(u);
}

/* MixKapila::computeMixtureVariables (MixKapila.cpp:147-172), MixMultiP (MixMultiP.cpp:142-168) */
static void mixture_vars(const Ctx *c, State *s) {
  int n = c->cfg->nph;
  Mix *m = &s->mx;
  if (c->cfg->model == ORC_EULER) return; /* MixEuler: no-op */
  int kap = (c->cfg->model == ORC_KAPILA);
  m->rho = 0.; m->p = 0.;
  for (int k = 0; k < n; k++) {
    m->rho += s->ph[k].alpha * s->ph[k].rho;
    m->p += s->ph[k].alpha * s->ph[k].p;
  }
  for (int k = 0; k < n; k++)
    s->ph[k].Y = kap ? s->ph[k].alpha * s->ph[k].rho / smax(m->rho, c->epsilon)
                     : s->ph[k].alpha * s->ph[k].rho / m->rho;
  m->e = 0.; m->cF = 0.; m->cW = 0.;
  for (int k = 0; k < n; k++) {
    Phase *p = &s->ph[k];
    m->e += p->Y * p->e;
    m->cF += p->Y * p->c * p->c;
    m->cW += kap ? p->alpha / smax((p->rho * p->c * p->c), c->epsilon) : p->alpha / (p->rho * p->c * p->c);
  }
  m->cF = sqrt(m->cF);
  m->cW = 1. / sqrt(m->rho * m->cW);
}

/* ThermalEq: the equations of state without the stiffened-gas offset for
   the ideal gas (Eos::getPInf is 0 for EosIG, Eos.h:149) */
static double te_pinf(const orc_eos *q) { return (q->type == ORC_EOS_SG) ? q->pinf : 0.; }

/* MixThermalEq::computeMixtureVariables (MixThermalEq.cpp:305-323): mixture
   density, mass fractions, energy and the sound speed sqrt(sum Y_k c_k^2) */
static void te_mixture_vars(const Ctx *c, State *s) {
  int n = c->cfg->nph;
  Mix *m = &s->mx;
  m->rho = 0.;
  for (int k = 0; k < n; k++) m->rho += s->ph[k].alpha * s->ph[k].rho;
  for (int k = 0; k < n; k++) s->ph[k].Y = s->ph[k].alpha * s->ph[k].rho / m->rho;
  m->e = 0.;
  m->cF = 0.;
  for (int k = 0; k < n; k++) {
    m->e += s->ph[k].Y * s->ph[k].e;
    m->cF += s->ph[k].Y * s->ph[k].c * s->ph[k].c;
  }
  m->cF = sqrt(m->cF);
}

/* ModThermalEq::fulfillState (ModThermalEq.cpp:78-89): every phase at the
   mixture pressure and temperature, PhaseThermalEq::extendedCalculusPhase
   (PhaseThermalEq.cpp:73-78), then the mixture variables */
static void te_fulfill(const Ctx *c, State *s) {
  for (int k = 0; k < c->cfg->nph; k++) {
    Phase *ph = &s->ph[k];
    const orc_eos *q = &c->cfg->eos[k];
    ph->p = s->mx.p;
    ph->rho = eos_rho_pt(q, s->mx.p, s->mx.T, c->epsilon);
    ph->e = eos_e(q, ph->rho, ph->p, c->epsilon);
    ph->c = eos_c(q, ph->rho, ph->p, c->epsilon);
  }
  te_mixture_vars(c, s);
}

/* Mixture::computeTsat (Mixture.cpp:63-104): the saturation temperature of
   the liquid and vapour phases at pressure p, Newton on the equality of their
   Gibbs energies, f = A + B/T + C log T - log(p + pinfV) + D log(p + pinfL);
   dTsat (may be NULL) receives dTsat/dp. */
static long g_sat_fail; /* computeTsat or computePressure reaching 50 iterations */
static double sat_tsat(const Ctx *c, double pressure, double *dTsat) {
  const orc_cfg *cfg = c->cfg;
  const orc_eos *qL = &cfg->eos[cfg->liq], *qV = &cfg->eos[cfg->vap];
  double gammaL = qL->gamma, pInfL = te_pinf(qL), cvL = qL->cv, e0L = qL->eref, s0L = cfg->sref[cfg->liq];
  double gammaV = qV->gamma, pInfV = te_pinf(qV), cvV = qV->cv, e0V = qV->eref, s0V = cfg->sref[cfg->vap];
  double A = (gammaL * cvL - gammaV * cvV + s0V - s0L) / (gammaV * cvV - cvV);
  double B = (e0L - e0V) / (gammaV * cvV - cvV);
  double C = (gammaV * cvV - gammaL * cvL) / (gammaV * cvV - cvV);
  double D = (gammaL * cvL - cvL) / (gammaV * cvV - cvV);
  int iteration = 0;
  double Tsat = 0.1 * B / C;
  double f = 0., df = 1.;
  do {
    Tsat -= f / df; iteration++;
    if (iteration > 50) {
#pragma omp atomic
      g_sat_fail++;
      break;
    }
    f = A + B / Tsat + C * log(Tsat) - log(pressure + pInfV) + D * log(pressure + pInfL);
    df = C / Tsat - B / (Tsat * Tsat);
  } while (fabs(f) > 1e-10);
  double dfdp = -1. / (pressure + pInfV) + D / (pressure + pInfL);
  if (dTsat) *dTsat = -dfdp / df;
  return Tsat;
}

/* EosSG/EosIG::computeDensitySaturation (EosSG.cpp:191-200, EosIG.cpp:188-197) */
static double eos_rho_sat(const orc_eos *q, double pressure, double Tsat, double dTsatdP, double *drhodp,
                          double eps) {
  double g = q->gamma, cv = q->cv;
  if (q->type == ORC_EOS_SG) {
    if (drhodp) {
      *drhodp = (g - 1.) * cv * Tsat - (pressure + q->pinf) * (g - 1.) * cv * dTsatdP;
      *drhodp /= smax((((g - 1.) * cv * Tsat) * ((g - 1.) * cv * Tsat)), eps);
    }
    return (pressure + q->pinf) / smax(((g - 1.) * cv * Tsat), eps);
  }
  if (drhodp) {
    *drhodp = (g - 1.) * cv * Tsat - pressure * (g - 1.) * cv * dTsatdP;
    *drhodp /= smax((((g - 1.) * cv * Tsat) * ((g - 1.) * cv * Tsat)), eps);
  }
  return pressure / smax(((g - 1.) * cv * Tsat), eps);
}

/* EosSG/EosIG::computeDensityEnergySaturation (EosSG.cpp:204-210, EosIG.cpp:201-207) */
static double eos_rhoe_sat(const orc_eos *q, double pressure, double rho, double drhodp, double *drhoedp) {
  double g = q->gamma;
  if (drhoedp) *drhoedp = 1. / (g - 1.) + drhodp * q->eref;
  if (q->type == ORC_EOS_SG) return (pressure + g * q->pinf) / (g - 1.) + rho * q->eref;
  return pressure / (g - 1.) + rho * q->eref;
}

/* PhaseEulerHomogeneous::extendedCalculusPhase (PhaseEulerHomogeneous.cpp:52-57)
   and MixEulerHomogeneous::computeMixtureVariables (MixEulerHomogeneous.cpp:
   130-149): the formulas of ThermalEq's */
static void eh_phases_mixture(const Ctx *c, State *s) {
  for (int k = 0; k < c->cfg->nph; k++) {
    Phase *ph = &s->ph[k];
    const orc_eos *q = &c->cfg->eos[k];
    ph->e = eos_e(q, ph->rho, ph->p, c->epsilon);
    ph->c = eos_c(q, ph->rho, ph->p, c->epsilon);
    ph->Et = ph->e + 0.5 * sqnorm(s->mx.u);
  }
  te_mixture_vars(c, s);
}

/* ModEulerHomogeneous::fulfillState (ModEulerHomogeneous.cpp:42-56): the
   saturation temperature of the mixture pressure, each phase at that pressure
   with its saturated density, then the mixture variables */
static void eh_fulfill(const Ctx *c, State *s) {
  double Tsat = sat_tsat(c, s->mx.p, NULL);
  for (int k = 0; k < c->cfg->nph; k++) {
    s->ph[k].p = s->mx.p;
    s->ph[k].rho = eos_rho_sat(&c->cfg->eos[k], s->mx.p, Tsat, Tsat, NULL, c->epsilon);
  }
  s->mx.T = Tsat;
  eh_phases_mixture(c, s);
}

/* Model::fulfillState: ModKapila.cpp:89-101, ModMultiP.cpp:83-91, ModEuler.cpp (phase only) */
static void fulfill(const Ctx *c, State *s) {
  if (c->cfg->model == ORC_THERMALEQ) { te_fulfill(c, s); return; }
  if (c->cfg->model == ORC_EULERHOM) { eh_fulfill(c, s); return; }
  for (int k = 0; k < c->cfg->nph; k++) extended_phase(c, k, &s->ph[k], s->mx.u);
  mixture_vars(c, s);
}

/* ------------------------------------------------------------------------ */
/* Conservative <-> primitive                                                */
/* ------------------------------------------------------------------------ */
/* FluxKapila::buildCons (FluxKapila.cpp:123-135), FluxMultiP::buildCons (FluxMultiP.cpp:122-140),
   FluxEuler::buildCons (FluxEuler.cpp:95-104) */
static void build_cons(const Ctx *c, const State *s, Flux *f) {
  const orc_cfg *cfg = c->cfg;
  if (cfg->model == ORC_EULER) {
    const Phase *p = &s->ph[0];
    f->masse[0] = p->rho;
    for (int d = 0; d < 3; d++) f->qdm[d] = p->u[d] * f->masse[0];
    f->energMix = f->masse[0] * p->Et;
    return;
  }
  if (cfg->model == ORC_EULERHOM) {
    /* FluxEulerHomogeneous::buildCons (FluxEulerHomogeneous.cpp:70-89): the
       mixture mass, momentum and total energy, the internal energy summed
       from the phases' energies at their own densities and pressures */
    double ein = 0.;
    f->masse[0] = 0.;
    for (int k = 0; k < cfg->nph; k++) {
      double ak = s->ph[k].alpha, rk = s->ph[k].rho;
      f->masse[0] += ak * rk;
      double ek = eos_e(&cfg->eos[k], rk, s->ph[k].p, c->epsilon);
      ein += ak * rk * ek;
    }
    for (int d = 0; d < 3; d++) f->qdm[d] = f->masse[0] * s->mx.u[d];
    f->energMix = ein + 0.5 * f->masse[0] * sqnorm(s->mx.u);
    return;
  }
  if (cfg->model == ORC_THERMALEQ) {
    /* FluxThermalEq::buildCons (FluxThermalEq.cpp:113-126) */
    for (int k = 0; k < cfg->nph; k++) f->masse[k] = s->ph[k].alpha * s->ph[k].rho;
    for (int d = 0; d < 3; d++) f->qdm[d] = s->mx.rho * s->mx.u[d];
    f->energMix = s->mx.rho * s->mx.E;
    return;
  }
  for (int k = 0; k < cfg->nph; k++) {
    const Phase *p = &s->ph[k];
    f->alpha[k] = p->alpha;
    f->masse[k] = p->alpha * p->rho;
    double ein = eos_e(&cfg->eos[k], p->rho, p->p, c->epsilon);
    f->energ[k] = p->alpha * p->rho * ein;
  }
  for (int d = 0; d < 3; d++) f->qdm[d] = s->mx.u[d] * s->mx.rho;
  f->energMix = s->mx.rho * s->mx.E;
}

static void zero_small_u(double *u) {
  if (fabs(u[0]) < 1.e-8) u[0] = 0.;
  if (fabs(u[1]) < 1.e-8) u[1] = 0.;
  if (fabs(u[2]) < 1.e-8) u[2] = 0.;
}

/* MixThermalEq::computePressure (MixThermalEq.cpp:121-150): the mixture
   pressure of two phases in pressure and temperature equilibrium from the
   mixture internal energy (a quadratic, Le Martelot 2013) */
static double te_pressure(const Ctx *c, const double *masses, double mixInternalEnerg) {
  const orc_eos *e1 = &c->cfg->eos[0], *e2 = &c->cfg->eos[1];
  double rhoMel = 0.;
  for (int k = 0; k < c->cfg->nph; k++) rhoMel += masses[k];
  double gamma1 = e1->gamma, pInf1 = te_pinf(e1), cv1 = e1->cv, e01 = e1->eref, Y1 = masses[0] / rhoMel;
  double gamma2 = e2->gamma, pInf2 = te_pinf(e2), cv2 = e2->cv, e02 = e2->eref, Y2 = masses[1] / rhoMel;
  double q = Y1 * e01 + Y2 * e02;
  double cvMel = Y1 * cv1 + Y2 * cv2;
  double A1 = Y1 * (gamma1 - 1.) * cv1 / cvMel * (rhoMel * (mixInternalEnerg - q) - pInf1);
  double A2 = Y2 * (gamma2 - 1.) * cv2 / cvMel * (rhoMel * (mixInternalEnerg - q) - pInf2);
  return 0.5 * (A1 + A2 - (pInf1 + pInf2)) +
         sqrt(0.25 * (A2 - A1 - (pInf2 - pInf1)) * (A2 - A1 - (pInf2 - pInf1)) + A1 * A2);
}

/* MixThermalEq::computeTemperature (MixThermalEq.cpp:154-178) */
static double te_temperature(const Ctx *c, const double *masses, double pressure) {
  double rhoMel = 0., T = 0.;
  for (int k = 0; k < c->cfg->nph; k++) rhoMel += masses[k];
  for (int k = 0; k < c->cfg->nph; k++) {
    const orc_eos *q = &c->cfg->eos[k];
    double Yk = masses[k] / rhoMel;
    T += Yk * (q->gamma - 1.) * q->cv / (pressure + te_pinf(q));
  }
  return 1. / (T * rhoMel);
}

/* FluxThermalEq::buildPrim (FluxThermalEq.cpp:130-162) */
static void te_build_prim(const Ctx *c, Flux *f, State *s) {
  int n = c->cfg->nph;
  double rhoMel = 0.;
  for (int k = 0; k < n; k++) rhoMel += f->masse[k];
  for (int d = 0; d < 3; d++) s->mx.u[d] = f->qdm[d] / rhoMel;
  zero_small_u(s->mx.u);
  double e = f->energMix / rhoMel - 0.5 * (s->mx.u[0] * s->mx.u[0] + s->mx.u[1] * s->mx.u[1] + s->mx.u[2] * s->mx.u[2]);
  double p = te_pressure(c, f->masse, e);
  double T = te_temperature(c, f->masse, p);
  s->mx.p = p;
  s->mx.T = T;
  for (int k = 0; k < n; k++) {
    Phase *ph = &s->ph[k];
    const orc_eos *q = &c->cfg->eos[k];
    ph->p = p;
    ph->rho = eos_rho_pt(q, p, T, c->epsilon);
    ph->alpha = f->masse[k] / ph->rho;
    ph->e = eos_e(q, ph->rho, ph->p, c->epsilon);
    ph->c = eos_c(q, ph->rho, ph->p, c->epsilon);
  }
  te_mixture_vars(c, s);
  s->mx.E = f->energMix / rhoMel;
}

/* MixEulerHomogeneous::computePressure (MixEulerHomogeneous.cpp:93-125): the
   pressure at which the saturated mixture of density masse holds the
   internal energy e, Newton from the destination state's liquid pressure */
static double eh_pressure(const Ctx *c, double masse, double internalEnergy, double p_guess) {
  const orc_cfg *cfg = c->cfg;
  const orc_eos *qL = &cfg->eos[cfg->liq], *qV = &cfg->eos[cfg->vap];
  double eps = c->epsilon;
  int iteration = 0;
  double pressure = p_guess;
  double f = 0., df = 1.;
  double alphaVap, dalphaVap, rhoLiq, drhoLiq, rhoVap, drhoVap, rhoeLiq, drhoeLiq, rhoeVap, drhoeVap, Tsat, dTsat;
  do {
    pressure -= f / df; iteration++;
    if (iteration > 50) {
#pragma omp atomic
      g_sat_fail++;
      break;
    }
    Tsat = sat_tsat(c, pressure, &dTsat);
    rhoVap = eos_rho_sat(qV, pressure, Tsat, dTsat, &drhoVap, eps);
    rhoLiq = eos_rho_sat(qL, pressure, Tsat, dTsat, &drhoLiq, eps);
    alphaVap = (masse - rhoLiq) / (rhoVap - rhoLiq);
    dalphaVap = (-drhoLiq * (rhoVap - rhoLiq) - (masse - rhoLiq) * (drhoVap - drhoLiq)) /
                ((rhoVap - rhoLiq) * (rhoVap - rhoLiq));
    rhoeVap = eos_rhoe_sat(qV, pressure, rhoVap, drhoVap, &drhoeVap);
    rhoeLiq = eos_rhoe_sat(qL, pressure, rhoLiq, drhoLiq, &drhoeLiq);
    f = masse * internalEnergy - alphaVap * (rhoeVap - rhoeLiq) - rhoeLiq;
    df = -dalphaVap * (rhoeVap - rhoeLiq) - alphaVap * (drhoeVap - drhoeLiq) - drhoeLiq;
  } while (fabs(f / (masse * internalEnergy)) > 1e-10);
  return pressure;
}

/* FluxEulerHomogeneous::buildPrim (FluxEulerHomogeneous.cpp:93-128) */
static void eh_build_prim(const Ctx *c, Flux *f, State *s) {
  const orc_cfg *cfg = c->cfg;
  int liq = cfg->liq, vap = cfg->vap;
  double masse = f->masse[0];
  for (int d = 0; d < 3; d++) s->mx.u[d] = f->qdm[d] / masse;
  zero_small_u(s->mx.u);
  double internalEnergy = f->energMix / masse - 0.5 * sqnorm(s->mx.u);
  double pressure = eh_pressure(c, masse, internalEnergy, s->ph[liq].p);
  double Tsat = sat_tsat(c, pressure, NULL);
  s->ph[vap].rho = eos_rho_sat(&cfg->eos[vap], pressure, Tsat, Tsat, NULL, c->epsilon);
  s->ph[liq].rho = eos_rho_sat(&cfg->eos[liq], pressure, Tsat, Tsat, NULL, c->epsilon);
  s->ph[vap].alpha = (masse - s->ph[liq].rho) / (s->ph[vap].rho - s->ph[liq].rho);
  s->ph[liq].alpha = 1. - s->ph[vap].alpha;
  for (int k = 0; k < cfg->nph; k++) s->ph[k].p = pressure;
  s->mx.p = pressure;
  s->mx.T = Tsat;
  eh_phases_mixture(c, s);
  s->mx.E = f->energMix / masse;
}

/* FluxKapila::buildPrim (FluxKapila.cpp:139-185), FluxMultiP::buildPrim (FluxMultiP.cpp:144-183),
   FluxEuler::buildPrim (FluxEuler.cpp:107-131) */
static void build_prim(const Ctx *c, Flux *f, State *s) {
  const orc_cfg *cfg = c->cfg;
  int n = cfg->nph;
  if (cfg->model == ORC_EULER) {
    Phase *p = &s->ph[0];
    const orc_eos *q = &cfg->eos[0];
    p->rho = f->masse[0];
    for (int d = 0; d < 3; d++) p->u[d] = f->qdm[d] / f->masse[0];
    zero_small_u(p->u);
    p->Et = f->energMix / f->masse[0];
    p->e = p->Et - 0.5 * sqnorm(p->u);
    p->p = eos_p(q, f->masse[0], p->e);
    p->c = eos_c(q, f->masse[0], p->p, c->epsilon);
    p->T = eos_T(q, f->masse[0], p->p, c->epsilon);
    return;
  }
  if (cfg->model == ORC_THERMALEQ) { te_build_prim(c, f, s); return; }
  if (cfg->model == ORC_EULERHOM) { eh_build_prim(c, f, s); return; }
  double un = 0.;
  if (cfg->model == ORC_KAPILA) {
    if (c->epsilon > 1.e-20) {
      for (int k = 0; k < n; k++) {
        if (f->alpha[k] < 0.) f->alpha[k] = 0.;
        if (f->alpha[k] > 1.) f->alpha[k] = 1.;
        un += f->alpha[k];
      }
    } else {
      for (int k = 0; k < n; k++) {
        if (f->alpha[k] <= 1.e-15) f->alpha[k] = 1e-15;
        if (f->alpha[k] >= 1. - 1.e-15) f->alpha[k] = 1.0 - 1e-15;
        un += f->alpha[k];
      }
    }
  } else {
    for (int k = 0; k < n; k++) {
      if (f->alpha[k] <= 1.e-10) f->alpha[k] = 1e-9;
      if (f->alpha[k] >= 1. - 1.e-10) f->alpha[k] = 1.0 - 1e-9;
      un += f->alpha[k];
    }
  }
  for (int k = 0; k < n; k++) f->alpha[k] /= un;
  double rhoMel = 0.;
  for (int k = 0; k < n; k++) {
    Phase *p = &s->ph[k];
    rhoMel = rhoMel + f->masse[k];
    p->alpha = f->alpha[k];
    double ein;
    if (cfg->model == ORC_KAPILA) {
      p->rho = f->masse[k] / smax(f->alpha[k], c->epsilon);
      ein = f->energ[k] / smax(f->masse[k], c->epsilon);
    } else {
      p->rho = f->masse[k] / f->alpha[k];
      ein = f->energ[k] / f->masse[k];
    }
    p->p = eos_p(&cfg->eos[k], p->rho, ein);
    verify_correct(c, k, p);
  }
  for (int d = 0; d < 3; d++) s->mx.u[d] = f->qdm[d] / rhoMel;
  zero_small_u(s->mx.u);
  for (int k = 0; k < n; k++) extended_phase(c, k, &s->ph[k], s->mx.u);
  mixture_vars(c, s);
  s->mx.E = f->energMix / rhoMel;
}

/* FluxMultiP::schemeCorrection (FluxMultiP.cpp:237-260); buf = U^n */
static void scheme_correction(const Ctx *c, Flux *f, const Flux *buf) {
  if (c->cfg->model != ORC_MULTIP) return;
  int n = c->cfg->nph;
  double rhoN = 0., rhoNp1 = 0., sum = 0., dek[MAXPH];
  for (int k = 0; k < n; k++) {
    rhoN += buf->masse[k];
    rhoNp1 += f->masse[k];
    dek[k] = f->energ[k] - buf->energ[k];
  }
  for (int k = 0; k < n; k++) sum += dek[k];
  double dE = f->energMix - buf->energMix;
  double dRU2 = sqnorm(f->qdm) / rhoNp1 - sqnorm(buf->qdm) / rhoN;
  if (fabs(sum) > 1.e-10 * f->energMix)
    for (int k = 0; k < n; k++) f->energ[k] = buf->energ[k] + dek[k] / sum * (dE - 0.5 * dRU2);
}

/* ------------------------------------------------------------------------ */
/* Relaxation and energy correction                                          */
/* ------------------------------------------------------------------------ */
/* Relaxations that reached the iteration limit in the last orc_run_from, and
   the (alpha, rho, p) per phase a first one started from (test statistics). */
static long g_relax_fail;
static double g_relax_fail_in[3 * MAXPH];
/* 1-based step of the last orc_run_from in which an injection or tank Newton
   solve did not converge (0: none); the run stops there */
static int g_inflow_fail_step;

/* RelaxationP::stiffRelaxation (RelaxationP.cpp:49-111). Returns iterations. */
static int relax_p(const Ctx *c, State *s) {
  const orc_cfg *cfg = c->cfg;
  int n = cfg->nph;
  int relax = 1;
  if (c->epsilon > 1.e-20)
    for (int k = 0; k < n; k++) if (s->ph[k].alpha > (1. - 1.e-5)) relax = 0;
  if (!relax) return 0;
  double ak[MAXPH], pk[MAXPH], rk[MAXPH], rkS[MAXPH], akS[MAXPH];
  double pStar = 0.;
  for (int k = 0; k < n; k++) {
    verify_correct(c, k, &s->ph[k]);
    ak[k] = s->ph[k].alpha; pk[k] = s->ph[k].p; rk[k] = s->ph[k].rho;
    pStar += ak[k] * pk[k];
  }
  int it = 0;
  double f = 0., df = 1.;
  do {
    it++;
    pStar -= f / df;
    for (int k = 0; k < n; k++) eos_fix_p(&cfg->eos[k], &pStar);
    f = -1.; df = 0.;
    for (int k = 0; k < n; k++) {
      double drho;
      rkS[k] = eos_rho_pfinal(&cfg->eos[k], pk[k], rk[k], pStar, &drho, c->epsilon);
      akS[k] = ak[k] * rk[k] / rkS[k];
      double dalpha = ak[k] * rk[k] * drho / (rkS[k] * rkS[k]);
      f += akS[k];
      df -= dalpha;
    }
    if (it > 100) break;
  } while (fabs(f) > 1e-10 && it < 100);
  if (it < 100) {
    for (int k = 0; k < n; k++) { s->ph[k].alpha = akS[k]; s->ph[k].rho = rkS[k]; s->ph[k].p = pStar; }
    s->mx.p = pStar;
  } else {
    long nfail;
#pragma omp atomic capture
    nfail = ++g_relax_fail;
    if (nfail == 1)
      for (int k = 0; k < n; k++) {
        g_relax_fail_in[3 * k] = ak[k]; g_relax_fail_in[3 * k + 1] = rk[k]; g_relax_fail_in[3 * k + 2] = pk[k];
      }
  }
  return it;
}

static int model_relaxes(const orc_cfg *cfg) {
  return cfg->model == ORC_KAPILA || (cfg->model == ORC_MULTIP && cfg->relax_p);
}

/* RelaxationPT::stiffRelaxation (RelaxationPT.cpp:49-84) with
   analyticalPressure (88-118) and analyticalTemperature (122-141): from the
   mixture energy, density and mass fractions the cell cached at its last
   computeMixtureVariables (before any relaxation), the pressure and
   temperature of two phases in equilibrium in closed form; the phases take
   them at constant partial densities alpha_k rho_k. */
static void relax_pt(const Ctx *c, State *s) {
  int n = c->cfg->nph;
  if (c->epsilon > 1.e-20)
    for (int k = 0; k < n; k++) if (s->ph[k].alpha > (1. - 1.e-5)) return;
  const orc_eos *e1 = &c->cfg->eos[0], *e2 = &c->cfg->eos[1];
  double e0 = s->mx.e, rho0 = s->mx.rho;
  double gamma1 = e1->gamma, pInf1 = te_pinf(e1), cv1 = e1->cv, e01 = e1->eref, Y1 = s->ph[0].Y;
  double gamma2 = e2->gamma, pInf2 = te_pinf(e2), cv2 = e2->cv, e02 = e2->eref, Y2 = s->ph[1].Y;
  double q = Y1 * e01 + Y2 * e02;
  double cvMel = Y1 * cv1 + Y2 * cv2;
  double A1 = Y1 * (gamma1 - 1.) * cv1 / cvMel * (rho0 * (e0 - q) - pInf1);
  double A2 = Y2 * (gamma2 - 1.) * cv2 / cvMel * (rho0 * (e0 - q) - pInf2);
  double pStar = 0.5 * (A1 + A2 - (pInf1 + pInf2)) +
                 sqrt(0.25 * (A2 - A1 - (pInf2 - pInf1)) * (A2 - A1 - (pInf2 - pInf1)) + A1 * A2);
  double T = 0.;
  for (int k = 0; k < n; k++) {
    const orc_eos *q2 = &c->cfg->eos[k];
    T += s->ph[k].Y * (q2->gamma - 1.) * q2->cv / (pStar + te_pinf(q2));
  }
  T = 1. / (T * rho0);
  double rkS[MAXPH], akS[MAXPH];
  for (int k = 0; k < n; k++) {
    rkS[k] = eos_rho_pt(&c->cfg->eos[k], pStar, T, c->epsilon);
    akS[k] = s->ph[k].alpha * s->ph[k].rho / rkS[k];
  }
  for (int k = 0; k < n; k++) { s->ph[k].alpha = akS[k]; s->ph[k].rho = rkS[k]; s->ph[k].p = pStar; }
  s->mx.p = pStar;
}

/* RelaxationPTMu::stiffRelaxation (RelaxationPTMu.cpp:51-157): the two phases
   (phase 0 the liquid, phase 1 the vapour, RelaxationPTMu.cpp:45-46) taken to
   saturation at the pressure where the saturated mixture of the cell's
   density holds its internal energy (Newton from sum alpha_k p_k of the
   phases of the relaxed state s, which are verified and corrected).  The
   density and energy are those of the cell's own mixture, and the result is
   written to the cell's own phases, then fulfilled (Cell::getMixture,
   getPhase and fulfillState without a type, RelaxationPTMu.cpp:47-48,
   138-149): in the predictor s is the predicted state and s_n the time-n
   state, which the predictor thus relaxes. */
static void relax_ptmu(const Ctx *c, State *s, State *s_n) {
  const orc_cfg *cfg = c->cfg;
  int n = cfg->nph, liq = 0, vap = 1;
  const orc_eos *qL = &cfg->eos[liq], *qV = &cfg->eos[vap];
  double eps = c->epsilon;
  double pStar = 0.;
  for (int k = 0; k < n; k++) {
    pStar += s->ph[k].alpha * s->ph[k].p;
    verify_correct(c, k, &s->ph[k]);
  }
  double rho = s_n->mx.rho;
  double rhoe = rho * s_n->mx.e;
  double dTsat = 0.;
  double Tsat = sat_tsat(c, pStar, &dTsat);
  double rhoLSat = 0., rhoVSat = 0., drhoLSat = 0., drhoVSat = 0.;
  double aLSat = 0., aVSat = 0., daLSat = 0., daVSat = 0.;
  double rhoeLSat, rhoeVSat, drhoeLSat, drhoeVSat;
  int iteration = 0;
  double f = 0., df = 1.;
  do {
    pStar -= f / df; iteration++;
    if (iteration > 50) {
#pragma omp atomic
      g_sat_fail++;
      break;
    }
    for (int k = 0; k < n; k++) eos_fix_p(&cfg->eos[k], &pStar);
    Tsat = sat_tsat(c, pStar, &dTsat);
    rhoLSat = eos_rho_sat(qL, pStar, Tsat, dTsat, &drhoLSat, eps);
    rhoVSat = eos_rho_sat(qV, pStar, Tsat, dTsat, &drhoVSat, eps);
    if (rhoLSat <= rho) { rhoLSat = rho + 1e-6; drhoLSat = 0.; }
    if (rhoVSat >= rho) { rhoVSat = rho - 1e-6; drhoVSat = 0.; }
    aLSat = (rho - rhoVSat) / (rhoLSat - rhoVSat);
    daLSat = (-drhoVSat * (rhoLSat - rhoVSat) - (rho - rhoVSat) * (drhoLSat - drhoVSat)) /
             ((rhoLSat - rhoVSat) * (rhoLSat - rhoVSat));
    if (aLSat <= 0.) aLSat = 1e-8;
    if (aLSat >= 1.) aLSat = 1. - 1e-8;
    aVSat = 1. - aLSat;
    daVSat = -daLSat;
    f = rhoe; df = 0.;
    rhoeLSat = eos_rhoe_sat(qL, pStar, rhoLSat, drhoLSat, &drhoeLSat);
    rhoeVSat = eos_rhoe_sat(qV, pStar, rhoVSat, drhoVSat, &drhoeVSat);
    f -= (aLSat * rhoeLSat + aVSat * rhoeVSat);
    df -= (daLSat * rhoeLSat + aLSat * drhoeLSat + daVSat * rhoeVSat + aVSat * drhoeVSat);
    f /= rhoe;
    df /= rhoe;
  } while (fabs(f) > 1e-10);
  (void)Tsat;
  s_n->ph[liq].alpha = aLSat; s_n->ph[liq].rho = rhoLSat; s_n->ph[liq].p = pStar;
  s_n->ph[vap].alpha = aVSat; s_n->ph[vap].rho = rhoVSat; s_n->ph[vap].p = pStar;
  fulfill(c, s_n);
}

/* FluxKapila::correctionEnergy (FluxKapila.cpp:239-268); the trailing
   cell->fulfillState() acts on the time-n state (Cell.cpp:282). */
static void correction_energy(const Ctx *c, State *s_type, State *s_n) {
  if (c->cfg->model != ORC_KAPILA) return;
  int n = c->cfg->nph;
  double ak[MAXPH], rk[MAXPH];
  for (int k = 0; k < n; k++) { ak[k] = s_type->ph[k].alpha; rk[k] = s_type->ph[k].rho; }
  double rhoe = s_type->mx.rho * s_type->mx.e;
  double p = rhoe, denom = 0.;
  for (int k = 0; k < n; k++) {
    double gpg, eref, og;
    eos_special(&c->cfg->eos[k], &gpg, &eref, &og);
    p -= ak[k] * (gpg + rk[k] * eref);
    denom += ak[k] * og;
  }
  p /= denom;
  for (int k = 0; k < n; k++) { s_type->ph[k].p = p; verify_correct(c, k, &s_type->ph[k]); }
  s_type->mx.p = p;
  fulfill(c, s_n);
}

/* Mixture::totalEnergyToInternalEnergy (MixKapila.cpp:188-195; MixEuler no-op) */
static void total_to_internal(const Ctx *c, State *s) {
  if (c->cfg->model == ORC_EULER) return;
  s->mx.e = s->mx.E - 0.5 * sqnorm(s->mx.u);
}

/* ------------------------------------------------------------------------ */
/* Riemann solvers in the face frame                                         */
/* ------------------------------------------------------------------------ */
/* ModKapila::solveRiemannIntern (ModKapila.cpp:107-218) and the identical
   ModMultiP::solveRiemannIntern (ModMultiP.cpp:101-214). */
static void hllc_multi(const Ctx *c, const State *L, const State *R, double dxL, double dxR,
                       double *dtMax, Flux *F) {
  int n = c->cfg->nph;
  double eps = c->epsilon;
  double uL = L->mx.u[0], cL = L->mx.cF, pL = L->mx.p, rhoL = L->mx.rho;
  double uR = R->mx.u[0], cR = R->mx.cF, pR = R->mx.p, rhoR = R->mx.rho;
  double sL = smin(uL - cL, uR - cR);
  double sR = smax(uR + cR, uL + cL);
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  if (fabs(sR) > 1.e-3) *dtMax = smin(*dtMax, dxR / fabs(sR));
  double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  if (sL >= 0. || sR <= 0.) {
    const State *S = (sL >= 0.) ? L : R;
    double uS = (sL >= 0.) ? uL : uR, pS = (sL >= 0.) ? pL : pR, rhoS = (sL >= 0.) ? rhoL : rhoR;
    for (int k = 0; k < n; k++) {
      double a = S->ph[k].alpha, d = S->ph[k].rho, e = S->ph[k].e;
      F->alpha[k] = a * sM;
      F->masse[k] = a * d * uS;
      F->energ[k] = a * d * e * uS;
    }
    double vY = S->mx.u[1], vZ = S->mx.u[2];
    double Et = S->mx.e + 0.5 * sqnorm(S->mx.u);
    F->qdm[0] = rhoS * uS * uS + pS;
    F->qdm[1] = rhoS * vY * uS;
    F->qdm[2] = rhoS * vZ * uS;
    F->energMix = (rhoS * Et + pS) * uS;
  } else {
    int left = (sM >= 0.);
    const State *S = left ? L : R;
    double uS = left ? uL : uR, pS = left ? pL : pR, mS = left ? mL : mR, sS = left ? sL : sR;
    double vY = S->mx.u[1], vZ = S->mx.u[2];
    double Et = S->mx.e + 0.5 * sqnorm(S->mx.u);
    double rhoStar = mS / (sS - sM);
    double EStar = Et + (sM - uS) * (sM + pS / mS);
    double pStar = mS * (sM - uS) + pS;
    for (int k = 0; k < n; k++) {
      double a = S->ph[k].alpha, d = S->ph[k].rho, pk = S->ph[k].p;
      double mk = d * (sS - uS);
      double rks = mk / (sS - sM);
      double pks = eos_isentrope(&c->cfg->eos[k], pk, d, rks, eps);
      double eks = eos_e(&c->cfg->eos[k], rks, pks, eps);
      F->alpha[k] = a * sM;
      F->masse[k] = a * rks * sM;
      F->energ[k] = a * rks * eks * sM;
    }
    F->qdm[0] = rhoStar * sM * sM + pStar;
    F->qdm[1] = rhoStar * vY * sM;
    F->qdm[2] = rhoStar * vZ * sM;
    F->energMix = (rhoStar * EStar + pStar) * sM;
  }
  F->sM = sM;
}

/* ModEuler::solveRiemannIntern (ModEuler.cpp:90-175) */
static void hllc_euler(const Ctx *c, const State *Ls, const State *Rs, double dxL, double dxR,
                       double *dtMax, Flux *F) {
  (void)c;
  const Phase *L = &Ls->ph[0], *R = &Rs->ph[0];
  double uL = L->u[0], vL = L->u[1], wL = L->u[2], pL = L->p, rhoL = L->rho, cL = L->c, EL = L->Et;
  double uR = R->u[0], vR = R->u[1], wR = R->u[2], pR = R->p, rhoR = R->rho, cR = R->c, ER = R->Et;
  double sL = smin(uL - cL, uR - cR);
  double sR = smax(uR + cR, uL + cL);
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  if (fabs(sR) > 1.e-3) *dtMax = smin(*dtMax, dxR / fabs(sR));
  double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  if (sL > 0.) {
    F->masse[0] = rhoL * uL;
    F->qdm[0] = rhoL * uL * uL + pL; F->qdm[1] = rhoL * vL * uL; F->qdm[2] = rhoL * wL * uL;
    F->energMix = (rhoL * EL + pL) * uL;
  } else if (sR < 0.) {
    F->masse[0] = rhoR * uR;
    F->qdm[0] = rhoR * uR * uR + pR; F->qdm[1] = rhoR * vR * uR; F->qdm[2] = rhoR * wR * uR;
    F->energMix = (rhoR * ER + pR) * uR;
  } else if (sM >= 0.) {
    double pStar = mL * (sM - uL) + pL;
    double rhoStar = mL / (sL - sM);
    double Estar = EL + (sM - uL) * (sM + pL / mL);
    F->masse[0] = rhoStar * sM;
    F->qdm[0] = rhoStar * sM * sM + pStar; F->qdm[1] = rhoStar * sM * vL; F->qdm[2] = rhoStar * sM * wL;
    F->energMix = (rhoStar * Estar + pStar) * sM;
  } else {
    double pStar = mR * (sM - uR) + pR;
    double rhoStar = mR / (sR - sM);
    double Estar = ER + (sM - uR) * (sM + pR / mR);
    F->masse[0] = rhoStar * sM;
    F->qdm[0] = rhoStar * sM * sM + pStar; F->qdm[1] = rhoStar * sM * vR; F->qdm[2] = rhoStar * sM * wR;
    F->energMix = (rhoStar * Estar + pStar) * sM;
  }
  F->sM = sM;
}

/* ModThermalEq::solveRiemannIntern (ModThermalEq.cpp:94-210): the mixture
   HLLC with the phases' partial densities */
static void hllc_te(const Ctx *c, const State *L, const State *R, double dxL, double dxR,
                    double *dtMax, Flux *F) {
  int n = c->cfg->nph;
  double uL = L->mx.u[0], cL = L->mx.cF, pL = L->mx.p, rhoL = L->mx.rho;
  double uR = R->mx.u[0], cR = R->mx.cF, pR = R->mx.p, rhoR = R->mx.rho;
  double sL = smin(uL - cL, uR - cR);
  double sR = smax(uR + cR, uL + cL);
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  if (fabs(sR) > 1.e-3) *dtMax = smin(*dtMax, dxR / fabs(sR));
  double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  if (sL >= 0. || sR <= 0.) {
    const State *S = (sL >= 0.) ? L : R;
    double uS = (sL >= 0.) ? uL : uR, pS = (sL >= 0.) ? pL : pR, rhoS = (sL >= 0.) ? rhoL : rhoR;
    for (int k = 0; k < n; k++) F->masse[k] = S->ph[k].alpha * S->ph[k].rho * uS;
    double vitY = S->mx.u[1], vitZ = S->mx.u[2];
    double totalEnergy = S->mx.e + 0.5 * sqnorm(S->mx.u);
    F->qdm[0] = rhoS * uS * uS + pS;
    F->qdm[1] = rhoS * vitY * uS;
    F->qdm[2] = rhoS * vitZ * uS;
    F->energMix = (rhoS * totalEnergy + pS) * uS;
  } else {
    int left = (sM >= 0.);
    const State *S = left ? L : R;
    double uS = left ? uL : uR, pS = left ? pL : pR, mS = left ? mL : mR, sS = left ? sL : sR;
    double vitY = S->mx.u[1], vitZ = S->mx.u[2];
    double totalEnergy = S->mx.e + 0.5 * sqnorm(S->mx.u);
    double rhoStar = mS / (sS - sM);
    double EStar = totalEnergy + (sM - uS) * (sM + pS / mS);
    double pStar = mS * (sM - uS) + pS;
    for (int k = 0; k < n; k++) {
      double mk = S->ph[k].alpha * S->ph[k].rho * (sS - uS);
      F->masse[k] = mk / (sS - sM) * sM;
    }
    F->qdm[0] = rhoStar * sM * sM + pStar;
    F->qdm[1] = rhoStar * vitY * sM;
    F->qdm[2] = rhoStar * vitZ * sM;
    F->energMix = (rhoStar * EStar + pStar) * sM;
  }
  F->sM = sM;
}

/* ModEulerHomogeneous::solveRiemannIntern (ModEulerHomogeneous.cpp:60-136):
   HLLC on the mixture (density, pressure, velocity, the mixture energy the
   cell holds plus |u|^2/2, sound speed sqrt(sum Y_k c_k^2)); the supersonic
   branches are strict, sL > 0 and sR < 0 */
static void hllc_eh(const Ctx *c, const State *L, const State *R, double dxL, double dxR,
                    double *dtMax, Flux *F) {
  (void)c;
  double uL = L->mx.u[0], vL = L->mx.u[1], wL = L->mx.u[2], cL = L->mx.cF, pL = L->mx.p, rhoL = L->mx.rho;
  double uR = R->mx.u[0], vR = R->mx.u[1], wR = R->mx.u[2], cR = R->mx.cF, pR = R->mx.p, rhoR = R->mx.rho;
  double EL = L->mx.e + 0.5 * sqnorm(L->mx.u), ER = R->mx.e + 0.5 * sqnorm(R->mx.u);
  double sL = smin(uL - cL, uR - cR);
  double sR = smax(uR + cR, uL + cL);
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  if (fabs(sR) > 1.e-3) *dtMax = smin(*dtMax, dxR / fabs(sR));
  double mL = rhoL * (sL - uL), mR = rhoR * (sR - uR);
  double sM = (pR - pL + mL * uL - mR * uR) / (mL - mR);
  if (fabs(sM) < 1.e-8) sM = 0.;
  memset(F->masse, 0, sizeof(F->masse));
  if (sL > 0.) {
    F->masse[0] = rhoL * uL;
    F->qdm[0] = rhoL * uL * uL + pL; F->qdm[1] = rhoL * vL * uL; F->qdm[2] = rhoL * wL * uL;
    F->energMix = (rhoL * EL + pL) * uL;
  } else if (sR < 0.) {
    F->masse[0] = rhoR * uR;
    F->qdm[0] = rhoR * uR * uR + pR; F->qdm[1] = rhoR * vR * uR; F->qdm[2] = rhoR * wR * uR;
    F->energMix = (rhoR * ER + pR) * uR;
  } else if (sM >= 0.) {
    double pStar = mL * (sM - uL) + pL;
    double rhoStar = mL / (sL - sM);
    double Estar = EL + (sM - uL) * (sM + pL / mL);
    F->masse[0] = rhoStar * sM;
    F->qdm[0] = rhoStar * sM * sM + pStar; F->qdm[1] = rhoStar * sM * vL; F->qdm[2] = rhoStar * sM * wL;
    F->energMix = (rhoStar * Estar + pStar) * sM;
  } else {
    double pStar = mR * (sM - uR) + pR;
    double rhoStar = mR / (sR - sM);
    double Estar = ER + (sM - uR) * (sM + pR / mR);
    F->masse[0] = rhoStar * sM;
    F->qdm[0] = rhoStar * sM * sM + pStar; F->qdm[1] = rhoStar * sM * vR; F->qdm[2] = rhoStar * sM * wR;
    F->energMix = (rhoStar * Estar + pStar) * sM;
  }
  F->sM = sM;
}

static void riemann(const Ctx *c, const State *L, const State *R, double dxL, double dxR,
                    double *dtMax, Flux *F) {
  if (c->cfg->model == ORC_EULER) hllc_euler(c, L, R, dxL, dxR, dtMax, F);
  else if (c->cfg->model == ORC_EULERHOM) hllc_eh(c, L, R, dxL, dxR, dtMax, F);
  else if (c->cfg->model == ORC_THERMALEQ) hllc_te(c, L, R, dxL, dxR, dtMax, F);
  else hllc_multi(c, L, R, dxL, dxR, dtMax, F);
}

/* MixThermalEq::computeTemperatureIsentrope and computeVolumeIsentrope
   (MixThermalEq.cpp:200-226, 256-277): the mixture isentrope at fixed mass
   fractions from (p0, T0) to p */
static double te_T_isentrope(const Ctx *c, const double *Yk, double p0, double T0, double p, double *dTdp) {
  int n = c->cfg->nph;
  double T = T0, cM = 0.;
  if (dTdp) *dTdp = 0.;
  for (int k = 0; k < n; k++) cM += Yk[k] * c->cfg->eos[k].gamma * c->cfg->eos[k].cv;
  for (int k = 0; k < n; k++) {
    const orc_eos *q = &c->cfg->eos[k];
    double puissance = (q->gamma - 1.) * Yk[k] * q->cv / cM;
    double fk = pow((p + te_pinf(q)) / ((p0 + te_pinf(q))), puissance);
    T *= fk;
    if (dTdp) *dTdp += puissance / (p + te_pinf(q));
  }
  if (dTdp) *dTdp *= T;
  return T;
}
static double te_v_isentrope(const Ctx *c, const double *Yk, double p0, double T0, double p) {
  double dTdp = 0.;
  double T = te_T_isentrope(c, Yk, p0, T0, p, &dTdp);
  double v = 0.;
  for (int k = 0; k < c->cfg->nph; k++) {
    const orc_eos *q = &c->cfg->eos[k];
    double vk = ((q->gamma - 1.) * q->cv * T) / (p + te_pinf(q));
    v += Yk[k] * vk;
  }
  return v;
}

/* ModThermalEq::solveRiemannOutflow (ModThermalEq.cpp:372-424): the left
   isentrope to the imposed p0 */
static void te_outflow(const Ctx *c, const State *L, double dxL, double *dtMax, double p0, Flux *F) {
  int n = c->cfg->nph;
  double uL = L->mx.u[0], cL = L->mx.cF, pL = L->mx.p, rhoL = L->mx.rho, TL = L->mx.T;
  double vL = L->mx.u[1], wL = L->mx.u[2];
  double Yk[MAXPH];
  for (int k = 0; k < n; k++) Yk[k] = L->ph[k].alpha * L->ph[k].rho / rhoL;
  double zL = rhoL * cL;
  double pStar = p0;
  double rhoStar = 1. / te_v_isentrope(c, Yk, pL, TL, p0);
  double vmv0 = 1. / rhoStar - 1. / rhoL, mL;
  if (fabs(vmv0) > 1e-10) mL = sqrt((pL - p0) / vmv0);
  else mL = zL;
  double sL = uL - mL / rhoL;
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  double uStar = uL + mL * vmv0;
  if (sL >= 0.) { uStar = uL; pStar = pL; rhoStar = rhoL; }
  double totalEnergy = L->mx.e + 0.5 * sqnorm(L->mx.u);
  double EStar = totalEnergy + (uStar - uL) * (uStar - pL / mL);
  for (int k = 0; k < n; k++) F->masse[k] = rhoStar * Yk[k] * uStar;
  F->qdm[0] = uStar * uStar * rhoStar + pStar;
  F->qdm[1] = uStar * vL * rhoStar;
  F->qdm[2] = uStar * wL * rhoStar;
  F->energMix = (EStar * rhoStar + pStar) * uStar;
  F->sM = uStar;
}

/* The outflow's half Riemann problem (imposed pressure p0) in the face frame:
   ModKapila::solveRiemannOutflow (ModKapila.cpp:479-549, isentropic left
   wave), ModMultiP::solveRiemannOutflow (ModMultiP.cpp:468-538, Hugoniot
   left wave), ModEuler::solveRiemannOutflow (ModEuler.cpp:427-475).  MultiP
   leaves mL unset when |v* - 1/rhoL| <= 1e-10; here it is rhoL cL, as in
   the Kapila form. */
static void riemann_outflow(const Ctx *c, const State *L, double dxL, double *dtMax, double p0, Flux *F) {
  const orc_cfg *cfg = c->cfg;
  double eps = c->epsilon;
  if (cfg->model == ORC_THERMALEQ) { te_outflow(c, L, dxL, dtMax, p0, F); return; }
  if (cfg->model == ORC_EULER) {
    const Phase *P = &L->ph[0];
    double uL = P->u[0], vL = P->u[1], wL = P->u[2], pL = P->p, rhoL = P->rho, cL = P->c;
    double zL = rhoL * cL;
    double sL = uL - cL;
    if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
    double pStar = p0;
    double rhoStar = eos_rho_hugoniot(&cfg->eos[0], pL, rhoL, pStar, 0.);
    double uStar = (pL + zL * uL - pStar) / zL;
    if (sL >= 0.) { uStar = uL; pStar = pL; rhoStar = rhoL; }
    else if (uStar < 0) { rhoStar = rhoL; }
    double eStar = eos_e(&cfg->eos[0], rhoStar, pStar, 0.);
    F->masse[0] = rhoStar * uStar;
    F->qdm[0] = rhoStar * uStar * uStar + pStar; F->qdm[1] = rhoStar * uStar * vL; F->qdm[2] = rhoStar * uStar * wL;
    F->energMix = (rhoStar * (eStar + 0.5 * (uStar * uStar + vL * vL + wL * wL)) + pStar) * uStar;
    F->sM = uStar;
    return;
  }
  int n = cfg->nph, kap = (cfg->model == ORC_KAPILA);
  double uL = L->mx.u[0], cL = L->mx.cF, pL = L->mx.p, rhoL = L->mx.rho;
  double uyL = L->mx.u[1], uzL = L->mx.u[2], EL = L->mx.e + 0.5 * sqnorm(L->mx.u);
  double zL = rhoL * cL;
  double pStar = p0, EStar = 0., vStar = 0., uStar = 0., rk[ORC_MAXPH];
  for (int k = 0; k < n; k++) {
    const Phase *P = &L->ph[k];
    if (kap) {
      rk[k] = eos_rho_isentropic(&cfg->eos[k], pL, P->rho, pStar, eps);
      vStar += P->alpha * P->rho / rhoL / smax(rk[k], eps);
    } else {
      rk[k] = eos_rho_hugoniot(&cfg->eos[k], pL, P->rho, pStar, eps);
      vStar += P->alpha * P->rho / rhoL / rk[k];
    }
  }
  double vSmvL = vStar - 1. / rhoL, mL = zL;
  if (fabs(vSmvL) > 1e-10) mL = sqrt((pL - pStar) / vSmvL);
  double sL = uL - mL / rhoL;
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  double sM = uL + mL * vSmvL;
  if (sL >= 0. || sM < 0) {  /* supersonic outflow, or inflow: the left state */
    uStar = uL; pStar = pL;
    for (int k = 0; k < n; k++) rk[k] = L->ph[k].rho;
    vStar = 1. / rhoL;
    EStar = EL;
  } else {
    uStar = sM;
    EStar = EL + (uStar - uL) * (uStar - pL / mL);
  }
  for (int k = 0; k < n; k++) {
    const Phase *P = &L->ph[k];
    double YkL = P->alpha * P->rho / rhoL;
    double ekStar = eos_e(&cfg->eos[k], rk[k], pStar, eps);
    F->alpha[k] = kap ? YkL / smax(rk[k], eps) / vStar * uStar : YkL / rk[k] / vStar * uStar;
    F->masse[k] = F->alpha[k] * rk[k];
    F->energ[k] = F->masse[k] * ekStar;
  }
  F->qdm[0] = uStar * uStar / vStar + pStar; F->qdm[1] = uStar * uyL / vStar; F->qdm[2] = uStar * uzL / vStar;
  F->energMix = (EStar / vStar + pStar) * uStar;
  F->sM = uStar;
}

/* Eos{IG,SG}::vfpfh, dvdpch, dvdhcp (EosIG.cpp:220-239, EosSG.cpp:223-242):
   specific volume from pressure and enthalpy and its two partial derivatives */
static double eos_vfpfh(const orc_eos *q, double p, double h, double eps) {
  double pp = (q->type == ORC_EOS_SG) ? q->gamma * (p + q->pinf) : q->gamma * p;
  return (q->gamma - 1.) * (h - q->eref) / smax(pp, eps);
}
static double eos_dvdpch(const orc_eos *q, double p, double h, double eps) {
  double pp = (q->type == ORC_EOS_SG) ? (p + q->pinf) * (p + q->pinf) : p * p;
  return (1. - q->gamma) / q->gamma * (h - q->eref) / smax(pp, eps);
}
static double eos_dvdhcp(const orc_eos *q, double p, double eps) {
  double pp = (q->type == ORC_EOS_SG) ? (p + q->pinf) : p;
  return (q->gamma - 1.) / q->gamma / smax(pp, eps);
}

/* The injection's half Riemann problem in the face frame (BoundCondInj,
   src/BoundConds/BoundCondInj.cpp:62, 198-201): a fluid of total enthalpy
   H_k0 (Eos::computeTotalEnthalpy, Eos.cpp:99-102) enters at the mass flow m0
   per unit area; the left wave is acoustic.  ModKapila::solveRiemannInflow
   (ModKapila.cpp:253-322) and ModMultiP::solveRiemannInflow
   (ModMultiP.cpp:249-318) are the same; ModEuler::solveRiemannInflow
   (ModEuler.cpp:215-303) is their one-phase form.  m0 is the deck's value; the
   reference changes its sign on reading.  Newton on the pressure from pL until
   |f| <= 1e-10; the reference aborts after 50 iterations, here the 50th
   iterate is kept and *nonconv set. */
static void riemann_inflow(const Ctx *c, const State *L, double dxL, double *dtMax, int side, Flux *F, int *nonconv) {
  const orc_cfg *cfg = c->cfg;
  double eps = c->epsilon;
  double m0 = -cfg->bc_m0[side];
  const double *ak0 = cfg->bc_ak0[side], *rhok0 = cfg->bc_rhok0[side], *pk0 = cfg->bc_pk0[side];
  if (cfg->model == ORC_EULER) {
    const orc_eos *q = &cfg->eos[0];
    const Phase *P = &L->ph[0];
    double uL = P->u[0], vL = P->u[1], wL = P->u[2], pL = P->p, rhoL = P->rho, cL = P->c;
    double zL = rhoL * cL;
    double sL = uL - cL;
    if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
    double u0 = m0 / rhok0[0];
    double H0 = eos_e(q, rhok0[0], pk0[0], 0.) + pk0[0] / rhok0[0] + 0.5 * u0 * u0;
    int it = 0;
    double p = pL, f = 0., df = 1., u, du, v, dv, h;
    do {
      p -= f / df; it++;
      if (it > 50) { *nonconv = 1; break; }
      eos_fix_p(q, &p);
      u = uL + (pL - p) / zL;
      if (u >= -1e-6) u = -1e-6;
      du = -1. / zL;
      v = u / m0; dv = du / m0;
      f = v; df = dv;
      h = H0 - 0.5 * u * u;
      v = eos_vfpfh(q, p, h, 0.);
      dv = eos_dvdpch(q, p, h, 0.) - eos_dvdhcp(q, p, 0.) * u * du;
      f -= v; df -= dv;
    } while (fabs(f) > 1e-10);
    double uStar = u, pStar = p, rhoStar = m0 / uStar;
    double eStar = eos_e(q, rhoStar, pStar, 0.);
    F->masse[0] = rhoStar * uStar;
    F->qdm[0] = rhoStar * uStar * uStar + pStar; F->qdm[1] = rhoStar * uStar * vL; F->qdm[2] = rhoStar * uStar * wL;
    F->energMix = (rhoStar * (eStar + 0.5 * (uStar * uStar + vL * vL + wL * wL)) + pStar) * uStar;
    F->sM = uStar;
    return;
  }
  int n = cfg->nph;
  double uL = L->mx.u[0], cL = L->mx.cF, pL = L->mx.p, rhoL = L->mx.rho;
  double vL = L->mx.u[1], wL = L->mx.u[2];
  double rho0 = 0., Hk0[MAXPH], Yk0[MAXPH], vk[MAXPH], a0[MAXPH], sum = 0.;
  for (int k = 0; k < n; k++) sum += ak0[k];                 /* BoundCondInj.cpp:108-117 */
  for (int k = 0; k < n; k++) a0[k] = ak0[k] / sum;
  for (int k = 0; k < n; k++) rho0 += a0[k] * rhok0[k];     /* Mix*::computeDensity */
  double u0 = m0 / rho0;
  for (int k = 0; k < n; k++) {
    Hk0[k] = eos_e(&cfg->eos[k], rhok0[k], pk0[k], eps) + pk0[k] / rhok0[k] + 0.5 * u0 * u0;
    Yk0[k] = a0[k] * rhok0[k] / rho0;
  }
  double sL = uL - cL;
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  double zL = rhoL * cL;
  int it = 0;
  double pStar = pL, f = 0., df = 1., u, du, v, dv, hk;
  do {
    pStar -= f / df; it++;
    if (it > 50) { *nonconv = 1; break; }
    for (int k = 0; k < n; k++) eos_fix_p(&cfg->eos[k], &pStar);
    u = uL + (pL - pStar) / zL;
    if (u >= -1e-6) u = -1e-6;
    du = -1. / zL;
    v = u / m0; dv = du / m0;
    f = v; df = dv;
    for (int k = 0; k < n; k++) {
      hk = Hk0[k] - 0.5 * u * u;
      vk[k] = eos_vfpfh(&cfg->eos[k], pStar, hk, eps);
      double dvk = eos_dvdpch(&cfg->eos[k], pStar, hk, eps) - eos_dvdhcp(&cfg->eos[k], pStar, eps) * u * du;
      f -= Yk0[k] * vk[k];
      df -= Yk0[k] * dvk;
    }
  } while (fabs(f) > 1e-10);
  double Estar = 0.5 * (u * u + vL * vL + wL * wL);
  for (int k = 0; k < n; k++) {
    double rhok = 1. / vk[k];
    double ek = eos_e(&cfg->eos[k], rhok, pStar, eps);
    Estar += Yk0[k] * ek;
    F->alpha[k] = Yk0[k] * vk[k] / v * u;
    F->masse[k] = F->alpha[k] * rhok;
    F->energ[k] = F->alpha[k] * rhok * ek;
  }
  F->qdm[0] = u * u / v + pStar; F->qdm[1] = u * vL / v; F->qdm[2] = u * wL / v;
  F->energMix = (Estar / v + pStar) * u;
  F->sM = u;
}

/* computeDensityIsentropic and computeDensityHugoniot with their derivative
   drho/dp (EosIG.cpp:146-162, EosSG.cpp:149-165), computeEnthalpyIsentropic
   (EosIG.cpp:177-184, EosSG.cpp:180-187) and computeDensity(p, T)
   (EosIG.cpp:111-114, EosSG.cpp:114-117) */
static double eos_rho_isentropic_d(const orc_eos *q, double p0, double rho0, double pf, double eps, double *drho) {
  double pi = (q->type == ORC_EOS_SG) ? q->pinf : 0.;
  double rf = eos_rho_isentropic(q, p0, rho0, pf, eps);
  *drho = rf / smax(q->gamma * (pf + pi), eps);
  return rf;
}
static double eos_rho_hugoniot_d(const orc_eos *q, double p0, double rho0, double pf, double eps, double *drho) {
  double pi = (q->type == ORC_EOS_SG) ? q->pinf : 0.;
  double denom = (q->type == ORC_EOS_SG) ? (q->gamma - 1.) * (pf + pi) + (q->gamma + 1.) * (p0 + pi)
                                         : (q->gamma - 1.) * pf + (q->gamma + 1.) * p0;
  *drho = (q->type == ORC_EOS_SG) ? rho0 * 4. * q->gamma * (p0 + pi) / smax(denom * denom, eps)
                                  : rho0 * 4. * q->gamma * p0 / smax(denom * denom, eps);
  return eos_rho_hugoniot(q, p0, rho0, pf, eps);
}
static double eos_h_isentropic(const orc_eos *q, double p0, double rho0, double pf, double eps, double *dh) {
  double drho, rf = eos_rho_isentropic_d(q, p0, rho0, pf, eps, &drho);
  double pp = (q->type == ORC_EOS_SG) ? pf + q->pinf : pf;
  *dh = q->gamma / (q->gamma - 1.) * (rf - pp * drho) / smax(rf * rf, eps);
  return q->gamma * pp / (q->gamma - 1.) / smax(rf, eps) + q->eref;
}
static double eos_rho_pt(const orc_eos *q, double p, double T, double eps) {
  double num = (q->type == ORC_EOS_SG) ? p + q->pinf : p;
  return num / smax((q->gamma - 1.) * q->cv * T, eps);
}

/* The tank's state (BoundCondTank.cpp:46-176): rho_k0 = rho_k(p0, T0); the
   fractions divided by their sum; from volume fractions the mass fractions
   alpha_k0 rho_k0 / rho, from mass fractions the volume fractions
   rho Y_k / rho_k0 with 1/rho = sum Y_k / rho_k0.  One phase: alpha 1. */
static void tank_state(const Ctx *c, int side, double *ak0, double *rhok0) {
  const orc_cfg *cfg = c->cfg;
  int n = cfg->nph;
  double p0 = cfg->bc_tank_p0[side], T0 = cfg->bc_tank_T0[side];
  for (int k = 0; k < n; k++) rhok0[k] = eos_rho_pt(&cfg->eos[k], p0, T0, c->epsilon);
  if (n == 1) { ak0[0] = 1.; return; }
  double fr[MAXPH], sum = 0.;
  for (int k = 0; k < n; k++) sum += cfg->bc_tank_frac[side][k];
  for (int k = 0; k < n; k++) fr[k] = cfg->bc_tank_frac[side][k] / sum;
  if (!cfg->bc_tank_massfrac[side]) {
    for (int k = 0; k < n; k++) ak0[k] = fr[k];
  } else {
    double rm = 0.;
    for (int k = 0; k < n; k++) rm += fr[k] / rhok0[k];
    rm = 1.0 / rm;
    for (int k = 0; k < n; k++) ak0[k] = rm * fr[k] / rhok0[k];
  }
}

/* The tank's half Riemann problem in the face frame (BoundCondTank): from
   the left wave at p* = p0, supersonic outflow keeps the left state and
   subsonic outflow the left star state at p0; otherwise fluid flows in from
   the tank, at rest with total enthalpy H0 and isentropic, and a Newton
   solve on p* matches the tank-side velocity to the left star velocity.
   ModKapila::solveRiemannTank (ModKapila.cpp:326-475, Hugoniot left wave,
   |f| <= 1e-2), ModMultiP::solveRiemannTank (ModMultiP.cpp:322-465, the same
   with |f| <= 1e-3, no epsilon guards, and sM of the first estimate as the
   contact speed), ModEuler::solveRiemannTank (ModEuler.cpp:305-423,
   isentropic left wave, |f| <= 1e-3).  Every Newton iterate's left wave
   enters dtMax, as in the reference; the reference stops after 50 iterations,
   here the last iterate is kept and *nonconv set. */
static void riemann_tank(const Ctx *c, const State *L, double dxL, double *dtMax, int side, Flux *F, int *nonconv) {
  const orc_cfg *cfg = c->cfg;
  double eps = c->epsilon;
  double p0 = cfg->bc_tank_p0[side];
  double ak0[MAXPH], rhok0[MAXPH];
  tank_state(c, side, ak0, rhok0);
  if (cfg->model == ORC_EULER) {
    const orc_eos *q = &cfg->eos[0];
    const Phase *P = &L->ph[0];
    double uL = P->u[0], vL = P->u[1], wL = P->u[2], pL = P->p, rhoL = P->rho, cL = P->c;
    double zL = rhoL * cL;
    double uStar, rhoStar, pStar = p0, vStar, wStar, mL, drho;
    double v = 1. / eos_rho_isentropic(q, pL, rhoL, pStar, 0.);
    double vmv0 = v - 1. / rhoL;
    if (fabs(vmv0) > 1e-10) mL = sqrt((pL - p0) / vmv0);
    else mL = zL;
    double sL = uL - mL / rhoL;
    if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
    double u = uL + mL * vmv0;
    if (sL >= 0.) { uStar = uL; pStar = pL; rhoStar = rhoL; vStar = vL; wStar = wL; }
    else if (u >= -1e-3) { uStar = u; pStar = p0; rhoStar = 1. / v; vStar = vL; wStar = wL; }
    else {
      double H0 = eos_e(q, rhok0[0], p0, 0.) + p0 / rhok0[0] + 0.5 * 0. * 0.;
      int it = 0;
      double p = 0.5 * p0, f = 0., df = 1., dv, h, dh;
      double uStarR = 0., duStarR = 0., uStarL = 0., duStarL = 0.;
      rhoStar = 0.;
      do {
        p -= f / df; it++;
        if (it > 50) { *nonconv = 1; break; }
        eos_fix_p(q, &p);
        if (p > p0) p = p0 - 1e-6;
        rhoStar = eos_rho_isentropic(q, p0, rhok0[0], p, 0.);
        h = eos_h_isentropic(q, p0, rhok0[0], p, 0., &dh);
        uStarR = -sqrt(2. * (H0 - h));
        duStarR = -dh / uStarR;
        double dmL;
        v = 1.0 / eos_rho_isentropic_d(q, pL, rhoL, p, 0., &drho);
        dv = -v * v * drho;
        vmv0 = v - 1. / rhoL;
        if (fabs(vmv0) > 1e-10) {
          mL = sqrt((pL - p) / vmv0);
          dmL = 0.5 * (-vmv0 + (p - pL) * dv) / (vmv0 * vmv0) / mL;
        } else { mL = zL; dmL = 0.; }
        sL = uL - mL / rhoL;
        if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
        uStarL = uL + mL * vmv0;
        duStarL = dmL * vmv0 + mL * dv;
        f = uStarR - uStarL;
        df = duStarR - duStarL;
      } while (fabs(f) > 1e-3);
      pStar = p;
      uStar = 0.5 * (uStarL + uStarR);
      vStar = 0.; wStar = 0.;
    }
    double eStar = eos_e(q, rhoStar, pStar, 0.);
    F->masse[0] = rhoStar * uStar;
    F->qdm[0] = rhoStar * uStar * uStar + pStar; F->qdm[1] = rhoStar * uStar * vStar; F->qdm[2] = rhoStar * uStar * wStar;
    F->energMix = (rhoStar * (eStar + 0.5 * (uStar * uStar + vStar * vStar + wStar * wStar)) + pStar) * uStar;
    F->sM = uStar;
    return;
  }
  int n = cfg->nph, kap = (cfg->model == ORC_KAPILA);
  double uL = L->mx.u[0], cL = L->mx.cF, pL = L->mx.p, rhoL = L->mx.rho;
  double uyL = L->mx.u[1], uzL = L->mx.u[2];
  double zL = rhoL * cL;
  double rkS[MAXPH], YkS[MAXPH], pStar = p0, vStar = 0., uStar, rhoStar, uyStar, uzStar, mL;
  for (int k = 0; k < n; k++) {
    const Phase *P = &L->ph[k];
    rkS[k] = eos_rho_hugoniot(&cfg->eos[k], P->p, P->rho, pStar, eps);
    vStar += kap ? P->alpha * P->rho / rhoL / smax(rkS[k], eps) : P->alpha * P->rho / rhoL / rkS[k];
  }
  double vmv0 = vStar - 1. / rhoL;
  if (fabs(vmv0) > 1e-10) mL = sqrt((pL - pStar) / vmv0);
  else mL = zL;
  double sL = uL - mL / rhoL;
  if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
  double sM = uL + mL * vmv0;
  if (sL >= 0.) {
    uStar = uL; pStar = pL;
    for (int k = 0; k < n; k++) { rkS[k] = L->ph[k].rho; YkS[k] = L->ph[k].alpha * L->ph[k].rho / rhoL; }
    rhoStar = rhoL; uyStar = uyL; uzStar = uzL;
  } else if (sM >= -1e-3) {
    uStar = sM; pStar = p0;
    for (int k = 0; k < n; k++) YkS[k] = L->ph[k].alpha * L->ph[k].rho / rhoL;
    rhoStar = 1. / vStar; uyStar = uyL; uzStar = uzL;
  } else {
    double H0 = 0., Yk0[MAXPH], rho0 = 0.;
    for (int k = 0; k < n; k++) rho0 += ak0[k] * rhok0[k];   /* Mix*::computeDensity */
    for (int k = 0; k < n; k++) {
      Yk0[k] = ak0[k] * rhok0[k] / rho0;
      H0 += Yk0[k] * (eos_e(&cfg->eos[k], rhok0[k], p0, eps) + p0 / rhok0[k] + 0.5 * 0. * 0.);
    }
    int it = 0;
    double p = 0.5 * p0, f = 0., df = 1., dhk, drhok;
    double uStarR = 0., duStarR = 0., uStarL = 0., duStarL = 0.;
    do {
      p -= f / df; it++;
      if (it > 50) { *nonconv = 1; break; }
      for (int k = 0; k < n; k++) eos_fix_p(&cfg->eos[k], &p);
      if (p > p0) p = p0 - 1e-6;
      uStarR = H0; duStarR = 0.;
      for (int k = 0; k < n; k++) {
        rkS[k] = eos_rho_isentropic(&cfg->eos[k], p0, rhok0[k], p, eps);
        double hk = eos_h_isentropic(&cfg->eos[k], p0, rhok0[k], p, eps, &dhk);
        uStarR -= Yk0[k] * hk;
        duStarR -= Yk0[k] * dhk;
      }
      uStarR = -sqrt(2. * uStarR);
      duStarR = duStarR / uStarR;
      double vStarL = 0., dvStarL = 0., dmL;
      for (int k = 0; k < n; k++) {
        const Phase *P = &L->ph[k];
        double rhok = eos_rho_hugoniot_d(&cfg->eos[k], P->p, P->rho, p, eps, &drhok);
        double YkL = P->alpha * P->rho / rhoL;
        if (kap) {
          vStarL += YkL / smax(rhok, eps);
          dvStarL -= YkL / smax(rhok * rhok, eps) * drhok;
        } else {
          vStarL += YkL / rhok;
          dvStarL -= YkL / (rhok * rhok) * drhok;
        }
      }
      vmv0 = vStarL - 1. / rhoL;
      if (fabs(vmv0) > 1e-10) {
        mL = sqrt((pL - p) / vmv0);
        dmL = 0.5 * (-vmv0 + (p - pL) * dvStarL) / (vmv0 * vmv0) / mL;
      } else { mL = zL; dmL = 0.; }
      sL = uL - mL / rhoL;
      if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
      uStarL = uL + mL * vmv0;
      duStarL = dmL * vmv0 + mL * dvStarL;
      f = uStarR - uStarL;
      df = duStarR - duStarL;
    } while (fabs(f) > (kap ? 1e-2 : 1e-3));
    pStar = p;
    uStar = 0.5 * (uStarL + uStarR);
    rhoStar = 0.;
    for (int k = 0; k < n; k++) {
      YkS[k] = Yk0[k];
      rhoStar += kap ? YkS[k] / smax(rkS[k], eps) : YkS[k] / rkS[k];
    }
    rhoStar = 1. / rhoStar;
    uyStar = 0.; uzStar = 0.;
  }
  double EStar = 0.5 * (uStar * uStar + uyStar * uyStar + uzStar * uzStar);
  for (int k = 0; k < n; k++) {
    double ek = eos_e(&cfg->eos[k], rkS[k], pStar, eps);
    EStar += YkS[k] * ek;
    F->alpha[k] = kap ? YkS[k] * rhoStar / smax(rkS[k], eps) * uStar : YkS[k] * rhoStar / rkS[k] * uStar;
    F->masse[k] = F->alpha[k] * rkS[k];
    F->energ[k] = F->masse[k] * ek;
  }
  F->qdm[0] = rhoStar * uStar * uStar + pStar; F->qdm[1] = rhoStar * uStar * uyStar; F->qdm[2] = rhoStar * uStar * uzStar;
  F->energMix = (rhoStar * EStar + pStar) * uStar;
  F->sM = kap ? uStar : sM;
}

/* The wall's half Riemann problem in the face frame: ModKapila::solveRiemannWall
   (ModKapila.cpp:224-249), ModMultiP::solveRiemannWall (ModMultiP.cpp:220-245),
   ModEuler::solveRiemannWall (ModEuler.cpp:181-210).  Only the normal
   momentum flux p* is nonzero and sM = 0 (F arrives zeroed).  The two
   products keep the reference's operand order, which differs by model. */
static void riemann_wall(const Ctx *c, const State *L, double dxL, double *dtMax, Flux *F) {
  double uL, cL, pL, rhoL, sL;
  if (c->cfg->model == ORC_EULER) {
    uL = L->ph[0].u[0]; pL = L->ph[0].p; rhoL = L->ph[0].rho; cL = L->ph[0].c;
    sL = smin(uL - cL, -uL - cL);
    if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
    F->qdm[0] = rhoL * uL * (uL - sL) + pL;
  } else {
    uL = L->mx.u[0]; cL = L->mx.cF; pL = L->mx.p; rhoL = L->mx.rho;
    sL = smin(uL - cL, -uL - cL);
    if (fabs(sL) > 1.e-3) *dtMax = smin(*dtMax, dxL / fabs(sL));
    F->qdm[0] = rhoL * (uL - sL) * uL + pL;
  }
}

/* ------------------------------------------------------------------------ */
/* Mesh: src/Meshes/MeshCartesian.cpp:154-160, 180-552                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int left, right;    /* right = -1 for boundary faces */
  int dir;            /* 0,1,2 */
  double nrm[3], tan[3], bin[3];
  double surface;
  double pos[3];
  int opp_left;       /* face opposite to this one across the left cell (-1: none/boundary) */
  int opp_right;      /* face opposite across the right cell */
  int side;           /* boundary faces: 0..5 = X-, X+, Y-, Y+, Z-, Z+ (interior: -1) */
  int bnd_left;       /* boundary face on the other side of the left cell (-1: none) */
  int bnd_right;      /* boundary face on the other side of the right cell */
  int opp_int;        /* boundary faces: the internal face across the cell (-1: none) */
} Face;

typedef struct {
  int ncell, nface, nint, dim;
  int n[3];
  double *posx, *posy, *posz, *dxs, *dys, *dzs;
  double volume_dummy;
  Face *faces;
  int *cell_faces;    /* per cell, up to 6 face indices in accumulation order */
  int *cell_nf;
  double *vol, *lcfl;
} Mesh;

static inline void cell_pos(const Mesh *m, int c, double *p) {
  int n0 = m->n[0], n1 = m->n[1];
  int i = c % n0, j = (c / n0) % n1, k = c / (n0 * n1);
  p[0] = m->posx[i]; p[1] = m->posy[j]; p[2] = m->posz[k];
}
static inline int cidx(const Mesh *m, int i, int j, int k) {
  int idx = 0;
  if (m->n[0] != 1) idx += i;
  if (m->n[1] != 1) idx += j * m->n[0];
  if (m->n[2] != 1) idx += k * m->n[0] * m->n[1];
  return idx;
}

static void set3(double *v, double a, double b, double c) { v[0] = a; v[1] = b; v[2] = c; }

static int mesh_build(Mesh *m, const orc_cfg *cfg) {
  memset(m, 0, sizeof(*m));
  for (int d = 0; d < 3; d++) m->n[d] = cfg->n[d];
  int nx = m->n[0], ny = m->n[1], nz = m->n[2];
  m->dim = (nx != 1) + (ny != 1) + (nz != 1);
  m->ncell = nx * ny * nz;
  double dX = cfg->L[0] / (double)nx, dY = cfg->L[1] / (double)ny, dZ = cfg->L[2] / (double)nz;
  m->posx = malloc(sizeof(double) * nx); m->posy = malloc(sizeof(double) * ny); m->posz = malloc(sizeof(double) * nz);
  for (int i = 0; i < nx; i++) m->posx[i] = (i + 0.5) * dX;
  for (int i = 0; i < ny; i++) m->posy[i] = (i + 0.5) * dY;
  for (int i = 0; i < nz; i++) m->posz[i] = (i + 0.5) * dZ;
  double vol = dX * dY * dZ;
  double lc = 1.e10;
  if (nx != 1) lc = smin(lc, dX);
  if (ny != 1) lc = smin(lc, dY);
  if (nz != 1) lc = smin(lc, dZ);
  if (m->dim > 1) lc *= 0.6;
  m->vol = malloc(sizeof(double)); m->vol[0] = vol;
  m->lcfl = malloc(sizeof(double)); m->lcfl[0] = lc;

  int nf = 0, nb = 0;
  if (nx != 1) { nf += (nx + 1) * ny * nz; nb += 2 * ny * nz; }
  if (ny != 1) { nf += (ny + 1) * nx * nz; nb += 2 * nx * nz; }
  if (nz != 1) { nf += (nz + 1) * nx * ny; nb += 2 * nx * ny; }
  m->nface = nf; m->nint = nf - nb;
  m->faces = calloc(nf, sizeof(Face));
  int f = 0;
  /* internal faces */
  for (int ix = 0; ix < nx - 1; ix++) for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++) {
    Face *F = &m->faces[f++];
    F->left = cidx(m, ix, iy, iz); F->right = cidx(m, ix + 1, iy, iz); F->dir = 0;
    set3(F->tan, 0., 1., 0.); set3(F->nrm, 1., 0., 0.); set3(F->bin, 0., 0., 1.);
    F->surface = dY * dZ;
    set3(F->pos, m->posx[ix] + 0.5 * dX, m->posy[iy], m->posz[iz]);
  }
  for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny - 1; iy++) for (int iz = 0; iz < nz; iz++) {
    Face *F = &m->faces[f++];
    F->left = cidx(m, ix, iy, iz); F->right = cidx(m, ix, iy + 1, iz); F->dir = 1;
    set3(F->tan, -1., 0., 0.); set3(F->nrm, 0., 1., 0.); set3(F->bin, 0., 0., 1.);
    F->surface = dX * dZ;
    set3(F->pos, m->posx[ix], m->posy[iy] + 0.5 * dY, m->posz[iz]);
  }
  for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz - 1; iz++) {
    Face *F = &m->faces[f++];
    F->left = cidx(m, ix, iy, iz); F->right = cidx(m, ix, iy, iz + 1); F->dir = 2;
    set3(F->tan, 1., 0., 0.); set3(F->nrm, 0., 0., 1.); set3(F->bin, 0., 1., 0.);
    F->surface = dX * dY;
    set3(F->pos, m->posx[ix], m->posy[iy], m->posz[iz] + 0.5 * dZ);
  }
  /* boundary faces (all transmissive), MeshCartesian.cpp:399-551 */
  if (nx != 1) {
    for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++) {
      Face *F = &m->faces[f++];
      F->left = cidx(m, 0, iy, iz); F->right = -1; F->dir = 0; F->side = 0;
      set3(F->tan, 0., -1., 0.); set3(F->nrm, -1., 0., 0.); set3(F->bin, 0., 0., 1.);
      F->surface = dY * dZ; set3(F->pos, 0., m->posy[iy], m->posz[iz]);
    }
    for (int iy = 0; iy < ny; iy++) for (int iz = 0; iz < nz; iz++) {
      Face *F = &m->faces[f++];
      F->left = cidx(m, nx - 1, iy, iz); F->right = -1; F->dir = 0; F->side = 1;
      set3(F->tan, 0., 1., 0.); set3(F->nrm, 1., 0., 0.); set3(F->bin, 0., 0., 1.);
      F->surface = dY * dZ; set3(F->pos, m->posx[nx - 1] + 0.5 * dX, m->posy[iy], m->posz[iz]);
    }
  }
  if (ny != 1) {
    for (int ix = 0; ix < nx; ix++) for (int iz = 0; iz < nz; iz++) {
      Face *F = &m->faces[f++];
      F->left = cidx(m, ix, 0, iz); F->right = -1; F->dir = 1; F->side = 2;
      set3(F->tan, 1., 0., 0.); set3(F->nrm, 0., -1., 0.); set3(F->bin, 0., 0., 1.);
      F->surface = dX * dZ; set3(F->pos, m->posx[ix], 0., m->posz[iz]);
    }
    for (int ix = 0; ix < nx; ix++) for (int iz = 0; iz < nz; iz++) {
      Face *F = &m->faces[f++];
      F->left = cidx(m, ix, ny - 1, iz); F->right = -1; F->dir = 1; F->side = 3;
      set3(F->tan, -1., 0., 0.); set3(F->nrm, 0., 1., 0.); set3(F->bin, 0., 0., 1.);
      F->surface = dX * dZ; set3(F->pos, m->posx[ix], m->posy[ny - 1] + 0.5 * dY, m->posz[iz]);
    }
  }
  if (nz != 1) {
    for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) {
      Face *F = &m->faces[f++];
      F->left = cidx(m, ix, iy, 0); F->right = -1; F->dir = 2; F->side = 4;
      set3(F->tan, -1., 0., 0.); set3(F->nrm, 0., 0., -1.); set3(F->bin, 0., 1., 0.);
      F->surface = dX * dY; set3(F->pos, m->posx[ix], m->posy[iy], 0.);
    }
    for (int ix = 0; ix < nx; ix++) for (int iy = 0; iy < ny; iy++) {
      Face *F = &m->faces[f++];
      F->left = cidx(m, ix, iy, nz - 1); F->right = -1; F->dir = 2; F->side = 5;
      set3(F->tan, 1., 0., 0.); set3(F->nrm, 0., 0., 1.); set3(F->bin, 0., 1., 0.);
      F->surface = dX * dY; set3(F->pos, m->posx[ix], m->posy[iy], m->posz[nz - 1] + 0.5 * dZ);
    }
  }
  /* per-cell face lists in global face order (accumulation order of Run.cpp:401) */
  m->cell_faces = malloc(sizeof(int) * 6 * m->ncell);
  m->cell_nf = calloc(m->ncell, sizeof(int));
  for (int fi = 0; fi < nf; fi++) {
    Face *F = &m->faces[fi];
    m->cell_faces[6 * F->left + m->cell_nf[F->left]++] = fi;
    if (F->right >= 0) m->cell_faces[6 * F->right + m->cell_nf[F->right]++] = fi;
  }
  /* opposite internal faces (the parallel face on the other side of a cell) */
  for (int fi = 0; fi < m->nint; fi++) {
    Face *F = &m->faces[fi];
    F->opp_left = -1; F->opp_right = -1;
    for (int q = 0; q < m->cell_nf[F->left]; q++) {
      int g = m->cell_faces[6 * F->left + q];
      if (g != fi && g < m->nint && m->faces[g].dir == F->dir) F->opp_left = g;
    }
    for (int q = 0; q < m->cell_nf[F->right]; q++) {
      int g = m->cell_faces[6 * F->right + q];
      if (g != fi && g < m->nint && m->faces[g].dir == F->dir) F->opp_right = g;
    }
  }
  /* boundary faces parallel to an internal face across its cell (walls at second order) */
  for (int fi = 0; fi < m->nint; fi++) { m->faces[fi].bnd_left = -1; m->faces[fi].bnd_right = -1; }
  for (int fi = m->nint; fi < nf; fi++) {
    Face *B = &m->faces[fi];
    B->opp_int = -1;
    for (int q = 0; q < m->cell_nf[B->left]; q++) {
      int g = m->cell_faces[6 * B->left + q];
      if (g < m->nint && m->faces[g].dir == B->dir) {
        B->opp_int = g;
        if (m->faces[g].left == B->left) m->faces[g].bnd_left = fi;
        else m->faces[g].bnd_right = fi;
      }
    }
  }
  return 0;
}

static void mesh_free(Mesh *m) {
  free(m->posx); free(m->posy); free(m->posz); free(m->faces); free(m->cell_faces); free(m->cell_nf);
  free(m->vol); free(m->lcfl);
}

static double dist3(const double *a, const double *b) {
  double v[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* Coord::localProjection / reverseProjection (Coord.cpp:132-150) */
static void proj(double *u, const Face *F) {
  double a = u[0] * F->nrm[0] + u[1] * F->nrm[1] + u[2] * F->nrm[2];
  double b = u[0] * F->tan[0] + u[1] * F->tan[1] + u[2] * F->tan[2];
  double c = u[0] * F->bin[0] + u[1] * F->bin[1] + u[2] * F->bin[2];
  u[0] = a; u[1] = b; u[2] = c;
}
static void rproj(double *u, const Face *F) {
  double a = F->nrm[0] * u[0] + F->tan[0] * u[1] + F->bin[0] * u[2];
  double b = F->nrm[1] * u[0] + F->tan[1] * u[1] + F->bin[1] * u[2];
  double c = F->nrm[2] * u[0] + F->tan[2] * u[1] + F->bin[2] * u[2];
  u[0] = a; u[1] = b; u[2] = c;
}
static void proj_state(const Ctx *c, State *s, const Face *F) {
  if (c->cfg->model == ORC_EULER) proj(s->ph[0].u, F);
  else proj(s->mx.u, F);
}
static void rproj_state(const Ctx *c, State *s, const Face *F) {
  if (c->cfg->model == ORC_EULER) rproj(s->ph[0].u, F);
  else rproj(s->mx.u, F);
}

/* ------------------------------------------------------------------------ */
/* Solver storage                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  Ctx c;
  Mesh m;
  State *sn, *s2;     /* vecPhases and vecPhasesO2 */
  Flux *cons;
  Flux *fflux;        /* per-face flux (after reverse projection) */
  double *fslope;     /* per-face primitive slopes (nvar per face) */
  int nvar;
  double *fdt;        /* per-face dtMax contributions */
  int inflow_nonconv; /* an injection or tank Newton solve hit its 50 iterations */
} Solver;

/* Primitive slope vector: phases (alpha, rho, p) or Euler (rho, p, u, v, w); mixture (u, v, w). */
static void prim_vec(const Ctx *c, const State *s, double *w) {
  int q = 0;
  if (c->cfg->model == ORC_EULER) {
    w[q++] = s->ph[0].rho; w[q++] = s->ph[0].p;
    w[q++] = s->ph[0].u[0]; w[q++] = s->ph[0].u[1]; w[q++] = s->ph[0].u[2];
    return;
  }
  if (c->cfg->model == ORC_EULERHOM) {
    /* PhaseEulerHomogeneous::computeSlopesPhase (alpha), MixEulerHomogeneous::
       computeSlopesMixture (p, u, v, w): PhaseEulerHomogeneous.cpp:139-142,
       MixEulerHomogeneous.cpp:331-337 */
    for (int k = 0; k < c->cfg->nph; k++) w[q++] = s->ph[k].alpha;
    w[q++] = s->mx.p;
    w[q++] = s->mx.u[0]; w[q++] = s->mx.u[1]; w[q++] = s->mx.u[2];
    return;
  }
  if (c->cfg->model == ORC_THERMALEQ) {
    /* PhaseThermalEq::computeSlopesPhase (alpha), MixThermalEq::computeSlopesMixture
       (p, T, u, v, w): PhaseThermalEq.cpp:139-142, MixThermalEq.cpp:457-464 */
    for (int k = 0; k < c->cfg->nph; k++) w[q++] = s->ph[k].alpha;
    w[q++] = s->mx.p; w[q++] = s->mx.T;
    w[q++] = s->mx.u[0]; w[q++] = s->mx.u[1]; w[q++] = s->mx.u[2];
    return;
  }
  for (int k = 0; k < c->cfg->nph; k++) { w[q++] = s->ph[k].alpha; w[q++] = s->ph[k].rho; w[q++] = s->ph[k].p; }
  w[q++] = s->mx.u[0]; w[q++] = s->mx.u[1]; w[q++] = s->mx.u[2];
  /* Transport::computeSlopeTransport (Transport.cpp:156-159) on the same distance */
  for (int t = 0; t < c->cfg->ntr; t++) w[q++] = s->tr[t];
}

/* CellInterfaceO2::computeSlopes (CellInterfaceO2.cpp:132-152) */
static void compute_face_slopes(Solver *S, State *st) {
  Mesh *m = &S->m;
  int nv = S->nvar;
#pragma omp parallel for schedule(static)
  for (int fi = 0; fi < m->nint; fi++) {
    Face *F = &m->faces[fi];
    double pl[3], pr[3], wl[32], wr[32];
    cell_pos(m, F->left, pl); cell_pos(m, F->right, pr);
    double dist = dist3(pl, pr);
    prim_vec(&S->c, &st[F->left], wl);
    prim_vec(&S->c, &st[F->right], wr);
    for (int v = 0; v < nv; v++) S->fslope[(size_t)fi * nv + v] = (wr[v] - wl[v]) / dist;
  }
  /* BoundCondWallO2::computeSlopes (BoundCondWallO2.cpp:81-111): a wall face's
     slope is the normal velocity over the signed centre-to-face distance
     (the mirror image), every other slope 0 */
  for (int fi = m->nint; fi < m->nface; fi++) {
    Face *F = &m->faces[fi];
    double *sl = &S->fslope[(size_t)fi * nv];
    for (int v = 0; v < nv; v++) sl[v] = 0.;
    if (S->c.cfg->bc[F->side] != ORC_BC_WALL) continue;
    double pl[3];
    cell_pos(m, F->left, pl);
    double dd = pl[F->dir] - F->pos[F->dir];
    if (fabs(dd) > 1.e-8) {
      int iu = (S->c.cfg->model == ORC_EULER) ? 2 + F->dir
             : (S->c.cfg->model == ORC_THERMALEQ) ? S->c.cfg->nph + 2 + F->dir : 3 * S->c.cfg->nph + F->dir;
      double w[32];
      prim_vec(&S->c, &st[F->left], w);
      sl[iu] = w[iu] / dd;
    }
  }
}

/* CellO2::computeLocalSlopes (CellO2.cpp:101-247) on a Cartesian mesh: slope1 is
   the reference face's slope, slope2 the opposite parallel internal face's
   (0 next to a boundary).  alpha_after: phase-0 alpha (time n) of the cell
   across the opposite face.  Returns the limited slope vector. */
static void local_slopes(Solver *S, int cell, int fref, int fopp, int cell_is_left,
                         double alpha_cell, double alpha_other, double *alpha_after, double *out) {
  const Ctx *c = &S->c;
  const orc_cfg *cfg = c->cfg;
  int nv = S->nvar;
  const double *s1 = &S->fslope[(size_t)fref * nv];
  double s2[32];
  double aafter = 0.;
  const Face *Fr = &S->m.faces[fref];
  int fb = cell_is_left ? Fr->bnd_left : Fr->bnd_right;
  if (fopp >= 0) {
    for (int v = 0; v < nv; v++) s2[v] = S->fslope[(size_t)fopp * nv + v];
    const Face *G = &S->m.faces[fopp];
    int other = (G->left == cell) ? G->right : G->left;
    aafter = (cfg->model == ORC_EULER) ? 0. : S->sn[other].ph[0].alpha;
  } else if (fb >= 0 && cfg->bc[S->m.faces[fb].side] == ORC_BC_WALL) {
    /* a BoundCondWallO2 face carries slopes; its "cell across" is the cell itself */
    for (int v = 0; v < nv; v++) s2[v] = S->fslope[(size_t)fb * nv + v];
    aafter = (cfg->model == ORC_EULER) ? 0. : S->sn[cell].ph[0].alpha;
  } else {
    for (int v = 0; v < nv; v++) s2[v] = 0.;
  }
  *alpha_after = aafter;
  double epsI = 1.e-4;
  int iface = (alpha_cell >= epsI) && (alpha_cell <= 1. - epsI) &&
              ((alpha_other - alpha_cell) * (alpha_cell - aafter) >= 1.e-8);
  int lg = iface ? cfg->lim[1] : cfg->lim[0];
  int lv = iface ? cfg->lim[3] : cfg->lim[2];
  if (cfg->model == ORC_EULER) {
    for (int v = 0; v < nv; v++) out[v] = limiter(lg, s1[v], s2[v]);
    return;
  }
  if (cfg->model == ORC_THERMALEQ || cfg->model == ORC_EULERHOM) {
    /* PhaseThermalEq::limitSlopes (volume-fraction limiter), MixThermalEq::limitSlopes;
       PhaseEulerHomogeneous::limitSlopes, MixEulerHomogeneous::limitSlopes
       (PhaseEulerHomogeneous.cpp:157-160, MixEulerHomogeneous.cpp:351-357) */
    for (int v = 0; v < nv; v++) out[v] = limiter(v < cfg->nph ? lv : lg, s1[v], s2[v]);
    return;
  }
  int q = 0;
  for (int k = 0; k < cfg->nph; k++) {
    out[q] = limiter(lv, s1[q], s2[q]); q++;
    out[q] = limiter(lg, s1[q], s2[q]); q++;
    out[q] = limiter(lg, s1[q], s2[q]); q++;
  }
  for (int d = 0; d < 3; d++, q++) out[q] = limiter(lg, s1[q], s2[q]);
  /* transports: the volume-fraction limiter of the same interface test (CellO2.cpp:232-246) */
  for (int t = 0; t < cfg->ntr; t++, q++) out[q] = limiter(lv, s1[q], s2[q]);
}

/* CellO2::computeLocalSlopesLimite (CellO2.cpp:250-354) for a wall face:
   slope1 is the wall's own slope, slope2 the internal face across the cell
   (0 if none); the interface test uses the cell's phase-0 alpha only. */
static void local_slopes_limite(Solver *S, int cell, int fwall, int fopp, double *out) {
  const orc_cfg *cfg = S->c.cfg;
  int nv = S->nvar;
  const double *s1 = &S->fslope[(size_t)fwall * nv];
  double s2[32];
  for (int v = 0; v < nv; v++) s2[v] = (fopp >= 0) ? S->fslope[(size_t)fopp * nv + v] : 0.;
  int iface = 0;
  if (cfg->model != ORC_EULER) {
    double a = S->sn[cell].ph[0].alpha;
    iface = (a >= 1.e-5) && (a <= 1. - 1.e-5);
  }
  int lg = iface ? cfg->lim[1] : cfg->lim[0];
  int lv = iface ? cfg->lim[3] : cfg->lim[2];
  if (cfg->model == ORC_EULER) {
    for (int v = 0; v < nv; v++) out[v] = limiter(lg, s1[v], s2[v]);
    return;
  }
  if (cfg->model == ORC_THERMALEQ || cfg->model == ORC_EULERHOM) {
    /* PhaseThermalEq::limitSlopes (volume-fraction limiter), MixThermalEq::limitSlopes;
       PhaseEulerHomogeneous::limitSlopes, MixEulerHomogeneous::limitSlopes
       (PhaseEulerHomogeneous.cpp:157-160, MixEulerHomogeneous.cpp:351-357) */
    for (int v = 0; v < nv; v++) out[v] = limiter(v < cfg->nph ? lv : lg, s1[v], s2[v]);
    return;
  }
  int q = 0;
  for (int k = 0; k < cfg->nph; k++) {
    out[q] = limiter(lv, s1[q], s2[q]); q++;
    out[q] = limiter(lg, s1[q], s2[q]); q++;
    out[q] = limiter(lg, s1[q], s2[q]); q++;
  }
  for (int d = 0; d < 3; d++, q++) out[q] = limiter(lg, s1[q], s2[q]);
  for (int t = 0; t < cfg->ntr; t++, q++) out[q] = limiter(lv, s1[q], s2[q]); /* CellO2.cpp:341-355 */
}

/* Phase*::extrapolate / Mix*::extrapolate without the phase verification
   (BoundCondWallO2::solveRiemann, BoundCondWallO2.cpp:115-150, calls none) */
static void extrapolate_raw(const Ctx *c, State *s, const double *sl, double dist) {
  int q = 0;
  if (c->cfg->model == ORC_THERMALEQ) {
    for (int k = 0; k < c->cfg->nph; k++, q++) s->ph[k].alpha += sl[q] * dist;
    s->mx.p += sl[q] * dist; q++;
    s->mx.T += sl[q] * dist; q++;
    for (int d = 0; d < 3; d++, q++) s->mx.u[d] = s->mx.u[d] + sl[q] * dist;
    return;
  }
  if (c->cfg->model == ORC_EULER) {
    Phase *p = &s->ph[0];
    p->rho += sl[0] * dist; p->p += sl[1] * dist;
    for (int d = 0; d < 3; d++) p->u[d] = p->u[d] + sl[2 + d] * dist;
    return;
  }
  for (int k = 0; k < c->cfg->nph; k++) {
    Phase *p = &s->ph[k];
    p->alpha += sl[q] * dist; q++;
    p->rho += sl[q] * dist; q++;
    p->p += sl[q] * dist; q++;
  }
  for (int d = 0; d < 3; d++, q++) s->mx.u[d] = s->mx.u[d] + sl[q] * dist;
  for (int t = 0; t < c->cfg->ntr; t++, q++) s->tr[t] += sl[q] * dist; /* BoundCondWallO2.cpp:142-144 */
}

static void extrapolate(const Ctx *c, State *s, const double *sl, double dist, double sign) {
  int q = 0;
  if (c->cfg->model == ORC_EULERHOM) {
    /* PhaseEulerHomogeneous::extrapolate + verifyAndCorrectPhase per phase, then
       MixEulerHomogeneous::extrapolate (MixEulerHomogeneous.cpp:341-347) */
    for (int k = 0; k < c->cfg->nph; k++, q++) {
      s->ph[k].alpha += (sign * sl[q]) * dist;
      verify_correct(c, k, &s->ph[k]);
    }
    s->mx.p += (sign * sl[q]) * dist; q++;
    for (int d = 0; d < 3; d++, q++) s->mx.u[d] = s->mx.u[d] + (sign * sl[q]) * dist;
    return;
  }
  if (c->cfg->model == ORC_THERMALEQ) {
    /* PhaseThermalEq::extrapolate + verifyAndCorrectPhase per phase, then
       MixThermalEq::extrapolate (MixThermalEq.cpp:471-478) */
    for (int k = 0; k < c->cfg->nph; k++, q++) {
      s->ph[k].alpha += (sign * sl[q]) * dist;
      verify_correct(c, k, &s->ph[k]);
    }
    s->mx.p += (sign * sl[q]) * dist; q++;
    s->mx.T += (sign * sl[q]) * dist; q++;
    for (int d = 0; d < 3; d++, q++) s->mx.u[d] = s->mx.u[d] + (sign * sl[q]) * dist;
    return;
  }
  if (c->cfg->model == ORC_EULER) {
    Phase *p = &s->ph[0];
    p->rho += (sign * sl[0]) * dist; p->p += (sign * sl[1]) * dist;
    p->u[0] = p->u[0] + (sign * sl[2]) * dist;
    p->u[1] = p->u[1] + (sign * sl[3]) * dist;
    p->u[2] = p->u[2] + (sign * sl[4]) * dist;
    verify_correct(c, 0, p);
    return;
  }
  for (int k = 0; k < c->cfg->nph; k++) {
    Phase *p = &s->ph[k];
    p->alpha += (sign * sl[q]) * dist; q++;
    p->rho += (sign * sl[q]) * dist; q++;
    p->p += (sign * sl[q]) * dist; q++;
    verify_correct(c, k, p);
  }
  for (int d = 0; d < 3; d++, q++) s->mx.u[d] = s->mx.u[d] + (sign * sl[q]) * dist;
  /* Transport::extrapolate (Transport.cpp:163-166), the slope negated for the
     right state (CellInterfaceO2.cpp:222-226, 254-257) */
  for (int t = 0; t < c->cfg->ntr; t++, q++) s->tr[t] += (sign * sl[q]) * dist;
}

/* THINC reconstruction, CellInterfaceO2.cpp:228-247 (left state) and 259-277
   (right state).  a_hi / a_lo are the phase-0 fractions on the two sides of
   the cell as the reference orders them: (right cell, left-left cell) for the
   left state, (right-right cell, left cell) for the right state. */
static void thinc(State *s, double acell, double a_hi, double a_lo, int right_side) {
  double beta = 1.6, epsI = 1.e-4, sign;
  if ((acell >= epsI) && (acell <= 1. - epsI) && ((a_hi - acell) * (acell - a_lo) >= 1.e-8)) {
    if (a_hi - a_lo > 0.) sign = 1.; else sign = -1.;
    double qmin = smin(a_hi, a_lo);
    double qmax = smax(a_hi, a_lo) - qmin;
    double C = (acell - qmin + 1.e-20) / (qmax + 1.e-20);
    double B = exp(sign * beta * (2. * C - 1.));
    double A = (B / cosh(beta) - 1.) / tanh(beta);
    double na;
    if (!right_side) na = qmin + 0.5 * qmax * (1. + sign * (tanh(beta) + A) / (1. + A * tanh(beta)));
    else na = qmin + 0.5 * qmax * (1. + sign * A);
    if (na < epsI) na = epsI;
    if (na > 1. - epsI) na = 1. - epsI;
    s->ph[0].alpha = na;
    s->ph[1].alpha = 1. - na;
  }
}

static int is_thinc(const orc_cfg *cfg) { return cfg->lim[2] == ORC_LIM_THINC || cfg->lim[3] == ORC_LIM_THINC; }

/* Transport fluxes of a face after its Riemann problem, sM known
   (CellInterface.cpp:139-140, BoundCond.cpp:95-96): Transport::solveRiemann
   on the two face states (Transport.cpp:68-72; a transmissive boundary passes
   its cell twice, BoundCondAbs.cpp:77-80), zero at a wall (solveRiemannWall,
   Transport.cpp:76-79), and at an outflow, injection or tank the cell's value
   leaving or the boundary's value entering (Transport.cpp:83-104). */
static void transport_flux(const Ctx *c, const State *L, const State *R, int bc, int side, Flux *F) {
  double sM = F->sM;
  for (int t = 0; t < c->cfg->ntr; t++) {
    if (R) F->tr[t] = (sM > 0.) ? L->tr[t] * sM : R->tr[t] * sM;
    else if (bc == ORC_BC_ABS) F->tr[t] = (sM > 0.) ? L->tr[t] * sM : L->tr[t] * sM;
    else if (bc == ORC_BC_WALL) F->tr[t] = 0.;
    else F->tr[t] = (sM > 0.) ? L->tr[t] * sM : c->cfg->bc_tr[side][t] * sM;
  }
}

/* Face Riemann problems of one pass (Run.cpp:356,381,401); fluxes stored per face. */
static void face_fluxes(Solver *S, State *st, int o2) {
  Mesh *m = &S->m;
  const Ctx *c = &S->c;
  double lcfl = m->lcfl[0];
#pragma omp parallel for schedule(static)
  for (int fi = 0; fi < m->nface; fi++) {
    Face *F = &m->faces[fi];
    double dtm = 1.e10;
    Flux *FL = &S->fflux[fi];
    memset(FL, 0, sizeof(Flux));
    if (F->right < 0) {
      /* BoundCond::solveRiemann (BoundCond.cpp:82-100) + BoundCondAbs (BoundCondAbs.cpp:70-73)
         or BoundCondWall (BoundCondWall.cpp:68-71) */
      State L = st[F->left];
      if (o2 && c->cfg->bc[F->side] == ORC_BC_WALL) {
        /* BoundCondWallO2::solveRiemann: the cell extrapolated to the wall
           (distance negated when the normal points to -x, -y or -z) */
        double sl[32], pl[3];
        local_slopes_limite(S, F->left, fi, F->opp_int, sl);
        cell_pos(m, F->left, pl);
        double dg = dist3(F->pos, pl);
        if (F->nrm[0] < 0. || F->nrm[1] < 0. || F->nrm[2] < 0.) dg = -dg;
        extrapolate_raw(c, &L, sl, dg);
      }
      proj_state(c, &L, F);
      fulfill(c, &L);
      if (c->cfg->bc[F->side] == ORC_BC_WALL) riemann_wall(c, &L, lcfl, &dtm, FL);
      else if (c->cfg->bc[F->side] == ORC_BC_OUTFLOW) riemann_outflow(c, &L, lcfl, &dtm, c->cfg->bc_p0[F->side], FL);
      else if (c->cfg->bc[F->side] == ORC_BC_INJ || c->cfg->bc[F->side] == ORC_BC_TANK) {
        int nc = 0;
        if (c->cfg->bc[F->side] == ORC_BC_INJ) riemann_inflow(c, &L, lcfl, &dtm, F->side, FL, &nc);
        else riemann_tank(c, &L, lcfl, &dtm, F->side, FL, &nc);
        if (nc) {
#pragma omp atomic write
          S->inflow_nonconv = 1;
        }
      }
      else riemann(c, &L, &L, lcfl, lcfl, &dtm, FL);
      transport_flux(c, &L, NULL, c->cfg->bc[F->side], F->side, FL);
    } else if (!o2) {
      /* CellInterface::solveRiemann (CellInterface.cpp:127-146): the cells' own cached state */
      State L = st[F->left], R = st[F->right];
      proj_state(c, &L, F); proj_state(c, &R, F);
      riemann(c, &L, &R, lcfl, lcfl, &dtm, FL);
      transport_flux(c, &L, &R, 0, -1, FL);
    } else {
      /* CellInterfaceO2::solveRiemann (CellInterfaceO2.cpp:197-295) */
      State L = st[F->left], R = st[F->right];
      double pl[3], pr[3];
      cell_pos(m, F->left, pl); cell_pos(m, F->right, pr);
      double dL = dist3(F->pos, pl), dR = dist3(F->pos, pr);
      int multi = (c->cfg->model != ORC_EULER);
      double aL = multi ? L.ph[0].alpha : 0., aR = multi ? R.ph[0].alpha : 0.;
      double slL[32], slR[32], aLL, aRR;
      local_slopes(S, F->left, fi, F->opp_left, 1, aL, aR, &aLL, slL);
      extrapolate(c, &L, slL, dL, 1.);
      if (is_thinc(c->cfg)) thinc(&L, aL, aR, aLL, 0);
      local_slopes(S, F->right, fi, F->opp_right, 0, aR, aL, &aRR, slR);
      extrapolate(c, &R, slR, dR, -1.);
      if (is_thinc(c->cfg)) thinc(&R, aR, aRR, aL, 1);
      proj_state(c, &L, F); proj_state(c, &R, F);
      fulfill(c, &L); fulfill(c, &R);
      riemann(c, &L, &R, lcfl, lcfl, &dtm, FL);
      transport_flux(c, &L, &R, 0, -1, FL);
    }
    rproj(FL->qdm, F);
    S->fdt[fi] = dtm;
  }
}

/* CellInterface::addFlux / subtractFlux (CellInterface.cpp:150-180) with the
   Flux*::addNonCons / subtractNonCons terms (FluxKapila.cpp:215-235). The
   non-conservative terms use the owning cell's time-n alpha_k, p_k. */
static void accumulate(Solver *S) {
  Mesh *m = &S->m;
  const Ctx *c = &S->c;
  int n = c->cfg->nph;
  /* FluxEulerHomogeneous::addFlux/subtractFlux (FluxEulerHomogeneous.cpp:
     25-39): one mass, as the Euler model, no non-conservative terms */
  int euler = (c->cfg->model == ORC_EULER || c->cfg->model == ORC_EULERHOM);
  int te = (c->cfg->model == ORC_THERMALEQ);
  double vol = m->vol[0];
#pragma omp parallel for schedule(static)
  for (int ci = 0; ci < m->ncell; ci++) {
    Flux *U = &S->cons[ci];
    memset(U, 0, sizeof(Flux));
    const State *sn = &S->sn[ci];
    for (int q = 0; q < m->cell_nf[ci]; q++) {
      const Face *F = &m->faces[m->cell_faces[6 * ci + q]];
      const Flux *FL = &S->fflux[m->cell_faces[6 * ci + q]];
      double coefA = F->surface / vol;
      coefA = coefA * 1.;
      if (F->right == ci) {
        if (euler) {
          U->masse[0] += coefA * FL->masse[0];
        } else if (te) {
          /* FluxThermalEq::addFlux (FluxThermalEq.cpp:62-69): no non-conservative terms */
          for (int k = 0; k < n; k++) U->masse[k] += coefA * FL->masse[k];
        } else {
          for (int k = 0; k < n; k++) {
            U->alpha[k] += coefA * FL->alpha[k];
            U->masse[k] += coefA * FL->masse[k];
            U->energ[k] += coefA * FL->energ[k];
          }
        }
        for (int d = 0; d < 3; d++) U->qdm[d] += FL->qdm[d] * coefA;
        U->energMix += coefA * FL->energMix;
        if (!euler && !te)
          for (int k = 0; k < n; k++) {
            U->alpha[k] += -coefA * sn->ph[k].alpha * FL->sM;
            U->energ[k] += coefA * sn->ph[k].alpha * sn->ph[k].p * FL->sM;
          }
        /* Transport::addFlux, addNonCons with the cell's time-n value
           (CellInterface.cpp:158-162, Transport.cpp:107-126) */
        for (int t = 0; t < c->cfg->ntr; t++) {
          U->tr[t] += coefA * FL->tr[t];
          U->tr[t] += -coefA * sn->tr[t] * FL->sM;
        }
      } else {
        if (euler) {
          U->masse[0] -= coefA * FL->masse[0];
        } else if (te) {
          for (int k = 0; k < n; k++) U->masse[k] -= coefA * FL->masse[k];
        } else {
          for (int k = 0; k < n; k++) {
            U->alpha[k] -= coefA * FL->alpha[k];
            U->masse[k] -= coefA * FL->masse[k];
            U->energ[k] -= coefA * FL->energ[k];
          }
        }
        for (int d = 0; d < 3; d++) U->qdm[d] -= FL->qdm[d] * coefA;
        U->energMix -= coefA * FL->energMix;
        if (!euler && !te)
          for (int k = 0; k < n; k++) {
            U->alpha[k] -= -coefA * sn->ph[k].alpha * FL->sM;
            U->energ[k] -= coefA * sn->ph[k].alpha * sn->ph[k].p * FL->sM;
          }
        /* Transport::subtractFlux, subtractNonCons (CellInterface.cpp:175-179,
           Transport.cpp:114-131) */
        for (int t = 0; t < c->cfg->ntr; t++) {
          U->tr[t] -= coefA * FL->tr[t];
          U->tr[t] -= -coefA * sn->tr[t] * FL->sM;
        }
      }
    }
  }
}

static double reduce_dt(Solver *S) {
  double d = 1.e10;
  for (int fi = 0; fi < S->m.nface; fi++) d = smin(d, S->fdt[fi]);
  return d;
}

/* U <- U*mult + Un, the Flux::multiply / addFlux(1.) pair of Cell::timeEvolution
   (Cell.cpp:214-227) and CellO2::predictionOrdre2 (CellO2.cpp:376-397). */
static void scale_add(const Ctx *c, Flux *U, double mult, const Flux *buf) {
  int n = c->cfg->nph;
  if (c->cfg->model == ORC_EULER || c->cfg->model == ORC_EULERHOM) {
    U->masse[0] *= mult;
  } else if (c->cfg->model == ORC_THERMALEQ) {
    for (int k = 0; k < n; k++) U->masse[k] *= mult;
  } else {
    for (int k = 0; k < n; k++) { U->alpha[k] *= mult; U->masse[k] *= mult; U->energ[k] *= mult; }
  }
  for (int d = 0; d < 3; d++) U->qdm[d] *= mult;
  U->energMix *= mult;
  if (c->cfg->model == ORC_EULER || c->cfg->model == ORC_EULERHOM) {
    U->masse[0] += 1. * buf->masse[0];
  } else if (c->cfg->model == ORC_THERMALEQ) {
    for (int k = 0; k < n; k++) U->masse[k] += 1. * buf->masse[k];
  } else {
    for (int k = 0; k < n; k++) {
      U->alpha[k] += 1. * buf->alpha[k];
      U->masse[k] += 1. * buf->masse[k];
      U->energ[k] += 1. * buf->energ[k];
    }
  }
  for (int d = 0; d < 3; d++) U->qdm[d] += buf->qdm[d] * 1.;
  U->energMix += 1. * buf->energMix;
}

/* Run::solveRelaxations (Run.cpp:457-474) on the time-n state. */
static void solve_relaxations(Solver *S) {
  const Ctx *c = &S->c;
  if (c->cfg->nph < 2 || c->cfg->model == ORC_EULER) return;
#pragma omp parallel for schedule(static)
  for (int ci = 0; ci < S->m.ncell; ci++) {
    State *s = &S->sn[ci];
    if (model_relaxes(c->cfg)) relax_p(c, s);
    if (c->cfg->relax_pt) relax_pt(c, s);
    if (c->cfg->relax_ptmu) relax_ptmu(c, s, s);
    total_to_internal(c, s);
    correction_energy(c, s, s);
  }
}

/* Transports of Cell::timeEvolution (Cell.cpp:222-226) and
   CellO2::predictionOrdre2 (CellO2.cpp:385-390): the accumulated value times
   dt (0.5*dt) added to the time-n value; no buildPrim, no relaxation. */
static void transport_update(const Ctx *c, Flux *U, double mult, const State *sn, State *out) {
  for (int t = 0; t < c->cfg->ntr; t++) {
    U->tr[t] *= mult;
    out->tr[t] = sn->tr[t] + U->tr[t];
  }
}

/* One time step of Run::solver (Run.cpp:204-277 / 281-412). Returns dtMax. */
static double step(Solver *S, double dt) {
  const Ctx *c = &S->c;
  Mesh *m = &S->m;
  double dtMax = 1.e10;
  if (c->cfg->order == 1) {
    face_fluxes(S, S->sn, 0);
    dtMax = smin(dtMax, reduce_dt(S));
    accumulate(S);
#pragma omp parallel for schedule(static)
    for (int ci = 0; ci < m->ncell; ci++) {
      Flux buf; memset(&buf, 0, sizeof(buf));
      build_cons(c, &S->sn[ci], &buf);
      scale_add(c, &S->cons[ci], dt, &buf);
      scheme_correction(c, &S->cons[ci], &buf);
      build_prim(c, &S->cons[ci], &S->sn[ci]);
      transport_update(c, &S->cons[ci], dt, &S->sn[ci], &S->sn[ci]);
    }
  } else {
    /* integrationProcedure: slopes on U^n (Run.cpp:293-295) */
    compute_face_slopes(S, S->sn);
    face_fluxes(S, S->sn, 1);
    dtMax = smin(dtMax, reduce_dt(S));
    accumulate(S);
    /* CellO2::predictionOrdre2 (CellO2.cpp:376-397) */
#pragma omp parallel for schedule(static)
    for (int ci = 0; ci < m->ncell; ci++) {
      Flux buf; memset(&buf, 0, sizeof(buf));
      State *sn = &S->sn[ci], *s2 = &S->s2[ci];
      build_cons(c, sn, &buf);
      scale_add(c, &S->cons[ci], 0.5 * dt, &buf);
      scheme_correction(c, &S->cons[ci], &buf);
      build_prim(c, &S->cons[ci], s2);
      transport_update(c, &S->cons[ci], 0.5 * dt, sn, s2);
      if (model_relaxes(c->cfg) && c->cfg->nph > 1) relax_p(c, s2);
      if (c->cfg->relax_pt) relax_pt(c, s2);
      if (c->cfg->relax_ptmu) relax_ptmu(c, s2, sn);
      total_to_internal(c, s2);
      correction_energy(c, s2, sn);
      fulfill(c, s2);
    }
    /* second pass on U* (Run.cpp:372-390) */
    compute_face_slopes(S, S->s2);
    face_fluxes(S, S->s2, 1);
    dtMax = smin(dtMax, reduce_dt(S));
    accumulate(S);
#pragma omp parallel for schedule(static)
    for (int ci = 0; ci < m->ncell; ci++) {
      Flux buf; memset(&buf, 0, sizeof(buf));
      build_cons(c, &S->sn[ci], &buf);
      scale_add(c, &S->cons[ci], dt, &buf);
      scheme_correction(c, &S->cons[ci], &buf);
      build_prim(c, &S->cons[ci], &S->sn[ci]);
      transport_update(c, &S->cons[ci], dt, &S->sn[ci], &S->sn[ci]);
    }
  }
  solve_relaxations(S);
  return dtMax;
}

/* Initial state: Cell::fill + Cell::completeFulfillState (Cell.cpp:127-155, 271-278). */
static void init_state(Solver *S, const double *in) {
  const Ctx *c = &S->c;
  int n = c->cfg->nph, nc = S->m.ncell;
  for (int ci = 0; ci < nc; ci++) {
    State *s = &S->sn[ci];
    memset(s, 0, sizeof(State));
    double u[3] = {in[(size_t)(3 * n) * nc + ci], in[(size_t)(3 * n + 1) * nc + ci], in[(size_t)(3 * n + 2) * nc + ci]};
    if (c->cfg->model == ORC_EULER) {
      s->ph[0].rho = in[(size_t)1 * nc + ci];
      s->ph[0].p = in[(size_t)2 * nc + ci];
      for (int d = 0; d < 3; d++) s->ph[0].u[d] = u[d];
      fulfill(c, s);
      continue;
    }
    if (c->cfg->model == ORC_EULERHOM) {
      /* Cell::fill + completeFulfillState: alpha_k, the mixture p, u; the
         phases take the saturated state of p (ModEulerHomogeneous::fulfillState) */
      for (int k = 0; k < n; k++) s->ph[k].alpha = in[(size_t)(3 * k) * nc + ci];
      s->mx.p = in[(size_t)2 * nc + ci];
      for (int d = 0; d < 3; d++) s->mx.u[d] = u[d];
      fulfill(c, s);
      s->mx.E = s->mx.e + 0.5 * sqnorm(s->mx.u);
      /* the predictor's state (CellO2::m_vecPhasesO2) starts as a copy of the
         domain's phases (CellO2::copyPhase, CellO2.cpp:93-97); its liquid
         pressure is the first predictor's Newton start */
      S->s2[ci] = *s;
      continue;
    }
    if (c->cfg->model == ORC_THERMALEQ) {
      /* Cell::fill + completeFulfillState: alpha_k, the mixture p and T, u */
      for (int k = 0; k < n; k++) s->ph[k].alpha = in[(size_t)(3 * k) * nc + ci];
      s->mx.p = in[(size_t)2 * nc + ci];
      s->mx.T = in[(size_t)(3 * n + 4) * nc + ci];
      for (int d = 0; d < 3; d++) s->mx.u[d] = u[d];
      fulfill(c, s);
      s->mx.E = s->mx.e + 0.5 * sqnorm(s->mx.u);
      continue;
    }
    for (int k = 0; k < n; k++) {
      s->ph[k].alpha = in[(size_t)(3 * k) * nc + ci];
      s->ph[k].rho = in[(size_t)(3 * k + 1) * nc + ci];
      s->ph[k].p = in[(size_t)(3 * k + 2) * nc + ci];
    }
    for (int d = 0; d < 3; d++) s->mx.u[d] = u[d];
    fulfill(c, s);
    s->mx.E = s->mx.e + 0.5 * sqnorm(s->mx.u);
    for (int t = 0; t < c->cfg->ntr; t++) s->tr[t] = in[(size_t)(3 * n + 4 + t) * nc + ci];
  }
}

static void write_state(Solver *S, double *out) {
  const Ctx *c = &S->c;
  int n = c->cfg->nph, nc = S->m.ncell;
  for (int ci = 0; ci < nc; ci++) {
    const State *s = &S->sn[ci];
    if (c->cfg->model == ORC_EULER) {
      out[(size_t)0 * nc + ci] = 1.;
      out[(size_t)1 * nc + ci] = s->ph[0].rho;
      out[(size_t)2 * nc + ci] = s->ph[0].p;
      for (int d = 0; d < 3; d++) out[(size_t)(3 + d) * nc + ci] = s->ph[0].u[d];
      out[(size_t)6 * nc + ci] = s->ph[0].Et;
      continue;
    }
    for (int k = 0; k < n; k++) {
      out[(size_t)(3 * k) * nc + ci] = s->ph[k].alpha;
      out[(size_t)(3 * k + 1) * nc + ci] = s->ph[k].rho;
      out[(size_t)(3 * k + 2) * nc + ci] = s->ph[k].p;
    }
    for (int d = 0; d < 3; d++) out[(size_t)(3 * n + d) * nc + ci] = s->mx.u[d];
    out[(size_t)(3 * n + 3) * nc + ci] = s->mx.E;
    for (int t = 0; t < c->cfg->ntr; t++) out[(size_t)(3 * n + 4 + t) * nc + ci] = s->tr[t];
    if (c->cfg->model == ORC_THERMALEQ || c->cfg->model == ORC_EULERHOM) out[(size_t)(3 * n + 4) * nc + ci] = s->mx.T;
  }
}

int orc_run(const orc_cfg *cfg, const double *state_in, double *state_out,
            int nsteps, double t_final, int max_steps,
            double *t_out, double *dt_next_out, double *dt_log, int nthreads) {
  return orc_run_from(cfg, state_in, state_out, nsteps, t_final, max_steps, 0., 1.e-15,
                      t_out, dt_next_out, dt_log, nthreads);
}

/* Run::resumeSimulation (src/Run.cpp:136-190): the state is initialised from
   the primitives as at t = 0 (Cell::completeFulfillState), the physical time
   starts at t0 and the first step uses dt0 (both read from infoCalcul.out by
   Output::readInfos, src/InputOutput/Output.cpp:370-394). */
int orc_run_from(const orc_cfg *cfg, const double *state_in, double *state_out,
                 int nsteps, double t_final, int max_steps, double t0, double dt0,
                 double *t_out, double *dt_next_out, double *dt_log, int nthreads) {
  if (!cfg || cfg->nph < 1 || cfg->nph > MAXPH) return -1;
  if (cfg->model == ORC_EULER && cfg->nph != 1) return -1;
  if (is_thinc(cfg) && cfg->nph != 2) return -1;
  /* the Euler model has no transport Riemann solver (Model.h:127, an error) */
  if (cfg->ntr < 0 || cfg->ntr > MAXTR || (cfg->ntr > 0 && cfg->model == ORC_EULER)) return -1;
  /* ThermalEq: two phases (MixThermalEq::computePressure), no transport
     solver, injection or tank (ModThermalEq has no solveRiemannInflow; its
     tank is not restated here) */
  /* RelaxationPTMu: two phases (RelaxationPTMu.cpp:30), here after Kapila's
     pressure relaxation, not together with RelaxationPT */
  if (cfg->relax_ptmu && (cfg->nph != 2 || cfg->model != ORC_KAPILA || cfg->relax_pt)) return -1;
  /* RelaxationPT: two phases (RelaxationPT.cpp:60), Kapila or MultiP */
  if (cfg->relax_pt && (cfg->nph != 2 || (cfg->model != ORC_KAPILA && cfg->model != ORC_MULTIP))) return -1;
  if (cfg->model == ORC_THERMALEQ) {
    if (cfg->nph != 2 || cfg->ntr > 0) return -1;
    for (int b = 0; b < 6; b++) if (cfg->bc[b] == ORC_BC_INJ || cfg->bc[b] == ORC_BC_TANK) return -1;
  }
  /* EulerHomogeneous: two phases, a liquid and a vapour (Input.cpp:479-483);
     only the interior Riemann solver (no wall, outflow, injection, tank or
     transport solver: Model.h:93-145 end the run); THINC sets two fractions
     of a phase-0/phase-1 pair as any model */
  if (cfg->model == ORC_EULERHOM) {
    if (cfg->nph != 2 || cfg->ntr > 0 || cfg->relax_pt || cfg->relax_ptmu) return -1;
    if (cfg->liq < 0 || cfg->liq > 1 || cfg->vap != 1 - cfg->liq) return -1;
    for (int b = 0; b < 6; b++) if (cfg->bc[b] != ORC_BC_ABS) return -1;
  }
  for (int b = 0; b < 6; b++) {
    if (cfg->bc[b] < ORC_BC_ABS || cfg->bc[b] > ORC_BC_TANK) return -1;

  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  Solver S;
  memset(&S, 0, sizeof(S));
  g_relax_fail = 0;
  g_sat_fail = 0;
  S.c.cfg = cfg;
  S.c.epsilon = (cfg->model == ORC_KAPILA && cfg->alpha_null) ? 1.e-15 : 0.;
  mesh_build(&S.m, cfg);
  S.nvar = (cfg->model == ORC_EULER) ? 5 : (cfg->model == ORC_THERMALEQ) ? cfg->nph + 5
         : (cfg->model == ORC_EULERHOM) ? cfg->nph + 4 : 3 * cfg->nph + 3 + cfg->ntr;
  S.sn = calloc(S.m.ncell, sizeof(State));
  S.s2 = calloc(S.m.ncell, sizeof(State));
  S.cons = calloc(S.m.ncell, sizeof(Flux));
  S.fflux = calloc(S.m.nface, sizeof(Flux));
  S.fslope = calloc((size_t)S.m.nface * S.nvar + 1, sizeof(double));
  S.fdt = calloc(S.m.nface, sizeof(double));
  if (!S.sn || !S.s2 || !S.cons || !S.fflux || !S.fslope || !S.fdt) return -2;
  init_state(&S, state_in);

  double dt = dt0, t = t0;
  int it = 0;
  float tf = (float)t_final;
  g_inflow_fail_step = 0;
  for (;;) {
    double dtMax = step(&S, dt);
    if (S.inflow_nonconv) {
      /* Errors::errorMessage ends the reference's process during this step
         (ModKapila.cpp:283, ModEuler.cpp, ModMultiP.cpp solveRiemannInflow) */
      g_inflow_fail_step = it + 1;
      break;
    }
    if (dt_log) dt_log[it] = dt;
    t += dt;
    it++;
    int done = (nsteps > 0) ? (it >= nsteps) : (t >= (double)tf);
    dt = cfg->cfl * dtMax;
    if (done || (max_steps > 0 && it >= max_steps)) break;
  }
  write_state(&S, state_out);
  if (t_out) *t_out = t;
  if (dt_next_out) *dt_next_out = dt;
  free(S.sn); free(S.s2); free(S.cons); free(S.fflux); free(S.fslope); free(S.fdt);
  mesh_free(&S.m);
  /* the reference stops with "solveRiemannInflow (or Tank) not converged" */
  return S.inflow_nonconv ? -3 : it;
}

/* ------------------------------------------------------------------------ */
/* Single-function entry points for known-answer tests                       */
/* ------------------------------------------------------------------------ */
double orc_limiter(int lim, double s1, double s2) { return limiter(lim, s1, s2); }
double orc_eos_energy(const orc_eos *e, double rho, double p, double eps) { return eos_e(e, rho, p, eps); }
double orc_eos_pressure(const orc_eos *e, double rho, double ein) { return eos_p(e, rho, ein); }
double orc_eos_sound(const orc_eos *e, double rho, double p, double eps) { return eos_c(e, rho, p, eps); }
double orc_eos_isentrope(const orc_eos *e, double p0, double rho0, double rho1, double eps) {
  return eos_isentrope(e, p0, rho0, rho1, eps);
}
double orc_eos_rho_pfinal(const orc_eos *e, double p0, double rho0, double pf, double *drho, double eps) {
  return eos_rho_pfinal(e, p0, rho0, pf, drho, eps);
}

int orc_inflow_failure_step(void) { return g_inflow_fail_step; }

long orc_relax_failures(double *first_inputs) {
  if (first_inputs) memcpy(first_inputs, g_relax_fail_in, sizeof(g_relax_fail_in));
  return g_relax_fail;
}

int orc_relax_p(const orc_cfg *cfg, double *alpha, double *rho, double *p) {
  Ctx c = {cfg, (cfg->model == ORC_KAPILA && cfg->alpha_null) ? 1.e-15 : 0.};
  State s; memset(&s, 0, sizeof(s));
  for (int k = 0; k < cfg->nph; k++) { s.ph[k].alpha = alpha[k]; s.ph[k].rho = rho[k]; s.ph[k].p = p[k]; }
  int it = relax_p(&c, &s);
  for (int k = 0; k < cfg->nph; k++) { alpha[k] = s.ph[k].alpha; rho[k] = s.ph[k].rho; p[k] = s.ph[k].p; }
  return it;
}

void orc_hllc(const orc_cfg *cfg, const double *wl, const double *wr, double dxl, double dxr,
              double *flux_out, double *dtmax) {
  Ctx c = {cfg, (cfg->model == ORC_KAPILA && cfg->alpha_null) ? 1.e-15 : 0.};
  State L, R; memset(&L, 0, sizeof(L)); memset(&R, 0, sizeof(R));
  int n = cfg->nph;
  const double *w[2] = {wl, wr};
  State *s[2] = {&L, &R};
  for (int side = 0; side < 2; side++) {
    if (cfg->model == ORC_EULER) {
      s[side]->ph[0].rho = w[side][1]; s[side]->ph[0].p = w[side][2];
      for (int d = 0; d < 3; d++) s[side]->ph[0].u[d] = w[side][3 + d];
    } else {
      for (int k = 0; k < n; k++) {
        s[side]->ph[k].alpha = w[side][3 * k]; s[side]->ph[k].rho = w[side][3 * k + 1]; s[side]->ph[k].p = w[side][3 * k + 2];
      }
      for (int d = 0; d < 3; d++) s[side]->mx.u[d] = w[side][3 * n + d];
    }
    fulfill(&c, s[side]);
  }
  Flux F; memset(&F, 0, sizeof(F));
  riemann(&c, &L, &R, dxl, dxr, dtmax, &F);
  int q = 0;
  if (cfg->model == ORC_EULER) {
    flux_out[q++] = F.masse[0];
  } else {
    for (int k = 0; k < n; k++) flux_out[q++] = F.alpha[k];
    for (int k = 0; k < n; k++) flux_out[q++] = F.masse[k];
    for (int k = 0; k < n; k++) flux_out[q++] = F.energ[k];
  }
  for (int d = 0; d < 3; d++) flux_out[q++] = F.qdm[d];
  flux_out[q++] = F.energMix;
  flux_out[q++] = F.sM;
}
