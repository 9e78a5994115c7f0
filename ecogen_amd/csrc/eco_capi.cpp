// C ABI of the explicit finite-volume step (include/ecogen.h).
//
// A context owns the device state of one solve: two time-level buffers
// (ping-pong, the state after k steps lives in buffer k&1), the MUSCL
// intermediate state U*, the per-direction geometry tables and the control
// block that carries the lagged dt, the physical time and the done flags on
// the device, so that steps are enqueued without host round trips.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/ecogen.h"
#include "eco_launch.h"
#include "eco_host.h"

using namespace eco;

struct eco_ctx {
  eco_config cfg;
  Geometry geo;
  SchemeParams sch;
  hipStream_t stream;
  int own_stream;
  int nprim;            // device primitive fields
  int nlevel;           // fields per time level (nprim, E, cached mixture density)
  int ntr;              // transported scalars: fields nlevel.. of a time level, nprim.. of U*
  double *buf[2];       // time levels, nlevel + ntr fields each
  double *bstar;        // U*, nprim + ntr fields
  double *bn2;          // RelaxationPTMu in MUSCL: the time-n state relaxed by the predictor
  double *smf[3];       // with transports: the faces' sM of the last stage, per direction
  double *tmp64;        // FP64 staging of set/get_state in the FP32-storage form
  size_t es;            // bytes per stored value (8 or 4)
  double *tables;       // geometry tables
  Control *ctl;
  int step;             // next step index to launch
  int have_state;
  hipStream_t cap;      // private stream for graph capture
  hipGraphExec_t gexec[6];  // kGraphSteps steps from a start slot of each residue mod 6
};

// Steps per captured graph: a multiple of 6, so that the slot-indexed ring
// (slot % 3) and ping-pong (slot & 1) entries of a replay repeat those of
// the capture whatever the start slot with the same residue.
static const int kGraphSteps = 48;

static int hipcode(hipError_t e) { return e == hipSuccess ? ECO_OK : ECO_ERR_HIP; }

extern "C" const char *eco_strerror(int code) {
  switch (code) {
    case ECO_OK: return "ok";
    case ECO_ERR_INVALID: return "invalid argument or configuration";
    case ECO_ERR_UNSUPPORTED: return "configuration not supported by the device path";
    case ECO_ERR_HIP: return "HIP runtime error";
    case ECO_ERR_NOMEM: return "device allocation failed";
    case ECO_ERR_STATE: return "call out of order";
    case ECO_ERR_SCRATCH: return "solve larger than the allocation-free entry point's static scratch";
    case ECO_ERR_NOT_CONVERGED: return "an injection or tank boundary solve did not converge: the run stopped";
  }
  return "unknown error";
}

extern "C" const char *eco_version(void) { return "ecogen-amd 0.1 (gfx950)"; }

static void lower_eos(const eco_eos &e, EosParams &q) {
  q.type = e.type;
  q.gamma = e.gamma;
  q.pinf = (e.type == ECO_EOS_SG) ? e.pinf : 0.;
  q.cv = e.cv;
  q.eref = e.eref;
  // sendSpecialMixtureEos (src/Eos/EosSG.cpp:217-222, EosIG.cpp:213-218)
  q.gpg = (e.type == ECO_EOS_SG) ? e.gamma * e.pinf / (e.gamma - 1.) : 0.;
  q.og = 1. / (e.gamma - 1.);
  q.gm1 = e.gamma - 1.;
  q.gpinf = (e.type == ECO_EOS_SG) ? e.gamma * e.pinf : 0.;
  q.pfloor = (e.type == ECO_EOS_SG) ? -(1. - 1.e-15) * e.pinf + 1.e-15 : 1.e-15;
}

// The constants of one tank side: rho_k0 = rho_k(p0, T0) (Eos*::computeDensity,
// EosIG.cpp:111-114, EosSG.cpp:114-117), the fractions divided by their sum,
// volume fractions from mass fractions through 1/rho = sum Y_k/rho_k0
// (BoundCondTank.cpp:46-176), then, as the tank branch of
// Mod{Kapila,MultiP}::solveRiemannTank computes them (ModKapila.cpp:386-390),
// Y_k0 = alpha_k0 rho_k0 / rho0 and H0 = sum Y_k0 (e_k + p0/rho_k0) of the
// fluid at rest; ModEuler::solveRiemannTank (ModEuler.cpp:358-360): H0 of
// the one phase.
static void lower_tank(const eco_config &c, int b, SchemeParams &S) {
  const double p0 = c.bc_tank_p0[b], T0 = c.bc_tank_T0[b];
  const double eps = S.eps;
  const int n = c.nphase;
  double rk[4], ak[4];
  for (int k = 0; k < n; k++) {
    const eco_eos &e = c.eos[k];
    const double num = (e.type == ECO_EOS_SG) ? p0 + e.pinf : p0;
    rk[k] = num / std::max((e.gamma - 1.) * e.cv * T0, eps);
  }
  auto energy = [&](const eco_eos &e, double rho, double p) {  // Eos{IG,SG}::computeEnergy
    const double num = (e.type == ECO_EOS_SG) ? p + e.gamma * e.pinf : p;
    return num / (e.gamma - 1.) / std::max(rho, eps) + e.eref;
  };
  S.tank_p0[b] = p0;
  if (c.model == ECO_MODEL_EULER) {
    S.tank_rho[b][0] = rk[0];
    S.tank_Y[b][0] = 1.;
    S.tank_H0[b] = energy(c.eos[0], rk[0], p0) + p0 / rk[0] + 0.5 * 0. * 0.;
    return;
  }
  double fr[4], sum = 0.;
  for (int k = 0; k < n; k++) sum += c.bc_tank_frac[b][k];
  for (int k = 0; k < n; k++) fr[k] = c.bc_tank_frac[b][k] / sum;
  if (!c.bc_tank_massfrac[b]) {
    for (int k = 0; k < n; k++) ak[k] = fr[k];
  } else {
    double rm = 0.;
    for (int k = 0; k < n; k++) rm += fr[k] / rk[k];
    rm = 1.0 / rm;
    for (int k = 0; k < n; k++) ak[k] = rm * fr[k] / rk[k];
  }
  double rho0 = 0., H0 = 0.;
  for (int k = 0; k < n; k++) rho0 += ak[k] * rk[k];
  for (int k = 0; k < n && k < kMaxPhases; k++) {
    S.tank_rho[b][k] = rk[k];
    S.tank_Y[b][k] = ak[k] * rk[k] / rho0;
    H0 += S.tank_Y[b][k] * (energy(c.eos[k], rk[k], p0) + p0 / rk[k] + 0.5 * 0. * 0.);
  }
  S.tank_H0[b] = H0;
}

// The constants of one injection side, as the first lines of
// Mod{Kapila,MultiP}::solveRiemannInflow (ModKapila.cpp:259-266) and
// ModEuler::solveRiemannInflow (ModEuler.cpp:240-241) compute them on each
// call: the signed flow (BoundCondInj.cpp:62), the injected density
// (Mix*::computeDensity), Y_k0 and H_k0 = e_k(rho_k0, p_k0) + p_k0/rho_k0 + u0^2/2
// (Eos::computeTotalEnthalpy, Eos.cpp:99-102).
static void lower_injection(const eco_config &c, int b, SchemeParams &S) {
  const double m0 = -c.bc_inj_m0[b];
  const double eps = S.eps;
  auto energy = [&](const eco_eos &e, double rho, double p) {  // Eos{IG,SG}::computeEnergy
    const double num = (e.type == ECO_EOS_SG) ? p + e.gamma * e.pinf : p;
    return num / (e.gamma - 1.) / std::max(rho, eps) + e.eref;
  };
  S.inj_m0[b] = m0;
  if (c.model == ECO_MODEL_EULER) {
    const double rho = c.bc_inj_rho[b][0], p = c.bc_inj_p[b][0];
    const double u0 = m0 / rho;
    S.inj_Y[b][0] = 1.;
    S.inj_H[b][0] = energy(c.eos[0], rho, p) + p / rho + 0.5 * u0 * u0;
    return;
  }
  // the reader divides the fractions by their sum (BoundCondInj.cpp:108-117)
  double ak0[4], sum = 0., rho0 = 0.;
  for (int k = 0; k < c.nphase; k++) sum += c.bc_inj_alpha[b][k];
  for (int k = 0; k < c.nphase; k++) ak0[k] = c.bc_inj_alpha[b][k] / sum;
  for (int k = 0; k < c.nphase; k++) rho0 += ak0[k] * c.bc_inj_rho[b][k];
  const double u0 = m0 / rho0;
  for (int k = 0; k < c.nphase && k < kMaxPhases; k++) {
    const double rho = c.bc_inj_rho[b][k], p = c.bc_inj_p[b][k];
    S.inj_H[b][k] = energy(c.eos[k], rho, p) + p / rho + 0.5 * u0 * u0;
    S.inj_Y[b][k] = ak0[k] * rho / rho0;
  }
}

int eco::validate_config(const eco_config *c) {
  if (!c) return ECO_ERR_INVALID;
  if (c->model < 0 || c->model > ECO_MODEL_EULERHOM) return ECO_ERR_INVALID;
  if (c->order != 1 && c->order != 2) return ECO_ERR_INVALID;
  if (c->cfl <= 0.) return ECO_ERR_INVALID;
  for (int d = 0; d < 3; d++) if (c->n[d] < 1 || !(c->L[d] > 0.)) return ECO_ERR_INVALID;
  for (int i = 0; i < 4; i++) if (c->lim[i] < 0 || c->lim[i] > 6) return ECO_ERR_INVALID;
  if (c->order == 2) {
    // Input.cpp:187-249: THINC only as a volume-fraction limiter; no "none" at order 2
    if (c->lim[0] == ECO_LIM_THINC || c->lim[1] == ECO_LIM_THINC) return ECO_ERR_INVALID;
    for (int i = 0; i < 4; i++) if (c->lim[i] == ECO_LIM_NONE) return ECO_ERR_INVALID;
  }
  for (int i = 0; i < 6; i++) {
    if (c->bc[i] < ECO_BC_ABS || c->bc[i] > ECO_BC_TANK) return ECO_ERR_INVALID;
    if (c->bc[i] == ECO_BC_TANK) {
      // BoundCondTank.cpp:46-144: p0 and T0; per phase a fraction in [0, 1], summing to 1
      if (!(c->bc_tank_p0[i] > 0.) || !(c->bc_tank_T0[i] > 0.)) return ECO_ERR_INVALID;
      if (c->model != ECO_MODEL_EULER) {
        double sum = 0.;
        for (int k = 0; k < c->nphase && k < 4; k++) {
          if (!(c->bc_tank_frac[i][k] >= 0. && c->bc_tank_frac[i][k] <= 1.)) return ECO_ERR_INVALID;
          sum += c->bc_tank_frac[i][k];
        }
        if (!(std::fabs(sum - 1.) <= 1.e-6)) return ECO_ERR_INVALID;
      }
    }
    if (c->bc[i] == ECO_BC_INJECTION) {
      // BoundCondInj.cpp:46-150 reads m0 and, per phase, alpha in [0, 1] summing to 1
      if (!(c->bc_inj_m0[i] != 0.)) return ECO_ERR_INVALID;
      const int np = (c->model == ECO_MODEL_EULER) ? 1 : c->nphase;
      double sum = 0.;
      for (int k = 0; k < np && k < 4; k++) {
        if (!(c->bc_inj_rho[i][k] > 0.)) return ECO_ERR_INVALID;
        if (np > 1 && !(c->bc_inj_alpha[i][k] >= 0. && c->bc_inj_alpha[i][k] <= 1.)) return ECO_ERR_INVALID;
        sum += c->bc_inj_alpha[i][k];
      }
      if (np > 1 && !(std::fabs(sum - 1.) <= 1.e-6)) return ECO_ERR_INVALID;
    }
    // walls, outflows and injections: FP64 path (the FP32-storage kernels keep transmissive boundaries)
    if (c->bc[i] != ECO_BC_ABS && c->precision == ECO_FP32S) return ECO_ERR_UNSUPPORTED;
  }
  if (c->model == ECO_MODEL_EULER) {
    if (c->nphase != 1) return ECO_ERR_INVALID;
  } else if (c->model == ECO_MODEL_EULERHOM) {
    // two phases, a liquid and a vapour (Input.cpp:479-483); ModEulerHomogeneous
    // has only the interior Riemann solver (Model.h:93-145 end the run for the
    // boundary and transport solvers); its MUSCL form starts the first
    // predictor's Newton solve from a pressure the reference leaves unset
    // (PhaseEulerHomogeneous.cpp:14-24, CellO2.cpp:93-97), so only first order
    if (c->nphase != 2 || c->ntransport > 0 || c->relax_pt || c->relax_ptmu || c->relax_p) return ECO_ERR_INVALID;
    if (c->liq < 0 || c->liq > 1 || c->vap != 1 - c->liq) return ECO_ERR_INVALID;
    for (int i = 0; i < 6; i++) if (c->bc[i] != ECO_BC_ABS) return ECO_ERR_INVALID;
    if (c->order != 1) return ECO_ERR_UNSUPPORTED;
    if (c->precision != ECO_FP64) return ECO_ERR_UNSUPPORTED;
  } else if (c->model == ECO_MODEL_THERMALEQ) {
    // two phases (MixThermalEq::computePressure, MixThermalEq.cpp:123); no
    // transport or injection solver (Model.h:127-135 end the run); the tank
    // (ModThermalEq::solveRiemannTank) and FP32 storage are not provided
    if (c->nphase != 2) return ECO_ERR_INVALID;
    if (c->ntransport > 0) return ECO_ERR_INVALID;
    for (int i = 0; i < 6; i++) {
      if (c->bc[i] == ECO_BC_INJECTION) return ECO_ERR_INVALID;
      if (c->bc[i] == ECO_BC_TANK) return ECO_ERR_UNSUPPORTED;
    }
    if (c->precision != ECO_FP64) return ECO_ERR_UNSUPPORTED;
  } else {
    if (c->nphase < 2 || c->nphase > 4) return ECO_ERR_INVALID;
    if (c->nphase > kMaxPhases) return ECO_ERR_UNSUPPORTED;
    // three or four phases: FP64 single solves; THINC is a two-phase reconstruction
    // (CellInterfaceO2.cpp:228-247 sets alpha_0 and 1 - alpha_0)
    if (c->nphase != 2 && (c->precision == ECO_FP32S || c->lim[2] == ECO_LIM_THINC || c->lim[3] == ECO_LIM_THINC))
      return ECO_ERR_UNSUPPORTED;
  }
  for (int k = 0; k < c->nphase; k++) {
    if (c->eos[k].type != ECO_EOS_IG && c->eos[k].type != ECO_EOS_SG) return ECO_ERR_INVALID;
    if (!(c->eos[k].gamma > 1.)) return ECO_ERR_INVALID;
  }
  // non-degenerate directions must come first (x, then y, then z)
  int dim = (c->n[0] > 1) + (c->n[1] > 1) + (c->n[2] > 1);
  for (int d = 0; d < dim; d++) if (c->n[d] < 2) return ECO_ERR_UNSUPPORTED;
  if (dim == 0) return ECO_ERR_UNSUPPORTED;
  if (c->precision != ECO_FP64 && c->precision != ECO_FP32S) return ECO_ERR_INVALID;
  // transports: the reference's Euler model has no transport Riemann solver
  // (Model::solveRiemannTransportIntern ends the run, src/Models/Model.h:127);
  // the FP32-storage kernels carry none
  if (c->ntransport < 0 || c->ntransport > ECO_MAX_TRANSPORTS) return ECO_ERR_INVALID;
  if (c->ntransport > 0 && c->model == ECO_MODEL_EULER) return ECO_ERR_INVALID;
  if (c->ntransport > 0 && c->precision == ECO_FP32S) return ECO_ERR_UNSUPPORTED;
  // RelaxationPTMu: two phases (RelaxationPTMu.cpp:30; phase 0 the liquid,
  // phase 1 the vapour), after Kapila's pressure relaxation, FP64 single
  // solves without transports
  if (c->relax_ptmu) {
    if (c->nphase != 2 || c->model != ECO_MODEL_KAPILA || c->relax_pt) return ECO_ERR_INVALID;
    if (c->precision != ECO_FP64 || c->ntransport > 0) return ECO_ERR_UNSUPPORTED;
  }
  // RelaxationPT: two phases (RelaxationPT.cpp:60), Kapila or MultiP, FP64
  if (c->relax_pt) {
    if (c->nphase != 2 || (c->model != ECO_MODEL_KAPILA && c->model != ECO_MODEL_MULTIP)) return ECO_ERR_INVALID;
    if (c->precision != ECO_FP64) return ECO_ERR_UNSUPPORTED;
  }
  // FP32 storage: MUSCL multiphase path (the batched/served configurations)
  if (c->precision == ECO_FP32S && (c->order != 2 || c->model == ECO_MODEL_EULER)) return ECO_ERR_UNSUPPORTED;
  return ECO_OK;
}

void eco::lower_config(const eco_config &c, Lowered &L, bool with_tables) {
  // geometry (src/Meshes/MeshCartesian.cpp:180-262, 300-397)
  Geometry &G = L.geo;
  std::memset(&G, 0, sizeof(G));
  int dim = 0;
  for (int d = 0; d < 3; d++) { G.n[d] = c.n[d]; dim += (c.n[d] > 1); }
  G.dim = dim;
  G.ncell = (long long)c.n[0] * c.n[1] * c.n[2];
  double dX = c.L[0] / (double)c.n[0], dY = c.L[1] / (double)c.n[1], dZ = c.L[2] / (double)c.n[2];
  double dd[3] = {dX, dY, dZ};
  double vol = dX * dY * dZ;
  double surf[3] = {dY * dZ, dX * dZ, dX * dY};
  for (int d = 0; d < 3; d++) G.coefA[d] = surf[d] / vol;
  double lc = 1.e10;
  for (int d = 0; d < 3; d++) if (c.n[d] != 1) lc = std::min(lc, dd[d]);
  if (dim > 1) lc *= 0.6;
  G.lcfl = lc;
  // per-direction distance tables (cell-centre and face positions as the reference builds them)
  size_t ntab = 0;
  for (int d = 0; d < 3; d++) ntab += 3 * (size_t)std::max(c.n[d], 1);
  if (with_tables) L.tab.assign(ntab, 0.);
  size_t o = 0;
  for (int d = 0; d < 3; d++) {
    L.off[d] = o;
    int n = c.n[d];
    for (int i = 0; with_tables && i + 1 < n; i++) {
      double xi = (i + 0.5) * dd[d], xj = (i + 1 + 0.5) * dd[d];
      double xf = xi + 0.5 * dd[d];
      double v = xi - xj;
      L.tab[o + i] = std::sqrt(v * v);                   // |x_i - x_{i+1}| (Element::distance)
      double a = xf - xi, b = xf - xj;
      L.tab[o + n + i] = std::sqrt(a * a);               // face - left cell (Face::distance)
      L.tab[o + 2 * n + i] = std::sqrt(b * b);           // face - right cell
    }
    o += 3 * (size_t)n;
    // boundary faces: the minus face at 0, the plus face at x_{n-1} + dX/2
    const double x0 = 0.5 * dd[d], xl = (n - 1 + 0.5) * dd[d];
    G.dwall[d][0] = x0 - 0.;
    G.dwall[d][1] = xl - (xl + 0.5 * dd[d]);
  }
  // scheme
  SchemeParams &S = L.sch;
  std::memset(&S, 0, sizeof(S));
  S.model = c.model;
  S.nph = c.nphase;
  S.order = c.order;
  for (int i = 0; i < 4; i++) S.lim[i] = c.lim[i];
  S.same_lim = (c.lim[0] == c.lim[1] && c.lim[2] == c.lim[3]) ? 1 : 0;
  S.thinc = (c.order == 2 && (c.lim[2] == ECO_LIM_THINC || c.lim[3] == ECO_LIM_THINC)) ? 1 : 0;
  S.alpha_null = c.alpha_null;
  S.eps = (c.model == ECO_MODEL_KAPILA && c.alpha_null) ? 1.e-15 : 0.;
  S.relax = (c.model == ECO_MODEL_KAPILA) || (c.model == ECO_MODEL_MULTIP && c.relax_p);
  S.relax_pt = c.relax_pt ? 1 : 0;
  S.relax_ptmu = (c.model == ECO_MODEL_KAPILA && c.relax_ptmu) ? 1 : 0;
  S.cfl = c.cfl;
  S.prec = c.precision;
  S.bc_wall = 0;
  S.bc_out = 0;
  S.bc_inj = 0;
  S.bc_tank = 0;
  for (int b = 0; b < 6; b++) {
    if (c.bc[b] == ECO_BC_WALL) S.bc_wall |= 1 << b;
    if (c.bc[b] == ECO_BC_OUTFLOW) S.bc_out |= 1 << b;
    if (c.bc[b] == ECO_BC_INJECTION) S.bc_inj |= 1 << b;
    if (c.bc[b] == ECO_BC_TANK) S.bc_tank |= 1 << b;
    S.bc_p0[b] = c.bc_p0[b];
  }
  for (int k = 0; k < c.nphase && k < kMaxPhases; k++) lower_eos(c.eos[k], S.eos[k]);
  for (int b = 0; b < 6; b++) if (c.bc[b] == ECO_BC_INJECTION) lower_injection(c, b, S);
  for (int b = 0; b < 6; b++) if (c.bc[b] == ECO_BC_TANK) lower_tank(c, b, S);
  // liquid-vapour saturation: Mixture::computeTsat's coefficients
  // (Mixture.cpp:69-85) for EulerHomogeneous's liquid and vapour and for
  // RelaxationPTMu's phases 0 and 1
  S.liq = 0;
  S.vap = 1;
  if (c.model == ECO_MODEL_EULERHOM) { S.liq = c.liq; S.vap = c.vap; }
  for (int i = 0; i < 4; i++) S.tsat[i] = 0.;
  if (c.nphase == 2 && (c.model == ECO_MODEL_EULERHOM || S.relax_ptmu)) {
    const eco_eos &qL = c.eos[S.liq], &qV = c.eos[S.vap];
    const double gammaL = qL.gamma, cvL = qL.cv, e0L = qL.eref, s0L = c.sref[S.liq];
    const double gammaV = qV.gamma, cvV = qV.cv, e0V = qV.eref, s0V = c.sref[S.vap];
    S.tsat[0] = (gammaL * cvL - gammaV * cvV + s0V - s0L) / (gammaV * cvV - cvV);
    S.tsat[1] = (e0L - e0V) / (gammaV * cvV - cvV);
    S.tsat[2] = (gammaV * cvV - gammaL * cvL) / (gammaV * cvV - cvV);
    S.tsat[3] = (gammaL * cvL - cvL) / (gammaV * cvV - cvV);
  }
  L.nprim = (c.model == ECO_MODEL_EULER) ? 2 + dim
          : (c.model == ECO_MODEL_THERMALEQ || c.model == ECO_MODEL_EULERHOM) ? c.nphase + 2 + dim : 3 * c.nphase + dim;
  S.ncache = (c.model == ECO_MODEL_MULTIP && (c.relax_p || c.relax_pt) && c.precision == ECO_FP64)
                 ? (c.order == 1 ? c.nphase + 3 : 1) : 0;
  // ThermalEq and EulerHomogeneous at first order: the cells' cached mixture energy
  if ((c.model == ECO_MODEL_THERMALEQ || c.model == ECO_MODEL_EULERHOM) && c.order == 1) S.ncache = 1;
  L.nlevel = L.nprim + 1 + S.ncache;
  L.ntr = c.ntransport;
  L.es = (c.precision == ECO_FP32S) ? sizeof(float) : sizeof(double);
}

void eco::bind_tables(const eco_config &c, Lowered &L, double *dev_tables) {
  for (int d = 0; d < 3; d++) {
    int n = c.n[d];
    L.geo.dcc[d] = dev_tables + L.off[d];
    L.geo.dfl[d] = dev_tables + L.off[d] + n;
    L.geo.dfr[d] = dev_tables + L.off[d] + 2 * n;
  }
}

// host layout (3N+4 fields, u,v,w) -> device layout (nprim fields)
void eco::host_field_map(const eco_config &c, int dim, std::vector<int> &map) {
  map.clear();
  if (c.model == ECO_MODEL_THERMALEQ || c.model == ECO_MODEL_EULERHOM) {
    // alpha_k, the mixture pressure (p_0), T (after E; EulerHomogeneous: not
    // read, the initialisation sets the saturation temperature), velocity
    for (int k = 0; k < c.nphase; k++) map.push_back(3 * k);
    map.push_back(2);
    map.push_back(ECO_NFIELDS(c.nphase));
    for (int d = 0; d < dim; d++) map.push_back(3 * c.nphase + d);
  } else if (c.model == ECO_MODEL_EULER) {
    map.push_back(1); map.push_back(2);
    for (int d = 0; d < dim; d++) map.push_back(3 + d);
  } else {
    for (int k = 0; k < c.nphase; k++) { map.push_back(3 * k); map.push_back(3 * k + 1); map.push_back(3 * k + 2); }
    for (int d = 0; d < dim; d++) map.push_back(3 * c.nphase + d);
  }
}

extern "C" int eco_create(const eco_config *cfg, void *hip_stream, eco_ctx **out) {
  if (!out) return ECO_ERR_INVALID;
  *out = nullptr;
  int rc = validate_config(cfg);
  if (rc) return rc;
  eco_ctx *x = new (std::nothrow) eco_ctx();
  if (!x) return ECO_ERR_NOMEM;
  x->cfg = *cfg;
  const eco_config &c = *cfg;
  Lowered L;
  lower_config(c, L);
  x->geo = L.geo;
  x->sch = L.sch;
  x->nprim = L.nprim;
  x->nlevel = L.nlevel;
  x->ntr = L.ntr;

  // stream
  if (hip_stream) { x->stream = (hipStream_t)hip_stream; x->own_stream = 0; }
  else { x->stream = nullptr; x->own_stream = 0; }

  // device memory
  const size_t nc = (size_t)x->geo.ncell;
  const size_t ntab = L.tab.size();
  hipError_t e = hipSuccess;
  x->buf[0] = x->buf[1] = x->bstar = x->bn2 = x->tables = x->tmp64 = nullptr;
  x->smf[0] = x->smf[1] = x->smf[2] = nullptr;
  x->ctl = nullptr;
  x->es = L.es;
  const size_t nlv = (size_t)x->nlevel + x->ntr;
  e = hipMalloc(&x->buf[0], nc * nlv * x->es);
  if (e == hipSuccess) e = hipMalloc(&x->buf[1], nc * nlv * x->es);
  if (e == hipSuccess && c.order == 2) e = hipMalloc(&x->bstar, nc * (x->nprim + x->ntr) * x->es);
  // RelaxationPTMu in MUSCL: the relaxed time-n state the predictor writes
  if (e == hipSuccess && c.order == 2 && x->sch.relax_ptmu) e = hipMalloc(&x->bn2, nc * nlv * x->es);
  for (int d = 0; d < x->geo.dim && x->ntr > 0 && e == hipSuccess; d++) {
    // faces of direction d: (n_d + 1) per line of cells
    const size_t nf = nc / (size_t)c.n[d] * (size_t)(c.n[d] + 1);
    e = hipMalloc(&x->smf[d], nf * sizeof(double));
  }
  if (e == hipSuccess && c.precision == ECO_FP32S) e = hipMalloc(&x->tmp64, nc * (x->nprim + 1) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&x->tables, ntab * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&x->ctl, sizeof(Control));
  if (e != hipSuccess) { eco_destroy(x); return ECO_ERR_NOMEM; }
  e = hipMemcpy(x->tables, L.tab.data(), ntab * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) { eco_destroy(x); return ECO_ERR_HIP; }
  bind_tables(c, L, x->tables);
  x->geo = L.geo;
  x->step = 0;
  x->have_state = 0;
  *out = x;
  return ECO_OK;
}

static int reset_control(eco_ctx *x) {
  Control h;
  std::memset(&h, 0, sizeof(h));
  h.t_final = -1.;
  h.max_steps = 0;
  return hipcode(hipMemcpyAsync(x->ctl, &h, sizeof(h), hipMemcpyHostToDevice, x->stream));
}

static StageArgs make_args(eco_ctx *x, int step) {
  StageArgs A;
  A.geo = x->geo;
  A.sch = x->sch;
  A.ctl = x->ctl;
  A.step = step;
  A.wn = x->buf[step & 1];
  A.wsrc = A.wn;
  A.wout = x->buf[(step + 1) & 1];
  for (int d = 0; d < 3; d++) A.smf[d] = x->smf[d];
  return A;
}

// The transport pass of a stage whose flow kernel ran with A (stage 0: first
// order, 1: predictor into U*, 2: corrector from U*).
static int launch_transport_stage(eco_ctx *x, const StageArgs &A, int stage) {
  const size_t nc = (size_t)x->geo.ncell;
  TransportArgs T;
  std::memset(&T, 0, sizeof(T));
  T.geo = x->geo;
  T.cfl = x->cfg.cfl;
  T.ntr = x->ntr;
  T.order = x->cfg.order;
  T.stage = stage;
  T.lim_g = x->cfg.lim[2];
  T.lim_i = x->cfg.lim[3];
  for (int b = 0; b < 6; b++) {
    const int k = x->cfg.bc[b];
    if (k == ECO_BC_WALL) T.bc_wall |= 1 << b;
    if (k == ECO_BC_OUTFLOW || k == ECO_BC_INJECTION || k == ECO_BC_TANK) T.bc_in |= 1 << b;
    for (int t = 0; t < x->ntr; t++) T.bc_tr[b][t] = x->cfg.bc_tr[b][t];
  }
  // phase-0 alpha is field 0 of a time level and of U*
  T.asrc = A.wsrc;
  T.an = A.wn;
  const size_t itr = (size_t)x->nlevel * nc;    // scalars in a time level
  const size_t istar = (size_t)x->nprim * nc;   // scalars in U*
  T.trn = A.wn + itr;
  T.trs = (stage == 2) ? A.wsrc + istar : A.wn + itr;
  T.tro = (stage == 1) ? A.wout + istar : A.wout + itr;
  for (int d = 0; d < 3; d++) T.sm[d] = x->smf[d];
  T.ctl = x->ctl;
  T.step = A.step;
  return launch_transport(T, x->stream);
}

// One stage's flow kernel and, with transports, its transport pass.
static int launch_stage_tr(eco_ctx *x, const StageArgs &A, int stage) {
  int rc = launch_stage(A, stage, x->stream);
  if (!rc && x->ntr > 0) rc = launch_transport_stage(x, A, stage);
  return rc;
}

static void host_to_dev_fields(const eco_ctx *x, std::vector<int> &map) { host_field_map(x->cfg, x->geo.dim, map); }

// device buffer the FP64 device-layout input is staged in
static double *staging(eco_ctx *x) { return x->tmp64 ? x->tmp64 : x->buf[0]; }

static int finish_set_state(eco_ctx *x) {
  int rc = reset_control(x);
  if (rc) return rc;
  StageArgs A = make_args(x, 0);
  A.wsrc = A.wn = staging(x);
  A.wout = staging(x);
  rc = launch_stage(A, 3, x->stream);
  if (rc) return rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;
  if (x->tmp64) {
    A.wout = x->buf[0];
    rc = launch_convert(A, 1, x->tmp64, x->stream);
    if (rc) return rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;
  }
  x->step = 0;
  x->have_state = 1;
  return hipcode(hipStreamSynchronize(x->stream));
}

// host (or device) fields of the transported scalars -> time level 0 (FP64 only)
static int copy_transports_in(eco_ctx *x, const double *state, hipMemcpyKind kind) {
  const size_t nc = (size_t)x->geo.ncell;
  const size_t hf = ECO_NFIELDS(x->cfg.nphase);
  for (int t = 0; t < x->ntr; t++) {
    hipError_t e = hipMemcpyAsync(x->buf[0] + ((size_t)x->nlevel + t) * nc, state + (hf + t) * nc, nc * sizeof(double),
                                  kind, x->stream);
    if (e != hipSuccess) return ECO_ERR_HIP;
  }
  return ECO_OK;
}

extern "C" int eco_set_state(eco_ctx *x, const double *host_state) {
  if (!x || !host_state) return ECO_ERR_INVALID;
  std::vector<int> map;
  host_to_dev_fields(x, map);
  size_t nc = (size_t)x->geo.ncell;
  for (size_t f = 0; f < map.size(); f++) {
    hipError_t e = hipMemcpyAsync(staging(x) + f * nc, host_state + (size_t)map[f] * nc, nc * sizeof(double),
                                  hipMemcpyHostToDevice, x->stream);
    if (e != hipSuccess) return ECO_ERR_HIP;
  }
  int rc = copy_transports_in(x, host_state, hipMemcpyHostToDevice);
  if (rc) return rc;
  return finish_set_state(x);
}

extern "C" int eco_set_clock(eco_ctx *x, double t, double dt) {
  if (!x || !(t >= 0.) || !(dt > 0.)) return ECO_ERR_INVALID;
  if (!x->have_state || x->step != 0) return ECO_ERR_STATE;
  // plain fields of the control block, read by the next launches on this stream
  hipError_t e = hipMemcpyAsync(&x->ctl->t0, &t, sizeof(double), hipMemcpyHostToDevice, x->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&x->ctl->dt0, &dt, sizeof(double), hipMemcpyHostToDevice, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  return hipcode(e);
}

extern "C" int eco_set_state_device(eco_ctx *x, const double *dev_state) {
  if (!x || !dev_state) return ECO_ERR_INVALID;
  std::vector<int> map;
  host_to_dev_fields(x, map);
  size_t nc = (size_t)x->geo.ncell;
  for (size_t f = 0; f < map.size(); f++) {
    hipError_t e = hipMemcpyAsync(staging(x) + f * nc, dev_state + (size_t)map[f] * nc, nc * sizeof(double),
                                  hipMemcpyDeviceToDevice, x->stream);
    if (e != hipSuccess) return ECO_ERR_HIP;
  }
  int rc = copy_transports_in(x, dev_state, hipMemcpyDeviceToDevice);
  if (rc) return rc;
  return finish_set_state(x);
}

void eco::fill_status(const Control &h, double lcfl, double cfl, eco_status *st) {
  st->steps = h.steps;
  st->t = (h.steps > 0) ? h.t_now : h.t0;
  st->dt_last = h.dt_last;
  double dtn = (h.dt0 > 0.) ? h.dt0 : 1.e-15;
  if (h.steps > 0) {
    // a solve's step k runs at launch slot k
    unsigned long long b = h.smax[(h.steps - 1) % 3];
    double sm;
    std::memcpy(&sm, &b, sizeof(sm));
    double dtMax = (sm > 0.) ? std::min(1.e10, lcfl / sm) : 1.e10;
    dtn = cfl * dtMax;
  }
  st->dt_next = dtn;
  st->flags = h.err;
  st->done = h.finished;
}

static int read_status(eco_ctx *x, eco_status *st, Control *hc) {
  Control h;
  hipError_t e = hipMemcpyAsync(&h, x->ctl, sizeof(Control), hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  if (e != hipSuccess) return ECO_ERR_HIP;
  if (hc) *hc = h;
  if (st) fill_status(h, x->geo.lcfl, x->cfg.cfl, st);
  return ECO_OK;
}

extern "C" int eco_get_state(eco_ctx *x, double *host_state) {
  if (!x || !host_state) return ECO_ERR_INVALID;
  if (!x->have_state) return ECO_ERR_STATE;
  Control h;
  int rc = read_status(x, nullptr, &h);
  if (rc) return rc;
  x->step = h.steps;
  const eco_config &c = x->cfg;
  size_t nc = (size_t)x->geo.ncell;
  const double *src = x->buf[h.steps & 1];
  if (x->tmp64) {
    StageArgs A = make_args(x, h.steps);
    A.wn = x->buf[h.steps & 1];
    int rc2 = launch_convert(A, 0, x->tmp64, x->stream);
    if (rc2) return rc2 < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;
    src = x->tmp64;
  }
  std::vector<int> map;
  host_to_dev_fields(x, map);
  int nf = ECO_NFIELDS(c.nphase);
  std::memset(host_state, 0, sizeof(double) * ECO_NFIELDS_CFG(c.model, c.nphase, x->ntr) * nc);
  hipError_t e = hipSuccess;
  for (size_t f = 0; f < map.size() && e == hipSuccess; f++)
    e = hipMemcpyAsync(host_state + (size_t)map[f] * nc, src + f * nc, nc * sizeof(double), hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_state + (size_t)(nf - 1) * nc, src + (size_t)x->nprim * nc, nc * sizeof(double),
                       hipMemcpyDeviceToHost, x->stream);
  for (int t = 0; t < x->ntr && e == hipSuccess; t++)
    e = hipMemcpyAsync(host_state + (size_t)(nf + t) * nc, src + ((size_t)x->nlevel + t) * nc, nc * sizeof(double),
                       hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  if (e != hipSuccess) return ECO_ERR_HIP;
  if (c.model == ECO_MODEL_EULER) for (size_t i = 0; i < nc; i++) host_state[i] = 1.;
  if (c.model == ECO_MODEL_THERMALEQ || c.model == ECO_MODEL_EULERHOM) {
    // every phase at the mixture pressure, rho_k = rho_k(p, T) as
    // Eos::computeDensity (EosSG.cpp:114-117, EosIG.cpp:111-114); for
    // EulerHomogeneous T is the saturation temperature and this the saturated
    // density (Eos*::computeDensitySaturation, EosSG.cpp:191-200)
    const double *p = host_state + 2 * nc, *T = host_state + (size_t)nf * nc;
    for (int k = 0; k < c.nphase; k++) {
      const eco_eos &q = c.eos[k];
      double *rk = host_state + (size_t)(3 * k + 1) * nc, *pk = host_state + (size_t)(3 * k + 2) * nc;
      for (size_t i = 0; i < nc; i++) {
        pk[i] = p[i];
        const double den = (q.gamma - 1.) * q.cv * T[i];
        rk[i] = ((q.type == ECO_EOS_SG) ? p[i] + q.pinf : p[i]) / ((den < 0.) ? 0. : den);
      }
    }
  }
  return ECO_OK;
}

static int launch_one(eco_ctx *x) {
  int k = x->step;
  int rc;
  if (x->cfg.order == 1) {
    rc = launch_stage_tr(x, make_args(x, k), 0);
  } else {
    StageArgs P = make_args(x, k);
    P.wout = x->bstar;
    P.wn2 = x->bn2;
    rc = launch_stage_tr(x, P, 1);
    if (!rc) {
      StageArgs Q = make_args(x, k);
      Q.wsrc = x->bstar;
      if (x->bn2) Q.wn = x->bn2;  // RelaxationPTMu: the time-n state the predictor relaxed
      rc = launch_stage_tr(x, Q, 2);
    }
  }
  if (rc) return rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;
  x->step = k + 1;
  return ECO_OK;
}

static int set_target(eco_ctx *x, int max_steps, double t_final) {
  // Control::t_final, max_steps, done[] and finished are plain fields written before the launches.
  hipError_t e = hipMemcpyAsync(&x->ctl->t_final, &t_final, sizeof(double), hipMemcpyHostToDevice, x->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&x->ctl->max_steps, &max_steps, sizeof(int), hipMemcpyHostToDevice, x->stream);
  int zero[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(&x->ctl->done[0], zero, sizeof(zero), hipMemcpyHostToDevice, x->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&x->ctl->finished, zero, sizeof(int), hipMemcpyHostToDevice, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  return hipcode(e);
}

// Graph of kGraphSteps steps starting at a slot congruent to r (mod 6),
// captured on first use.  Its kernel arguments depend on the slot only
// through its parity (buffer roles) and its residues mod 2 and 3.
static int graph_for(eco_ctx *x, int r, hipGraphExec_t *out) {
  if (x->gexec[r]) { *out = x->gexec[r]; return ECO_OK; }
  if (!x->cap && hipStreamCreateWithFlags(&x->cap, hipStreamNonBlocking) != hipSuccess) return ECO_ERR_HIP;
  if (hipStreamBeginCapture(x->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return ECO_ERR_HIP;
  const int save_step = x->step;
  hipStream_t save_stream = x->stream;
  x->step = r;
  x->stream = x->cap;
  int rc = ECO_OK;
  for (int i = 0; i < kGraphSteps && rc == ECO_OK; i++) rc = launch_one(x);
  x->stream = save_stream;
  x->step = save_step;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(x->cap, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return ECO_ERR_HIP;
  e = hipGraphInstantiate(&x->gexec[r], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { x->gexec[r] = nullptr; return ECO_ERR_HIP; }
  *out = x->gexec[r];
  return ECO_OK;
}

// Enqueues nsteps steps: whole graphs of kGraphSteps, then single launches.
static int launch_steps(eco_ctx *x, int nsteps) {
  while (nsteps >= kGraphSteps) {
    hipGraphExec_t ge;
    int rc = graph_for(x, x->step % 6, &ge);
    if (rc) return rc;
    if (hipGraphLaunch(ge, x->stream) != hipSuccess) return ECO_ERR_HIP;
    x->step += kGraphSteps;
    nsteps -= kGraphSteps;
  }
  for (; nsteps > 0; nsteps--) {
    int rc = launch_one(x);
    if (rc) return rc;
  }
  return ECO_OK;
}

extern "C" int eco_launch_steps(eco_ctx *x, int nsteps) {
  if (!x || nsteps < 0) return ECO_ERR_INVALID;
  if (!x->have_state) return ECO_ERR_STATE;
  for (int i = 0; i < nsteps; i++) {
    int rc = launch_one(x);
    if (rc) return rc;
  }
  return ECO_OK;
}

extern "C" int eco_sync(eco_ctx *x, eco_status *st) {
  if (!x) return ECO_ERR_INVALID;
  Control h;
  int rc = read_status(x, st, &h);
  if (rc) return rc;
  x->step = h.steps;
  return ECO_OK;
}

// ECO_ERR_NOT_CONVERGED once a boundary solve stopped the run (status kept)
static int stop_code(const eco_status &s) { return (s.flags & ECO_WARN_INFLOW_NOT_CONVERGED) ? ECO_ERR_NOT_CONVERGED : ECO_OK; }

extern "C" int eco_step(eco_ctx *x, eco_status *st) {
  if (!x) return ECO_ERR_INVALID;
  if (!x->have_state) return ECO_ERR_STATE;
  int rc = set_target(x, x->step + 1, -1.);
  if (!rc) rc = launch_one(x);
  eco_status s;
  if (!rc) rc = eco_sync(x, &s);
  if (rc) return rc;
  if (st) *st = s;
  return stop_code(s);
}

extern "C" int eco_run(eco_ctx *x, int nsteps, double t_final, eco_status *st) {
  if (!x) return ECO_ERR_INVALID;
  if (!x->have_state) return ECO_ERR_STATE;
  int rc;
  if (nsteps > 0) {
    rc = set_target(x, x->step + nsteps, -1.);
    if (!rc) rc = launch_steps(x, nsteps);
    if (rc) return rc;
    eco_status s;
    rc = eco_sync(x, &s);
    if (rc) return rc;
    if (st) *st = s;
    return stop_code(s);
  }
  rc = set_target(x, nsteps < 0 ? x->step - nsteps : 0, t_final);
  if (rc) return rc;
  // Time control: enqueue chunks of captured steps, check the device done
  // flag between chunks (steps past the target are no-ops on the device).
  int chunk = kGraphSteps;
  eco_status s;
  for (;;) {
    rc = launch_steps(x, chunk);
    if (rc) return rc;
    rc = eco_sync(x, &s);
    if (rc) return rc;
    if (s.done) break;
    chunk = std::min(chunk * 2, 4 * kGraphSteps);
  }
  if (st) *st = s;
  return stop_code(s);
}

extern "C" int eco_time_steps(eco_ctx *x, int nsteps, double *stage_ms, int *launches, eco_status *st) {
  if (!x || nsteps <= 0) return ECO_ERR_INVALID;
  if (!x->have_state) return ECO_ERR_STATE;
  int rc = set_target(x, x->step + nsteps, -1.);
  if (rc) return rc;
  const int nst = (x->cfg.order == 1) ? 1 : 2;
  std::vector<hipEvent_t> ev((size_t)nsteps * nst * 2, nullptr);
  for (size_t i = 0; i < ev.size(); i++) {
    if (hipEventCreate(&ev[i]) != hipSuccess) {
      for (size_t j = 0; j < i; j++) (void)hipEventDestroy(ev[j]);
      return ECO_ERR_HIP;
    }
  }
  for (int i = 0; i < nsteps && !rc; i++) {
    int k = x->step;
    if (nst == 1) {
      (void)hipEventRecord(ev[2 * i], x->stream);
      rc = launch_stage_tr(x, make_args(x, k), 0);
      (void)hipEventRecord(ev[2 * i + 1], x->stream);
    } else {
      StageArgs P = make_args(x, k);
      P.wout = x->bstar;
      P.wn2 = x->bn2;
      (void)hipEventRecord(ev[4 * i], x->stream);
      rc = launch_stage_tr(x, P, 1);
      (void)hipEventRecord(ev[4 * i + 1], x->stream);
      if (!rc) {
        StageArgs Q = make_args(x, k);
        Q.wsrc = x->bstar;
        if (x->bn2) Q.wn = x->bn2;
        (void)hipEventRecord(ev[4 * i + 2], x->stream);
        rc = launch_stage_tr(x, Q, 2);
        (void)hipEventRecord(ev[4 * i + 3], x->stream);
      }
    }
    if (rc) rc = rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;
    else x->step = k + 1;
  }
  if (!rc) rc = hipcode(hipStreamSynchronize(x->stream));
  double ms[3] = {0., 0., 0.};
  int cnt[3] = {0, 0, 0};
  for (int i = 0; i < nsteps && !rc; i++) {
    for (int s = 0; s < nst; s++) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, ev[(size_t)(i * nst + s) * 2], ev[(size_t)(i * nst + s) * 2 + 1]) != hipSuccess) {
        rc = ECO_ERR_HIP;
        break;
      }
      int slot = (nst == 1) ? 0 : s + 1;
      ms[slot] += t;
      cnt[slot]++;
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  if (stage_ms) for (int s = 0; s < 3; s++) stage_ms[s] = ms[s];
  if (launches) for (int s = 0; s < 3; s++) launches[s] = cnt[s];
  return eco_sync(x, st);
}

namespace eco { int launch_copy(void *dst, const void *src, unsigned long long nbytes, hipStream_t st); }

extern "C" int eco_copy_calibration(void *dst, const void *src, unsigned long long nbytes, void *hip_stream) {
  if (!dst || !src || (nbytes % 8)) return ECO_ERR_INVALID;
  return launch_copy(dst, src, nbytes, (hipStream_t)hip_stream) ? ECO_ERR_HIP : ECO_OK;
}

extern "C" const double *eco_state_device(eco_ctx *x, int *nfields_device) {
  if (nfields_device) *nfields_device = 0;
  // FP32-storage contexts hold a packed format (FP32 fields, the alpha pair
  // as one FP64 value in two 32-bit words, a reduced energy): no FP64 view
  if (!x || x->es != sizeof(double)) return nullptr;
  if (nfields_device) *nfields_device = x->nprim + 1;
  return x->buf[x->step & 1];
}

extern "C" void eco_destroy(eco_ctx *x) {
  if (!x) return;
  for (auto &g : x->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  if (x->cap) (void)hipStreamDestroy(x->cap);
  if (x->buf[0]) (void)hipFree(x->buf[0]);
  if (x->buf[1]) (void)hipFree(x->buf[1]);
  if (x->bstar) (void)hipFree(x->bstar);
  if (x->bn2) (void)hipFree(x->bn2);
  for (auto &p : x->smf)
    if (p) (void)hipFree(p);
  if (x->tables) (void)hipFree(x->tables);
  if (x->ctl) (void)hipFree(x->ctl);
  if (x->tmp64) (void)hipFree(x->tmp64);
  delete x;
}
