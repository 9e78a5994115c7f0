// Plain-old-data parameter blocks shared by the host descriptors and the
// gfx950 kernels of the explicit finite-volume step.
#pragma once
#include <stdint.h>

namespace eco {

enum ModelKind : int { MODEL_EULER = 0, MODEL_KAPILA = 1, MODEL_MULTIP = 2, MODEL_THERMALEQ = 3, MODEL_EULERHOM = 4 };
// ThermalEq and EulerHomogeneous store the same fields (alpha_k, the mixture
// pressure and temperature, velocity; the phases at that pressure and
// temperature) and share the flux layout (per-phase masses: EulerHomogeneous
// carries its one mass in phase 0's slot, the others stay 0)
__host__ __device__ constexpr bool te_like(int m) { return m == MODEL_THERMALEQ || m == MODEL_EULERHOM; }
enum LimiterKind : int { LIM_NONE = 0, LIM_MINMOD = 1, LIM_VANLEER = 2, LIM_VANALBADA = 3,
                         LIM_SUPERBEE = 4, LIM_MC = 5, LIM_THINC = 6 };
enum EosKind : int { EOS_IG = 0, EOS_SG = 1 };
enum Precision : int { PREC_FP64 = 0, PREC_FP32S = 1 };

constexpr int kMaxPhases = 4;  // phases handled by the device path (FP32 storage and THINC: 2)

// Equation of state constants (src/Eos/EosSG.h, src/Eos/EosIG.h) plus the
// mixture-EOS coefficients of sendSpecialMixtureEos, computed once on the host
// with the same expressions the reference evaluates per call.
struct EosParams {
  int type;
  double gamma, pinf, cv, eref;
  double gpg;  // gamma*pinf/(gamma-1)   (0 for IG)
  double og;   // 1/(gamma-1)
  double gm1;  // gamma-1
  double gpinf;  // gamma*pinf            (0 for IG)
  double pfloor;  // -(1-1e-15)*pinf + 1e-15 (SG) or 1e-15 (IG)
};

// Uniform Cartesian grid geometry.  Per-direction 1-D tables reproduce the
// reference's cell-centre and face positions (src/Meshes/MeshCartesian.cpp:
// 180-197, 300-397): dcc[i] = |x_i - x_{i+1}|, dfl[i] = |xf_{i+1/2} - x_i|,
// dfr[i] = |xf_{i+1/2} - x_{i+1}|.
struct Geometry {
  int n[3];
  int dim;
  long long ncell;
  double coefA[3];  // face surface / cell volume per direction
  double lcfl;      // CFL length (x0.6 in multi-D)
  const double *dcc[3];
  const double *dfl[3];
  const double *dfr[3];
  double dwall[3][2];  // signed centre-to-face distance (cell - face) of the boundary cells, minus / plus side
};

struct SchemeParams {
  int model, nph, order;
  int lim[4];        // global, interface, global VF, interface VF
  int same_lim;      // interface detection cannot change the limiter
  int thinc;
  int alpha_null;
  double eps;        // 1e-15 when alphaNull, else 0 (src/Eos/Eos.cpp:114-122)
  int relax;         // stiff pressure relaxation applied
  int relax_pt;      // pressure-temperature relaxation after it (RelaxationPT, two phases)
  // Values of the reference's cell cache a time level stores after E (MultiP
  // with relaxation, FP64): FluxMultiP::buildPrim sets them, the relaxation
  // changes alpha_k, rho_k, p_k but leaves them (RelaxationP.cpp:49-111), so
  // they cannot be recomputed from the stored state.  0: none; 1 (MUSCL):
  // the mixture density, by which the next step's buildCons multiplies
  // velocity and energy (FluxMultiP.cpp:122-140); NPH + 3 (first order,
  // whose faces take the cells' cached states, CellInterface.cpp:127-146):
  // the mixture density, the phase energies e_k, the frozen sound speed and
  // the mixture pressure (the relaxed p*, or sum alpha_k p_k when the
  // relaxation did not converge).  ThermalEq at first order: 1, the cells'
  // mixture internal energy (sum Y_k e_k initially, then E - |u|^2/2).
  int ncache;
  double cfl;
  int prec;          // PREC_FP64 or PREC_FP32S (storage)
  int bc_wall;       // bit 2*dir+side set: wall boundary (BoundCondWall), else transmissive
  int bc_out;        // bit 2*dir+side set: outflow boundary (BoundCondOutflow) at bc_p0[2*dir+side]
  int bc_inj;        // bit 2*dir+side set: injection boundary (BoundCondInj)
  double bc_p0[6];
  // injection constants per side (BoundCondInj + the first lines of
  // Mod*::solveRiemannInflow): signed mass flow (negative: into the domain),
  // injected mass fractions Y_k0 and phase total enthalpies H_k0
  double inj_m0[6];
  double inj_Y[6][kMaxPhases];
  double inj_H[6][kMaxPhases];
  int bc_tank;       // bit 2*dir+side set: tank boundary (BoundCondTank)
  // tank constants per side (BoundCondTank.cpp:46-176 and the tank branch of
  // Mod*::solveRiemannTank): pressure, phase densities rho_k(p0, T0), mass
  // fractions Y_k0 and the total enthalpy H0 of the fluid at rest
  double tank_p0[6];
  double tank_rho[6][kMaxPhases];
  double tank_Y[6][kMaxPhases];
  double tank_H0[6];
  EosParams eos[kMaxPhases];
  // liquid-vapour saturation (EulerHomogeneous, RelaxationPTMu): the liquid
  // and vapour phases and the coefficients A, B, C, D of the saturation
  // temperature's Gibbs-energy equality (Mixture::computeTsat,
  // Mixture.cpp:81-85), computed on the host with the reference's expressions
  int relax_ptmu;    // RelaxationPTMu after the pressure relaxation (Kapila, two phases)
  int liq, vap;
  double tsat[4];
};

// Device-resident control block of one solve (one per context / pool case).
// dt handling reproduces Run::solver (src/Run.cpp:204-277): dt of a case's
// first step is 1e-15, dt of step k is cfl * dtMax(k-1); dtMax is min over
// all Riemann problems of lcfl/|s|, kept here as max |s| (division is
// monotonic).  Entries are indexed by the launch slot (ping-pong slot&1, ring
// slot%3), so a case only needs to see consecutive slots; the number of steps
// a case has taken is cnt[], independent of the slot it started at.
struct Control {
  unsigned long long smax[3];  // max wave speed bits, ring entry slot%3
  double t[2];                 // physical time after the step of slot s in t[s&1]
  int done[2];                 // target reached after the step of slot s in done[s&1]
  int cnt[2];                  // steps taken after the step of slot s in cnt[s&1]
  int steps;                   // steps taken (host view)
  int finished;                // target reached (host view; cleared by the host)
  unsigned int err;            // soft-error flags (ECO_WARN_*)
  int max_steps;               // iteration mode: step count; time mode: cap (0 = none)
  double dt_last;              // dt used by the last step taken
  double t_now;                // physical time after the last step taken
  double t_final;              // <0: iteration mode
  double t0;                   // physical time before the first step (0, or a resumed run's)
  double dt0;                  // dt of the first step (0: the reference's 1e-15)
};

}  // namespace eco
