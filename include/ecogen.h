/*
 * ecogen.h -- C ABI of the MI355X explicit finite-volume step.
 *
 * Drop-in boundary for the reference ECOGEN time step on Cartesian meshes.
 * The reference has no FFI: its boundary is the C++ virtual API driven by
 * Run::solver (src/Run.cpp:204-277).  Each entry point below names the
 * reference routine(s) it replaces.  Plain pointers and sizes only; the host
 * owns host arrays, a context owns its device memory.  Errors are returned as
 * codes (never exit), soft numerical events as flags.
 *
 * State layout (host and device SoA, field-major, ncell = n[0]*n[1]*n[2]
 * doubles per field, cell index i + j*n[0] + k*n[0]*n[1] as in
 * MeshCartesian::construitIGlobal, src/Meshes/MeshCartesian.cpp:154-160):
 *     for each phase k: alpha_k, rho_k, p_k     (Euler: 1, rho, p)
 *     u, v, w
 *     E  (mixture specific total energy; written by eco_get_state, ignored on input)
 *     the cfg->ntransport transported scalars (Cell::m_vecTransports), if any
 * => ECO_NFIELDS(nphase) = 3*nphase + 4 fields, ECO_NFIELDS_TR(nphase, ntransport)
 *    with the transports.
 */
#ifndef ECOGEN_H
#define ECOGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ECO_NFIELDS(nph) (3 * (nph) + 4)
#define ECO_NFIELDS_TR(nph, ntr) (3 * (nph) + 4 + (ntr))
/* fields of a configuration's host state: transports after E, and for
   ECO_MODEL_THERMALEQ and ECO_MODEL_EULERHOM the mixture temperature last */
#define ECO_NFIELDS_CFG(model, nph, ntr) (3 * (nph) + 4 + (ntr) + ((model) == 3 || (model) == 4))
#define ECO_MAX_TRANSPORTS 4

/* ECO_MODEL_THERMALEQ: two phases in pressure and temperature equilibrium
   (ModThermalEq, src/Models/ThermalEq); host state per phase alpha_k, rho_k,
   p_k (p_k the mixture pressure, rho_k = rho_k(p, T)), u, v, w, E, then the
   mixture temperature T: on input alpha_k, p_0, u and T are read */
/* ECO_MODEL_EULERHOM: the homogeneous liquid-vapour model (ModEulerHomogeneous,
   src/Models/EulerHomogeneous): one mass, momentum and energy equation, the
   liquid (phase liq) and vapour (phase vap) at saturation; first order,
   transmissive boundaries.  Host state as ECO_MODEL_THERMALEQ's: per phase
   alpha_k, rho_k (the saturated density), p_k (the mixture pressure), u, v,
   w, E, then the saturation temperature; on input alpha_k, p_0 and u are read */
enum eco_model { ECO_MODEL_EULER = 0, ECO_MODEL_KAPILA = 1, ECO_MODEL_MULTIP = 2, ECO_MODEL_THERMALEQ = 3,
                 ECO_MODEL_EULERHOM = 4 };
enum eco_limiter { ECO_LIM_NONE = 0, ECO_LIM_MINMOD = 1, ECO_LIM_VANLEER = 2, ECO_LIM_VANALBADA = 3,
                   ECO_LIM_SUPERBEE = 4, ECO_LIM_MC = 5, ECO_LIM_THINC = 6 };
enum eco_eos_type { ECO_EOS_IG = 0, ECO_EOS_SG = 1 };
enum eco_precision { ECO_FP64 = 0, ECO_FP32S = 1 };
/* Boundary conditions: transmissive (BoundCondAbs) and wall (BoundCondWall at
   first order, which the reference also uses for a first-order SYMMETRY
   boundary; BoundCondWallO2 at second order; src/InputOutput/Input.cpp:849-852),
   outflow at an imposed pressure (BoundCondOutflow, at both orders; the
   pressure in bc_p0), injection at an imposed mass flow (BoundCondInj, at both
   orders; the flow and the injected fluid in bc_inj_*; Input.cpp:853) and a
   tank at rest (BoundCondTank, at both orders; its pressure, temperature and
   fractions in bc_tank_*; Input.cpp:854).  All but ECO_BC_ABS run in the FP64
   path only. */
enum eco_bc { ECO_BC_ABS = 0, ECO_BC_WALL = 1, ECO_BC_OUTFLOW = 2, ECO_BC_INJECTION = 3, ECO_BC_TANK = 4 };

enum eco_status_code {
  ECO_OK = 0,
  ECO_ERR_INVALID = -1,      /* bad argument or configuration */
  ECO_ERR_UNSUPPORTED = -2,  /* valid reference option not provided by the device path */
  ECO_ERR_HIP = -3,          /* HIP runtime error (launch, allocation, copy) */
  ECO_ERR_NOMEM = -4,        /* device allocation failed */
  ECO_ERR_STATE = -5,        /* call out of order (e.g. step before set_state) */
  ECO_ERR_SCRATCH = -8,      /* eco_run_noalloc: the solve exceeds the static scratch */
  ECO_ERR_NOT_CONVERGED = -9 /* an injection or tank boundary's Newton solve did not converge in 50
                                iterations: the run stopped there, as the reference's process ends
                                there (Errors::errorMessage; src/Models/Kapila/ModKapila.cpp:283, 405);
                                the status still reports the steps completed */
};

/* Soft-error flags (mirror the reference's non-fatal errors vector). */
#define ECO_WARN_RELAX_NOT_CONVERGED 1u /* RelaxationP Newton hit 100 iterations (src/Relaxations/RelaxationP.cpp:92-109) */
#define ECO_WARN_NAN 2u                 /* non-finite pressure or density produced */
#define ECO_WARN_INFLOW_NOT_CONVERGED 4u /* injection or tank Newton hit 50 iterations; the reference
                                            stops there (src/Models/Kapila/ModKapila.cpp:283, 405), and
                                            so does the device: later steps are no-ops, eco_step and
                                            eco_run return ECO_ERR_NOT_CONVERGED */

/* Equation of state (src/Eos/EosIG.h, src/Eos/EosSG.h; parameters of libEOS/*.xml). */
typedef struct {
  int type;                       /* ECO_EOS_IG | ECO_EOS_SG */
  double gamma, pinf, cv, eref;   /* pinf ignored for IG */
} eco_eos;

/* Solve configuration: what Input::entreeMain/entreeMesh/entreeModel read
   from the deck (src/InputOutput/Input.cpp:85-625) for this path. */
typedef struct {
  int model;        /* enum eco_model */
  int nphase;       /* 1 for Euler; 2, or 3-4 in FP64 without THINC, for Kapila/MultiP; 2 for ThermalEq, EulerHomogeneous */
  int order;        /* 1 = FIRSTORDER, 2 = SECONDORDER (MUSCL predictor-corrector) */
  int lim[4];       /* global, interface, global volume-fraction, interface volume-fraction limiter */
  int alpha_null;   /* Kapila flowModel alphaNull attribute */
  int relax_p;      /* MultiP <relaxation type="P"/> (Kapila always relaxes) */
  int precision;    /* enum eco_precision */
  int n[3];         /* numberCells x, y, z */
  double L[3];      /* dimensions x, y, z (m) */
  double cfl;       /* computationControl CFL */
  int bc[6];        /* X-, X+, Y-, Y+, Z-, Z+ : enum eco_bc (all but ECO_BC_ABS: FP64 only) */
  eco_eos eos[4];   /* per phase */
  double bc_p0[6];  /* <dataOutflow p0> of each ECO_BC_OUTFLOW side (Pa) */
  /* each ECO_BC_INJECTION side (BoundCondInj.cpp:46-150): <dataInjection m0>,
     the mass flow per unit area into the domain (kg/s/m^2, as in the deck),
     and the injected fluid's alpha_k, rho_k, p_k (Euler: rho and p in [0]) */
  double bc_inj_m0[6];
  double bc_inj_alpha[6][4], bc_inj_rho[6][4], bc_inj_p[6][4];
  /* each ECO_BC_TANK side (BoundCondTank.cpp:46-176): <dataTank p0 T0> and, for
     several phases, the tank's volume fractions (or, with bc_tank_massfrac
     set, mass fractions) per phase */
  double bc_tank_p0[6], bc_tank_T0[6];
  double bc_tank_frac[6][4];
  int bc_tank_massfrac[6];
  /* transport equations (<flowModel numberTransports>, Input.cpp:472-540;
     src/Transport/Transport.cpp): passive scalars carried by the contact
     velocity, 0..ECO_MAX_TRANSPORTS, Kapila and MultiP in FP64 single solves
     (the reference's Euler model has no transport Riemann solver: Model.h:127
     ends the run).  bc_tr: each outflow, injection or tank side's
     <transport value> per scalar, the value entering through it
     (Transport.cpp:83-104). */
  int ntransport;
  double bc_tr[6][ECO_MAX_TRANSPORTS];
  /* <relaxation type="PT"/> (RelaxationPT, src/Relaxations/RelaxationPT.cpp:
     49-141) after the pressure relaxation: Kapila and MultiP, two phases,
     FP64 */
  int relax_pt;
  /* <relaxation type="PTMu"> (RelaxationPTMu, src/Relaxations/
     RelaxationPTMu.cpp:51-157) after Kapila's pressure relaxation: the two
     phases (phase 0 the liquid, phase 1 the vapour) relaxed to saturation;
     FP64 single solves without transports */
  int relax_ptmu;
  /* ECO_MODEL_EULERHOM: <flowModel liquid vapor>, the liquid and vapour phase
     indices (0 and 1 in either order) */
  int liq, vap;
  /* each phase's EOS entropyRef (libEOS/*.xml), used by the saturation
     temperature of EulerHomogeneous and RelaxationPTMu (Mixture::computeTsat) */
  double sref[4];
} eco_config;

typedef struct eco_ctx eco_ctx;

typedef struct {
  int steps;          /* steps taken since eco_set_state */
  double t;           /* physical time */
  double dt_last;     /* dt of the last step taken */
  double dt_next;     /* CFL*dtMax for the next step */
  uint32_t flags;     /* ECO_WARN_* accumulated */
  int done;           /* the last eco_run/eco_launch target was reached */
} eco_status;

/* Replaces Run::Run + Run::initialize's mesh/array allocation
   (src/Run.cpp:43,55-108): allocates the device SoA state of one solve.
   hip_stream: hipStream_t to run on (NULL = default stream). */
int eco_create(const eco_config *cfg, void *hip_stream, eco_ctx **out);

/* Replaces Cell::fill + Cell::completeFulfillState (src/Cell.cpp:127-155,
   271-278): uploads the initial primitive state (layout above, E ignored),
   resets time, step counter and the lagged dt (first step uses 1e-15). */
int eco_set_state(eco_ctx *ctx, const double *host_state);

/* Same from a device pointer (same layout), no host copy. */
int eco_set_state_device(eco_ctx *ctx, const double *dev_state);

/* Replaces the clock part of Run::resumeSimulation (src/Run.cpp:164-190,
   Output::readInfos, src/InputOutput/Output.cpp:370-405): after eco_set_state
   and before the first step, start the physical time at t and run the first
   step with dt (the reference reads both from infoCalcul.out; its first
   resumed step uses that dt, not cfl * dtMax of the resumed state).  Later
   steps take the lagged dt as usual; status->steps counts from 0. */
int eco_set_clock(eco_ctx *ctx, double t, double dt);

/* Downloads the current state (layout above, E included). Synchronises. */
int eco_get_state(eco_ctx *ctx, double *host_state);

/* One full time step of Run::solver (src/Run.cpp:220-272: face fluxes,
   cell update, relaxation, energy correction, dtMax reduction). Synchronises
   and reports the status. */
int eco_step(eco_ctx *ctx, eco_status *status);

/* The time loop of Run::solver: nsteps > 0 runs exactly nsteps more steps
   (iteration control), nsteps <= 0 runs until t >= t_final compared as float
   (physicalTime control, src/Run.cpp:239-244, Input.cpp:163); nsteps < 0
   also stops after -nsteps more steps (a guard the reference does not have;
   status->t tells whether t_final was reached).  Synchronises. */
int eco_run(eco_ctx *ctx, int nsteps, double t_final, eco_status *status);

/* Asynchronous form: enqueue nsteps steps on the context stream, no host sync.
   Steps past a reached t_final are no-ops on the device. */
int eco_launch_steps(eco_ctx *ctx, int nsteps);

/* Waits for the context stream and fills status. */
int eco_sync(eco_ctx *ctx, eco_status *status);

/* Measurement form of eco_run(nsteps > 0): the same launches, each stage
   kernel bracketed by HIP events on the context stream.  stage_ms[3] receives
   the summed kernel time of the first-order / predictor / corrector kernels
   (with transports, each stage's transport pass included), launches[3] the
   number of launches of each.  Synchronises. */
int eco_time_steps(eco_ctx *ctx, int nsteps, double *stage_ms, int *launches, eco_status *status);

/* Measurement utility (no reference counterpart): a plain device copy of
   nbytes (multiple of 8) with a streaming kernel, used to calibrate HBM byte
   counters against a kernel of known traffic. */
int eco_copy_calibration(void *dst, const void *src, unsigned long long nbytes, void *hip_stream);

/* Device pointer of the current state (device layout: per phase alpha, rho, p,
   then the nondegenerate velocity components, then E) for zero-copy access.
   FP64 contexts only: an ECO_FP32S context keeps a packed storage format and
   returns NULL with *nfields_device = 0 (eco_get_state converts it). */
const double *eco_state_device(eco_ctx *ctx, int *nfields_device);

/* Allocation-free form.  No counterpart in the reference: for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  One call does what eco_create +
   eco_set_state_device + eco_run + reading the state back do, from state_in
   to state_out, both device arrays in the layout above (E ignored on input,
   written on output; they may not overlap), with the same results bit for
   bit.  nsteps > 0: exactly nsteps steps; nsteps <= 0: until t >= t_final
   (as float), at most -nsteps steps when nsteps < 0.  FP64 only, no transports,
   not ECO_MODEL_THERMALEQ.
   Every intermediate buffer (second time level, MUSCL U*, distance tables,
   control block) lives in a static device array of the library,
   ECO_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; no
   hipMalloc, hipFree or graph is used.  A solve needing more scratch than
   that (eco_noalloc_bytes) is refused with ECO_ERR_SCRATCH before any GPU
   work.  All calls on one GPU share its scratch: calls that overlap in time
   on one GPU race (each call synchronises hip_stream before returning).
   hip_stream: hipStream_t (NULL = default stream). */
#define ECO_NOALLOC_SCRATCH_BYTES (256ull << 20)
int eco_run_noalloc(const eco_config *cfg, const double *state_in, double *state_out, int nsteps, double t_final,
                    void *hip_stream, eco_status *status);

/* Scratch bytes eco_run_noalloc needs for cfg (0 for a cfg it refuses). */
unsigned long long eco_noalloc_bytes(const eco_config *cfg);

/* Releases device memory (Run::finalize, src/Run.cpp:496-530). */
void eco_destroy(eco_ctx *ctx);

const char *eco_strerror(int code);

/* Library version string. */
const char *eco_version(void);

#ifdef __cplusplus
}
#endif
#endif
