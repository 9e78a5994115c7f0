/*
 * CPU oracle for the explicit finite-volume step (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C, FP64 restatement of the reference ECOGEN step on Cartesian
 * meshes (Euler / Kapila / MultiP models, IG and SG equations of state,
 * first order or MUSCL predictor-corrector, the six limiters and THINC,
 * stiff pressure relaxation, energy correction, transmissive boundaries and
 * the lagged CFL time step).  It keeps the reference's full per-cell state
 * (including the derived quantities the reference caches between calls) so
 * that it reproduces the reference's arithmetic operation by operation.
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * use this library, and only as the checker.  The product path never calls
 * it.
 */
#ifndef ECO_ORACLE_H
#define ECO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_EULER = 0, ORC_KAPILA = 1, ORC_MULTIP = 2, ORC_THERMALEQ = 3, ORC_EULERHOM = 4 };
/* ORC_EULERHOM: the homogeneous liquid-vapour model (ModEulerHomogeneous,
   src/Models/EulerHomogeneous): one mass, momentum and energy equation, two
   phases at saturation (pressure, saturation temperature, saturated
   densities), transmissive boundaries only */
enum { ORC_LIM_NONE = 0, ORC_LIM_MINMOD = 1, ORC_LIM_VANLEER = 2, ORC_LIM_VANALBADA = 3,
       ORC_LIM_SUPERBEE = 4, ORC_LIM_MC = 5, ORC_LIM_THINC = 6 };
enum { ORC_EOS_IG = 0, ORC_EOS_SG = 1 };
#define ORC_MAXPH 4
#define ORC_MAXTR 4 /* transport equations (Transport, src/Transport/Transport.cpp) */

/* boundary conditions: BoundCondAbs, and the wall: BoundCondWall at first
   order (also the reference's first-order SYMMETRY), BoundCondWallO2 at second
   order (src/InputOutput/Input.cpp:850-852) */
enum { ORC_BC_ABS = 0, ORC_BC_WALL = 1, ORC_BC_OUTFLOW = 2, ORC_BC_INJ = 3, ORC_BC_TANK = 4 };
/* outflow: BoundCondOutflow, imposed p0; injection: BoundCondInj, imposed mass
   flow m0 of a fluid of given alpha_k, rho_k, p_k; tank: BoundCondTank, a
   reservoir at rest at p0, T0 holding the phases in given volume (or mass)
   fractions */

typedef struct {
  int type;
  double gamma, pinf, cv, eref;
} orc_eos;

typedef struct {
  int model;         /* ORC_EULER / ORC_KAPILA / ORC_MULTIP / ORC_THERMALEQ */
  int nph;           /* 1 for Euler */
  int order;         /* 1 or 2 */
  int lim[4];        /* global, interface, global volume fraction, interface volume fraction */
  int alpha_null;    /* Kapila alphaNull option (epsilon = 1e-15 if set, else 0) */
  int relax_p;       /* MultiP: stiff pressure relaxation requested (Kapila always relaxes) */
  int n[3];
  double L[3];
  double cfl;
  orc_eos eos[ORC_MAXPH];
  int bc[6];         /* X-, X+, Y-, Y+, Z-, Z+: ORC_BC_ABS, ORC_BC_WALL or ORC_BC_OUTFLOW */
  double bc_p0[6];   /* outflow pressure per side */
  double bc_m0[6];   /* injection: <dataInjection m0>, mass flow per unit area into the domain */
  double bc_ak0[6][ORC_MAXPH], bc_rhok0[6][ORC_MAXPH], bc_pk0[6][ORC_MAXPH];  /* injected fluid per phase */
  double bc_tank_p0[6], bc_tank_T0[6];   /* tank: <dataTank p0 T0> */
  double bc_tank_frac[6][ORC_MAXPH];     /* tank: <dataFluid alpha> (or massFrac) per phase */
  int bc_tank_massfrac[6];               /* tank: 1 if bc_tank_frac holds mass fractions */
  int ntr;                               /* transported scalars (<flowModel numberTransports>); not Euler */
  double bc_tr[6][ORC_MAXTR];            /* outflow, injection, tank: <transport value> per transport */
  int relax_pt;      /* Kapila, MultiP: <relaxation type="PT"/> after the P relaxation (two phases) */
  int relax_ptmu;    /* Kapila: <relaxation type="PTMu"/> after the P relaxation (two phases, liq/vap below) */
  int liq, vap;      /* liquid and vapour phase indices: <flowModel liquid vapor> (EulerHomogeneous),
                        <dataPTMu liquid vapor> (PTMu) */
  double sref[ORC_MAXPH]; /* each EOS's entropyRef (Mixture::computeTsat, Mixture.cpp:63-104) */
} orc_cfg;

/*
 * State layout (field-major SoA, ncell = n[0]*n[1]*n[2] doubles per field,
 * cell index i + j*n0 + k*n0*n1):
 *   for each phase k: alpha_k, rho_k, p_k      (3*nph fields; Euler: alpha ignored)
 *   u, v, w                                    (3 fields)
 *   E  (mixture specific total energy; output only)
 *   the ntr transported scalars
 * => nfields = 3*nph + 4 + ntr.
 * ThermalEq (two phases in pressure and temperature equilibrium, ModThermalEq):
 * per phase alpha_k, rho_k, p_k (p_k the mixture pressure, rho_k = rho_k(p, T);
 * on input alpha_k and p_0 are read), u, v, w, E, then the mixture
 * temperature T (read on input) => nfields = 3*nph + 5, no transports.
 * EulerHomogeneous: the same fields (p_k the mixture pressure, rho_k the
 * saturated densities, T the saturation temperature); on input alpha_k, p_0
 * and u are read.
 *
 * orc_run: initialise from (alpha, rho, p, u) exactly as the reference does
 * at t = 0, then march.  nsteps > 0: iteration-controlled run of nsteps
 * steps; nsteps <= 0: time-controlled run to t_final (compared as float, as
 * the reference reads totalTime with QueryFloatAttribute).  dt_log (optional,
 * length >= steps taken) receives the dt used by each step.
 * Returns the number of steps taken, or a negative error code (-1 invalid
 * configuration, -2 allocation, -3 an injection or tank Newton solve did not
 * converge in 50 iterations, where the reference stops).
 */
int orc_run(const orc_cfg *cfg, const double *state_in, double *state_out,
            int nsteps, double t_final, int max_steps,
            double *t_out, double *dt_next_out, double *dt_log, int nthreads);

/* orc_run resumed: physical time from t0, first step with dt0 (orc_run is
   t0 = 0, dt0 = 1e-15). */
int orc_run_from(const orc_cfg *cfg, const double *state_in, double *state_out,
                 int nsteps, double t_final, int max_steps, double t0, double dt0,
                 double *t_out, double *dt_next_out, double *dt_log, int nthreads);

/* Single-function checks used by the known-answer tests. */
double orc_limiter(int lim, double s1, double s2);
double orc_eos_energy(const orc_eos *e, double rho, double p, double epsilon);
double orc_eos_pressure(const orc_eos *e, double rho, double ein);
double orc_eos_sound(const orc_eos *e, double rho, double p, double epsilon);
double orc_eos_isentrope(const orc_eos *e, double p0, double rho0, double rho1, double epsilon);
double orc_eos_rho_pfinal(const orc_eos *e, double p0, double rho0, double pf, double *drho, double epsilon);
/* Relaxation on one cell: in/out alpha, rho, p (nph each). Returns iterations (100 = not applied). */
int orc_relax_p(const orc_cfg *cfg, double *alpha, double *rho, double *p);
/* Relaxations of the last orc_run/orc_run_from that reached the iteration
   limit (no update); first_inputs (3*MAXPH doubles, may be NULL) receives the
   alpha, rho, p per phase (after verifyAndCorrectPhase) of the first one. */
long orc_relax_failures(double *first_inputs);
/* The 1-based step of the last orc_run/orc_run_from in which an injection or
   tank Newton solve reached 50 iterations (the run stopped there, returning
   -3, as the reference's process ends there); 0 when none did. */
int orc_inflow_failure_step(void);
/* One HLLC solve in the face frame, inputs are primitive states (alpha, rho, p per phase; u,v,w).
   flux_out: alpha[nph], mass[nph], energy[nph], qdm[3], E, sM  (Euler: mass, qdm[3], E, sM). */
void orc_hllc(const orc_cfg *cfg, const double *wl, const double *wr, double dxl, double dxr,
              double *flux_out, double *dtmax);

#ifdef __cplusplus
}
#endif
#endif
