#pragma once
// gfx950 kernels of the explicit finite-volume step on Cartesian meshes
// (template definitions; instantiated per model in eco_k_*.hip).
//
// One fused kernel per stage: the first-order step (Run::solveHyperbolic,
// src/Run.cpp:396-412, then Run::solveRelaxations 457-474), the MUSCL
// predictor (Run.cpp:347-364 with CellO2::predictionOrdre2, CellO2.cpp:376-397)
// and the MUSCL corrector (Run.cpp:372-390 then solveRelaxations).
//
// Layout: field-major SoA in HBM, one double per cell per field.  Each block
// owns a tile of cells; for every direction the block computes the Riemann
// fluxes of the faces its tile touches (each interior face of the tile once),
// stages them in LDS, and each thread accumulates its own cell's faces in the
// reference order (x-, x+, y-, y+, z-, z+, then boundary faces X-, X+, Y-, Y+,
// Z-, Z+) -- owner computes, no atomics on the state.  The cell update, the
// buildPrim, the stiff pressure relaxation (FP64 Newton) and the energy
// correction are fused behind the accumulation.  The CFL reduction keeps the
// largest wave speed: wave shuffle, LDS across waves, one 64-bit atomic max
// per block into a 3-slot ring of the control block (src/Run.cpp:225,270).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "eco_physics.hpp"
#include "eco_launch.h"

namespace eco {

// Stored primitives: per phase alpha, rho, p and the velocity (Euler: rho,
// p, velocity; ThermalEq: per phase alpha, the mixture p and T, velocity).
template <int MODEL, int NPH, int DIM>
struct Layout {
  static constexpr bool TE = te_like(MODEL);
  static constexpr int NPRIM = (MODEL == MODEL_EULER) ? 2 + DIM : TE ? NPH + 2 + DIM : 3 * NPH + DIM;
  static constexpr int IE = NPRIM;    // total energy field
  static constexpr int NVAR = NPRIM;  // slope variables
  // flux values + sM
  static constexpr int NFV = (MODEL == MODEL_EULER) ? DIM + 3 : TE ? NPH + DIM + 2 : 3 * NPH + DIM + 2;
  static constexpr int NACC = NFV - 1;  // accumulated values
  // first velocity component in the slope vector
  static constexpr int IU0 = (MODEL == MODEL_EULER) ? 2 : TE ? NPH + 2 : 3 * NPH;
  // slope variable v is a volume fraction (limited by the volume-fraction limiters)
  static constexpr bool is_alpha(int v) {
    return (MODEL == MODEL_EULER) ? false : TE ? (v < NPH) : (v < 3 * NPH && v % 3 == 0);
  }
};

// Device-memory pointers taken from an argument table are generic pointers
// to the compiler (only kernel arguments are known to be global), and
// accesses through them become flat memory operations.  Every state and
// table pointer of the stage kernels is global memory: the kernels carry it
// in pointers of the global address space, so that loads and stores are
// global ones.
#define ECO_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ ECO_GLOBAL T *gmem(T *p) {
  return (ECO_GLOBAL T *)p;
}

// Tile shapes and the occupancy request are build-time parameters chosen by
// the autotuner (ecogen_amd/autotune.py writes eco_tuning.h).
#include "eco_tuning.h"
template <int DIM> struct Tile;
template <> struct Tile<1> {
  static constexpr int X = ECO_TILE1_X, Y = 1, Z = 1, W = ECO_WAVES1, LS = ECO_LIMSPEC1;
  static constexpr int WS = W, WS32 = W;
  static constexpr int C = X * Y * Z, T = ECO_THREADS1 > 0 ? ECO_THREADS1 : C;
  static constexpr int WB = ECO_WAVESB1 > 0 ? ECO_WAVESB1 : W;
  static constexpr int WB32 = ECO_WAVESB32_1 > 0 ? ECO_WAVESB32_1 : WB;
};
template <> struct Tile<2> {
  static constexpr int X = ECO_TILE2_X, Y = ECO_TILE2_Y, Z = 1, W = ECO_WAVES2, LS = ECO_LIMSPEC2;
  static constexpr int WS = ECO_WAVESS2, WS32 = ECO_WAVESS32_2;  // staged kernels
  static constexpr int C = X * Y * Z, T = ECO_THREADS2 > 0 ? ECO_THREADS2 : C;
  static constexpr int WB = ECO_WAVESB2 > 0 ? ECO_WAVESB2 : W;
  static constexpr int WB32 = ECO_WAVESB32_2 > 0 ? ECO_WAVESB32_2 : WB;
};
template <> struct Tile<3> {
  static constexpr int X = ECO_TILE3_X, Y = ECO_TILE3_Y, Z = ECO_TILE3_Z, W = ECO_WAVES3, LS = ECO_LIMSPEC3;
  static constexpr int WS = W, WS32 = W;
  static constexpr int C = X * Y * Z, T = ECO_THREADS3 > 0 ? ECO_THREADS3 : C;
  static constexpr int WB = ECO_WAVESB3 > 0 ? ECO_WAVESB3 : W;
  static constexpr int WB32 = ECO_WAVESB32_3 > 0 ? ECO_WAVESB32_3 : WB;
};
// The FP32-storage staged 2-D kernel's tile: its own height, the rest as Tile.
template <int DIM> struct Tile32S : Tile<DIM> {};
template <> struct Tile32S<2> : Tile<2> {
  static constexpr int Y = ECO_TILE32S_2_Y, C = Tile<2>::X * ECO_TILE32S_2_Y;
};
static_assert(Tile<1>::T >= Tile<1>::C && Tile<1>::T % 64 == 0, "1-D workgroup");
static_assert(Tile<2>::T >= Tile<2>::C && Tile<2>::T % 64 == 0, "2-D workgroup");
static_assert(Tile<3>::T >= Tile<3>::C && Tile<3>::T % 64 == 0, "3-D workgroup");
static_assert(Tile32S<2>::T >= Tile32S<2>::C, "2-D FP32-storage staged workgroup");

template <int DIM>
struct TileFaces {
  static constexpr int T[3] = {Tile<DIM>::X, Tile<DIM>::Y, Tile<DIM>::Z};
  static constexpr int nf(int d) { return d < DIM ? (T[0] + (d == 0)) * (T[1] + (d == 1)) * (T[2] + (d == 2)) : 0; }
  static constexpr int MAXF = (nf(0) > nf(1) ? (nf(0) > nf(2) ? nf(0) : nf(2)) : (nf(1) > nf(2) ? nf(1) : nf(2)));
};

// ---------------------------------------------------------------------------
// state access
// ---------------------------------------------------------------------------
template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void load_prim(PW W, long long nc, long long idx, CellState<NPH> &s) {
  using Ly = Layout<MODEL, NPH, DIM>;
  if constexpr (MODEL == MODEL_EULER) {
    s.r[0] = W[idx];
    s.p[0] = W[nc + idx];
#pragma unroll
    for (int d = 0; d < 3; d++) s.u[d] = (d < DIM) ? W[(2 + d) * nc + idx] : 0.;
    s.a[0] = 1.;
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) s.a[k] = W[k * nc + idx];
    s.pm = W[NPH * nc + idx];
    s.T = W[(NPH + 1) * nc + idx];
#pragma unroll
    for (int d = 0; d < 3; d++) s.u[d] = (d < DIM) ? W[(NPH + 2 + d) * nc + idx] : 0.;
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      s.a[k] = W[(3 * k) * nc + idx];
      s.r[k] = W[(3 * k + 1) * nc + idx];
      s.p[k] = W[(3 * k + 2) * nc + idx];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) s.u[d] = (d < DIM) ? W[(3 * NPH + d) * nc + idx] : 0.;
  }
  (void)Ly::NPRIM;
}

template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void store_prim(PW W, long long nc, long long idx, const CellState<NPH> &s) {
  if constexpr (MODEL == MODEL_EULER) {
    W[idx] = s.r[0];
    W[nc + idx] = s.p[0];
#pragma unroll
    for (int d = 0; d < DIM; d++) W[(2 + d) * nc + idx] = s.u[d];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) W[k * nc + idx] = s.a[k];
    W[NPH * nc + idx] = s.pm;
    W[(NPH + 1) * nc + idx] = s.T;
#pragma unroll
    for (int d = 0; d < DIM; d++) W[(NPH + 2 + d) * nc + idx] = s.u[d];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      W[(3 * k) * nc + idx] = s.a[k];
      W[(3 * k + 1) * nc + idx] = s.r[k];
      W[(3 * k + 2) * nc + idx] = s.p[k];
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) W[(3 * NPH + d) * nc + idx] = s.u[d];
  }
}

// primitive vector used for slopes: (alpha,rho,p)_k, u_0..u_{DIM-1}  (Euler:
// rho, p, u; ThermalEq: alpha_k, p, T, u as PhaseThermalEq::computeSlopesPhase
// and MixThermalEq::computeSlopesMixture, PhaseThermalEq.cpp:139-142,
// MixThermalEq.cpp:457-464)
template <int MODEL, int NPH, int DIM>
__device__ __forceinline__ void prim_to_vec(const CellState<NPH> &s, double *w) {
  if constexpr (MODEL == MODEL_EULER) {
    w[0] = s.r[0]; w[1] = s.p[0];
#pragma unroll
    for (int d = 0; d < DIM; d++) w[2 + d] = s.u[d];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) w[k] = s.a[k];
    w[NPH] = s.pm; w[NPH + 1] = s.T;
#pragma unroll
    for (int d = 0; d < DIM; d++) w[NPH + 2 + d] = s.u[d];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { w[3 * k] = s.a[k]; w[3 * k + 1] = s.r[k]; w[3 * k + 2] = s.p[k]; }
#pragma unroll
    for (int d = 0; d < DIM; d++) w[3 * NPH + d] = s.u[d];
  }
}

template <int MODEL, int NPH, int DIM>
__device__ __forceinline__ void vec_to_prim(const double *w, CellState<NPH> &s) {
  if constexpr (MODEL == MODEL_EULER) {
    s.r[0] = w[0]; s.p[0] = w[1];
#pragma unroll
    for (int d = 0; d < DIM; d++) s.u[d] = w[2 + d];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) s.a[k] = w[k];
    s.pm = w[NPH]; s.T = w[NPH + 1];
#pragma unroll
    for (int d = 0; d < DIM; d++) s.u[d] = w[NPH + 2 + d];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { s.a[k] = w[3 * k]; s.r[k] = w[3 * k + 1]; s.p[k] = w[3 * k + 2]; }
#pragma unroll
    for (int d = 0; d < DIM; d++) s.u[d] = w[3 * NPH + d];
  }
}

// THINC volume-fraction reconstruction (CellInterfaceO2.cpp:228-247, 259-277).
template <int NPH>
__device__ __forceinline__ void thinc(CellState<NPH> &s, double acell, double a_hi, double a_lo, bool right_side) {
  if constexpr (NPH == 2) {
    const double beta = 1.6, epsI = 1.e-4;
    if ((acell >= epsI) && (acell <= 1. - epsI) && ((a_hi - acell) * (acell - a_lo) >= 1.e-8)) {
      double sign = (a_hi - a_lo > 0.) ? 1. : -1.;
      double qmin = smin(a_hi, a_lo);
      double qmax = smax(a_hi, a_lo) - qmin;
      double C = (acell - qmin + 1.e-20) / (qmax + 1.e-20);
      // exp as the C library computes it; cosh(1.6) and tanh(1.6) are the
      // library's values (checked against it on the host)
      constexpr double COSH_BETA = 0x1.49ea5b153125dp+1, TANH_BETA = 0x1.d7e4f0d5dad30p-1;
      double B = exp_lib(sign * beta * (2. * C - 1.));
      double A = (B / COSH_BETA - 1.) / TANH_BETA;
      double na;
      if (!right_side) na = qmin + 0.5 * qmax * (1. + sign * (TANH_BETA + A) / (1. + A * TANH_BETA));
      else na = qmin + 0.5 * qmax * (1. + sign * A);
      if (na < epsI) na = epsI;
      if (na > 1. - epsI) na = 1. - epsI;
      s.a[0] = na;
      s.a[1] = 1. - na;
    }
  }
}

// pack/unpack the face flux through LDS (structure of arrays, one row per value)
template <int MODEL, int NPH, int DIM, int MAXF>
__device__ __forceinline__ void put_flux(double (*L)[MAXF], int f, const FluxVec<NPH> &F) {
  int q = 0;
  if constexpr (MODEL == MODEL_EULER) {
    L[q++][f] = F.m[0];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) L[q++][f] = F.m[k];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { L[q++][f] = F.a[k]; L[q++][f] = F.m[k]; L[q++][f] = F.en[k]; }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) L[q++][f] = F.q[d];
  L[q++][f] = F.E;
  L[q++][f] = F.sM;
}

template <int MODEL, int NPH, int DIM, int MAXF>
__device__ __forceinline__ void get_flux(double (*L)[MAXF], int f, FluxVec<NPH> &F) {
  int q = 0;
  if constexpr (MODEL == MODEL_EULER) {
    F.m[0] = L[q++][f];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) F.m[k] = L[q++][f];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { F.a[k] = L[q++][f]; F.m[k] = L[q++][f]; F.en[k] = L[q++][f]; }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) F.q[d] = (d < DIM) ? L[q++][f] : 0.;
  F.E = L[q++][f];
  F.sM = L[q++][f];
}

// CellInterface::addFlux/subtractFlux (CellInterface.cpp:150-180) with the
// non-conservative terms of FluxKapila/FluxMultiP::addNonCons (owning cell's
// time-n alpha_k, p_k).  sgn = +1 for the face on the cell's low side.
template <int MODEL, int NPH>
__device__ __forceinline__ void add_face(FluxVec<NPH> &U, const FluxVec<NPH> &F, double coefA, bool add,
                                         const CellState<NPH> &own) {
  if (add) {
    if constexpr (MODEL == MODEL_EULER) {
      U.m[0] += coefA * F.m[0];
    } else if constexpr (te_like(MODEL)) {
      // FluxThermalEq::addFlux (FluxThermalEq.cpp:62-69): no non-conservative terms
#pragma unroll
      for (int k = 0; k < NPH; k++) U.m[k] += coefA * F.m[k];
    } else {
#pragma unroll
      for (int k = 0; k < NPH; k++) { U.a[k] += coefA * F.a[k]; U.m[k] += coefA * F.m[k]; U.en[k] += coefA * F.en[k]; }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) U.q[d] += F.q[d] * coefA;
    U.E += coefA * F.E;
    if constexpr (MODEL != MODEL_EULER && !te_like(MODEL)) {
#pragma unroll
      for (int k = 0; k < NPH; k++) {
        U.a[k] += -coefA * own.a[k] * F.sM;
        U.en[k] += coefA * own.a[k] * own.p[k] * F.sM;
      }
    }
  } else {
    if constexpr (MODEL == MODEL_EULER) {
      U.m[0] -= coefA * F.m[0];
    } else if constexpr (te_like(MODEL)) {
#pragma unroll
      for (int k = 0; k < NPH; k++) U.m[k] -= coefA * F.m[k];
    } else {
#pragma unroll
      for (int k = 0; k < NPH; k++) { U.a[k] -= coefA * F.a[k]; U.m[k] -= coefA * F.m[k]; U.en[k] -= coefA * F.en[k]; }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) U.q[d] -= F.q[d] * coefA;
    U.E -= coefA * F.E;
    if constexpr (MODEL != MODEL_EULER && !te_like(MODEL)) {
#pragma unroll
      for (int k = 0; k < NPH; k++) {
        U.a[k] -= -coefA * own.a[k] * F.sM;
        U.en[k] -= coefA * own.a[k] * own.p[k] * F.sM;
      }
    }
  }
}

template <int NPH>
__device__ __forceinline__ void zero_flux(FluxVec<NPH> &U) {
#pragma unroll
  for (int k = 0; k < NPH; k++) { U.a[k] = 0.; U.m[k] = 0.; U.en[k] = 0.; }
  U.q[0] = U.q[1] = U.q[2] = 0.;
  U.E = 0.;
  U.sM = 0.;
}

// per-thread accumulator parked in LDS between directions
template <int MODEL, int NPH, int DIM, int NT>
__device__ __forceinline__ void park_acc(double (*A)[NT], int t, const FluxVec<NPH> &U) {
  int q = 0;
  if constexpr (MODEL == MODEL_EULER) {
    A[q++][t] = U.m[0];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) A[q++][t] = U.m[k];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { A[q++][t] = U.a[k]; A[q++][t] = U.m[k]; A[q++][t] = U.en[k]; }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) A[q++][t] = U.q[d];
  A[q++][t] = U.E;
}

template <int MODEL, int NPH, int DIM, int NT>
__device__ __forceinline__ void unpark_acc(double (*A)[NT], int t, FluxVec<NPH> &U) {
  int q = 0;
  zero_flux(U);
  if constexpr (MODEL == MODEL_EULER) {
    U.m[0] = A[q++][t];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) U.m[k] = A[q++][t];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) { U.a[k] = A[q++][t]; U.m[k] = A[q++][t]; U.en[k] = A[q++][t]; }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) U.q[d] = A[q++][t];
  U.E = A[q++][t];
}

// The cell-cache values of a MultiP-with-relaxation state (SchemeParams::
// ncache): after E the mixture density, then at first order the phase
// energies, the frozen sound speed and the mixture pressure.
template <int MODEL, int NPH, typename PW>
__device__ __forceinline__ void store_cache(PW W, long long nc, int ie, long long idx, int ncache,
                                            const CellState<NPH> &s) {
  W[(ie + 1) * nc + idx] = s.rho;
  if (ncache > 1) {
#pragma unroll
    for (int k = 0; k < NPH; k++) W[(ie + 2 + k) * nc + idx] = s.e[k];
    W[(ie + 2 + NPH) * nc + idx] = s.cF;
    W[(ie + 3 + NPH) * nc + idx] = s.pm;
  }
}
template <int MODEL, int NPH, typename PW>
__device__ __forceinline__ void load_cache(PW W, long long nc, int ie, long long idx, CellState<NPH> &s) {
  s.rho = W[(ie + 1) * nc + idx];
#pragma unroll
  for (int k = 0; k < NPH; k++) s.e[k] = W[(ie + 2 + k) * nc + idx];
  s.cF = W[(ie + 2 + NPH) * nc + idx];
  s.pm = W[(ie + 3 + NPH) * nc + idx];
}

// time-n alpha_k, p_k of one cell: the non-conservative terms' coefficients
template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void load_nc_coeffs(PW W, long long nc, long long idx, CellState<NPH> &s) {
  if constexpr (MODEL != MODEL_EULER && !te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) { s.a[k] = W[(3 * k) * nc + idx]; s.p[k] = W[(3 * k + 2) * nc + idx]; }
  }
}

// Step bookkeeping of Run::solver (src/Run.cpp:220-272) for one launch slot.
struct StepInfo {
  double dt;
  int cnt;  // steps the case has taken before this one
};

// Returns false when the case already reached its target (the launch is a
// no-op for it; its last kernel carries time, count and flag to this slot).
__device__ __forceinline__ bool step_begin(Control *ctl, const Geometry &G, double cfl, int slot, int blk,
                                           bool first_kernel, bool last_kernel, StepInfo &si) {
  const int prv = (slot + 1) & 1;
  const int cnt = ctl->cnt[prv];
  // an injection or tank solve that did not converge stops the run, as the
  // reference's process ends there (Errors::errorMessage, ModKapila.cpp:283)
  const bool stopped = (ctl->err & 4u) != 0;
  if (ctl->done[prv] || stopped) {
    if (last_kernel && blk == 0 && threadIdx.x == 0) {
      const int cur = slot & 1;
      ctl->done[cur] = 1;
      ctl->t[cur] = ctl->t[prv];
      ctl->cnt[cur] = cnt;
      if (stopped) ctl->finished = 1;
    }
    return false;
  }
  // m_dt of Run::Run (src/Run.cpp:43), or the dt a resumed run reads back
  // (Output::readInfos, src/InputOutput/Output.cpp:370-394)
  double dt = (ctl->dt0 > 0.) ? ctl->dt0 : 1.e-15;
  if (cnt > 0) {
    // written by the previous step's kernels: visible at this launch
    const unsigned long long bits = ctl->smax[(slot + 2) % 3];
    double sm = __longlong_as_double((long long)bits);
    double dtMax = (sm > 0.) ? smin(1.e10, G.lcfl / sm) : 1.e10;
    dt = cfl * dtMax;
  }
  if (first_kernel && blk == 0 && threadIdx.x == 0) ctl->smax[(slot + 1) % 3] = 0ull;
  si.dt = dt;
  si.cnt = cnt;
  return true;
}

// Called by one thread per tile with the tile's largest wave speed m.
__device__ __forceinline__ void step_end(Control *ctl, double m, int slot, int blk, bool last_kernel,
                                         const StepInfo &si) {
  // Positive doubles order like their bit patterns.  The running maximum only
  // grows during a step, so a tile whose value does not exceed what a relaxed
  // read already shows leaves it unchanged: the atomic (one per tile, all on
  // one address) is issued only by tiles that can raise it.
  if (m > 0.) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(m);
    if (b > __atomic_load_n(&ctl->smax[slot % 3], __ATOMIC_RELAXED)) atomicMax(&ctl->smax[slot % 3], b);
  }
  if (last_kernel && blk == 0) {
    const int cur = slot & 1, prv = cur ^ 1;
    const double t = ((si.cnt == 0) ? ctl->t0 : ctl->t[prv]) + si.dt;
    const int n = si.cnt + 1;
    ctl->t[cur] = t;
    ctl->cnt[cur] = n;
    ctl->steps = n;
    ctl->dt_last = si.dt;
    ctl->t_now = t;
    // iteration or physical-time control (src/Run.cpp:239-244, tFinal is a
    // float); a time-controlled run may also carry a step cap (max_steps > 0)
    const int done = (ctl->t_final < 0.) ? (n >= ctl->max_steps)
                                          : (t >= (double)(float)ctl->t_final || (ctl->max_steps > 0 && n >= ctl->max_steps));
    ctl->done[cur] = done;
    if (done) ctl->finished = 1;
  }
}

template <int MODEL, int NPH, int DIM>
__device__ __forceinline__ void put_flux_s(double *L, int stride, int f, const FluxVec<NPH> &F) {
  int q = 0;
  if constexpr (MODEL == MODEL_EULER) {
    L[(q++) * stride + f] = F.m[0];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) L[(q++) * stride + f] = F.m[k];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      L[(q++) * stride + f] = F.a[k]; L[(q++) * stride + f] = F.m[k]; L[(q++) * stride + f] = F.en[k];
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) L[(q++) * stride + f] = F.q[d];
  L[(q++) * stride + f] = F.E;
  L[(q++) * stride + f] = F.sM;
}

template <int MODEL, int NPH, int DIM>
__device__ __forceinline__ void get_flux_s(const double *L, int stride, int f, FluxVec<NPH> &F) {
  int q = 0;
  if constexpr (MODEL == MODEL_EULER) {
    F.m[0] = L[(q++) * stride + f];
  } else if constexpr (te_like(MODEL)) {
#pragma unroll
    for (int k = 0; k < NPH; k++) F.m[k] = L[(q++) * stride + f];
  } else {
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      F.a[k] = L[(q++) * stride + f]; F.m[k] = L[(q++) * stride + f]; F.en[k] = L[(q++) * stride + f];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) F.q[d] = (d < DIM) ? L[(q++) * stride + f] : 0.;
  F.E = L[(q++) * stride + f];
  F.sM = L[(q++) * stride + f];
}

// index of the face of direction dir whose high-side cell is g (g[dir] in
// [0, n_dir]: the minus boundary face at 0, the plus one at n_dir)
__device__ __forceinline__ long long face_index(const Geometry &G, int dir, const int *g) {
  const long long n0 = G.n[0] + (dir == 0), n1 = G.n[1] + (dir == 1);
  return g[0] + n0 * (g[1] + n1 * (long long)g[2]);
}

// the cell's two faces of one direction from the LDS flux rows (low face added,
// high face subtracted, in that order: CellInterface.cpp:100-116).  smo
// (transports only): the stage's face sM array of this direction, where the
// cell writes its low face's sM (each interior face has one high-side cell).
template <int MODEL, int NPH, int DIM, int dir>
__device__ __forceinline__ void accumulate_dir(FluxVec<NPH> &acc, const double *sF, int stride, int base, int lx, int ly,
                                               int lz, const int *ci, const Geometry &G, const CellState<NPH> &own,
                                               double *smo = nullptr) {
  constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y;
  const int T0 = TX + (dir == 0), T1 = TY + (dir == 1);
  int fl[3] = {lx, ly, lz};
  const int flo = fl[0] + T0 * (fl[1] + T1 * fl[2]);
  fl[dir] += 1;
  const int fhi = fl[0] + T0 * (fl[1] + T1 * fl[2]);
  FluxVec<NPH> F;
  if (ci[dir] >= 1) {
    get_flux_s<MODEL, NPH, DIM>(sF, stride, base + flo, F);
    add_face<MODEL, NPH>(acc, F, G.coefA[dir], true, own);
    if (smo) gmem(smo)[face_index(G, dir, ci)] = F.sM;
  }
  if (ci[dir] <= G.n[dir] - 2) {
    get_flux_s<MODEL, NPH, DIM>(sF, stride, base + fhi, F);
    add_face<MODEL, NPH>(acc, F, G.coefA[dir], false, own);
  }
}

// Limited extrapolation of every variable to the face (CellInterfaceO2::
// computeSlopes, CellO2::computeLocalSlopes, CellO2.cpp:120-175): slopes at
// i-1/2, i+1/2, i+3/2, the low cell limited from (i+1/2, i-1/2), the high
// cell from (i+1/2, i+3/2).  LV/LA: compile-time limiter of the non-alpha /
// alpha variables, or -1 for the runtime choice from S.lim and the interface
// flags.  RECIP: the FP32-storage form (slopes and limiters in FP32);
// otherwise the reference's arithmetic in FP64.
template <int MODEL, int NPH, int DIM, int LV, int LA, bool RECIP, typename TW>
__device__ __forceinline__ void muscl_vars(const SchemeParams &S, const TW *wl_, const TW *wr_, const TW *wll,
                                           const TW *whh, bool has_m, bool has_p, double dcc, double dcm, double dcp,
                                           double dL, double dR, bool ifl, bool ifr, double *xl, double *xr,
                                           int iuw = -1, double swm = 0., double swp = 0.) {
  using Ly = Layout<MODEL, NPH, DIM>;
  if constexpr (RECIP) {
    // FP32-storage form: differences of the (decoded) cell values in FP64,
    // slopes, limiters and the extrapolation in FP32 (the volume fractions'
    // increment added to their FP64 value)
    const float idc = __builtin_amdgcn_rcpf((float)dcc), idm = __builtin_amdgcn_rcpf((float)dcm),
                idp = __builtin_amdgcn_rcpf((float)dcp);
    const float fL = (float)dL, fR = (float)dR;
#pragma unroll
    for (int v = 0; v < Ly::NVAR; v++) {
      const double wl = wl_[v], wr = wr_[v];
      const float sc = (float)(wr - wl) * idc;
      const float sm = has_m ? (float)(wl - (double)wll[v]) * idm : 0.f;
      const float sp = has_p ? (float)((double)whh[v] - wr) * idp : 0.f;
      const bool isa = Ly::is_alpha(v);
      float gl, gr;
      if constexpr (LV >= 0) {
        gl = isa ? limiter_kt<LA, float>(sc, sm) : limiter_kt<LV, float>(sc, sm);
        gr = isa ? limiter_kt<LA, float>(sc, sp) : limiter_kt<LV, float>(sc, sp);
      } else {
        const int llim = isa ? (ifl ? S.lim[3] : S.lim[2]) : (ifl ? S.lim[1] : S.lim[0]);
        const int rlim = isa ? (ifr ? S.lim[3] : S.lim[2]) : (ifr ? S.lim[1] : S.lim[0]);
        gl = limiter_f(llim, sc, sm);
        gr = limiter_f(rlim, sc, sp);
      }
      if (isa) {
        xl[v] = wl + (double)(gl * fL);
        xr[v] = wr + (double)(-gr * fR);
      } else {
        xl[v] = (double)((float)wl + gl * fL);
        xr[v] = (double)((float)wr + (-gr) * fR);
      }
    }
  } else {
    // the reference's correctly rounded quotients through shared reciprocals
    const double ydc = 1. / dcc, ydm = 1. / dcm, ydp = 1. / dcp;
#pragma unroll
    for (int v = 0; v < Ly::NVAR; v++) {
      const double wl = wl_[v], wr = wr_[v];
      const double sc = qdiv(wr - wl, dcc, ydc);
      // a missing neighbour gives slope 0, or behind a second-order wall the
      // wall face's slope for the normal velocity (variable iuw)
      const double sm = has_m ? qdiv(wl - (double)wll[v], dcm, ydm) : ((v == iuw) ? swm : 0.);
      const double sp = has_p ? qdiv((double)whh[v] - wr, dcp, ydp) : ((v == iuw) ? swp : 0.);
      const bool isa = Ly::is_alpha(v);
      double gl, gr;
      if constexpr (LV >= 0) {
        gl = isa ? limiter_k<LA>(sc, sm) : limiter_k<LV>(sc, sm);
        gr = isa ? limiter_k<LA>(sc, sp) : limiter_k<LV>(sc, sp);
      } else {
        const int llim = isa ? (ifl ? S.lim[3] : S.lim[2]) : (ifl ? S.lim[1] : S.lim[0]);
        const int rlim = isa ? (ifr ? S.lim[3] : S.lim[2]) : (ifr ? S.lim[1] : S.lim[0]);
        gl = limiter(llim, sc, sm);
        gr = limiter(rlim, sc, sp);
      }
      xl[v] = wl + gl * dL;
      xr[v] = wr + (-gr) * dR;
    }
  }
}

// one limiter for the non-alpha variables and one for alpha (no interface
// detection): compile-time limiter pair for the common choices
template <int MODEL, int NPH, int DIM, bool SAMELIM, bool RECIP, typename TW>
__device__ __forceinline__ void muscl_dispatch(const SchemeParams &S, const TW *wl_, const TW *wr_, const TW *wll,
                                               const TW *whh, bool has_m, bool has_p, double dcc, double dcm,
                                               double dcp, double dL, double dR, bool ifl, bool ifr, double *xl,
                                               double *xr, int iuw = -1, double swm = 0., double swp = 0.) {
#define ECO_MUSCL(LV, LA) \
  muscl_vars<MODEL, NPH, DIM, LV, LA, RECIP>(S, wl_, wr_, wll, whh, has_m, has_p, dcc, dcm, dcp, dL, dR, ifl, ifr, xl, xr, \
                                             iuw, swm, swp)
  if constexpr (SAMELIM && Tile<DIM>::LS != 0) {
    const int lv = S.lim[0];
    const int la = (MODEL == MODEL_EULER) ? lv : ((S.lim[2] == LIM_THINC) ? LIM_MINMOD : S.lim[2]);
    if (lv == la) {
      switch (lv) {
        case LIM_MINMOD: ECO_MUSCL(LIM_MINMOD, LIM_MINMOD); return;
        case LIM_VANLEER: ECO_MUSCL(LIM_VANLEER, LIM_VANLEER); return;
        case LIM_VANALBADA: ECO_MUSCL(LIM_VANALBADA, LIM_VANALBADA); return;
        case LIM_MC: ECO_MUSCL(LIM_MC, LIM_MC); return;
        case LIM_SUPERBEE: ECO_MUSCL(LIM_SUPERBEE, LIM_SUPERBEE); return;
        default: break;
      }
    } else if (la == LIM_MINMOD) {
      switch (lv) {
        case LIM_VANLEER: ECO_MUSCL(LIM_VANLEER, LIM_MINMOD); return;
        case LIM_SUPERBEE: ECO_MUSCL(LIM_SUPERBEE, LIM_MINMOD); return;
        default: break;
      }
    }
  }
  ECO_MUSCL(-1, -1);
#undef ECO_MUSCL
}

// BoundCondWallO2::solveRiemann (BoundCondWallO2.cpp:115-150) before the
// wall's Riemann problem: the boundary cell extrapolated to the wall, with
// slopes limited between the wall's own slope (normal velocity over the
// signed centre-to-wall distance, every other slope 0) and the internal face
// across the cell (CellO2::computeLocalSlopesLimite, CellO2.cpp:250-354:
// interface limiters when the cell's time-n alpha lies in [1e-5, 1-1e-5]).
// No phase verification follows the extrapolation there.
template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void wall_state_o2(const SchemeParams &S, const Geometry &G, PW Wsrc, PW Wn,
                                              long long nc, long long cidx, long long sd, int dir, int side,
                                              CellState<NPH> &L) {
  using Ly = Layout<MODEL, NPH, DIM>;
  const int iu = Ly::IU0 + dir;
  CellState<NPH> N;
  const long long in = side == 0 ? cidx + sd : cidx - sd;  // the cell across the internal face
  load_prim<MODEL, NPH, DIM>(Wsrc, nc, in, N);
  double w[Ly::NVAR], wn[Ly::NVAR];
  prim_to_vec<MODEL, NPH, DIM>(L, w);
  prim_to_vec<MODEL, NPH, DIM>(N, wn);
  const int i = side == 0 ? 0 : G.n[dir] - 2;  // index of the internal face along dir
  const double dcc = gmem(G.dcc[dir])[i];
  bool iface = false;
  if constexpr (MODEL != MODEL_EULER) {
    const double a = Wn[cidx];
    iface = (a >= 1.e-5) && (a <= 1. - 1.e-5);
  }
  const int lg = iface ? S.lim[1] : S.lim[0];
  const int lv = iface ? S.lim[3] : S.lim[2];
  const double dw = G.dwall[dir][side];
  const double dg = side == 0 ? -fabs(dw) : fabs(dw);
#pragma unroll
  for (int v = 0; v < Ly::NVAR; v++) {
    const double s1 = (v == iu) ? w[v] / dw : 0.;
    const double s2 = side == 0 ? (wn[v] - w[v]) / dcc : (w[v] - wn[v]) / dcc;
    const bool isa = Ly::is_alpha(v);
    w[v] = w[v] + limiter(isa ? lv : lg, s1, s2) * dg;
  }
  vec_to_prim<MODEL, NPH, DIM>(w, L);
}

// One interior face of the tile: reconstruction (MUSCL) or cached cell states
// (first order), HLLC, flux to LDS slot ``slot`` (row stride ``stride``).
// Returns the largest wave speed of the Riemann problem (0 if no face).
template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM, int dir, bool WALL = true>
__device__ __forceinline__ double face_fp64(const StageArgs &A, const int *o, const long long *strd, int f,
                                            double *sF, int stride, int slot) {
  using Ly = Layout<MODEL, NPH, DIM>;
  constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y;
  constexpr bool O2 = (STAGE != 0);
  const Geometry &G = A.geo;
  const SchemeParams &S = A.sch;
  const long long nc = G.ncell;
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
  ECO_GLOBAL const double *__restrict__ Wsrc = gmem(A.wsrc);
  ECO_GLOBAL const double *__restrict__ Wn = gmem(A.wn);
  const int T0 = TX + (dir == 0), T1 = TY + (dir == 1);
  const int ndir = G.n[dir];
    const int fx = f % T0, fy = (f / T0) % T1, fz = f / (T0 * T1);
    int g[3] = {o[0] + fx, o[1] + fy, o[2] + fz};  // cell on the high side of the face
    if (g[0] >= nx + (dir == 0) || g[1] >= ny + (dir == 1) || g[2] >= nz + (dir == 2)) return 0.;
    const int gi = g[dir];
    if (gi < 1 || gi > ndir - 1) return 0.;  // boundary faces handled per cell below
    const long long ihi = g[0] + strd[1] * g[1] + strd[2] * g[2];
    const long long ilo = ihi - strd[dir];
    const int i = gi - 1;  // index along dir of the low-side cell
    CellState<NPH> L, R;
    load_prim<MODEL, NPH, DIM>(Wsrc, nc, ilo, L);
    load_prim<MODEL, NPH, DIM>(Wsrc, nc, ihi, R);
    if constexpr (!O2) {
      if constexpr (MODEL == MODEL_EULER) {
        // first order, cached FluxEuler::buildPrim state: E stored, c from (rho, p)
        L.E = Wsrc[Ly::IE * nc + ilo]; R.E = Wsrc[Ly::IE * nc + ihi];
        L.c[0] = eos_c(S.eos[0], L.r[0], L.p[0], S.eps);
        R.c[0] = eos_c(S.eos[0], R.r[0], R.p[0], S.eps);
      } else {
        if constexpr (MODEL == MODEL_MULTIP) {
          if (S.ncache > 1) {
            // after a relaxation the cells' cached energies, sound speed,
            // density and pressure are those of buildPrim (and p*)
            load_cache<MODEL, NPH>(Wsrc, nc, Ly::IE, ilo, L);
            load_cache<MODEL, NPH>(Wsrc, nc, Ly::IE, ihi, R);
          } else {
            fulfill<MODEL, NPH>(S, L);
            fulfill<MODEL, NPH>(S, R);
          }
          // MultiP keeps the mixture energy of totalEnergyToInternalEnergy (Run.cpp:471)
          L.em = Wsrc[Ly::IE * nc + ilo] - 0.5 * sqn(L.u);
          R.em = Wsrc[Ly::IE * nc + ihi] - 0.5 * sqn(R.u);
        } else if constexpr (te_like(MODEL)) {
          // the cells' cached mixture energy: sum Y_k e_k of the initial
          // state, then E - |u|^2/2 of Cell::correctionEnergy (Cell.cpp:254-258)
          fulfill<MODEL, NPH>(S, L);
          fulfill<MODEL, NPH>(S, R);
          L.em = Wsrc[(Ly::IE + 1) * nc + ilo];
          R.em = Wsrc[(Ly::IE + 1) * nc + ihi];
        } else {
          fulfill<MODEL, NPH>(S, L);
          fulfill<MODEL, NPH>(S, R);
        }
      }
    } else {
      // MUSCL: limited slopes of both cells from the face slopes at
      // i-1/2, i+1/2, i+3/2 (CellInterfaceO2::computeSlopes + CellO2::
      // computeLocalSlopes), then extrapolation to the face, one variable at a time.
      const bool has_m = (i >= 1), has_p = (i + 2 <= ndir - 1);
      const long long sd = strd[dir];
      // a second-order wall behind a boundary cell (BoundCondWallO2) carries
      // the slope u_n / (x_cell - x_wall) of the normal velocity and counts
      // the cell itself as the cell across it (CellO2.cpp:121-205)
      const bool wm = WALL && !has_m && ((S.bc_wall >> (2 * dir)) & 1);
      const bool wp = WALL && !has_p && ((S.bc_wall >> (2 * dir + 1)) & 1);
      // interface detection: alpha of the face's cells from the stage state,
      // alpha across the opposite face from the time-n state (CellO2.cpp:144-150)
      double aL = 0., aR = 0., aLL = 0., aRR = 0.;
      if constexpr (MODEL != MODEL_EULER) {
        aL = L.a[0]; aR = R.a[0];
        if (!SAMELIM || S.thinc) {
          aLL = has_m ? Wn[ilo - sd] : (wm ? Wn[ilo] : 0.);
          aRR = has_p ? Wn[ihi + sd] : (wp ? Wn[ihi] : 0.);
        }
      }
      bool ifl = false, ifr = false;
      if constexpr (!SAMELIM && MODEL != MODEL_EULER) {
        ifl = (aL >= 1.e-4) && (aL <= 1. - 1.e-4) && ((aR - aL) * (aL - aLL) >= 1.e-8);
        ifr = (aR >= 1.e-4) && (aR <= 1. - 1.e-4) && ((aL - aR) * (aR - aRR) >= 1.e-8);
      }
      ECO_GLOBAL const double *Gc = gmem(G.dcc[dir]);
      const double dcc = Gc[i];
      const double dcm = has_m ? Gc[i - 1] : 1.;
      const double dcp = has_p ? Gc[i + 1] : 1.;
      const double dL = gmem(G.dfl[dir])[i], dR = gmem(G.dfr[dir])[i];
      double wl_[Ly::NVAR], wr_[Ly::NVAR], wll[Ly::NVAR], whh[Ly::NVAR];
      prim_to_vec<MODEL, NPH, DIM>(L, wl_);
      prim_to_vec<MODEL, NPH, DIM>(R, wr_);
      // all neighbour loads issued before any arithmetic, unconditionally (a
      // missing neighbour reads the face's own cell, in bounds; its slope is 0)
      const long long jm = has_m ? ilo - sd : ilo, jp = has_p ? ihi + sd : ihi;
#pragma unroll
      for (int v = 0; v < Ly::NVAR; v++) {
        ECO_GLOBAL const double *Fv = Wsrc + (long long)v * nc;
        wll[v] = Fv[jm];
        whh[v] = Fv[jp];
      }
      constexpr int IU = Ly::IU0 + dir;  // normal velocity in the slope vector
      double swm = 0., swp = 0.;
      if (wm) swm = wl_[IU] / G.dwall[dir][0];
      if (wp) swp = wr_[IU] / G.dwall[dir][1];
      double xl[Ly::NVAR], xr[Ly::NVAR];
      muscl_dispatch<MODEL, NPH, DIM, SAMELIM, false>(S, wl_, wr_, wll, whh, has_m, has_p, dcc, dcm, dcp, dL, dR, ifl,
                                                      ifr, xl, xr, IU, swm, swp);
      vec_to_prim<MODEL, NPH, DIM>(xl, L);
      vec_to_prim<MODEL, NPH, DIM>(xr, R);
      if constexpr (MODEL == MODEL_EULER) {
        verify_correct<MODEL, NPH>(S, L, 0);
        verify_correct<MODEL, NPH>(S, R, 0);
      } else {
#pragma unroll
        for (int k = 0; k < NPH; k++) { verify_correct<MODEL, NPH>(S, L, k); verify_correct<MODEL, NPH>(S, R, k); }
      }
      if (S.thinc) {
        thinc<NPH>(L, aL, aR, aLL, false);
        thinc<NPH>(R, aR, aRR, aL, true);
      }
    }
    to_face_frame(dir, 0, L.u);
    to_face_frame(dir, 0, R.u);
    if constexpr (O2) {
      fulfill<MODEL, NPH>(S, L);
      fulfill<MODEL, NPH>(S, R);
    }
    FluxVec<NPH> F;
    const double smx = hllc<MODEL, NPH>(S, L, R, F);
    from_face_frame(dir, 0, F.q);
    put_flux_s<MODEL, NPH, DIM>(sF, stride, slot, F);
    return smx;
}

template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM, int dir, int NT, bool WALL = true>
__device__ __forceinline__ void stage_dir_3d(const StageArgs &A, const int *o, const long long *strd, int tid, bool active,
                                             int lx, int ly, int lz, const int *ci, long long cidx, double *sF, int stride,
                                             double (*sAcc)[NT], FluxVec<NPH> &acc, double &smx) {
  if constexpr (dir < DIM) {
    constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y, TZ = Tile<DIM>::Z;
    const int nfd = (TX + (dir == 0)) * (TY + (dir == 1)) * (TZ + (dir == 2));
    for (int f = tid; f < nfd; f += NT)
      smx = smax(smx, face_fp64<MODEL, NPH, DIM, STAGE, SAMELIM, dir, WALL>(A, o, strd, f, sF, stride, f));
    __syncthreads();
    if (active) {
      if (dir > 0) unpark_acc<MODEL, NPH, DIM, NT>(sAcc, tid, acc);
      CellState<NPH> own;
      load_nc_coeffs<MODEL, NPH, DIM>(gmem(A.wn), A.geo.ncell, cidx, own);
      accumulate_dir<MODEL, NPH, DIM, dir>(acc, sF, stride, 0, lx, ly, lz, ci, A.geo, own, A.smf[dir]);
      if (dir < DIM - 1) park_acc<MODEL, NPH, DIM, NT>(sAcc, tid, acc);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Staged MUSCL face pass of 2-D tiles (one limiter for the non-alpha and one
// for the alpha variables, no walls, no THINC): the tile's stage state and a
// two-cell halo are loaded into LDS once, each cell's limited slope is
// computed once per direction (CellO2::computeLocalSlopes, CellO2.cpp:101-247:
// on a uniform Cartesian grid the slope of a cell towards either face of a
// direction is the limiter of the same two face slopes, and the limiters are
// symmetric in their arguments), and each face extrapolates its two cells
// from the staged values and slopes.  The face slopes, limiters and
// extrapolations are those of face_fp64 operation for operation, so the
// fluxes are the same bit for bit; a cell's zero limited slope may carry
// the other sign, which adds nothing to a stored value (stored states are
// never -0: velocities below 1e-8 are stored as +0).
// ---------------------------------------------------------------------------
template <int MODEL, int NPH, int DIM>
struct StagedLds {
  using Ly = Layout<MODEL, NPH, DIM>;
  static constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y, NT = Tile<DIM>::T;
  static constexpr int SX = TX + 4, SY = TY + 4, NS = SX * SY;      // state: tile + two-cell halo
  static constexpr int NGX = (TX + 2) * TY, NGY = TX * (TY + 2);    // slopes: tile + one-cell halo along dir
  static constexpr int NG = NGX > NGY ? NGX : NGY;
  static constexpr int NF0 = (TX + 1) * TY, NF1 = TX * (TY + 1);    // faces of one direction
  static constexpr int NFM = NF0 > NF1 ? NF0 : NF1;
  static_assert(NFM <= NT, "staged faces: at most one face per thread and direction");
  // Kapila's HLLC takes an isentropic power per phase and face: with
  // ECO_POWTAB_LDS2 the power's log and exp tables are gathered per lane
  // from LDS rather than through the vector cache (5 KB per workgroup)
  static constexpr bool PTAB = (MODEL == MODEL_KAPILA) && (ECO_POWTAB_LDS2 != 0);
  double S[Ly::NVAR][NS];
  double ptab[PTAB ? POWTAB_WORDS : 1];
  // the limited slopes of one direction, then (after every thread has read
  // them) the fluxes of that direction's faces: the two share one block, so
  // that four workgroups fit a CU's LDS
  union {
    double G[Ly::NVAR][NG];
    double F[Ly::NFV][NFM];
  };
  double sRed[NT / 64];
};

// The stage state of the tile and its halo into LDS (cells beyond the domain
// read the nearest cell: their slopes and faces are never used).
template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void stage_load_state(StagedLds<MODEL, NPH, DIM> &L, PW W, long long nc, const int *o, int nx,
                                                 int ny, int tid) {
  using LD = StagedLds<MODEL, NPH, DIM>;
  using Ly = Layout<MODEL, NPH, DIM>;
  constexpr int NT = LD::NT, NS = LD::NS, SX = LD::SX;
  constexpr int NIT = (NS + NT - 1) / NT;
  long long idx[NIT];
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int c = tid + it * NT;
    const int cc = c < NS ? c : NS - 1;
    const int gx = min(max(o[0] - 2 + cc % SX, 0), nx - 1);
    const int gy = min(max(o[1] - 2 + cc / SX, 0), ny - 1);
    idx[it] = gx + (long long)gy * nx;
  }
#pragma unroll
  for (int v = 0; v < Ly::NVAR; v++) {
    double val[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) val[it] = W[(long long)v * nc + idx[it]];
#pragma unroll
    for (int it = 0; it < NIT; it++)
      if (tid + it * NT < NS) L.S[v][tid + it * NT] = val[it];
  }
}

// Limited slopes of the cells of direction dir's slope region: for cell j,
// limiter(s_{j+1/2}, s_{j-1/2}) with s_{k+1/2} = (w_{k+1} - w_k) / dcc_k
// (CellInterfaceO2::computeSlopes, CellInterfaceO2.cpp:132-152) and a missing
// face slope (next to a boundary) taken as 0.  LV/LA: compile-time limiters
// of the non-alpha / alpha variables, or -1 for the runtime choice.
template <int MODEL, int NPH, int DIM, int dir, int LV, int LA>
__device__ __forceinline__ void stage_slopes(StagedLds<MODEL, NPH, DIM> &L, const SchemeParams &S, const Geometry &G,
                                             const int *o, int tid) {
  using LD = StagedLds<MODEL, NPH, DIM>;
  using Ly = Layout<MODEL, NPH, DIM>;
  constexpr int TX = LD::TX, SX = LD::SX, NT = LD::NT;
  constexpr int NGD = (dir == 0) ? LD::NGX : LD::NGY;
  constexpr int RW = (dir == 0) ? TX + 2 : TX;  // slope-region row width
  constexpr int SD = (dir == 0) ? 1 : SX;       // state-stage stride along dir
  ECO_GLOBAL const double *Gc = gmem(G.dcc[dir]);
  const int ndir = G.n[dir];
  for (int c = tid; c < NGD; c += NT) {
    const int cx = (dir == 0) ? c % RW - 1 : c % RW;
    const int cy = (dir == 0) ? c / RW : c / RW - 1;
    const int gi = o[dir] + (dir == 0 ? cx : cy);
    const bool has_m = (gi >= 1) && (gi <= ndir - 1), has_p = (gi >= 0) && (gi + 1 <= ndir - 1);
    const double dcm = has_m ? Gc[gi - 1] : 1., dcp = has_p ? Gc[gi] : 1.;
    const double ydm = 1. / dcm, ydp = 1. / dcp;
    const int s0 = (cx + 2) + (cy + 2) * SX;
#pragma unroll
    for (int v = 0; v < Ly::NVAR; v++) {
      const double wm = L.S[v][s0 - SD], w = L.S[v][s0], wp = L.S[v][s0 + SD];
      const double sm = has_m ? qdiv(w - wm, dcm, ydm) : 0.;
      const double sp = has_p ? qdiv(wp - w, dcp, ydp) : 0.;
      const bool isa = Ly::is_alpha(v);
      double g;
      if constexpr (LV >= 0) {
        g = isa ? limiter_k<LA>(sp, sm) : limiter_k<LV>(sp, sm);
      } else {
        g = limiter(isa ? S.lim[2] : S.lim[0], sp, sm);
      }
      L.G[v][c] = g;
    }
  }
}

template <int MODEL, int NPH, int DIM, int dir>
__device__ __forceinline__ void stage_slopes_dispatch(StagedLds<MODEL, NPH, DIM> &L, const SchemeParams &S,
                                                      const Geometry &G, const int *o, int tid) {
  const int lv = S.lim[0];
  const int la = (MODEL == MODEL_EULER) ? lv : S.lim[2];
  if (lv == la) {
    switch (lv) {
      case LIM_MINMOD: stage_slopes<MODEL, NPH, DIM, dir, LIM_MINMOD, LIM_MINMOD>(L, S, G, o, tid); return;
      case LIM_VANLEER: stage_slopes<MODEL, NPH, DIM, dir, LIM_VANLEER, LIM_VANLEER>(L, S, G, o, tid); return;
      case LIM_VANALBADA: stage_slopes<MODEL, NPH, DIM, dir, LIM_VANALBADA, LIM_VANALBADA>(L, S, G, o, tid); return;
      case LIM_MC: stage_slopes<MODEL, NPH, DIM, dir, LIM_MC, LIM_MC>(L, S, G, o, tid); return;
      case LIM_SUPERBEE: stage_slopes<MODEL, NPH, DIM, dir, LIM_SUPERBEE, LIM_SUPERBEE>(L, S, G, o, tid); return;
      default: break;
    }
  } else if (la == LIM_MINMOD && lv == LIM_VANLEER) {
    stage_slopes<MODEL, NPH, DIM, dir, LIM_VANLEER, LIM_MINMOD>(L, S, G, o, tid);
    return;
  }
  stage_slopes<MODEL, NPH, DIM, dir, -1, -1>(L, S, G, o, tid);
}

// The interior faces of direction dir from the staged states and slopes:
// extrapolation (CellInterfaceO2::solveRiemann, CellInterfaceO2.cpp:197-295),
// verify/correct, face frame, fulfillState, HLLC.  Thread f computes face f;
// after a barrier (every thread has read the slopes, whose LDS block the
// fluxes take over) it writes the flux to slot f.  Every thread of the
// workgroup calls this (it holds a barrier).
template <int MODEL, int NPH, int DIM, int dir>
__device__ __forceinline__ double stage_faces(StagedLds<MODEL, NPH, DIM> &L, const SchemeParams &S, const Geometry &G,
                                              const int *o, int tid) {
  using LD = StagedLds<MODEL, NPH, DIM>;
  using Ly = Layout<MODEL, NPH, DIM>;
  constexpr int TX = LD::TX, SX = LD::SX;
  constexpr int NFD = (dir == 0) ? LD::NF0 : LD::NF1;
  constexpr int RW = (dir == 0) ? TX + 1 : TX;  // face rows
  constexpr int GW = (dir == 0) ? TX + 2 : TX;  // slope-region row width
  constexpr int SD = (dir == 0) ? 1 : SX, GD = (dir == 0) ? 1 : TX;
  const int nx = G.n[0], ny = G.n[1], ndir = G.n[dir];
  ECO_GLOBAL const double *Dl = gmem(G.dfl[dir]);
  ECO_GLOBAL const double *Dr = gmem(G.dfr[dir]);
  double smx = 0.;
  FluxVec<NPH> F;
  bool have = false;
  const int f = tid;
  const int fx = f % RW, fy = f / RW;
  const int g0 = o[0] + fx, g1 = o[1] + fy;  // cell on the face's high side
  const int gi = (dir == 0) ? g0 : g1;
  // faces of the tile inside the domain; boundary faces are handled per cell
  if (f < NFD && g0 < nx + (dir == 0) && g1 < ny + (dir == 1) && gi >= 1 && gi <= ndir - 1) {
    const int lx = fx - (dir == 0), ly = fy - (dir == 1);  // tile coordinates of the low-side cell
    const int sl = (lx + 2) + (ly + 2) * SX, sr = sl + SD;
    const int gl = (dir == 0) ? (lx + 1) + ly * GW : lx + (ly + 1) * GW, gr = gl + GD;
    const double dL = Dl[gi - 1], dR = Dr[gi - 1];
    double xl[Ly::NVAR], xr[Ly::NVAR];
#pragma unroll
    for (int v = 0; v < Ly::NVAR; v++) {
      xl[v] = L.S[v][sl] + L.G[v][gl] * dL;
      xr[v] = L.S[v][sr] + (-L.G[v][gr]) * dR;
    }
    CellState<NPH> Lc, Rc;
    Lc.u[0] = Lc.u[1] = Lc.u[2] = 0.;
    Rc.u[0] = Rc.u[1] = Rc.u[2] = 0.;
    if constexpr (MODEL == MODEL_EULER) { Lc.a[0] = 1.; Rc.a[0] = 1.; }
    vec_to_prim<MODEL, NPH, DIM>(xl, Lc);
    vec_to_prim<MODEL, NPH, DIM>(xr, Rc);
    if constexpr (MODEL == MODEL_EULER) {
      verify_correct<MODEL, NPH>(S, Lc, 0);
      verify_correct<MODEL, NPH>(S, Rc, 0);
    } else {
#pragma unroll
      for (int k = 0; k < NPH; k++) { verify_correct<MODEL, NPH>(S, Lc, k); verify_correct<MODEL, NPH>(S, Rc, k); }
    }
    to_face_frame(dir, 0, Lc.u);
    to_face_frame(dir, 0, Rc.u);
    fulfill<MODEL, NPH>(S, Lc);
    fulfill<MODEL, NPH>(S, Rc);
    smx = smax(smx, hllc<MODEL, NPH>(S, Lc, Rc, F, LD::PTAB ? L.ptab : nullptr));
    from_face_frame(dir, 0, F.q);
    have = true;
  }
  __syncthreads();  // the slopes are read: their block takes the fluxes
  if (have) put_flux_s<MODEL, NPH, DIM>(&L.F[0][0], LD::NFM, f, F);
  return smx;
}

// ---------------------------------------------------------------------------
// one tile of a stage
// STAGE 0: first-order step; 1: MUSCL predictor (writes U*); 2: MUSCL corrector
// ---------------------------------------------------------------------------
template <int MODEL, int NPH, int DIM>
struct StageLds {
  using Ly = Layout<MODEL, NPH, DIM>;
  using TF = TileFaces<DIM>;
  static constexpr int NT = Tile<DIM>::T;  // threads
  static constexpr bool MERGED = (DIM <= 2);
  static constexpr int NF0 = TF::nf(0), NF1 = TF::nf(1);
  static constexpr int NFT = MERGED ? NF0 + NF1 : TF::MAXF;  // LDS flux slots
  double sF[Ly::NFV][NFT];
  double sAcc[MERGED ? 1 : Ly::NACC][NT];
  double sRed[NT / 64];
};

// LDS of a stage kernel: the staged 2-D MUSCL pass or the direct one.
template <int MODEL, int NPH, int DIM, bool STAGED>
using StageLdsT = typename std::conditional<STAGED, StagedLds<MODEL, NPH, DIM>, StageLds<MODEL, NPH, DIM>>::type;

// Tile `blk` of the case described by A, launch slot `slot`.  STAGED: the
// staged face pass of 2-D MUSCL stages (single limiter form, no walls, no THINC).
template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM, bool WALL = true, bool STAGED = false>
__device__ __forceinline__ void stage_tile(const StageArgs &A, const int slot, const int blk,
                                           StageLdsT<MODEL, NPH, DIM, STAGED> &lds) {
  static_assert(!STAGED || (DIM == 2 && STAGE != 0 && SAMELIM && !WALL), "staged pass: 2-D MUSCL, one limiter, no walls");
  using Ly = Layout<MODEL, NPH, DIM>;
  using LD = StageLds<MODEL, NPH, DIM>;
  constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y, TZ = Tile<DIM>::Z;
  constexpr int NT = LD::NT;
  constexpr bool MERGED = LD::MERGED;
  constexpr int NF0 = LD::NF0, NF1 = LD::NF1;
  constexpr int NFT = LD::NFT;
  auto &sRed = lds.sRed;

  const Geometry &G = A.geo;
  const SchemeParams &S = A.sch;
  Control *ctl = A.ctl;
  const int tid = threadIdx.x;
  const bool first_kernel = (STAGE != 2);
  const bool last_kernel = (STAGE != 1);

  StepInfo si;
  if (!step_begin(ctl, G, S.cfl, slot, blk, first_kernel, last_kernel, si)) return;
  const double dt = si.dt;

  const long long nc = G.ncell;
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
  // tile origin (tile index linearised over the tile grid, x fastest)
  const int tgx = (nx + TX - 1) / TX, tgy = (ny + TY - 1) / TY;
  int b = blk;
  const int bx = b % tgx; b /= tgx;
  const int by = b % tgy; b /= tgy;
  const int bz = b;
  const int o[3] = {bx * TX, by * TY, bz * TZ};
  const int lx = tid % TX, ly = (tid / TX) % TY, lz = tid / (TX * TY);
  const int ci[3] = {o[0] + lx, o[1] + ly, o[2] + lz};
  const bool active = tid < Tile<DIM>::C && ci[0] < nx && ci[1] < ny && ci[2] < nz;
  const long long strd[3] = {1, (long long)nx, (long long)nx * ny};
  const long long cidx = active ? ci[0] + strd[1] * ci[1] + strd[2] * ci[2] : 0;

  ECO_GLOBAL const double *__restrict__ Wsrc = gmem(A.wsrc);  // state the fluxes are built from (U^n or U*)
  ECO_GLOBAL const double *__restrict__ Wn = gmem(A.wn);      // time-n state
  ECO_GLOBAL double *__restrict__ Wout = gmem(A.wout);

  FluxVec<NPH> acc;
  zero_flux(acc);
  double smx = 0.;

  if constexpr (STAGED) {
    // staged 2-D pass: state and halo into LDS; per direction the cells'
    // limited slopes, then the faces, then each cell's two faces
    using SL = StagedLds<MODEL, NPH, DIM>;
    stage_load_state<MODEL, NPH, DIM>(lds, Wsrc, nc, o, nx, ny, tid);
    if constexpr (SL::PTAB) {
      for (int i = tid; i < POWTAB_WORDS; i += NT)
        lds.ptab[i] = (i < POWTAB_EXP) ? powtab::LOG[i / 3][i % 3] : as_double(powtab::EXP[i - POWTAB_EXP]);
    }
    __syncthreads();
    stage_slopes_dispatch<MODEL, NPH, DIM, 0>(lds, S, G, o, tid);
    __syncthreads();
    smx = smax(smx, stage_faces<MODEL, NPH, DIM, 0>(lds, S, G, o, tid));
    __syncthreads();
    CellState<NPH> ownc;
    if (active) {
      load_nc_coeffs<MODEL, NPH, DIM>(Wn, nc, cidx, ownc);
      accumulate_dir<MODEL, NPH, DIM, 0>(acc, &lds.F[0][0], SL::NFM, 0, lx, ly, lz, ci, G, ownc, A.smf[0]);
    }
    __syncthreads();  // the x fluxes are read: the block takes the y slopes
    stage_slopes_dispatch<MODEL, NPH, DIM, 1>(lds, S, G, o, tid);
    __syncthreads();
    smx = smax(smx, stage_faces<MODEL, NPH, DIM, 1>(lds, S, G, o, tid));
    __syncthreads();
    if (active) accumulate_dir<MODEL, NPH, DIM, 1>(acc, &lds.F[0][0], SL::NFM, 0, lx, ly, lz, ci, G, ownc, A.smf[1]);
  } else if constexpr (MERGED) {
    auto &sF = lds.sF;
    // all faces of the tile (both directions) in one pass, one barrier
    for (int w = tid; w < NF0 + NF1; w += NT) {
      if (w < NF0) smx = smax(smx, face_fp64<MODEL, NPH, DIM, STAGE, SAMELIM, 0, WALL>(A, o, strd, w, &sF[0][0], NFT, w));
      else if constexpr (DIM > 1)
        smx = smax(smx, face_fp64<MODEL, NPH, DIM, STAGE, SAMELIM, 1, WALL>(A, o, strd, w - NF0, &sF[0][0], NFT, w));
    }
    __syncthreads();
    if (active) {
      CellState<NPH> own;
      load_nc_coeffs<MODEL, NPH, DIM>(Wn, nc, cidx, own);
      accumulate_dir<MODEL, NPH, DIM, 0>(acc, &sF[0][0], NFT, 0, lx, ly, lz, ci, G, own, A.smf[0]);
      if constexpr (DIM > 1) accumulate_dir<MODEL, NPH, DIM, 1>(acc, &sF[0][0], NFT, NF0, lx, ly, lz, ci, G, own, A.smf[1]);
    }
  } else {
    auto &sF = lds.sF;
    auto &sAcc = lds.sAcc;
    // 3-D: one direction at a time, the accumulator parked in LDS in between
    stage_dir_3d<MODEL, NPH, DIM, STAGE, SAMELIM, 0, NT, WALL>(A, o, strd, tid, active, lx, ly, lz, ci, cidx, &sF[0][0], NFT, sAcc, acc, smx);
    stage_dir_3d<MODEL, NPH, DIM, STAGE, SAMELIM, 1, NT, WALL>(A, o, strd, tid, active, lx, ly, lz, ci, cidx, &sF[0][0], NFT, sAcc, acc, smx);
    stage_dir_3d<MODEL, NPH, DIM, STAGE, SAMELIM, 2, NT, WALL>(A, o, strd, tid, active, lx, ly, lz, ci, cidx, &sF[0][0], NFT, sAcc, acc, smx);
  }

  if (active) {
    // own time-n state (non-conservative terms of the boundary faces, buildCons)
    CellState<NPH> own;
    load_prim<MODEL, NPH, DIM>(Wn, nc, cidx, own);
    own.E = Wn[Ly::IE * nc + cidx];
    // boundary faces (BoundCond::computeFlux, BoundCond.cpp:68-100): transmissive
    // (BoundCondAbs, the interior solver on (L, L)), wall (BoundCondWall /
    // BoundCondWallO2), outflow (BoundCondOutflow), injection (BoundCondInj)
    // or tank (BoundCondTank), the last three on the cell's own state at both orders
#pragma unroll
    for (int dir = 0; dir < DIM; dir++) {
#pragma unroll
      for (int side = 0; side < 2; side++) {
        const bool on = side == 0 ? (ci[dir] == 0) : (ci[dir] == G.n[dir] - 1);
        if (!on) continue;
        CellState<NPH> L;
        load_prim<MODEL, NPH, DIM>(Wsrc, nc, cidx, L);
        const bool wall = WALL && ((S.bc_wall >> (2 * dir + side)) & 1);
        const bool outf = WALL && ((S.bc_out >> (2 * dir + side)) & 1);
        const bool inj = WALL && ((S.bc_inj >> (2 * dir + side)) & 1);
        const bool tank = WALL && ((S.bc_tank >> (2 * dir + side)) & 1);
        if constexpr (STAGE != 0 && WALL) {
          if (wall) wall_state_o2<MODEL, NPH, DIM>(S, G, Wsrc, Wn, nc, cidx, strd[dir], dir, side, L);
        }
        to_face_frame(dir, side == 0, L.u);
        fulfill<MODEL, NPH>(S, L);
        FluxVec<NPH> F;
        if (wall) smx = smax(smx, wall_flux<MODEL, NPH>(L, F));
        else if (outf) smx = smax(smx, outflow_flux<MODEL, NPH>(S, L, S.bc_p0[2 * dir + side], F));
        else if (inj) {
          const int b = 2 * dir + side;
          bool nonconv = false;
          smx = smax(smx, inflow_flux<MODEL, NPH>(S, L, S.inj_m0[b], S.inj_Y[b], S.inj_H[b], F, nonconv));
          if (nonconv) atomicOr(&ctl->err, 4u);
        } else if (tank) {
          const int b = 2 * dir + side;
          bool nonconv = false;
          smx = smax(smx, tank_flux<MODEL, NPH>(S, L, S.tank_p0[b], S.tank_rho[b], S.tank_Y[b], S.tank_H0[b], F, nonconv));
          if (nonconv) atomicOr(&ctl->err, 4u);
        }
        else smx = smax(smx, hllc<MODEL, NPH>(S, L, L, F));
        from_face_frame(dir, side == 0, F.q);
        add_face<MODEL, NPH>(acc, F, G.coefA[dir], false, own);
        if (A.smf[dir]) {
          int g[3] = {ci[0], ci[1], ci[2]};
          g[dir] += side;
          gmem(A.smf[dir])[face_index(G, dir, g)] = F.sM;
        }
      }
    }

    // cell update
    CellState<NPH> sn = own;
    if constexpr (te_like(MODEL)) {
      // phase densities at (p, T) and the mixture density of the last buildPrim
      fulfill<MODEL, NPH>(S, sn);
    } else if constexpr (MODEL != MODEL_EULER) {
      sn.rho = 0.;
#pragma unroll
      for (int k = 0; k < NPH; k++) sn.rho += sn.a[k] * sn.r[k];
      // MultiP with relaxation: the mixture density cached before the relaxation
      if (MODEL == MODEL_MULTIP && S.ncache) sn.rho = Wn[(Ly::IE + 1) * nc + cidx];
    }
    FluxVec<NPH> buf;
    build_cons<MODEL, NPH>(S, sn, buf);
    scale_add<MODEL, NPH>(acc, (STAGE == 1) ? 0.5 * dt : dt, buf);
    scheme_correction<MODEL, NPH>(acc, buf);
    CellState<NPH> s1;
    // EulerHomogeneous: buildPrim's pressure Newton starts at the cell's last
    // pressure (MixEulerHomogeneous.cpp:103)
    if constexpr (MODEL == MODEL_EULERHOM) s1.pm = sn.pm;
    build_prim<MODEL, NPH>(S, acc, s1);
    bool ok = true;
    if constexpr (MODEL != MODEL_EULER) {
      if (S.relax) ok = relax_p<MODEL, NPH>(S, s1);
      if (S.relax_pt) relax_pt<MODEL, NPH>(S, s1);
      if constexpr (MODEL == MODEL_KAPILA && NPH == 2) {
        if (S.relax_ptmu) {
          // RelaxationPTMu: from the cell's own mixture into the cell's own
          // phases; the predictor thus relaxes the time-n state, which it
          // writes for the corrector (A.wn2)
          if constexpr (STAGE == 1) {
            CellState<NPH> tn = own;
            fulfill<MODEL, NPH>(S, tn);
            relax_ptmu<MODEL, NPH>(S, s1, tn, tn.rho, tn.em);
            ECO_GLOBAL double *__restrict__ Wn2 = gmem(A.wn2);
            store_prim<MODEL, NPH, DIM>(Wn2, nc, cidx, tn);
            Wn2[Ly::IE * nc + cidx] = own.E;
          } else {
            relax_ptmu<MODEL, NPH>(S, s1, s1, s1.rho, s1.em);
            fulfill<MODEL, NPH>(S, s1);
          }
        }
      }
      correction_energy<MODEL, NPH>(S, s1);
    }
    if (!ok) atomicOr(&ctl->err, 1u);
    bool bad = false;
    if constexpr (MODEL != MODEL_EULER) {
#pragma unroll
      for (int k = 0; k < NPH; k++) bad |= !(s1.p[k] == s1.p[k]) || !(s1.r[k] == s1.r[k]);
    } else {
      bad = !(s1.p[0] == s1.p[0]);
    }
    if (bad) atomicOr(&ctl->err, 2u);
    store_prim<MODEL, NPH, DIM>(Wout, nc, cidx, s1);
    if constexpr (STAGE != 1) {
      Wout[Ly::IE * nc + cidx] = s1.E;
      // buildPrim's mixture density (the relaxation leaves it unchanged)
      if (MODEL == MODEL_MULTIP && S.ncache) store_cache<MODEL, NPH>(Wout, nc, Ly::IE, cidx, S.ncache, s1);
      // ThermalEq first order: the mixture energy after Cell::correctionEnergy
      if (te_like(MODEL) && S.ncache) Wout[(Ly::IE + 1) * nc + cidx] = s1.E - 0.5 * sqn(s1.u);
    }
  }

  // CFL reduction: wave shuffle, LDS, one atomic per block
  for (int off = 32; off > 0; off >>= 1) smx = smax(smx, __shfl_xor(smx, off));
  if ((tid & 63) == 0) sRed[tid >> 6] = smx;
  __syncthreads();
  if (tid == 0) {
    double m = sRed[0];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) m = smax(m, sRed[w]);
    step_end(ctl, m, slot, blk, last_kernel, si);
  }
}

// Workgroups are dispatched to the XCDs round robin (workgroup i on XCD i mod
// 8), each XCD with its own L2.  Workgroup i takes the tile that makes every
// XCD walk one contiguous band of the tiles (x fastest): tiles next to each
// other, which read each other's halo cells, then share an L2.  A bijection
// of [0, n) for any n (the first n mod 8 XCDs take one tile more).  On 2-D
// tiles only: it halves the MUSCL stages' HBM reads there at the same time,
// while on the 3-D 16x4x3 tiles (whose y and z neighbours are 24 and 2304
// tiles away) it made the MUSCL step 9 % slower.
template <int DIM>
__device__ __forceinline__ int xcd_band_tile(int wg, int n) {
  if constexpr (ECO_XCD_BANDS == 0 || DIM != 2) return wg;
  constexpr int NX = 8;
  const int q = n / NX, r = n % NX;
  const int x = wg % NX, l = wg / NX;
  return (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// WALL = false: the MUSCL stages of a solve without walls (no wall code in
// the face pass or the boundary faces; the register allocation of the hot
// path is the transmissive one)
template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM, bool WALL = true, bool STAGED = false>
__global__ void __launch_bounds__(Tile<DIM>::T, STAGED ? Tile<DIM>::WS : Tile<DIM>::W)
k_stage(const StageArgs A) {
  __shared__ StageLdsT<MODEL, NPH, DIM, STAGED> lds;
  stage_tile<MODEL, NPH, DIM, STAGE, SAMELIM, WALL, STAGED>(A, A.step, xcd_band_tile<DIM>(blockIdx.x, gridDim.x), lds);
}

// Persistent form over a pool: every workgroup walks the class's tile table
// (stride gridDim.x); each entry names a case and a tile of it.  The tile
// count is read from device memory, so a captured launch stays valid while
// the table changes between replays.
template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM>
__global__ void __launch_bounds__(Tile<DIM>::T, Tile<DIM>::WB)
k_stage_batch(const StageArgs *__restrict__ args, const TileRef *__restrict__ tiles,
                                   const int *__restrict__ ntile, const int slot) {
  __shared__ StageLds<MODEL, NPH, DIM> lds;
  const int ntiles = __builtin_amdgcn_readfirstlane(*ntile);
  const int kidx = (STAGE == 2) ? 1 : 0;
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    // uniform by construction; readfirstlane keeps the entry (and every
    // address derived from it) in scalar registers
    TileRef t = tiles[w];
    t.c = __builtin_amdgcn_readfirstlane(t.c);
    t.blk = __builtin_amdgcn_readfirstlane(t.blk);
    const StageArgs &A = args[(size_t)t.c * 4 + (slot & 1) * 2 + kidx];
    stage_tile<MODEL, NPH, DIM, STAGE, SAMELIM>(A, slot, t.blk, lds);
    __syncthreads();  // LDS reuse by the next tile
  }
}

// E = e + |u|^2/2 at t = 0 (Cell::completeFulfillState, Cell.cpp:271-278)
template <int MODEL, int NPH, int DIM>
__global__ void k_init(const StageArgs A) {
  using Ly = Layout<MODEL, NPH, DIM>;
  const long long nc = A.geo.ncell;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nc) return;
  CellState<NPH> s;
  load_prim<MODEL, NPH, DIM>(A.wsrc, nc, i, s);
  if constexpr (MODEL == MODEL_EULERHOM) {
    // the saturation temperature of the initial pressure (ModEulerHomogeneous::fulfillState)
    s.T = sat_tsat(A.sch, s.pm, nullptr);
    A.wout[(NPH + 1) * nc + i] = s.T;
  }
  fulfill<MODEL, NPH>(A.sch, s);
  if constexpr (MODEL != MODEL_EULER) s.E = s.em + 0.5 * sqn(s.u);
  A.wout[Ly::IE * nc + i] = s.E;
  if (MODEL == MODEL_MULTIP && A.sch.ncache) store_cache<MODEL, NPH>(A.wout, nc, Ly::IE, i, A.sch.ncache, s);
  // ThermalEq first order: the initial cells' mixture energy sum Y_k e_k
  // (Cell::completeFulfillState, Cell.cpp:271-278)
  if (te_like(MODEL) && A.sch.ncache) A.wout[(Ly::IE + 1) * nc + i] = s.em;
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
template <int MODEL, int NPH, int DIM>
int launch_t(const StageArgs &A, int stage, hipStream_t st) {
  constexpr int NT = Tile<DIM>::T;
  const int tgx = (A.geo.n[0] + Tile<DIM>::X - 1) / Tile<DIM>::X;
  const int tgy = (A.geo.n[1] + Tile<DIM>::Y - 1) / Tile<DIM>::Y;
  const int tgz = (A.geo.n[2] + Tile<DIM>::Z - 1) / Tile<DIM>::Z;
  dim3 grid((unsigned)(tgx * tgy * tgz)), block(NT);
  const bool same = A.sch.same_lim != 0;
  const bool wall = (A.sch.bc_wall | A.sch.bc_out | A.sch.bc_inj | A.sch.bc_tank) != 0;  // kernels with boundary code
  // 2-D MUSCL with one limiter form, no walls and no THINC: the staged face pass
  constexpr bool CAN_STAGE = (DIM == 2);
  const bool staged = CAN_STAGE && same && !wall && !A.sch.thinc && ECO_STAGED2 != 0;
  if constexpr (CAN_STAGE) {
    if (staged && stage == 1) {
      hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 1, true, false, true>), grid, block, 0, st, A);
      return (int)hipGetLastError();
    }
    if (staged && stage == 2) {
      hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 2, true, false, true>), grid, block, 0, st, A);
      return (int)hipGetLastError();
    }
  }
  if (stage == 0 && wall) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 0, true>), grid, block, 0, st, A);
  else if (stage == 0) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 0, true, false>), grid, block, 0, st, A);
  else if (stage == 1 && same && wall) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 1, true>), grid, block, 0, st, A);
  else if (stage == 1 && same) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 1, true, false>), grid, block, 0, st, A);
  else if (stage == 1 && wall) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 1, false>), grid, block, 0, st, A);
  else if (stage == 1) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 1, false, false>), grid, block, 0, st, A);
  else if (stage == 2 && same && wall) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 2, true>), grid, block, 0, st, A);
  else if (stage == 2 && same) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 2, true, false>), grid, block, 0, st, A);
  else if (stage == 2 && wall) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 2, false>), grid, block, 0, st, A);
  else if (stage == 2) hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 2, false, false>), grid, block, 0, st, A);
  else if (stage == 3) {
    const long long nc = A.geo.ncell;
    hipLaunchKernelGGL((k_init<MODEL, NPH, DIM>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, A);
  }
  return (int)hipGetLastError();
}

// First-order-only models (EulerHomogeneous: transmissive boundaries, first
// order): the first-order stage and the initialisation
template <int MODEL, int NPH, int DIM>
int launch_t_fo(const StageArgs &A, int stage, hipStream_t st) {
  constexpr int NT = Tile<DIM>::T;
  const int tgx = (A.geo.n[0] + Tile<DIM>::X - 1) / Tile<DIM>::X;
  const int tgy = (A.geo.n[1] + Tile<DIM>::Y - 1) / Tile<DIM>::Y;
  const int tgz = (A.geo.n[2] + Tile<DIM>::Z - 1) / Tile<DIM>::Z;
  const bool wall = (A.sch.bc_wall | A.sch.bc_out | A.sch.bc_inj | A.sch.bc_tank) != 0;
  if (stage == 0 && !wall) {
    hipLaunchKernelGGL((k_stage<MODEL, NPH, DIM, 0, true, false>), dim3((unsigned)(tgx * tgy * tgz)), dim3(NT), 0, st, A);
  } else if (stage == 3) {
    const long long nc = A.geo.ncell;
    hipLaunchKernelGGL((k_init<MODEL, NPH, DIM>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, A);
  } else {
    return -1;
  }
  return (int)hipGetLastError();
}

template <int MODEL, int NPH>
int launch_d_fo(const StageArgs &A, int stage, hipStream_t st) {
  switch (A.geo.dim) {
    case 1: return launch_t_fo<MODEL, NPH, 1>(A, stage, st);
    case 2: return launch_t_fo<MODEL, NPH, 2>(A, stage, st);
    case 3: return launch_t_fo<MODEL, NPH, 3>(A, stage, st);
  }
  return -1;
}

template <int MODEL, int NPH>
int launch_d(const StageArgs &A, int stage, hipStream_t st) {
  switch (A.geo.dim) {
    case 1: return launch_t<MODEL, NPH, 1>(A, stage, st);
    case 2: return launch_t<MODEL, NPH, 2>(A, stage, st);
    case 3: return launch_t<MODEL, NPH, 3>(A, stage, st);
  }
  return -1;
}

template <int MODEL, int NPH, int DIM, int STAGE, bool SAME>
int launch_tb1(const BatchLaunch &L, hipStream_t st) {
  constexpr int NT = Tile<DIM>::T;
  if (L.occupancy)
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(L.occupancy, k_stage_batch<MODEL, NPH, DIM, STAGE, SAME>,
                                                             NT, 0);
  hipLaunchKernelGGL((k_stage_batch<MODEL, NPH, DIM, STAGE, SAME>), dim3((unsigned)L.grid), dim3(NT), 0, st, L.args,
                     L.tiles, L.ntiles, L.slot);
  return (int)hipGetLastError();
}

template <int MODEL, int NPH, int DIM>
int launch_tb(const BatchLaunch &L, hipStream_t st) {
  switch (L.stage) {
    case 0: return launch_tb1<MODEL, NPH, DIM, 0, true>(L, st);
    case 1: return L.same ? launch_tb1<MODEL, NPH, DIM, 1, true>(L, st) : launch_tb1<MODEL, NPH, DIM, 1, false>(L, st);
    case 2: return L.same ? launch_tb1<MODEL, NPH, DIM, 2, true>(L, st) : launch_tb1<MODEL, NPH, DIM, 2, false>(L, st);
  }
  return -1;
}

template <int MODEL, int NPH>
int launch_db(const BatchLaunch &L, hipStream_t st) {
  switch (L.dim) {
    case 1: return launch_tb<MODEL, NPH, 1>(L, st);
    case 2: return launch_tb<MODEL, NPH, 2>(L, st);
    case 3: return launch_tb<MODEL, NPH, 3>(L, st);
  }
  return -1;
}

}  // namespace eco
