// Transport equations of the explicit finite-volume step (src/Transport/
// Transport.cpp): passive scalars carried by the contact velocity sM of each
// face's Riemann problem.  They feed nothing back into the flow, so they run
// as one light pass after each stage's flow kernel, from the faces' sM that
// kernel wrote (StageArgs::smf): one thread per cell, the cell's faces in
// the reference's order, owner computes (no atomics).  FP64, the reference's
// arithmetic operation for operation.
//
// Per face (CellInterface.cpp:139-140, 158-162, 175-179):
//   flux      F = (sM > 0 ? left : right) * sM         Transport::solveRiemann (68-72)
//   high cell acc += coefA*F;  acc += -coefA*T_n*sM    addFlux, addNonCons (107-126)
//   low cell  acc -= coefA*F;  acc -= -coefA*T_n*sM    subtractFlux, subtractNonCons
// with T_n the cell's time-n value; boundary faces only subtract: a wall
// carries no flux (solveRiemannWall), a transmissive side passes its cell
// twice (BoundCondAbs.cpp:77-80), an outflow, injection or tank side takes
// the cell's value leaving and its own value entering (83-104).  MUSCL
// extrapolates the face values with slopes limited by the volume-fraction
// limiters of the same interface test as the phases (CellO2.cpp:101-247,
// CellInterfaceO2.cpp:197-257).  Update (Cell.cpp:222-226, CellO2.cpp:
// 385-390): T = T_n + acc*dt (predictor: acc*(0.5*dt)).
#include <hip/hip_runtime.h>

#include "eco_stage.inl"

namespace eco {

constexpr int kMaxTr = 4;

// Whether this launch slot's stage took its step, and its dt.  Predictor:
// the decision and dt of the stage kernels' step_begin (a read-only call);
// last stage: the step was taken if the stage's last kernel advanced the
// step count, with the dt it recorded.
__device__ __forceinline__ bool transport_step(const TransportArgs &T, double &dt) {
  Control *ctl = T.ctl;
  if (T.stage == 1) {
    StepInfo si;
    if (!step_begin(ctl, T.geo, T.cfl, T.step, 1, false, false, si)) return false;
    dt = si.dt;
    return true;
  }
  const int cur = T.step & 1, prv = cur ^ 1;
  if (ctl->cnt[cur] == ctl->cnt[prv]) return false;
  dt = ctl->dt_last;
  return true;
}

// Face values of the scalars on the face of direction d between cells lo and
// hi (index i = position of lo along d): the stage values at first order,
// the limited extrapolations at second order.
template <bool O2>
__device__ __forceinline__ void face_values(const TransportArgs &T, int d, long long lo, long long hi, int i,
                                            long long sd, double *xl, double *xr) {
  const Geometry &G = T.geo;
  const long long nc = G.ncell;
  ECO_GLOBAL const double *__restrict__ Ts = gmem(T.trs);
  if constexpr (!O2) {
    for (int t = 0; t < T.ntr; t++) { xl[t] = Ts[t * nc + lo]; xr[t] = Ts[t * nc + hi]; }
  } else {
    const int nd = G.n[d];
    const bool has_m = (i >= 1), has_p = (i + 2 <= nd - 1);
    // a second-order wall behind a boundary cell counts the cell itself as
    // the cell across it; its scalar slopes are 0 (BoundCondWallO2.cpp:74-95)
    const bool wm = !has_m && ((T.bc_wall >> (2 * d)) & 1);
    const bool wp = !has_p && ((T.bc_wall >> (2 * d + 1)) & 1);
    bool ifl = false, ifr = false;
    if (T.lim_g != T.lim_i) {
      ECO_GLOBAL const double *As = gmem(T.asrc);
      ECO_GLOBAL const double *An = gmem(T.an);
      const double aL = As[lo], aR = As[hi];
      const double aLL = has_m ? An[lo - sd] : (wm ? An[lo] : 0.);
      const double aRR = has_p ? An[hi + sd] : (wp ? An[hi] : 0.);
      ifl = (aL >= 1.e-4) && (aL <= 1. - 1.e-4) && ((aR - aL) * (aL - aLL) >= 1.e-8);
      ifr = (aR >= 1.e-4) && (aR <= 1. - 1.e-4) && ((aL - aR) * (aR - aRR) >= 1.e-8);
    }
    const int ll = ifl ? T.lim_i : T.lim_g, lr = ifr ? T.lim_i : T.lim_g;
    ECO_GLOBAL const double *Gc = gmem(G.dcc[d]);
    const double dcc = Gc[i];
    const double dL = gmem(G.dfl[d])[i], dR = gmem(G.dfr[d])[i];
    for (int t = 0; t < T.ntr; t++) {
      ECO_GLOBAL const double *F = Ts + t * nc;
      const double wl = F[lo], wr = F[hi];
      // Transport::computeSlopeTransport (Transport.cpp:156-159) per face
      const double sc = (wr - wl) / dcc;
      const double sm = has_m ? (wl - F[lo - sd]) / Gc[i - 1] : 0.;
      const double sp = has_p ? (F[hi + sd] - wr) / Gc[i + 1] : 0.;
      const double gl = limiter(ll, sc, sm), gr = limiter(lr, sc, sp);
      // Transport::extrapolate (Transport.cpp:163-166), the right slope negated
      xl[t] = wl + gl * dL;
      xr[t] = wr + (-gr) * dR;
    }
  }
}

template <int DIM, bool O2>
__global__ void __launch_bounds__(256) k_transport(const TransportArgs T) {
  const Geometry &G = T.geo;
  const long long nc = G.ncell;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double dt;
  if (!transport_step(T, dt)) return;
  const int nx = G.n[0], ny = G.n[1];
  const int ci[3] = {(int)(c % nx), (int)((c / nx) % ny), (int)(c / ((long long)nx * ny))};
  const long long strd[3] = {1, (long long)nx, (long long)nx * ny};
  ECO_GLOBAL const double *__restrict__ Tn = gmem(T.trn);
  double tn[kMaxTr], acc[kMaxTr];
  for (int t = 0; t < T.ntr; t++) { tn[t] = Tn[t * nc + c]; acc[t] = 0.; }
  // interior faces: per direction the low face (added) then the high face (subtracted)
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const double coefA = G.coefA[d];
    ECO_GLOBAL const double *Sm = gmem(T.sm[d]);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const bool interior = side == 0 ? (ci[d] >= 1) : (ci[d] <= G.n[d] - 2);
      if (!interior) continue;
      const long long lo = side == 0 ? c - strd[d] : c, hi = side == 0 ? c : c + strd[d];
      int g[3] = {ci[0], ci[1], ci[2]};
      g[d] += side;
      const double sM = Sm[face_index(G, d, g)];
      double xl[kMaxTr], xr[kMaxTr];
      face_values<O2>(T, d, lo, hi, ci[d] - 1 + side, strd[d], xl, xr);
      for (int t = 0; t < T.ntr; t++) {
        const double F = (sM > 0.) ? xl[t] * sM : xr[t] * sM;
        if (side == 0) {
          acc[t] += coefA * F;
          acc[t] += -coefA * tn[t] * sM;
        } else {
          acc[t] -= coefA * F;
          acc[t] -= -coefA * tn[t] * sM;
        }
      }
    }
  }
  // boundary faces X-, X+, Y-, Y+, Z-, Z+ on the cell's stage value
  ECO_GLOBAL const double *__restrict__ Ts = gmem(T.trs);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const double coefA = G.coefA[d];
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const bool on = side == 0 ? (ci[d] == 0) : (ci[d] == G.n[d] - 1);
      if (!on) continue;
      const int b = 2 * d + side;
      int g[3] = {ci[0], ci[1], ci[2]};
      g[d] += side;
      const double sM = gmem(T.sm[d])[face_index(G, d, g)];
      const bool wall = (T.bc_wall >> b) & 1, in = (T.bc_in >> b) & 1;
      for (int t = 0; t < T.ntr; t++) {
        const double xl = Ts[t * nc + c];
        const double F = wall ? 0. : ((sM > 0.) ? xl * sM : (in ? T.bc_tr[b][t] : xl) * sM);
        acc[t] -= coefA * F;
        acc[t] -= -coefA * tn[t] * sM;
      }
    }
  }
  const double mult = (T.stage == 1) ? 0.5 * dt : dt;
  ECO_GLOBAL double *__restrict__ To = gmem(T.tro);
  for (int t = 0; t < T.ntr; t++) {
    acc[t] *= mult;
    To[t * nc + c] = tn[t] + acc[t];
  }
}

int launch_transport(const TransportArgs &T, hipStream_t st) {
  if (T.ntr < 1 || T.ntr > kMaxTr) return -1;
  const long long nc = T.geo.ncell;
  const dim3 grid((unsigned)((nc + 255) / 256)), block(256);
  const bool o2 = T.order == 2;
  auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, block, 0, st, T); };
  switch (T.geo.dim) {
    case 1: if (o2) go(k_transport<1, true>); else go(k_transport<1, false>); break;
    case 2: if (o2) go(k_transport<2, true>); else go(k_transport<2, false>); break;
    case 3: if (o2) go(k_transport<3, true>); else go(k_transport<3, false>); break;
    default: return -1;
  }
  return (int)hipGetLastError();
}

}  // namespace eco
