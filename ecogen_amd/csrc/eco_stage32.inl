// FP32-storage form of the MUSCL stage kernels (multiphase models).
//
// State in HBM: the volume-fraction pair in two 32-bit words (one FP64
// value, below), rho_k, p_k and the velocity in FP32, plus the reduced
// mixture energy R = rho*E - sum_k alpha_k (gamma_k pinf_k/(gamma_k-1) + rho_k eref_k)
// in FP32.  R is a pressure-scale quantity: the stiffened-gas offset is
// re-added in FP64 from the stored alpha_k, rho_k, so the (alpha, E) pair
// stays self-consistent after rounding and the mixture pressure recovered by
// the energy correction keeps its accuracy.
//
// Compute: wave-speed estimates, the face slopes, the limiters and the
// extrapolation increments in FP32 (each increment added to the cell's
// decoded FP64 value); every flux component, the accumulation, buildPrim,
// the stiff pressure relaxation (Newton) and the energy correction in FP64.
#pragma once
#include "eco_stage.inl"

namespace eco {

// Two-phase volume fractions in the two 32-bit words of the alpha_0 and
// alpha_1 planes: the smaller fraction as one FP64 value (high word in the
// alpha_0 plane, low word in the alpha_1 plane), negated when it is phase
// 1's; the larger is its FP64 complement.  The pair costs the bytes of two
// FP32 values and stores no rounding: alpha rounded to FP32 every step
// (6e-8 of the smaller fraction) was the largest source of FP32-storage
// error on long runs, through the stiffened-gas energy
// alpha*gamma*pinf/(gamma-1) and the pressure equilibrium.  The words are
// only moved as bits, never used as FP32 numbers.
__device__ __forceinline__ void decode_alpha(float w0, float w1, double &a0, double &a1) {
  const double d = __hiloint2double(__float_as_int(w0), __float_as_int(w1));
  if (__builtin_signbit(d)) { a1 = -d; a0 = 1. - a1; }
  else { a0 = d; a1 = 1. - a0; }
}

__device__ __forceinline__ void encode_alpha(double a0, double a1, float &w0, float &w1) {
  const double d = (a0 <= a1) ? a0 : -a1;
  w0 = __int_as_float(__double2hiint(d));
  w1 = __int_as_float(__double2loint(d));
}

template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void load_prim_f(PW W, long long nc, long long idx, CellState<NPH> &s) {
  static_assert(NPH == 2, "FP32 storage holds two-phase states");
  decode_alpha(W[idx], W[3 * nc + idx], s.a[0], s.a[1]);
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    s.r[k] = (double)W[(3 * k + 1) * nc + idx];
    s.p[k] = (double)W[(3 * k + 2) * nc + idx];
  }
#pragma unroll
  for (int d = 0; d < 3; d++) s.u[d] = (d < DIM) ? (double)W[(3 * NPH + d) * nc + idx] : 0.;
}

template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void store_prim_f(PW W, long long nc, long long idx, const CellState<NPH> &s) {
  float w0, w1;
  encode_alpha(s.a[0], s.a[1], w0, w1);
  W[idx] = w0;
  W[3 * nc + idx] = w1;
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    W[(3 * k + 1) * nc + idx] = (float)s.r[k];
    W[(3 * k + 2) * nc + idx] = (float)s.p[k];
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) W[(3 * NPH + d) * nc + idx] = (float)s.u[d];
}

// energy offset sum_k alpha_k (gamma pinf/(gamma-1) + rho_k eref) of a state
template <int NPH>
__device__ __forceinline__ double energy_offset(const SchemeParams &S, const CellState<NPH> &s) {
  double o = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) o += s.a[k] * (S.eos[k].gpg + s.r[k] * S.eos[k].eref);
  return o;
}

// the FP32-rounded state as stored (for the offset of R)
template <int NPH>
__device__ __forceinline__ void round_state(CellState<NPH> &s) {
  float w0, w1;
  encode_alpha(s.a[0], s.a[1], w0, w1);
  decode_alpha(w0, w1, s.a[0], s.a[1]);
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    s.r[k] = (double)(float)s.r[k];
    s.p[k] = (double)(float)s.p[k];
  }
}

// Specific internal energy with one reciprocal: the FP32-storage path needs
// the fluxes mutually consistent in FP64, not the reference's rounding.
__device__ __forceinline__ double eos_e_fast(const EosParams &q, double rho, double p, double eps) {
  return (p + q.gpinf) * q.og * rcp_fast(smax(rho, eps)) + q.eref;
}

// energies and mixture variables in FP64 (no sound speeds): the part of
// fulfillState the fluxes need at full precision
template <int MODEL, int NPH>
__device__ __forceinline__ void fulfill_energy(const SchemeParams &S, CellState<NPH> &s) {
  const double eps = S.eps;
#pragma unroll
  for (int k = 0; k < NPH; k++) s.e[k] = eos_e_fast(S.eos[k], s.r[k], s.p[k], eps);
  s.rho = 0.; s.pm = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) { s.rho += s.a[k] * s.r[k]; s.pm += s.a[k] * s.p[k]; }
  const double irho = rcp_fast((MODEL == MODEL_KAPILA) ? smax(s.rho, eps) : s.rho);
#pragma unroll
  for (int k = 0; k < NPH; k++) s.Y[k] = s.a[k] * s.r[k] * irho;
  s.em = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) s.em += s.Y[k] * s.e[k];
}

// buildCons with one reciprocal per phase
template <int MODEL, int NPH>
__device__ __forceinline__ void build_cons_fast(const SchemeParams &S, const CellState<NPH> &s, FluxVec<NPH> &f) {
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    f.a[k] = s.a[k];
    f.m[k] = s.a[k] * s.r[k];
    f.en[k] = s.a[k] * (s.p[k] + S.eos[k].gpinf) * S.eos[k].og + f.m[k] * S.eos[k].eref;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) f.q[d] = s.u[d] * s.rho;
  f.E = s.rho * s.E;
}

// RelaxationP::stiffRelaxation (RelaxationP.cpp:49-111) with the Newton
// residual in the form sum ak denom/num - 1 (one division per phase per
// iteration), same skip rule, floors, tolerance and 100-iteration cap.
template <int MODEL, int NPH>
__device__ __forceinline__ bool relax_p_fast(const SchemeParams &S, CellState<NPH> &s) {
  if (S.eps > 1.e-20) {
#pragma unroll
    for (int k = 0; k < NPH; k++) if (s.a[k] > (1. - 1.e-5)) return true;
  }
  double ak[NPH], pk[NPH], rk[NPH];
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
    pStar -= f * rcp_fast(df);
#pragma unroll
    for (int k = 0; k < NPH; k++) pStar = eos_fix_p(S.eos[k], pStar);
    f = -1.; df = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) {
      const EosParams &q = S.eos[k];
      const double num = q.gamma * (pStar + q.pinf);
      const double inum = rcp_fast(num);
      f += ak[k] * (num + pk[k] - pStar) * inum;
      df -= ak[k] * q.gamma * (pk[k] + q.pinf) * inum * inum;
    }
  } while (fabs(f) > 1e-10 && it < 100);
  if (it >= 100) return false;
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    const EosParams &q = S.eos[k];
    const double num = q.gamma * (pStar + q.pinf);
    const double denom = num + pk[k] - pStar;
    s.a[k] = ak[k] * denom * rcp_fast(num);
    s.r[k] = rk[k] * num * rcp_fast(denom);
    s.p[k] = pStar;
  }
  s.pm = pStar;
  return true;
}

__device__ __forceinline__ float sminf_(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smaxf_(float a, float b) { return (a < b) ? b : a; }

// frozen sound speed in FP32 (MixKapila.cpp:147-172)
template <int NPH>
__device__ __forceinline__ float frozen_c_f(const SchemeParams &S, const CellState<NPH> &s) {
  // hardware reciprocals and square root (1 ulp each)
  float rho = 0.f, ck2[NPH];
#pragma unroll
  for (int k = 0; k < NPH; k++) rho += (float)s.a[k] * (float)s.r[k];
  const float irho = __builtin_amdgcn_rcpf(rho);
  float c2 = 0.f;
#pragma unroll
  for (int k = 0; k < NPH; k++) {
    const EosParams &q = S.eos[k];
    const float r = (float)s.r[k], p = (float)s.p[k];
    const float ir = __builtin_amdgcn_rcpf(r);
    ck2[k] = (q.type == EOS_SG) ? (float)q.gamma * (p + (float)q.pinf) * ir : (float)q.gamma * p * ir;
    c2 += ((float)s.a[k] * r) * irho * ck2[k];
  }
  return __builtin_amdgcn_sqrtf(c2);
}

// Davis estimates (ModKapila.cpp:117-121) in FP32; returns max |s| above the guard
template <int NPH>
__device__ __forceinline__ float speeds_f(const SchemeParams &S, const CellState<NPH> &L, const CellState<NPH> &R,
                                          double &sL, double &sR) {
  const float cL = frozen_c_f<NPH>(S, L), cR = frozen_c_f<NPH>(S, R);
  const float uL = (float)L.u[0], uR = (float)R.u[0];
  const float a = sminf_(uL - cL, uR - cR);
  const float b = smaxf_(uR + cR, uL + cL);
  float smx = 0.f;
  if (fabsf(a) > 1.e-3f) smx = fabsf(a);
  if (fabsf(b) > 1.e-3f) smx = smaxf_(smx, fabsf(b));
  sL = (double)a;
  sR = (double)b;
  return smx;
}

// time-n alpha_k, p_k of the owning cell (non-conservative terms)
template <int NPH, typename PW>
__device__ __forceinline__ void load_nc_coeffs_f(PW Wn, long long nc, long long cidx, CellState<NPH> &own) {
  decode_alpha(Wn[cidx], Wn[3 * nc + cidx], own.a[0], own.a[1]);
#pragma unroll
  for (int k = 0; k < NPH; k++) own.p[k] = (double)Wn[(3 * k + 2) * nc + cidx];
}

// One interior face of the tile in the FP32-storage form: FP32 loads,
// reconstruction and limiting in FP64 (slopes through one reciprocal per
// cell distance), FP32 wave-speed estimates, FP64 HLLC fluxes to LDS slot
// ``slot``.  Returns the largest wave speed above the guard (0 if no face).
template <int MODEL, int NPH, int DIM, bool SAMELIM, int dir>
__device__ __forceinline__ float face_fp32s(const StageArgs &A, const int *o, const long long *strd, int f, double *sF,
                                            int stride, int slot) {
  using Ly = Layout<MODEL, NPH, DIM>;
  constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y;
  const Geometry &G = A.geo;
  const SchemeParams &S = A.sch;
  const long long nc = G.ncell;
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
  ECO_GLOBAL const float *__restrict__ Wsrc = (ECO_GLOBAL const float *)gmem(A.wsrc);
  ECO_GLOBAL const float *__restrict__ Wn = (ECO_GLOBAL const float *)gmem(A.wn);
  const int T0 = TX + (dir == 0), T1 = TY + (dir == 1);
  const int ndir = G.n[dir];
  const int fx = f % T0, fy = (f / T0) % T1, fz = f / (T0 * T1);
  int g[3] = {o[0] + fx, o[1] + fy, o[2] + fz};
  if (g[0] >= nx + (dir == 0) || g[1] >= ny + (dir == 1) || g[2] >= nz + (dir == 2)) return 0.f;
  const int gi = g[dir];
  if (gi < 1 || gi > ndir - 1) return 0.f;
  const long long ihi = g[0] + strd[1] * g[1] + strd[2] * g[2];
  const long long ilo = ihi - strd[dir];
  const long long sd = strd[dir];
  const int i = gi - 1;
  const bool has_m = (i >= 1), has_p = (i + 2 <= ndir - 1);
  // every load unconditional (a missing neighbour reads the face's own cell,
  // in bounds, and its slope is replaced by 0): no branch, no wait between loads
  const long long jm = has_m ? ilo - sd : ilo, jp = has_p ? ihi + sd : ihi;
  float wl_[Ly::NVAR], wr_[Ly::NVAR], wll[Ly::NVAR], whh[Ly::NVAR];
#pragma unroll
  for (int v = 0; v < Ly::NVAR; v++) {
    ECO_GLOBAL const float *Fv = Wsrc + (long long)v * nc;
    wl_[v] = Fv[ilo];
    wr_[v] = Fv[ihi];
    wll[v] = Fv[jm];
    whh[v] = Fv[jp];
  }
  double wlv[Ly::NVAR], wrv[Ly::NVAR], wlm[Ly::NVAR], whp[Ly::NVAR];
#pragma unroll
  for (int v = 0; v < Ly::NVAR; v++) { wlv[v] = wl_[v]; wrv[v] = wr_[v]; wlm[v] = wll[v]; whp[v] = whh[v]; }
  decode_alpha(wl_[0], wl_[3], wlv[0], wlv[3]);
  decode_alpha(wr_[0], wr_[3], wrv[0], wrv[3]);
  decode_alpha(wll[0], wll[3], wlm[0], wlm[3]);
  decode_alpha(whh[0], whh[3], whp[0], whp[3]);
  const double aL = wlv[0], aR = wrv[0];
  double aLL = 0., aRR = 0., tmp;
  if (!SAMELIM || S.thinc) {
    if (has_m) decode_alpha(Wn[ilo - sd], Wn[3 * nc + ilo - sd], aLL, tmp);
    if (has_p) decode_alpha(Wn[ihi + sd], Wn[3 * nc + ihi + sd], aRR, tmp);
  }
  bool ifl = false, ifr = false;
  if constexpr (!SAMELIM) {
    ifl = (aL >= 1.e-4) && (aL <= 1. - 1.e-4) && ((aR - aL) * (aL - aLL) >= 1.e-8);
    ifr = (aR >= 1.e-4) && (aR <= 1. - 1.e-4) && ((aL - aR) * (aR - aRR) >= 1.e-8);
  }
  ECO_GLOBAL const double *Gc = gmem(G.dcc[dir]);
  const double dcc = Gc[i];
  const double dcm = has_m ? Gc[i - 1] : 1.;
  const double dcp = has_p ? Gc[i + 1] : 1.;
  const double dL = gmem(G.dfl[dir])[i], dR = gmem(G.dfr[dir])[i];
  double xl[Ly::NVAR], xr[Ly::NVAR];
  muscl_dispatch<MODEL, NPH, DIM, SAMELIM, true>(S, wlv, wrv, wlm, whp, has_m, has_p, dcc, dcm, dcp, dL, dR, ifl, ifr,
                                                 xl, xr);
  CellState<NPH> L, R;
  vec_to_prim<MODEL, NPH, DIM>(xl, L);
  vec_to_prim<MODEL, NPH, DIM>(xr, R);
#pragma unroll
  for (int d = DIM; d < 3; d++) { L.u[d] = 0.; R.u[d] = 0.; }
#pragma unroll
  for (int k = 0; k < NPH; k++) { verify_correct<MODEL, NPH>(S, L, k); verify_correct<MODEL, NPH>(S, R, k); }
  if (S.thinc) {
    thinc<NPH>(L, aL, aR, aLL, false);
    thinc<NPH>(R, aR, aRR, aL, true);
  }
  to_face_frame(dir, 0, L.u);
  to_face_frame(dir, 0, R.u);
  double sL, sR;
  const float smx = speeds_f<NPH>(S, L, R, sL, sR);
  fulfill_energy<MODEL, NPH>(S, L);
  fulfill_energy<MODEL, NPH>(S, R);
  FluxVec<NPH> F;
  hllc_mixed<MODEL, NPH>(S, L, R, sL, sR, F);
  from_face_frame(dir, 0, F.q);
  put_flux_s<MODEL, NPH, DIM>(sF, stride, slot, F);
  return smx;
}

template <int MODEL, int NPH, int DIM, bool SAMELIM, int dir, int NT>
__device__ __forceinline__ void stage32_dir_3d(const StageArgs &A, const int *o, const long long *strd, int tid,
                                               bool active, int lx, int ly, int lz, const int *ci, long long cidx,
                                               double *sF, int stride, double (*sAcc)[NT], FluxVec<NPH> &acc,
                                               float &smx) {
  if constexpr (dir < DIM) {
    constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y, TZ = Tile<DIM>::Z;
    const int nfd = (TX + (dir == 0)) * (TY + (dir == 1)) * (TZ + (dir == 2));
    for (int f = tid; f < nfd; f += NT) smx = smaxf_(smx, face_fp32s<MODEL, NPH, DIM, SAMELIM, dir>(A, o, strd, f, sF, stride, f));
    __syncthreads();
    if (active) {
      if (dir > 0) unpark_acc<MODEL, NPH, DIM, NT>(sAcc, tid, acc);
      CellState<NPH> own;
      load_nc_coeffs_f<NPH>((ECO_GLOBAL const float *)gmem(A.wn), A.geo.ncell, cidx, own);
      accumulate_dir<MODEL, NPH, DIM, dir>(acc, sF, stride, 0, lx, ly, lz, ci, A.geo, own);
      if (dir < DIM - 1) park_acc<MODEL, NPH, DIM, NT>(sAcc, tid, acc);
    }
    __syncthreads();
  }
}

// Staged face pass of 2-D tiles in the FP32-storage form (one limiter form,
// no THINC): the stored 32-bit words of the tile and a two-cell halo in LDS,
// one FP32 limited slope per cell and direction, then the faces.  The face
// slopes, limiters and extrapolation increments are those of face_fp32s
// operation for operation (the limiters are symmetric in their arguments),
// so the fluxes are the same bit for bit as the direct pass's.
template <int MODEL, int NPH, int DIM>
struct StagedLds32 {
  using Ly = Layout<MODEL, NPH, DIM>;
  static constexpr int TX = Tile32S<DIM>::X, TY = Tile32S<DIM>::Y, NT = Tile32S<DIM>::T;
  static constexpr int SX = TX + 4, SY = TY + 4, NS = SX * SY;
  static constexpr int NGX = (TX + 2) * TY, NGY = TX * (TY + 2);
  static constexpr int NG = NGX > NGY ? NGX : NGY;
  static constexpr int NF0 = (TX + 1) * TY, NF1 = TX * (TY + 1);
  static constexpr int NFM = NF0 > NF1 ? NF0 : NF1;
  static_assert(NFM <= NT, "staged faces: at most one face per thread and direction");
  float S[Ly::NVAR][NS];
  // the limited slopes of one direction, then (after every thread has read
  // them) the fluxes of its faces
  union {
    float G[Ly::NVAR][NG];
    double F[Ly::NFV][NFM];
  };
  double sRed[NT / 64];
};

// decoded FP64 values of one staged cell (the volume-fraction pair decoded,
// the other words widened)
template <int MODEL, int NPH, int DIM>
__device__ __forceinline__ void staged32_vec(const StagedLds32<MODEL, NPH, DIM> &L, int c, double *w) {
  using Ly = Layout<MODEL, NPH, DIM>;
#pragma unroll
  for (int v = 0; v < Ly::NVAR; v++) w[v] = (double)L.S[v][c];
  decode_alpha(L.S[0][c], L.S[3][c], w[0], w[3]);
}

template <int MODEL, int NPH, int DIM, typename PW>
__device__ __forceinline__ void stage32_load_state(StagedLds32<MODEL, NPH, DIM> &L, PW W, long long nc, const int *o,
                                                   int nx, int ny, int tid) {
  using LD = StagedLds32<MODEL, NPH, DIM>;
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
    float val[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) val[it] = W[(long long)v * nc + idx[it]];
#pragma unroll
    for (int it = 0; it < NIT; it++)
      if (tid + it * NT < NS) L.S[v][tid + it * NT] = val[it];
  }
}

template <int MODEL, int NPH, int DIM, int dir, int LV, int LA>
__device__ __forceinline__ void stage32_slopes(StagedLds32<MODEL, NPH, DIM> &L, const SchemeParams &S,
                                               const Geometry &G, const int *o, int tid) {
  using LD = StagedLds32<MODEL, NPH, DIM>;
  using Ly = Layout<MODEL, NPH, DIM>;
  constexpr int TX = LD::TX, SX = LD::SX, NT = LD::NT;
  constexpr int NGD = (dir == 0) ? LD::NGX : LD::NGY;
  constexpr int RW = (dir == 0) ? TX + 2 : TX;
  constexpr int SD = (dir == 0) ? 1 : SX;
  ECO_GLOBAL const double *Gc = gmem(G.dcc[dir]);
  const int ndir = G.n[dir];
  for (int c = tid; c < NGD; c += NT) {
    const int cx = (dir == 0) ? c % RW - 1 : c % RW;
    const int cy = (dir == 0) ? c / RW : c / RW - 1;
    const int gi = o[dir] + (dir == 0 ? cx : cy);
    const bool has_m = (gi >= 1) && (gi <= ndir - 1), has_p = (gi >= 0) && (gi + 1 <= ndir - 1);
    const float idm = __builtin_amdgcn_rcpf((float)(has_m ? Gc[gi - 1] : 1.));
    const float idp = __builtin_amdgcn_rcpf((float)(has_p ? Gc[gi] : 1.));
    const int s0 = (cx + 2) + (cy + 2) * SX;
    // differences of the FP32-stored variables in FP32 (exact operands, one
    // rounding, as the FP64 difference rounded to FP32); of the volume
    // fractions in FP64 from their exact pairs
    static_assert(NPH == 2, "FP32 storage: two phases");
    double am[2], ac[2], ap[2];
    decode_alpha(L.S[0][s0 - SD], L.S[3][s0 - SD], am[0], am[1]);
    decode_alpha(L.S[0][s0], L.S[3][s0], ac[0], ac[1]);
    decode_alpha(L.S[0][s0 + SD], L.S[3][s0 + SD], ap[0], ap[1]);
#pragma unroll
    for (int v = 0; v < Ly::NVAR; v++) {
      float dm, dp;
      if (v == 0 || v == 3) {
        dm = (float)(ac[v / 3] - am[v / 3]);
        dp = (float)(ap[v / 3] - ac[v / 3]);
      } else {
        const float c = L.S[v][s0];
        dm = c - L.S[v][s0 - SD];
        dp = L.S[v][s0 + SD] - c;
      }
      const float sm = has_m ? dm * idm : 0.f;
      const float sp = has_p ? dp * idp : 0.f;
      const bool isa = (v < 3 * NPH) && (v % 3 == 0);
      float g;
      if constexpr (LV >= 0) {
        g = isa ? limiter_kt<LA, float>(sp, sm) : limiter_kt<LV, float>(sp, sm);
      } else {
        g = limiter_f(isa ? S.lim[2] : S.lim[0], sp, sm);
      }
      L.G[v][c] = g;
    }
  }
}

template <int MODEL, int NPH, int DIM, int dir>
__device__ __forceinline__ void stage32_slopes_dispatch(StagedLds32<MODEL, NPH, DIM> &L, const SchemeParams &S,
                                                        const Geometry &G, const int *o, int tid) {
  const int lv = S.lim[0], la = S.lim[2];
  if (lv == la) {
    switch (lv) {
      case LIM_MINMOD: stage32_slopes<MODEL, NPH, DIM, dir, LIM_MINMOD, LIM_MINMOD>(L, S, G, o, tid); return;
      case LIM_VANLEER: stage32_slopes<MODEL, NPH, DIM, dir, LIM_VANLEER, LIM_VANLEER>(L, S, G, o, tid); return;
      case LIM_VANALBADA: stage32_slopes<MODEL, NPH, DIM, dir, LIM_VANALBADA, LIM_VANALBADA>(L, S, G, o, tid); return;
      case LIM_MC: stage32_slopes<MODEL, NPH, DIM, dir, LIM_MC, LIM_MC>(L, S, G, o, tid); return;
      case LIM_SUPERBEE: stage32_slopes<MODEL, NPH, DIM, dir, LIM_SUPERBEE, LIM_SUPERBEE>(L, S, G, o, tid); return;
      default: break;
    }
  } else if (la == LIM_MINMOD && lv == LIM_VANLEER) {
    stage32_slopes<MODEL, NPH, DIM, dir, LIM_VANLEER, LIM_MINMOD>(L, S, G, o, tid);
    return;
  }
  stage32_slopes<MODEL, NPH, DIM, dir, -1, -1>(L, S, G, o, tid);
}

// Thread f computes face f of direction dir; after a barrier (every thread
// has read the slopes, whose LDS block the fluxes take over) it writes the
// flux to slot f.  Every thread of the workgroup calls this.
template <int MODEL, int NPH, int DIM, int dir>
__device__ __forceinline__ float stage32_faces(StagedLds32<MODEL, NPH, DIM> &L, const SchemeParams &S,
                                               const Geometry &G, const int *o, int tid) {
  using LD = StagedLds32<MODEL, NPH, DIM>;
  using Ly = Layout<MODEL, NPH, DIM>;
  constexpr int TX = LD::TX, SX = LD::SX;
  constexpr int NFD = (dir == 0) ? LD::NF0 : LD::NF1;
  constexpr int RW = (dir == 0) ? TX + 1 : TX;
  constexpr int GW = (dir == 0) ? TX + 2 : TX;
  constexpr int SD = (dir == 0) ? 1 : SX, GD = (dir == 0) ? 1 : TX;
  const int nx = G.n[0], ny = G.n[1], ndir = G.n[dir];
  ECO_GLOBAL const double *Dl = gmem(G.dfl[dir]);
  ECO_GLOBAL const double *Dr = gmem(G.dfr[dir]);
  float smx = 0.f;
  FluxVec<NPH> F;
  bool have = false;
  const int f = tid;
  const int fx = f % RW, fy = f / RW;
  const int g0 = o[0] + fx, g1 = o[1] + fy;
  const int gi = (dir == 0) ? g0 : g1;
  if (f < NFD && g0 < nx + (dir == 0) && g1 < ny + (dir == 1) && gi >= 1 && gi <= ndir - 1) {
    const int lx = fx - (dir == 0), ly = fy - (dir == 1);
    const int sl = (lx + 2) + (ly + 2) * SX, sr = sl + SD;
    const int gl = (dir == 0) ? (lx + 1) + ly * GW : lx + (ly + 1) * GW, gr = gl + GD;
    const float fL = (float)Dl[gi - 1], fR = (float)Dr[gi - 1];
    // the FP32-stored variables extrapolated in FP32 (one rounding of the
    // face value, below the storage's own); the volume fractions in FP64
    // from their exact pairs
    double xl[Ly::NVAR], xr[Ly::NVAR];
    double al[2], ar[2];
    decode_alpha(L.S[0][sl], L.S[3][sl], al[0], al[1]);
    decode_alpha(L.S[0][sr], L.S[3][sr], ar[0], ar[1]);
#pragma unroll
    for (int v = 0; v < Ly::NVAR; v++) {
      if (v == 0 || v == 3) {
        xl[v] = al[v / 3] + (double)(L.G[v][gl] * fL);
        xr[v] = ar[v / 3] + (double)(-L.G[v][gr] * fR);
      } else {
        xl[v] = (double)(L.S[v][sl] + L.G[v][gl] * fL);
        xr[v] = (double)(L.S[v][sr] + (-L.G[v][gr]) * fR);
      }
    }
    CellState<NPH> Lc, Rc;
    vec_to_prim<MODEL, NPH, DIM>(xl, Lc);
    vec_to_prim<MODEL, NPH, DIM>(xr, Rc);
#pragma unroll
    for (int d = DIM; d < 3; d++) { Lc.u[d] = 0.; Rc.u[d] = 0.; }
#pragma unroll
    for (int k = 0; k < NPH; k++) { verify_correct<MODEL, NPH>(S, Lc, k); verify_correct<MODEL, NPH>(S, Rc, k); }
    to_face_frame(dir, 0, Lc.u);
    to_face_frame(dir, 0, Rc.u);
    double sL, sR;
    smx = smaxf_(smx, speeds_f<NPH>(S, Lc, Rc, sL, sR));
    fulfill_energy<MODEL, NPH>(S, Lc);
    fulfill_energy<MODEL, NPH>(S, Rc);
    hllc_mixed<MODEL, NPH>(S, Lc, Rc, sL, sR, F);
    from_face_frame(dir, 0, F.q);
    have = true;
  }
  __syncthreads();  // the slopes are read: their block takes the fluxes
  if (have) put_flux_s<MODEL, NPH, DIM>(&L.F[0][0], LD::NFM, f, F);
  return smx;
}

template <int MODEL, int NPH, int DIM, bool STAGED>
using Stage32LdsT = typename std::conditional<STAGED, StagedLds32<MODEL, NPH, DIM>, StageLds<MODEL, NPH, DIM>>::type;

// Tile `blk` of a MUSCL stage in the FP32-storage form (STAGE 1 predictor,
// 2 corrector), launch slot `slot`.
template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM, bool STAGED = false>
__device__ __forceinline__ void stage32_tile(const StageArgs &A, const int slot, const int blk,
                                             Stage32LdsT<MODEL, NPH, DIM, STAGED> &lds) {
  static_assert(!STAGED || (DIM == 2 && SAMELIM), "staged pass: 2-D, one limiter form");
  using Ly = Layout<MODEL, NPH, DIM>;
  using LD = StageLds<MODEL, NPH, DIM>;
  using TL = typename std::conditional<STAGED, Tile32S<DIM>, Tile<DIM>>::type;  // this kernel's tile
  constexpr int TX = TL::X, TY = TL::Y, TZ = TL::Z;
  constexpr int NT = LD::NT;
  constexpr bool MERGED = LD::MERGED;
  constexpr int NF0 = LD::NF0, NF1 = LD::NF1;
  constexpr int NFT = LD::NFT;
  auto &sRed = lds.sRed;

  const Geometry &G = A.geo;
  const SchemeParams &S = A.sch;
  Control *ctl = A.ctl;
  const int tid = threadIdx.x;
  const bool first_kernel = (STAGE == 1);
  const bool last_kernel = (STAGE == 2);

  StepInfo si;
  if (!step_begin(ctl, G, S.cfl, slot, blk, first_kernel, last_kernel, si)) return;
  const double dt = si.dt;

  const long long nc = G.ncell;
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
  const int tgx = (nx + TX - 1) / TX, tgy = (ny + TY - 1) / TY;
  int b = blk;
  const int bx = b % tgx; b /= tgx;
  const int by = b % tgy; b /= tgy;
  const int bz = b;
  const int o[3] = {bx * TX, by * TY, bz * TZ};
  const int lx = tid % TX, ly = (tid / TX) % TY, lz = tid / (TX * TY);
  const int ci[3] = {o[0] + lx, o[1] + ly, o[2] + lz};
  const bool active = tid < TL::C && ci[0] < nx && ci[1] < ny && ci[2] < nz;
  const long long strd[3] = {1, (long long)nx, (long long)nx * ny};
  const long long cidx = active ? ci[0] + strd[1] * ci[1] + strd[2] * ci[2] : 0;

  ECO_GLOBAL const float *__restrict__ Wsrc = (ECO_GLOBAL const float *)gmem(A.wsrc);
  ECO_GLOBAL const float *__restrict__ Wn = (ECO_GLOBAL const float *)gmem(A.wn);

  FluxVec<NPH> acc;
  zero_flux(acc);
  float smx = 0.f;

  if constexpr (STAGED) {
    using SL = StagedLds32<MODEL, NPH, DIM>;
    stage32_load_state<MODEL, NPH, DIM>(lds, Wsrc, nc, o, nx, ny, tid);
    __syncthreads();
    stage32_slopes_dispatch<MODEL, NPH, DIM, 0>(lds, S, G, o, tid);
    __syncthreads();
    smx = smaxf_(smx, stage32_faces<MODEL, NPH, DIM, 0>(lds, S, G, o, tid));
    __syncthreads();
    CellState<NPH> ownc;
    if (active) {
      load_nc_coeffs_f<NPH>(Wn, nc, cidx, ownc);
      accumulate_dir<MODEL, NPH, DIM, 0>(acc, &lds.F[0][0], SL::NFM, 0, lx, ly, lz, ci, G, ownc);
    }
    __syncthreads();  // the x fluxes are read: the block takes the y slopes
    stage32_slopes_dispatch<MODEL, NPH, DIM, 1>(lds, S, G, o, tid);
    __syncthreads();
    smx = smaxf_(smx, stage32_faces<MODEL, NPH, DIM, 1>(lds, S, G, o, tid));
    __syncthreads();
    if (active) accumulate_dir<MODEL, NPH, DIM, 1>(acc, &lds.F[0][0], SL::NFM, 0, lx, ly, lz, ci, G, ownc);
  } else if constexpr (MERGED) {
    auto &sF = lds.sF;
    // all faces of the tile (both directions) in one pass, one barrier
    for (int w = tid; w < NF0 + NF1; w += NT) {
      if (w < NF0) smx = smaxf_(smx, face_fp32s<MODEL, NPH, DIM, SAMELIM, 0>(A, o, strd, w, &sF[0][0], NFT, w));
      else if constexpr (DIM > 1)
        smx = smaxf_(smx, face_fp32s<MODEL, NPH, DIM, SAMELIM, 1>(A, o, strd, w - NF0, &sF[0][0], NFT, w));
    }
    __syncthreads();
    if (active) {
      CellState<NPH> own;
      load_nc_coeffs_f<NPH>(Wn, nc, cidx, own);
      accumulate_dir<MODEL, NPH, DIM, 0>(acc, &sF[0][0], NFT, 0, lx, ly, lz, ci, G, own);
      if constexpr (DIM > 1) accumulate_dir<MODEL, NPH, DIM, 1>(acc, &sF[0][0], NFT, NF0, lx, ly, lz, ci, G, own);
    }
  } else {
    auto &sF = lds.sF;
    auto &sAcc = lds.sAcc;
    stage32_dir_3d<MODEL, NPH, DIM, SAMELIM, 0>(A, o, strd, tid, active, lx, ly, lz, ci, cidx, &sF[0][0], NFT, sAcc, acc, smx);
    stage32_dir_3d<MODEL, NPH, DIM, SAMELIM, 1>(A, o, strd, tid, active, lx, ly, lz, ci, cidx, &sF[0][0], NFT, sAcc, acc, smx);
    stage32_dir_3d<MODEL, NPH, DIM, SAMELIM, 2>(A, o, strd, tid, active, lx, ly, lz, ci, cidx, &sF[0][0], NFT, sAcc, acc, smx);
  }

  if (active) {
    CellState<NPH> own;
    load_prim_f<MODEL, NPH, DIM>(Wn, nc, cidx, own);
    const double rhoE = (double)Wn[Ly::IE * nc + cidx] + energy_offset<NPH>(S, own);
    // transmissive boundary faces on the stage state
#pragma unroll
    for (int dir = 0; dir < DIM; dir++) {
#pragma unroll
      for (int side = 0; side < 2; side++) {
        const bool on = side == 0 ? (ci[dir] == 0) : (ci[dir] == G.n[dir] - 1);
        if (!on) continue;
        CellState<NPH> L;
        load_prim_f<MODEL, NPH, DIM>(Wsrc, nc, cidx, L);
        to_face_frame(dir, side == 0, L.u);
        double sL, sR;
        smx = smaxf_(smx, speeds_f<NPH>(S, L, L, sL, sR));
        fulfill_energy<MODEL, NPH>(S, L);
        FluxVec<NPH> F;
        hllc_mixed<MODEL, NPH>(S, L, L, sL, sR, F);
        from_face_frame(dir, side == 0, F.q);
        add_face<MODEL, NPH>(acc, F, G.coefA[dir], false, own);
      }
    }
    // cell update in FP64
    CellState<NPH> sn = own;
    sn.rho = 0.;
#pragma unroll
    for (int k = 0; k < NPH; k++) sn.rho += sn.a[k] * sn.r[k];
    sn.E = rhoE * rcp_fast(sn.rho);
    FluxVec<NPH> buf;
    build_cons_fast<MODEL, NPH>(S, sn, buf);
    buf.E = rhoE;
    scale_add<MODEL, NPH>(acc, (STAGE == 1) ? 0.5 * dt : dt, buf);
    scheme_correction<MODEL, NPH>(acc, buf);
    const double rhoE_new = acc.E;
    CellState<NPH> s1;
    build_prim<MODEL, NPH, true>(S, acc, s1);
    bool ok = true;
    if (S.relax) ok = relax_p_fast<MODEL, NPH>(S, s1);
    correction_energy<MODEL, NPH, true>(S, s1);
    if (!ok) atomicOr(&ctl->err, 1u);
    bool bad = false;
#pragma unroll
    for (int k = 0; k < NPH; k++) bad |= !(s1.p[k] == s1.p[k]) || !(s1.r[k] == s1.r[k]);
    if (bad) atomicOr(&ctl->err, 2u);
    ECO_GLOBAL float *W = (ECO_GLOBAL float *)gmem(A.wout);
    store_prim_f<MODEL, NPH, DIM>(W, nc, cidx, s1);
    if constexpr (STAGE == 2) {
      CellState<NPH> st = s1;
      round_state<NPH>(st);
      W[Ly::IE * nc + cidx] = (float)(rhoE_new - energy_offset<NPH>(S, st));
    }
  }

  float m = smx;
  for (int off = 32; off > 0; off >>= 1) m = smaxf_(m, __shfl_xor(m, off));
  if ((tid & 63) == 0) sRed[tid >> 6] = (double)m;
  __syncthreads();
  if (tid == 0) {
    double mm = sRed[0];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) mm = smax(mm, sRed[w]);
    step_end(ctl, mm, slot, blk, last_kernel, si);
  }
}

template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM, bool STAGED = false>
__global__ void __launch_bounds__(Tile<DIM>::T, STAGED ? Tile<DIM>::WS32 : Tile<DIM>::W)
k_stage32(const StageArgs A) {
  __shared__ Stage32LdsT<MODEL, NPH, DIM, STAGED> lds;
  stage32_tile<MODEL, NPH, DIM, STAGE, SAMELIM, STAGED>(A, A.step, xcd_band_tile<DIM>(blockIdx.x, gridDim.x), lds);
}

// persistent form over a pool (as k_stage_batch)
template <int MODEL, int NPH, int DIM, int STAGE, bool SAMELIM>
__global__ void __launch_bounds__(Tile<DIM>::T, Tile<DIM>::WB32)
k_stage32_batch(const StageArgs *__restrict__ args, const TileRef *__restrict__ tiles,
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
    stage32_tile<MODEL, NPH, DIM, STAGE, SAMELIM>(A, slot, t.blk, lds);
    __syncthreads();  // LDS reuse by the next tile
  }
}

// FP64 device state (nprim fields + E) -> FP32 storage (nprim fields + R)
template <int MODEL, int NPH, int DIM>
__global__ void k_to_f32(const StageArgs A, const double *__restrict__ src) {
  using Ly = Layout<MODEL, NPH, DIM>;
  const long long nc = A.geo.ncell;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nc) return;
  CellState<NPH> s;
  load_prim<MODEL, NPH, DIM>(src, nc, i, s);
  double rho = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) rho += s.a[k] * s.r[k];
  const double rhoE = rho * src[Ly::IE * nc + i];
  float *W = (float *)A.wout;
  store_prim_f<MODEL, NPH, DIM>(W, nc, i, s);
  round_state<NPH>(s);
  W[Ly::IE * nc + i] = (float)(rhoE - energy_offset<NPH>(A.sch, s));
}

// FP32 storage -> FP64 device layout with the specific total energy E
template <int MODEL, int NPH, int DIM>
__global__ void k_from_f32(const StageArgs A, double *__restrict__ dst) {
  using Ly = Layout<MODEL, NPH, DIM>;
  const long long nc = A.geo.ncell;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const float *W = (const float *)A.wn;
  CellState<NPH> s;
  load_prim_f<MODEL, NPH, DIM>(W, nc, i, s);
  double rho = 0.;
#pragma unroll
  for (int k = 0; k < NPH; k++) rho += s.a[k] * s.r[k];
  const double rhoE = (double)W[Ly::IE * nc + i] + energy_offset<NPH>(A.sch, s);
  store_prim<MODEL, NPH, DIM>(dst, nc, i, s);
  dst[Ly::IE * nc + i] = rhoE / rho;
}

template <int MODEL, int NPH, int DIM>
int launch_t32(const StageArgs &A, int stage, hipStream_t st, double *tmp) {
  constexpr int NT = Tile<DIM>::T;
  const int tgx = (A.geo.n[0] + Tile<DIM>::X - 1) / Tile<DIM>::X;
  const int tgy = (A.geo.n[1] + Tile<DIM>::Y - 1) / Tile<DIM>::Y;
  const int tgz = (A.geo.n[2] + Tile<DIM>::Z - 1) / Tile<DIM>::Z;
  dim3 grid((unsigned)(tgx * tgy * tgz)), block(NT);
  const bool same = A.sch.same_lim != 0;
  const long long nc = A.geo.ncell;
  const dim3 g1((unsigned)((nc + 255) / 256)), b1(256);
  // 2-D with one limiter form and no THINC: the staged face pass
  if constexpr (DIM == 2) {
    if (same && !A.sch.thinc && ECO_STAGED2 != 0 && (stage == 1 || stage == 2)) {
      // the staged kernel's own tile
      const int sgx = (A.geo.n[0] + Tile32S<DIM>::X - 1) / Tile32S<DIM>::X;
      const int sgy = (A.geo.n[1] + Tile32S<DIM>::Y - 1) / Tile32S<DIM>::Y;
      const dim3 sgrid((unsigned)(sgx * sgy));
      if (stage == 1) hipLaunchKernelGGL((k_stage32<MODEL, NPH, DIM, 1, true, true>), sgrid, block, 0, st, A);
      else hipLaunchKernelGGL((k_stage32<MODEL, NPH, DIM, 2, true, true>), sgrid, block, 0, st, A);
      return (int)hipGetLastError();
    }
  }
  switch (stage) {
    case 1:
      if (same) hipLaunchKernelGGL((k_stage32<MODEL, NPH, DIM, 1, true>), grid, block, 0, st, A);
      else hipLaunchKernelGGL((k_stage32<MODEL, NPH, DIM, 1, false>), grid, block, 0, st, A);
      break;
    case 2:
      if (same) hipLaunchKernelGGL((k_stage32<MODEL, NPH, DIM, 2, true>), grid, block, 0, st, A);
      else hipLaunchKernelGGL((k_stage32<MODEL, NPH, DIM, 2, false>), grid, block, 0, st, A);
      break;
    case 4: hipLaunchKernelGGL((k_to_f32<MODEL, NPH, DIM>), g1, b1, 0, st, A, (const double *)tmp); break;
    case 5: hipLaunchKernelGGL((k_from_f32<MODEL, NPH, DIM>), g1, b1, 0, st, A, tmp); break;
    default: return -1;
  }
  return (int)hipGetLastError();
}

template <int MODEL, int NPH>
int launch_d32(const StageArgs &A, int stage, hipStream_t st, double *tmp) {
  switch (A.geo.dim) {
    case 1: return launch_t32<MODEL, NPH, 1>(A, stage, st, tmp);
    case 2: return launch_t32<MODEL, NPH, 2>(A, stage, st, tmp);
    case 3: return launch_t32<MODEL, NPH, 3>(A, stage, st, tmp);
  }
  return -1;
}

template <int MODEL, int NPH, int DIM, int STAGE, bool SAME>
int launch_tb32_1(const BatchLaunch &L, hipStream_t st) {
  constexpr int NT = Tile<DIM>::T;
  if (L.occupancy)
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(L.occupancy,
                                                             k_stage32_batch<MODEL, NPH, DIM, STAGE, SAME>, NT, 0);
  hipLaunchKernelGGL((k_stage32_batch<MODEL, NPH, DIM, STAGE, SAME>), dim3((unsigned)L.grid), dim3(NT), 0, st, L.args,
                     L.tiles, L.ntiles, L.slot);
  return (int)hipGetLastError();
}

template <int MODEL, int NPH, int DIM>
int launch_tb32(const BatchLaunch &L, hipStream_t st) {
  switch (L.stage) {
    case 1:
      return L.same ? launch_tb32_1<MODEL, NPH, DIM, 1, true>(L, st) : launch_tb32_1<MODEL, NPH, DIM, 1, false>(L, st);
    case 2:
      return L.same ? launch_tb32_1<MODEL, NPH, DIM, 2, true>(L, st) : launch_tb32_1<MODEL, NPH, DIM, 2, false>(L, st);
  }
  return -1;
}

template <int MODEL, int NPH>
int launch_db32(const BatchLaunch &L, hipStream_t st) {
  switch (L.dim) {
    case 1: return launch_tb32<MODEL, NPH, 1>(L, st);
    case 2: return launch_tb32<MODEL, NPH, 2>(L, st);
    case 3: return launch_tb32<MODEL, NPH, 3>(L, st);
  }
  return -1;
}

}  // namespace eco
