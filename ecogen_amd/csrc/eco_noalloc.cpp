// Allocation-free form of a single solve (eco_run_noalloc, include/ecogen.h).
//
// The same launches as eco_create + eco_set_state_device + eco_run with the
// state read back, for callers that cannot allocate device memory.  Every
// buffer beyond the caller's input and output lives in `scratch`, a static
// device array of the library: the control block, the per-direction distance
// tables, the two time levels and the MUSCL intermediate state.  The tables
// are computed on the device from the same expressions the host lowering
// uses, the control block is set by a one-thread kernel and the steps are
// direct launches, so the entry point calls no hipMalloc, hipFree, graph or
// host-to-device copy.
//
// Each GPU holds its own copy of `scratch` until the process ends.  Its
// device address is looked up once per device, with that device current,
// and kept; host code passes that address, never the array's name, to the
// kernels and copies.  Calls on one GPU share the array: calls that overlap
// in time on one GPU race, so a caller runs them one at a time (each call
// synchronises its stream before it returns).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>

#include "../../include/ecogen.h"
#include "eco_host.h"
#include "eco_launch.h"

using namespace eco;

static __device__ __align__(256) double scratch[ECO_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` on each device (null: not looked up yet)

// The device address of this device's scratch, looked up on its first use
// with the device current.
int scratch_ptr(void **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ECO_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return ECO_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ECO_ERR_HIP;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return ECO_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Distance tables of lower_config (cell-centre and face positions as the
// reference builds them, MeshCartesian.cpp:180-197, 300-397), one thread per
// entry, the same operations in the same order.
__global__ void k_tables(double *tab, int n0, int n1, int n2, double L0, double L1, double L2) {
  const int n[3] = {n0, n1, n2};
  const double Ls[3] = {L0, L1, L2};
  long long o = 0;
  for (int d = 0; d < 3; d++) {
    const int nd = n[d];
    const double dd = Ls[d] / (double)nd;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += gridDim.x * blockDim.x) {
      double vc = 0., va = 0., vb = 0.;
      if (i + 1 < nd) {
        const double xi = (i + 0.5) * dd, xj = (i + 1 + 0.5) * dd;
        const double xf = xi + 0.5 * dd;
        const double v = xi - xj;
        vc = sqrt(v * v);
        const double a = xf - xi, b = xf - xj;
        va = sqrt(a * a);
        vb = sqrt(b * b);
      }
      tab[o + i] = vc;
      tab[o + nd + i] = va;
      tab[o + 2 * nd + i] = vb;
    }
    o += 3 * (long long)nd;
  }
}

// Control block of a fresh solve (reset_control and set_target of eco_capi.cpp).
__global__ void k_control(Control *ctl, int max_steps, double t_final) {
  Control h = {};
  h.t_final = t_final;
  h.max_steps = max_steps;
  *ctl = h;
}

__global__ void k_fill(double *p, long long n, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

int launch_fill(double *p, long long n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, st, p, n, v);
  return hipGetLastError() == hipSuccess ? ECO_OK : ECO_ERR_HIP;
}

}  // namespace

extern "C" unsigned long long eco_noalloc_bytes(const eco_config *cfg) {
  if (validate_config(cfg) != ECO_OK) return 0;
  // configurations eco_run_noalloc refuses
  if (cfg->precision != ECO_FP64 || cfg->ntransport > 0 || cfg->model == ECO_MODEL_THERMALEQ ||
      cfg->model == ECO_MODEL_EULERHOM || cfg->relax_ptmu) return 0;
  const eco_config &c = *cfg;
  int dim = (c.n[0] > 1) + (c.n[1] > 1) + (c.n[2] > 1);
  const size_t nc = (size_t)c.n[0] * c.n[1] * c.n[2];
  const int nprim = (c.model == ECO_MODEL_EULER) ? 2 + dim : 3 * c.nphase + dim;
  const int ncache = (c.model == ECO_MODEL_MULTIP && (c.relax_p || c.relax_pt)) ? (c.order == 1 ? c.nphase + 3 : 1) : 0;
  const size_t nlevel = (size_t)nprim + 1 + ncache;
  size_t ntab = 0;
  for (int d = 0; d < 3; d++) ntab += 3 * (size_t)std::max(c.n[d], 1);
  size_t b = align256(sizeof(Control)) + align256(ntab * sizeof(double));
  b += 2 * align256(nc * nlevel * sizeof(double));
  if (c.order == 2) b += align256(nc * (size_t)nprim * sizeof(double));
  return b;
}

extern "C" int eco_run_noalloc(const eco_config *cfg, const double *state_in, double *state_out, int nsteps,
                               double t_final, void *hip_stream, eco_status *status) {
  // everything checked on the host before the first HIP call
  int rc = validate_config(cfg);
  if (rc) return rc;
  if (!state_in || !state_out || state_in == state_out) return ECO_ERR_INVALID;
  if (cfg->precision != ECO_FP64) return ECO_ERR_UNSUPPORTED;
  // transported scalars need a transport pass and face sM arrays: contexts only
  if (cfg->ntransport > 0) return ECO_ERR_UNSUPPORTED;
  // the thermal-equilibrium model keeps its own state layout: contexts only
  if (cfg->model == ECO_MODEL_THERMALEQ || cfg->model == ECO_MODEL_EULERHOM || cfg->relax_ptmu) return ECO_ERR_UNSUPPORTED;
  if (nsteps <= 0 && !(t_final > 0.)) return ECO_ERR_INVALID;
  if (nsteps == INT_MIN) return ECO_ERR_INVALID;
  const unsigned long long need = eco_noalloc_bytes(cfg);
  if (need == 0) return ECO_ERR_INVALID;
  if (need > (unsigned long long)ECO_NOALLOC_SCRATCH_BYTES) return ECO_ERR_SCRATCH;

  const eco_config &c = *cfg;
  Lowered L;
  lower_config(c, L, /*with_tables=*/false);
  hipStream_t st = (hipStream_t)hip_stream;
  void *base = nullptr;
  rc = scratch_ptr(&base);
  if (rc) return rc;

  // scratch layout: control | tables | level 0 | level 1 | U*
  const size_t nc = (size_t)L.geo.ncell;
  char *b = (char *)base;
  Control *ctl = (Control *)b;
  b += align256(sizeof(Control));
  double *tables = (double *)b;
  size_t ntab = 0;
  for (int d = 0; d < 3; d++) ntab += 3 * (size_t)std::max(c.n[d], 1);
  b += align256(ntab * sizeof(double));
  double *lev[2];
  lev[0] = (double *)b;
  b += align256(nc * L.nlevel * sizeof(double));
  lev[1] = (double *)b;
  b += align256(nc * L.nlevel * sizeof(double));
  double *bstar = (c.order == 2) ? (double *)b : nullptr;
  bind_tables(c, L, tables);

  hipLaunchKernelGGL(k_tables, dim3(64), dim3(256), 0, st, tables, c.n[0], c.n[1], c.n[2], c.L[0], c.L[1], c.L[2]);
  if (hipGetLastError() != hipSuccess) return ECO_ERR_HIP;
  // iteration control: exactly nsteps; time control: up to t_final, with at
  // most -nsteps steps when nsteps < 0 (as eco_run)
  const int nlaunch = nsteps > 0 ? nsteps : 0;
  hipLaunchKernelGGL(k_control, dim3(1), dim3(1), 0, st, ctl, nsteps > 0 ? nsteps : -nsteps, nsteps > 0 ? -1. : t_final);
  if (hipGetLastError() != hipSuccess) return ECO_ERR_HIP;

  // input (host layout on the device) -> level 0 (device layout), then E
  // (and the cached mixture density) at t = 0 (k_init)
  int map[3 * kMaxPhases + 3];
  int nmap = 0;
  if (c.model == ECO_MODEL_EULER) {
    map[nmap++] = 1;
    map[nmap++] = 2;
  } else {
    for (int k = 0; k < c.nphase; k++) { map[nmap++] = 3 * k; map[nmap++] = 3 * k + 1; map[nmap++] = 3 * k + 2; }
  }
  const int unit0 = (c.model == ECO_MODEL_EULER) ? 3 : 3 * c.nphase;
  for (int d = 0; d < L.geo.dim; d++) map[nmap++] = unit0 + d;
  for (int f = 0; f < nmap; f++)
    if (hipMemcpyAsync(lev[0] + (size_t)f * nc, state_in + (size_t)map[f] * nc, nc * sizeof(double),
                       hipMemcpyDeviceToDevice, st) != hipSuccess)
      return ECO_ERR_HIP;
  StageArgs A;
  A.geo = L.geo;
  A.sch = L.sch;
  A.ctl = ctl;
  A.step = 0;
  A.wsrc = A.wn = lev[0];
  A.wout = lev[0];
  rc = launch_stage(A, 3, st);
  if (rc) return rc < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP;

  // steps: iteration control launches exactly nsteps; time control launches
  // chunks and reads the done flag between them (steps past the target are
  // no-ops on the device)
  auto launch_step = [&](int k) -> int {
    StageArgs S = A;
    S.step = k;
    S.wn = lev[k & 1];
    S.wsrc = S.wn;
    S.wout = lev[(k + 1) & 1];
    int r;
    if (c.order == 1) {
      r = launch_stage(S, 0, st);
    } else {
      StageArgs P = S;
      P.wout = bstar;
      r = launch_stage(P, 1, st);
      if (!r) {
        StageArgs Q = S;
        Q.wsrc = bstar;
        r = launch_stage(Q, 2, st);
      }
    }
    return r ? (r < 0 ? ECO_ERR_UNSUPPORTED : ECO_ERR_HIP) : ECO_OK;
  };
  Control h;
  int k = 0;
  if (nlaunch > 0) {
    for (; k < nlaunch; k++)
      if ((rc = launch_step(k))) return rc;
    if (hipMemcpyAsync(&h, ctl, sizeof(Control), hipMemcpyDeviceToHost, st) != hipSuccess) return ECO_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return ECO_ERR_HIP;
  } else {
    int chunk = 48;
    for (;;) {
      for (int i = 0; i < chunk; i++, k++)
        if ((rc = launch_step(k))) return rc;
      if (hipMemcpyAsync(&h, ctl, sizeof(Control), hipMemcpyDeviceToHost, st) != hipSuccess) return ECO_ERR_HIP;
      if (hipStreamSynchronize(st) != hipSuccess) return ECO_ERR_HIP;
      if (h.finished) break;
      chunk = std::min(chunk * 2, 192);
    }
  }

  // level of the last step taken -> output (host layout, E included)
  const double *src = lev[h.steps & 1];
  const int nf = ECO_NFIELDS(c.nphase);
  for (int f = 0; f < nf; f++) {
    bool mapped = (f == nf - 1);
    for (int g = 0; g < nmap; g++) mapped |= (map[g] == f);
    if (!mapped && (rc = launch_fill(state_out + (size_t)f * nc, (long long)nc,
                                     (c.model == ECO_MODEL_EULER && f == 0) ? 1. : 0., st)))
      return rc;
  }
  for (int f = 0; f < nmap; f++)
    if (hipMemcpyAsync(state_out + (size_t)map[f] * nc, src + (size_t)f * nc, nc * sizeof(double),
                       hipMemcpyDeviceToDevice, st) != hipSuccess)
      return ECO_ERR_HIP;
  if (hipMemcpyAsync(state_out + (size_t)(nf - 1) * nc, src + (size_t)L.nprim * nc, nc * sizeof(double),
                     hipMemcpyDeviceToDevice, st) != hipSuccess)
    return ECO_ERR_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) return ECO_ERR_HIP;
  if (status) fill_status(h, L.geo.lcfl, c.cfl, status);
  // a boundary solve that stopped the run: as eco_run
  return (h.err & ECO_WARN_INFLOW_NOT_CONVERGED) ? ECO_ERR_NOT_CONVERGED : ECO_OK;
}
