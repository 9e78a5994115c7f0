// Launch interface between the C-ABI layer and the stage kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "eco_params.h"

namespace eco {

// State pointers are typed double; in the FP32-storage form they point to
// float arrays and the FP32 kernels reinterpret them.
struct StageArgs {
  Geometry geo;
  SchemeParams sch;
  const double *wsrc;  // state the fluxes are built from (U^n, or U* in the corrector)
  const double *wn;    // time-n state
  double *wout;        // output state
  Control *ctl;
  int step;            // 0-based step index of the solve
  // with transports: per direction the contact velocity sM of every face of
  // the stage (faces of direction d indexed g0 + (n0 + (d==0)) * (g1 +
  // (n1 + (d==1)) * g2), g_d in [0, n_d]), written for the transport pass
  double *smf[3] = {nullptr, nullptr, nullptr};
  // RelaxationPTMu in MUSCL: the time level where the predictor writes the
  // time-n state it relaxes (the corrector's time-n state)
  double *wn2 = nullptr;
};

// Transport pass of one stage (src/Transport/Transport.cpp): after the flow
// kernel of the stage, from the faces' sM it wrote.  Scalars are stored
// field-major, ncell doubles per scalar.
struct TransportArgs {
  Geometry geo;
  double cfl;
  int ntr, order;
  int stage;           // 0 first order, 1 MUSCL predictor, 2 MUSCL corrector
  int lim_g, lim_i;    // global / interface volume-fraction limiters
  int bc_wall, bc_in;  // bit 2*dir+side: wall; outflow, injection or tank (imposed entering value)
  double bc_tr[6][4];  // entering values per side and scalar
  const double *asrc;  // phase-0 alpha of the stage state (interface detection)
  const double *an;    // phase-0 alpha at time n (cells across the opposite faces)
  const double *trs;   // scalars of the stage state (time n, or U*)
  const double *trn;   // scalars at time n (non-conservative terms, update base)
  double *tro;         // updated scalars
  const double *sm[3];
  Control *ctl;
  int step;
};
int launch_transport(const TransportArgs &T, hipStream_t stream);

// One tile of one case in a pool's persistent launch.
struct TileRef {
  int c;    // case slot in the class's argument table
  int blk;  // tile index within the case (x fastest)
};

// Device-resident descriptor of one kernel class of a pool: per case four
// argument blocks ([slot parity][first/second kernel of the step]), the tile
// table and its length.
struct BatchDesc {
  const StageArgs *args;
  const TileRef *tiles;
  int ntiles;
};

// Launch of a pool's persistent stage kernel for one class.  The tables are
// passed as kernel pointer arguments (device pointers): the backend then
// knows them, and the state pointers read from them, to be global memory.
struct BatchLaunch {
  int dim, stage, same;      // class within the model (stage 0, 1, 2 as launch_stage)
  const StageArgs *args;     // four per case (BatchDesc::args)
  const TileRef *tiles;      // BatchDesc::tiles
  const int *ntiles;         // &BatchDesc::ntiles on the device
  int slot, grid;
  int *occupancy;            // non-null: report workgroups per CU for this kernel, no launch
};
int launch_batch(int model, int prec, int nph, const BatchLaunch &L, hipStream_t stream);

// Tile shape (cells per workgroup along x, y, z) of the stage kernels for a dimension.
void tile_shape(int dim, int t[3]);

// stage: 0 first-order step, 1 MUSCL predictor, 2 MUSCL corrector, 3 initial total energy.
// Returns a hipError_t value (0 on success) or -1 for an unsupported configuration.
int launch_stage(const StageArgs &A, int stage, hipStream_t stream);

// FP32-storage conversions: to_f32 = 1: FP64 device layout in tmp -> A.wout;
// to_f32 = 0: A.wn -> FP64 device layout in tmp (E included).
int launch_convert(const StageArgs &A, int to_f32, double *tmp, hipStream_t stream);

}  // namespace eco
