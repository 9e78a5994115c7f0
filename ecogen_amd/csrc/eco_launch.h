// Launch interface between the C-ABI layer and the stage kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "eco_params.h"

namespace eco {

struct StageArgs {
  Geometry geo;
  SchemeParams sch;
  const double *wsrc;  // state the fluxes are built from (U^n, or U* in the corrector)
  const double *wn;    // time-n state
  double *wout;        // output state
  Control *ctl;
  int step;            // 0-based step index of the solve
};

// stage: 0 first-order step, 1 MUSCL predictor, 2 MUSCL corrector, 3 initial total energy.
// Returns a hipError_t value (0 on success) or -1 for an unsupported configuration.
int launch_stage(const StageArgs &A, int stage, hipStream_t stream);

}  // namespace eco
