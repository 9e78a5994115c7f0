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
};

// stage: 0 first-order step, 1 MUSCL predictor, 2 MUSCL corrector, 3 initial total energy.
// Returns a hipError_t value (0 on success) or -1 for an unsupported configuration.
int launch_stage(const StageArgs &A, int stage, hipStream_t stream);

// FP32-storage conversions: to_f32 = 1: FP64 device layout in tmp -> A.wout;
// to_f32 = 0: A.wn -> FP64 device layout in tmp (E included).
int launch_convert(const StageArgs &A, int to_f32, double *tmp, hipStream_t stream);

}  // namespace eco
