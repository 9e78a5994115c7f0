// Model dispatch of the stage kernels.
#include "eco_launch.h"

namespace eco {
int launch_euler(const StageArgs &A, int stage, hipStream_t st);
int launch_kapila(const StageArgs &A, int stage, hipStream_t st);
int launch_multip(const StageArgs &A, int stage, hipStream_t st);

int launch_stage(const StageArgs &A, int stage, hipStream_t st) {
  if (A.sch.model == MODEL_EULER) return launch_euler(A, stage, st);
  if (A.sch.nph != 2) return -1;
  if (A.sch.model == MODEL_KAPILA) return launch_kapila(A, stage, st);
  if (A.sch.model == MODEL_MULTIP) return launch_multip(A, stage, st);
  return -1;
}
}  // namespace eco
