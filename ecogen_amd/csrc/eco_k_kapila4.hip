// Instantiation of the stage kernels for the kapila model with four phases
// (FP64 storage, single solves and pools).
#include "eco_stage.inl"

namespace eco {
int launch_kapila4(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_KAPILA, 4>(A, stage, st); }
int batch_kapila4(const BatchLaunch &L, hipStream_t st) { return launch_db<MODEL_KAPILA, 4>(L, st); }
}  // namespace eco
