// Instantiation of the stage kernels for the kapila model with three phases
// (FP64 storage, single solves and pools).
#include "eco_stage.inl"

namespace eco {
int launch_kapila3(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_KAPILA, 3>(A, stage, st); }
int batch_kapila3(const BatchLaunch &L, hipStream_t st) { return launch_db<MODEL_KAPILA, 3>(L, st); }
}  // namespace eco
