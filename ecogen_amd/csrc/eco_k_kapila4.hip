// Instantiation of the stage kernels for the kapila model with four phases
// (FP64 storage, single solves).
#include "eco_stage.inl"

namespace eco {
int launch_kapila4(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_KAPILA, 4>(A, stage, st); }
}  // namespace eco
