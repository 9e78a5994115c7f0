// Instantiation of the stage kernels for the kapila model.
#include "eco_stage.inl"

namespace eco {
int launch_kapila(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_KAPILA, 2>(A, stage, st); }
}  // namespace eco
