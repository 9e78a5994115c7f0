// Instantiation of the stage kernels for the euler model.
#include "eco_stage.inl"

namespace eco {
int launch_euler(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_EULER, 1>(A, stage, st); }
}  // namespace eco
