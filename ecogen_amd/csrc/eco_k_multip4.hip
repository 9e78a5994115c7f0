// Instantiation of the stage kernels for the multip model with four phases
// (FP64 storage, single solves).
#include "eco_stage.inl"

namespace eco {
int launch_multip4(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_MULTIP, 4>(A, stage, st); }
}  // namespace eco
