// Instantiation of the stage kernels for the multip model with four phases
// (FP64 storage, single solves and pools).
#include "eco_stage.inl"

namespace eco {
int launch_multip4(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_MULTIP, 4>(A, stage, st); }
int batch_multip4(const BatchLaunch &L, hipStream_t st) { return launch_db<MODEL_MULTIP, 4>(L, st); }
}  // namespace eco
