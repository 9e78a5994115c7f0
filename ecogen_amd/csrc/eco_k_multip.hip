// Instantiation of the stage kernels for the multip model.
#include "eco_stage.inl"

namespace eco {
int launch_multip(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_MULTIP, 2>(A, stage, st); }
int batch_multip(const BatchLaunch &L, hipStream_t st) { return launch_db<MODEL_MULTIP, 2>(L, st); }
}  // namespace eco
