// Instantiation of the stage kernels for the euler model.
#include "eco_stage.inl"

namespace eco {
int launch_euler(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_EULER, 1>(A, stage, st); }
int batch_euler(const BatchLaunch &L, hipStream_t st) { return launch_db<MODEL_EULER, 1>(L, st); }
}  // namespace eco
