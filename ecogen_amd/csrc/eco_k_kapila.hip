// Instantiation of the stage kernels for the kapila model.
#include "eco_stage.inl"

namespace eco {
int launch_kapila(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_KAPILA, 2>(A, stage, st); }
int batch_kapila(const BatchLaunch &L, hipStream_t st) { return launch_db<MODEL_KAPILA, 2>(L, st); }
}  // namespace eco
