// Instantiation of the stage kernels for the thermal-equilibrium model (two phases).
#include "eco_stage.inl"

namespace eco {
int launch_thermaleq(const StageArgs &A, int stage, hipStream_t st) { return launch_d<MODEL_THERMALEQ, 2>(A, stage, st); }
}  // namespace eco
