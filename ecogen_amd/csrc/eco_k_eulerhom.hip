// Instantiation of the stage kernels for the homogeneous liquid-vapour model
// (two phases, first order).
#include "eco_stage.inl"

namespace eco {
int launch_eulerhom(const StageArgs &A, int stage, hipStream_t st) { return launch_d_fo<MODEL_EULERHOM, 2>(A, stage, st); }
}  // namespace eco
