// Instantiation of the FP32-storage stage kernels for the multip model.
#include "eco_stage32.inl"

namespace eco {
int launch_multip32(const StageArgs &A, int stage, hipStream_t st, double *tmp) {
  return launch_d32<MODEL_MULTIP, 2>(A, stage, st, tmp);
}
int batch_multip32(const BatchLaunch &L, hipStream_t st) { return launch_db32<MODEL_MULTIP, 2>(L, st); }
}  // namespace eco
