// Model dispatch of the stage kernels.
#include "eco_launch.h"
#include "eco_tuning.h"

namespace eco {
__global__ void k_copy(double *__restrict__ d, const double *__restrict__ s, unsigned long long n) {
  unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = s[i];
}

int launch_copy(void *dst, const void *src, unsigned long long nbytes, hipStream_t st) {
  const unsigned long long n = nbytes / 8;
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, (double *)dst, (const double *)src, n);
  return (int)hipGetLastError();
}

int launch_euler(const StageArgs &A, int stage, hipStream_t st);
int launch_kapila(const StageArgs &A, int stage, hipStream_t st);
int launch_multip(const StageArgs &A, int stage, hipStream_t st);
int launch_kapila3(const StageArgs &A, int stage, hipStream_t st);
int launch_multip3(const StageArgs &A, int stage, hipStream_t st);
int launch_kapila4(const StageArgs &A, int stage, hipStream_t st);
int launch_multip4(const StageArgs &A, int stage, hipStream_t st);
int launch_kapila32(const StageArgs &A, int stage, hipStream_t st, double *tmp);
int launch_multip32(const StageArgs &A, int stage, hipStream_t st, double *tmp);
int launch_thermaleq(const StageArgs &A, int stage, hipStream_t st);
int launch_eulerhom(const StageArgs &A, int stage, hipStream_t st);

int launch_stage(const StageArgs &A, int stage, hipStream_t st) {
  if (A.sch.model == MODEL_THERMALEQ) return (A.sch.nph == 2 && A.sch.prec == PREC_FP64) ? launch_thermaleq(A, stage, st) : -1;
  if (A.sch.model == MODEL_EULERHOM) return (A.sch.nph == 2 && A.sch.prec == PREC_FP64) ? launch_eulerhom(A, stage, st) : -1;
  if (A.sch.prec == PREC_FP32S && stage != 3) {
    if (A.sch.model == MODEL_KAPILA) return launch_kapila32(A, stage, st, nullptr);
    if (A.sch.model == MODEL_MULTIP) return launch_multip32(A, stage, st, nullptr);
    return -1;
  }
  if (A.sch.model == MODEL_EULER) return launch_euler(A, stage, st);
  if (A.sch.nph == 3) {
    if (A.sch.model == MODEL_KAPILA) return launch_kapila3(A, stage, st);
    if (A.sch.model == MODEL_MULTIP) return launch_multip3(A, stage, st);
    return -1;
  }
  if (A.sch.nph == 4) {
    if (A.sch.model == MODEL_KAPILA) return launch_kapila4(A, stage, st);
    if (A.sch.model == MODEL_MULTIP) return launch_multip4(A, stage, st);
    return -1;
  }
  if (A.sch.nph != 2) return -1;
  if (A.sch.model == MODEL_KAPILA) return launch_kapila(A, stage, st);
  if (A.sch.model == MODEL_MULTIP) return launch_multip(A, stage, st);
  return -1;
}
int batch_euler(const BatchLaunch &L, hipStream_t st);
int batch_kapila(const BatchLaunch &L, hipStream_t st);
int batch_multip(const BatchLaunch &L, hipStream_t st);
int batch_kapila32(const BatchLaunch &L, hipStream_t st);
int batch_multip32(const BatchLaunch &L, hipStream_t st);
int batch_kapila3(const BatchLaunch &L, hipStream_t st);
int batch_multip3(const BatchLaunch &L, hipStream_t st);
int batch_kapila4(const BatchLaunch &L, hipStream_t st);
int batch_multip4(const BatchLaunch &L, hipStream_t st);

int launch_batch(int model, int prec, int nph, const BatchLaunch &L, hipStream_t st) {
  if (prec == PREC_FP32S) {
    if (nph != 2) return -1;
    if (model == MODEL_KAPILA) return batch_kapila32(L, st);
    if (model == MODEL_MULTIP) return batch_multip32(L, st);
    return -1;
  }
  if (model == MODEL_EULER) return batch_euler(L, st);
  if (nph == 3) return model == MODEL_KAPILA ? batch_kapila3(L, st) : model == MODEL_MULTIP ? batch_multip3(L, st) : -1;
  if (nph == 4) return model == MODEL_KAPILA ? batch_kapila4(L, st) : model == MODEL_MULTIP ? batch_multip4(L, st) : -1;
  if (nph != 2) return -1;
  if (model == MODEL_KAPILA) return batch_kapila(L, st);
  if (model == MODEL_MULTIP) return batch_multip(L, st);
  return -1;
}

void tile_shape(int dim, int t[3]) {
  if (dim == 1) { t[0] = ECO_TILE1_X; t[1] = 1; t[2] = 1; }
  else if (dim == 2) { t[0] = ECO_TILE2_X; t[1] = ECO_TILE2_Y; t[2] = 1; }
  else { t[0] = ECO_TILE3_X; t[1] = ECO_TILE3_Y; t[2] = ECO_TILE3_Z; }
}

int launch_convert(const StageArgs &A, int to_f32, double *tmp, hipStream_t st) {
  const int stage = to_f32 ? 4 : 5;
  if (A.sch.model == MODEL_KAPILA) return launch_kapila32(A, stage, st, tmp);
  if (A.sch.model == MODEL_MULTIP) return launch_multip32(A, stage, st, tmp);
  return -1;
}
}  // namespace eco
