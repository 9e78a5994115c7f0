// Build-time kernel shape parameters (overridable with -D).  Chosen by the
// tile/occupancy sweep of ecogen_amd/autotune.py on MI355X.
#pragma once
#ifndef ECO_TILE1_X
#define ECO_TILE1_X 504
#endif
#ifndef ECO_TILE2_X
#define ECO_TILE2_X 16
#endif
#ifndef ECO_TILE2_Y
#define ECO_TILE2_Y 13
#endif
#ifndef ECO_TILE3_X
#define ECO_TILE3_X 16
#endif
#ifndef ECO_TILE3_Y
#define ECO_TILE3_Y 4
#endif
#ifndef ECO_TILE3_Z
#define ECO_TILE3_Z 3
#endif
#ifndef ECO_MIN_WAVES_PER_EU
#define ECO_MIN_WAVES_PER_EU 1
#endif
#ifndef ECO_WAVES1
#define ECO_WAVES1 1
#endif
#ifndef ECO_WAVES2
#define ECO_WAVES2 2
#endif
// waves per SIMD requested for the staged 2-D kernels (FP64, FP32 storage)
#ifndef ECO_WAVESS2
#define ECO_WAVESS2 4
#endif
#ifndef ECO_WAVESS32_2
#define ECO_WAVESS32_2 4
#endif
#ifndef ECO_WAVES3
#define ECO_WAVES3 3
#endif
// staged 2-D MUSCL face pass (tile state and halo in LDS, one limited slope
// per cell and direction) for single-limiter runs without walls or THINC
#ifndef ECO_STAGED2
#define ECO_STAGED2 1
#endif
// compile-time (branch-free) limiter for the single-limiter configurations
#ifndef ECO_LIMSPEC1
#define ECO_LIMSPEC1 1
#endif
#ifndef ECO_LIMSPEC2
#define ECO_LIMSPEC2 1
#endif
#ifndef ECO_LIMSPEC3
#define ECO_LIMSPEC3 1
#endif
// workgroup size (threads) per dimension; 0 = one thread per tile cell.  A
// workgroup larger than its tile lets the face pass (more faces than cells)
// finish in fewer iterations; the extra threads idle in the cell update.
#ifndef ECO_THREADS1
#define ECO_THREADS1 512
#endif
#ifndef ECO_THREADS2
#define ECO_THREADS2 256
#endif
#ifndef ECO_THREADS3
#define ECO_THREADS3 256
#endif
// minimum waves per SIMD requested for the persistent pool kernels (0 = as
// the single-solve kernels)
#ifndef ECO_WAVESB1
#define ECO_WAVESB1 0
#endif
#ifndef ECO_WAVESB2
#define ECO_WAVESB2 0
#endif
#ifndef ECO_WAVESB3
#define ECO_WAVESB3 0
#endif
// minimum waves per SIMD requested for the FP32-storage pool kernels (0 = as
// the FP64 pool kernels), chosen by the autotuner's pool target.  At their
// natural register counts (179 and 196 VGPRs) the 1-D and 2-D kernels run
// one 512-thread or two 256-thread workgroups per CU; capping them at 128
// and 168 VGPRs (a few spilled values) doubles that and took the 2048-case
// batched run from 32.2 to about 28 s on MI355X.  The FP64 pool kernels
// would spill 300 bytes per lane at those caps and keep their own.
#ifndef ECO_WAVESB32_1
#define ECO_WAVESB32_1 4
#endif
#ifndef ECO_WAVESB32_2
#define ECO_WAVESB32_2 3
#endif
#ifndef ECO_WAVESB32_3
#define ECO_WAVESB32_3 0
#endif

// single-solve stage kernels: workgroup i takes the tile that gives every XCD
// one contiguous band of tiles (1), or tile i (0)
#ifndef ECO_XCD_BANDS
#define ECO_XCD_BANDS 1
#endif

// the FP32-storage staged 2-D kernel's tile height (its width is
// ECO_TILE2_X): 16x14 there, 16x13 for every other 2-D kernel
#ifndef ECO_TILE32S_2_Y
#define ECO_TILE32S_2_Y 14
#endif

// the FP64 staged 2-D Kapila kernel copies the isentropic power's tables to
// LDS (1) or reads them through the cache (0).  At 16x13 tiles the copy's
// 5 KB would leave three workgroups per CU instead of four.
#ifndef ECO_POWTAB_LDS2
#define ECO_POWTAB_LDS2 0
#endif
