// Internal interface of the fast FP64 stage kernels (hf2d_fast.hip), which
// are compiled in a translation unit of their own, with FMA contraction
// and the fast division and square root of hf_scalar.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

// Default tile of the fast FP64 kind: 32 x 12 cells, 512 threads, one
// workgroup per tile. Its LDS tile (79 KB) and 72-84 VGPRs let two
// workgroups share a CU, so one workgroup's loads and barriers overlap the
// other's arithmetic. Measured at 4096 x 4096 (flood workload): 8.4 G
// cell-updates/s against 7.85 G for the bitwise kind's 32 x 24 tile with
// 768 threads (one workgroup per CU, first stage persistent).
constexpr int HF2D_FAST_TILE_W = 32, HF2D_FAST_TILE_H = 12, HF2D_FAST_THREADS = 512;

// Launch the stage kernel of the fast FP64 kind: args points at a
// StageArgs<double> (args_size bytes, checked against the kernels' own
// layout), mode 0 = first stage, 1 = final SSP-RK2 stage, shape 0/1/2 = the
// fast default (HF2D_FAST_TILE_*) / the mid / the small FP64 tile, tiles =
// the launch's tile count (one workgroup each). Returns a hipError_t.
hipError_t hf2d_fast_stage_launch(const void* args, size_t args_size, int mode, int shape, unsigned tiles,
                                  hipStream_t stream);
