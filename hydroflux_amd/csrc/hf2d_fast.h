// Internal interface of the fast FP64 stage kernels (hf2d_fast.hip), which
// are compiled in a translation unit of their own, with FMA contraction
// and the fast division and square root of hf_scalar.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

// Launch the stage kernel of the fast FP64 kind: args points at a
// StageArgs<double> (args_size bytes, checked against the kernels' own
// layout), mode 0 = first stage, 1 = final SSP-RK2 stage, shape 0/1/2 = the
// default / mid / small FP64 tile, tiles = the launch's tile count,
// persist_grid = workgroups of a persistent launch (CUs x resident
// workgroups per CU). Returns a hipError_t.
hipError_t hf2d_fast_stage_launch(const void* args, size_t args_size, int mode, int shape, unsigned tiles,
                                  unsigned persist_grid, hipStream_t stream);

// Resident workgroups per CU of the fast kind's persistent stage kernel.
hipError_t hf2d_fast_stage_occupancy(int* blocks_per_cu);
