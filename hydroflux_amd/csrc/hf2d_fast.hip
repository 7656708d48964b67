// Stage kernels of the fast FP64 kind (HF2D_F64_FAST): the FP64 tile code of
// hf2d_tile.h compiled with FMA contraction (this file's build flag
// -ffp-contract=fast) and with hf_scalar.h's fast division (reciprocal and
// two Newton steps, shared per denominator), square root and power. Same
// state layout, driver and launches as HF2D_F64; results differ from the
// bitwise kind in the last bits of each operation, and those differences
// grow along the run as any perturbation does.
//
// The tile code's namespaces are renamed in this translation unit so that
// its inline functions, built with other arithmetic, are distinct symbols
// from the bitwise build's.
#define HF_FAST_F64 1
#define hf hf_fast
#define hfm hfm_fast
#include "hf2d_device.h"
#undef hf
#undef hfm
#include "hf2d_fast.h"

#include <cstring>
#include <type_traits>

namespace {
// The fast kind's first stage runs persistent; its final stage runs per
// tile: every thread-index mask tried for a persistent final stage either
// spills 15-36 VGPRs to scratch (+0.13 GB written per launch) or is slower
// (flood workload 4096 x 4096: per tile 1.32 ms, persistent 1.31-1.54 ms).
constexpr bool kPersistFinal = false;

template <int MODE, class Sh> hipError_t launch(const hf_fast::StageArgs<double>& a, unsigned grid, hipStream_t st) {
  hipLaunchKernelGGL((stage_kernel<double, Sh::TW, Sh::TH, Sh::NT, MODE>), dim3(grid), dim3(Sh::NT), 0, st, a);
  return hipGetLastError();
}
template <int MODE>
hipError_t launch_mode(const hf_fast::StageArgs<double>& a, int shape, unsigned tiles, unsigned persist_grid,
                       hipStream_t st) {
  if (shape == 1) return launch<MODE, ShapeMid<double>>(a, tiles, st);
  if (shape == 2) return launch<MODE, ShapeSmall<double>>(a, tiles, st);
  if constexpr (Persist<double>::on && (MODE == 0 || kPersistFinal)) {
    using Sh = Shape<double>;
    const unsigned grid = tiles < persist_grid ? tiles : persist_grid;
    hipLaunchKernelGGL((stage_persistent_kernel<double, Sh::TW, Sh::TH, Sh::NT, MODE>), dim3(grid), dim3(Sh::NT), 0,
                       st, a);
    return hipGetLastError();
  } else {
    (void)persist_grid;
    return launch<MODE, Shape<double>>(a, tiles, st);
  }
}
}  // namespace

hipError_t hf2d_fast_stage_launch(const void* args, size_t args_size, int mode, int shape, unsigned tiles,
                                  unsigned persist_grid, hipStream_t stream) {
  static_assert(std::is_trivially_copyable<hf_fast::StageArgs<double>>::value, "launch arguments are plain data");
  if (args_size != sizeof(hf_fast::StageArgs<double>)) return hipErrorInvalidValue;
  hf_fast::StageArgs<double> a;
  std::memcpy(&a, args, sizeof a);
  return mode == 0 ? launch_mode<0>(a, shape, tiles, persist_grid, stream)
                   : launch_mode<1>(a, shape, tiles, persist_grid, stream);
}

hipError_t hf2d_fast_stage_occupancy(int* blocks_per_cu) {
  using Sh = Shape<double>;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                      stage_persistent_kernel<double, Sh::TW, Sh::TH, Sh::NT, 0>,
                                                      Sh::NT, 0);
}
