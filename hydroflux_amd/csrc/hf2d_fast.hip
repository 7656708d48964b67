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
template <int MODE, class Sh> hipError_t launch(const hf_fast::StageArgs<double>& a, unsigned grid, hipStream_t st) {
  hipLaunchKernelGGL((stage_kernel<double, Sh::TW, Sh::TH, Sh::NT, MODE>), dim3(grid), dim3(Sh::NT), 0, st, a);
  return hipGetLastError();
}
template <int MODE> hipError_t launch_mode(const hf_fast::StageArgs<double>& a, int shape, unsigned grid, hipStream_t st) {
  if (shape == 1) return launch<MODE, ShapeMid<double>>(a, grid, st);
  if (shape == 2) return launch<MODE, ShapeSmall<double>>(a, grid, st);
  if (!Persist<double>::on) return launch<MODE, Shape<double>>(a, grid, st);
  using Sh = Shape<double>;
  hipLaunchKernelGGL((stage_persistent_kernel<double, Sh::TW, Sh::TH, Sh::NT, MODE>), dim3(grid), dim3(Sh::NT), 0, st,
                     a);
  return hipGetLastError();
}
}  // namespace

hipError_t hf2d_fast_stage_launch(const void* args, size_t args_size, int mode, int shape, unsigned grid,
                                  hipStream_t stream) {
  static_assert(std::is_trivially_copyable<hf_fast::StageArgs<double>>::value, "launch arguments are plain data");
  if (args_size != sizeof(hf_fast::StageArgs<double>)) return hipErrorInvalidValue;
  hf_fast::StageArgs<double> a;
  std::memcpy(&a, args, sizeof a);
  return mode == 0 ? launch_mode<0>(a, shape, grid, stream) : launch_mode<1>(a, shape, grid, stream);
}

hipError_t hf2d_fast_stage_occupancy(int* blocks_per_cu) {
  using Sh = Shape<double>;
  int p0 = 0, p1 = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &p0, stage_persistent_kernel<double, Sh::TW, Sh::TH, Sh::NT, 0>, Sh::NT, 0);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&p1, stage_persistent_kernel<double, Sh::TW, Sh::TH, Sh::NT, 1>,
                                                   Sh::NT, 0);
  if (e != hipSuccess) return e;
  *blocks_per_cu = p0 < p1 ? p0 : p1;
  return hipSuccess;
}
