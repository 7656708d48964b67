// Stage kernels of the fast FP64 kind (HF2D_F64_FAST): the FP64 tile code of
// hf2d_tile.h compiled with FMA contraction (this file's build flag
// -ffp-contract=fast) and with hf_scalar.h's fast division (reciprocal and
// two Newton steps, shared per denominator), square root and power, and with
// interior faces evaluated in primitive variables (prim_face in
// hf2d_tile.h). Same state layout and driver as HF2D_F64; its default tile
// is 32 x 12 with 512 threads, one workgroup per tile (hf2d_fast.h).
// Results differ from the bitwise kind in the last bits of each operation,
// and those differences grow along the run as any perturbation does.
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
struct FastShape {
  static constexpr int TW = HF2D_FAST_TILE_W, TH = HF2D_FAST_TILE_H, NT = HF2D_FAST_THREADS;
};

template <int MODE, class Sh> hipError_t launch(const hf_fast::StageArgs<double>& a, unsigned grid, hipStream_t st) {
  hipLaunchKernelGGL((stage_kernel<double, Sh::TW, Sh::TH, Sh::NT, MODE>), dim3(grid), dim3(Sh::NT), 0, st, a);
  return hipGetLastError();
}
template <int MODE>
hipError_t launch_mode(const hf_fast::StageArgs<double>& a, int shape, unsigned tiles, hipStream_t st) {
  if (shape == 1) return launch<MODE, ShapeMid<double>>(a, tiles, st);
  if (shape == 2) return launch<MODE, ShapeSmall<double>>(a, tiles, st);
  return launch<MODE, FastShape>(a, tiles, st);
}
}  // namespace

hipError_t hf2d_fast_stage_launch(const void* args, size_t args_size, int mode, int shape, unsigned tiles,
                                  hipStream_t stream) {
  static_assert(std::is_trivially_copyable<hf_fast::StageArgs<double>>::value, "launch arguments are plain data");
  if (args_size != sizeof(hf_fast::StageArgs<double>)) return hipErrorInvalidValue;
  hf_fast::StageArgs<double> a;
  std::memcpy(&a, args, sizeof a);
  return mode == 0 ? launch_mode<0>(a, shape, tiles, stream) : launch_mode<1>(a, shape, tiles, stream);
}
