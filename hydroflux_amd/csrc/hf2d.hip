// MI355X (gfx950) implementation of the C ABI in include/hydroflux_amd.h.
//
// Device layout: SoA float64 planes (h, hu, hv) x 2 state buffers, bed z,
// Manning n, plus tangent planes for the Dual kind. One SSP-RK2 step is
// three launches on the handle's stream:
//   stage kernel (first)  : U^n            -> U^(1)   (dt from the device clock)
//   stage kernel (final)  : U^(1), U^n     -> U^{n+1} (+combine, friction, rain,
//                                                       point sources, CFL maxima)
//   finalize (1 thread)   : t += dt, steps += 1, flip the CFL slot
// so the host never reads dt back; run_until polls the clock every chunk.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hydroflux_amd.h"
#include "hf2d_device.h"
#include "hf2d_fast.h"


// ============================================================== handle
struct hf2d_sim {
  int64_t nr = 0, nc = 0;
  int device = 0;  // the HIP device current at creation: its buffers and stream
  double dx = 1.0, dy = 1.0;
  int kind = HF2D_F64;
  int n_tan = 0;  // tangents per value: Dual 1, DualN<N> N, else 0
  uint32_t host_seq = 0;  // host-clocked launches so far (StageCommon::host_seq)
  hipStream_t stream = nullptr;
  void* st[2][3] = {{nullptr}};   // state buffers (double, or float for HF2D_F32)
  void* tg[2][3] = {{nullptr}};   // tangents (Dual; DualN: N per cell, [cell][k])
  void *man = nullptr, *mant = nullptr;
  double *z = nullptr, *zt = nullptr;  // bed (double / Dual kinds)
  double* datum = nullptr;             // FP32 kind: FP64 datum per 64x64 block
  float* zdev = nullptr;               // FP32 kind: FP32 deviation from the datum
  int64_t datum_cols = 0;
  bool has_zt = false;
  int cur = 0;  // index of the buffer holding the current state
  Ctrl* ctrl = nullptr;
  PointSrc* srcs = nullptr;
  int n_src = 0, cap_src = 0;
  int bc_kind[4] = {HF2D_BC_WALL, HF2D_BC_WALL, HF2D_BC_WALL, HF2D_BC_WALL};
  double bc_param[4] = {0, 0, 0, 0};
  hf2d_config cfg{};
  bool ws_valid = false;  // ctrl->ws[parity] holds the current state's maxima
  // enqueue_steps left driver steps in flight: once they finish, the slots
  // hold the maxima of the state they leave (hf2d_synchronize records that)
  bool driver_async = false;
  bool ws_with_bcs = false;
  // active-tile list (f64 and f32 kinds): per-tile wet flags of U^n, quiet
  // run lengths, pinned tiles (edge, point source), scratch, the list and
  // its length (tile_list[n_tiles]).
  unsigned char *tile_wet = nullptr, *tile_run = nullptr, *tile_pin = nullptr, *tile_act = nullptr;
  int* tile_list = nullptr;
  unsigned long long* ws_part = nullptr;  // CFL slots of the final stage (WS_SLOTS x WS_STRIDE)
  int tw = 0, th = 0;  // stage tile shape: Shape<T>, ShapeMid<T> or ShapeSmall<T> (by grid size)
  int shape = 0;       // 0 Shape, 1 ShapeMid, 2 ShapeSmall
  int tiles_x = 0, tiles_y = 0;
  // workgroups of the persistent stage kernels (FP64 kinds, default tile):
  // CUs x resident workgroups per CU, a multiple of the 8 XCDs
  int64_t persist_grid = 1;
  bool list_valid = false;  // tile_wet / tile_run follow the current state
  // driver forcing series ([0] rain, [1] point sources) and gauges
  SeriesDesc series_host[2] = {};
  double* series_buf[2] = {nullptr, nullptr};  // device t[n] then v[n]
  SeriesDesc* series_dev = nullptr;
  int64_t* gauge_cells = nullptr;
  double* gauge_buf = nullptr;
  long long* gauge_count = nullptr;
  int n_gauges = 0;
  long long gauge_cap = 0;
  // GRAPH_STEPS device-clocked steps captured once and replayed; dropped
  // whenever a launch argument baked into it changes.
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_cur = -1;
  bool graph_list = false;
  std::string err;
  // kernel timing
  bool timing = false;
  std::vector<hipEvent_t> ev;  // pairs: [s1 start, s1 end, s2 start, s2 end] per step
  size_t ev_used = 0;
  double timed_ms[2] = {0, 0};
  int64_t timed_count = 0;
};

namespace {

constexpr int64_t GRAPH_STEPS = 16;

// The state, mesh, boundaries or configuration changed: the CFL slots and
// the active-tile flags no longer describe it.
void state_changed(hf2d_sim* s) {
  s->ws_valid = false;
  s->list_valid = false;
  s->driver_async = false;
}

// Destroy the captured step graph. Launches of it may still be in flight
// (hf2d_enqueue_steps returns without waiting, and so do the setters that
// call this), and HIP does not defer the destruction of an executable graph
// that is still running, so the stream is drained first.
void drop_graph(hf2d_sim* s) {
  if (s->gexec && s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
}

int fail(hf2d_sim* s, int code, const std::string& msg) {
  if (s) s->err = msg;
  return code;
}

#define HIPCHK(sim, expr)                                                                 \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(sim, HF2D_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

hf2d_config default_config() {
  hf2d_config c{};
  c.cfl = 0.4;
  c.integrator = HF2D_SSPRK2;
  c.cfl_with_bcs = 1;
  c.max_dt = 60.0;
  c.friction = 1;
  c.friction_dry_tol = hf::H_DRY;
  c.max_steps = 10000000;
  c.rain_rate = 0.0;
  return c;
}



// Arguments for a launch reading state buffer `in`, writing buffer `out`,
// with U^n in buffer `un`.
// Grids of at most this many tiles run every tile: they fit the device in
// one or two rounds of workgroups, so tiles left out save no time, while the
// list costs a compaction launch per step and a dependent load of the list
// in every workgroup (UK EA Test 4, 169 tiles: 41.8 vs 36.9 us/step).
#ifndef HF_LIST_MIN_TILES
#define HF_LIST_MIN_TILES 512
#endif

// The active-tile list is used when it cannot change a bit (hf2d_tile.h,
// tile_skipped): f64/f32 state, SSP-RK2, no rain, no friction on dry
// cells, not switched off; and where it can save time: more than
// HF_LIST_MIN_TILES tiles, unless HF2D_OPT_ACTIVE_TILES_ANY_SIZE is set.
bool list_usable(const hf2d_sim* s) {
  const bool big = (int64_t)s->tiles_x * s->tiles_y > HF_LIST_MIN_TILES ||
                   (s->cfg.options & HF2D_OPT_ACTIVE_TILES_ANY_SIZE) != 0;
  return s->tile_list != nullptr && s->cfg.integrator == HF2D_SSPRK2 && s->cfg.rain_rate == 0.0 &&
         s->series_host[0].n == 0 &&
         (!s->cfg.friction || s->cfg.friction_dry_tol >= hf::H_DRY) &&
         (s->cfg.options & HF2D_OPT_NO_ACTIVE_TILES) == 0 && big;
}

template <class T> StageArgs<T> make_args(hf2d_sim* s, int in, int out, int un) {
  StageArgs<T> a{};
  a.nr = s->nr;
  a.nc = s->nc;
  a.dx = s->dx;
  a.dy = s->dy;
  a.ih = fld<T>(s->st[in][0], s->tg[in][0]);
  a.ihu = fld<T>(s->st[in][1], s->tg[in][1]);
  a.ihv = fld<T>(s->st[in][2], s->tg[in][2]);
  a.oh = fldw<T>(s->st[out][0], s->tg[out][0]);
  a.ohu = fldw<T>(s->st[out][1], s->tg[out][1]);
  a.ohv = fldw<T>(s->st[out][2], s->tg[out][2]);
  a.nh = fld<T>(s->st[un][0], s->tg[un][0]);
  a.nhu = fld<T>(s->st[un][1], s->tg[un][1]);
  a.nhv = fld<T>(s->st[un][2], s->tg[un][2]);
  a.z = fld<T>(s->z, s->has_zt ? s->zt : nullptr);
  a.man = fld<T>(s->man, s->mant);
  a.datum = s->datum;
  a.zdev = s->zdev;
  a.datum_log2 = DATUM_LOG2;
  a.datum_cols = s->datum_cols;
  for (int k = 0; k < 4; ++k) {
    a.bc_kind[k] = s->bc_kind[k];
    a.bc_param[k] = s->bc_param[k];
  }
  a.ctrl = s->ctrl;
  a.cfl = s->cfg.cfl;
  a.max_dt = s->cfg.max_dt;
  a.cfl_with_bcs = s->cfg.cfl_with_bcs;
  a.srcs = s->srcs;
  a.tiles_x = (int)((s->nc + s->tw - 1) / s->tw);
  a.tiles_y = (int)((s->nr + s->th - 1) / s->th);
  return a;
}

template <class T> struct WideDualN {
  static constexpr int n = 0;
};
template <int N> struct WideDualN<hf::dualn<N>> {
  static constexpr int n = N;
};

// Component group c (tangents [G c, G c + G)) of a DualN<N> launch as a
// DualN<G> launch on the DualN<G> kernel's tiles (G = 4: 16 x 8). Every
// group computes the same values, bit for bit (a DualN value never depends on its
// tangents, and each tangent component follows the same rule whatever N
// is); only the last group stores values and reduces the CFL wave speeds,
// so the in-place final stage reads U^n values that no earlier group has
// overwritten.
template <int G, int N> StageArgs<hf::dualn<G>> group_args(const StageArgs<hf::dualn<N>>& a, int c) {
  using GT = hf::dualn<G>;
  static_assert(N % G == 0 && N > G, "whole groups of a wider DualN");
  const bool last = c == N / G - 1;
  StageArgs<GT> b{};
  static_cast<hf::StageCommon&>(b) = a;
  auto r = [&](const Fld<hf::dualn<N>>& f) { return Fld<GT>{f.v, f.d ? f.d + G * c : nullptr, f.ds}; };
  auto w = [&](const FldW<hf::dualn<N>>& f) { return FldW<GT>{f.v, f.d + G * c, f.ds, last ? f.wv : 0}; };
  b.ih = r(a.ih);
  b.ihu = r(a.ihu);
  b.ihv = r(a.ihv);
  b.oh = w(a.oh);
  b.ohu = w(a.ohu);
  b.ohv = w(a.ohv);
  b.nh = r(a.nh);
  b.nhu = r(a.nhu);
  b.nhv = r(a.nhv);
  b.z = r(a.z);
  b.man = r(a.man);
  if (!last) {
    b.wave_speed = 0;
    b.tile_wet = nullptr;
  }
  return b;
}

// Workgroups of a persistent stage launch over `tiles` tiles: all that stay
// resident at once (s->persist_grid), fewer on a grid with fewer tiles.
unsigned persistent_grid(const hf2d_sim* s, int64_t tiles) {
  return (unsigned)(tiles < s->persist_grid ? tiles : s->persist_grid);
}

template <class T, class Sh> int launch_shaped(hf2d_sim* s, const StageArgs<T>& a, int mode) {
  const unsigned grid = (unsigned)((int64_t)a.tiles_x * a.tiles_y);
  if constexpr (Persist<T>::on && std::is_same<Sh, Shape<T>>::value) {
    const unsigned pg = persistent_grid(s, grid);
    if (mode == hf::MODE_STAGE)
      hipLaunchKernelGGL((stage_persistent_kernel<T, Sh::TW, Sh::TH, Sh::NT, hf::MODE_STAGE>), dim3(pg),
                         dim3(Sh::NT), 0, s->stream, a);
    else
      hipLaunchKernelGGL((stage_persistent_kernel<T, Sh::TW, Sh::TH, Sh::NT, hf::MODE_FINAL_RK2>), dim3(pg),
                         dim3(Sh::NT), 0, s->stream, a);
    HIPCHK(s, hipGetLastError());
    return HF2D_OK;
  }
  if (mode == hf::MODE_STAGE)
    hipLaunchKernelGGL((stage_kernel<T, Sh::TW, Sh::TH, Sh::NT, hf::MODE_STAGE>), dim3(grid), dim3(Sh::NT), 0,
                       s->stream, a);
  else
    hipLaunchKernelGGL((stage_kernel<T, Sh::TW, Sh::TH, Sh::NT, hf::MODE_FINAL_RK2>), dim3(grid), dim3(Sh::NT), 0,
                       s->stream, a);
  HIPCHK(s, hipGetLastError());
  return HF2D_OK;
}

template <class T> int launch_stage_kernel(hf2d_sim* s, const StageArgs<T>& a, int mode) {
  if constexpr (std::is_same<T, double>::value) {
    if (s->kind == HF2D_F64_FAST) {  // the fast FP64 build of the same kernels (hf2d_fast.hip)
      const unsigned tiles = (unsigned)((int64_t)a.tiles_x * a.tiles_y);
      HIPCHK(s, hf2d_fast_stage_launch(&a, sizeof a, mode == hf::MODE_STAGE ? 0 : 1, s->shape, tiles, s->stream));
      return HF2D_OK;
    }
  }
  if constexpr (ShapeMid<T>::on) {
    if (s->shape == 1) return launch_shaped<T, ShapeMid<T>>(s, a, mode);
  }
  if constexpr (ShapeSmall<T>::on) {
    if (s->shape == 2) return launch_shaped<T, ShapeSmall<T>>(s, a, mode);
  }
  return launch_shaped<T, Shape<T>>(s, a, mode);
}

long long tiles_of(const hf2d_sim* s, int tw, int th) { return ((s->nr + th - 1) / th) * ((s->nc + tw - 1) / tw); }

// The grid's stage tile shape, fixed at creation.
template <class T> int choose_tiles(hf2d_sim* s) {
  s->shape = 0;
  s->tw = Shape<T>::TW;
  s->th = Shape<T>::TH;
  if (s->kind == HF2D_F64_FAST) {  // the fast kind's own default tile (hf2d_fast.h)
    s->tw = HF2D_FAST_TILE_W;
    s->th = HF2D_FAST_TILE_H;
  }
  if (tiles_of(s, s->tw, s->th) >= SMALL_GRID_TILES) return HF2D_OK;
  if constexpr (ShapeMid<T>::on) {
    if (tiles_of(s, ShapeMid<T>::TW, ShapeMid<T>::TH) >= MID_GRID_TILES) {
      s->shape = 1;
      s->tw = ShapeMid<T>::TW;
      s->th = ShapeMid<T>::TH;
      return HF2D_OK;
    }
  }
  if constexpr (ShapeSmall<T>::on) {
    s->shape = 2;
    s->tw = ShapeSmall<T>::TW;
    s->th = ShapeSmall<T>::TH;
  }
  return HF2D_OK;
}

// s->persist_grid: the device's CUs times the persistent stage kernels'
// resident workgroups per CU (the smaller of the two modes'), rounded down
// to a multiple of 8 so that every workgroup keeps its XCD's tile band.
template <class T> int choose_persist(hf2d_sim* s) {
  s->persist_grid = 1;
  if (s->kind == HF2D_F64_FAST) return HF2D_OK;  // one workgroup per tile (hf2d_fast.h)
  if constexpr (Persist<T>::on) {
    using Sh = Shape<T>;
    int ncu = 0, per0 = 0, per1 = 0;
    HIPCHK(s, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, s->device));
    HIPCHK(s, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per0, stage_persistent_kernel<T, Sh::TW, Sh::TH, Sh::NT, hf::MODE_STAGE>, Sh::NT, 0));
    HIPCHK(s, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per1, stage_persistent_kernel<T, Sh::TW, Sh::TH, Sh::NT, hf::MODE_FINAL_RK2>, Sh::NT, 0));
    const int64_t per = per0 < per1 ? per0 : per1;
    const int64_t g = (int64_t)ncu * (per > 0 ? per : 1);
    s->persist_grid = g >= 8 ? g / 8 * 8 : g;
  }
  return HF2D_OK;
}

template <class T> int launch_stage(hf2d_sim* s, StageArgs<T> a, int mode, int timing_slot) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->timing && timing_slot >= 0) {
    if (s->ev_used + 2 > s->ev.size()) {
      for (int k = 0; k < 64; ++k) {
        hipEvent_t e;
        HIPCHK(s, hipEventCreate(&e));
        s->ev.push_back(e);
      }
    }
    e0 = s->ev[s->ev_used];
    e1 = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    HIPCHK(s, hipEventRecord(e0, s->stream));
  }
  constexpr int N = WideDualN<T>::n;
  if constexpr (N > 4) {
    for (int c = 0; c < N / 4; ++c) {
      const int rc = launch_stage_kernel(s, group_args<4>(a, c), mode);
      if (rc) return rc;
    }
  } else {
    const int rc = launch_stage_kernel(s, a, mode);
    if (rc) return rc;
  }
  if (e1) {
    HIPCHK(s, hipEventRecord(e1, s->stream));
    (void)timing_slot;
  }
  return HF2D_OK;
}

// Harvest finished timing event pairs (order: stage-1 pair, final pair).
int harvest_timing(hf2d_sim* s) {
  if (!s->timing || s->ev_used == 0) return HF2D_OK;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (size_t k = 0; k + 3 < s->ev_used; k += 4) {
    float ms0 = 0, ms1 = 0;
    HIPCHK(s, hipEventElapsedTime(&ms0, s->ev[k], s->ev[k + 1]));
    HIPCHK(s, hipEventElapsedTime(&ms1, s->ev[k + 2], s->ev[k + 3]));
    s->timed_ms[0] += ms0;
    s->timed_ms[1] += ms1;
    s->timed_count += 1;
  }
  s->ev_used = 0;
  return HF2D_OK;
}

unsigned tail_grid(hf2d_sim* s) {
  int64_t n = s->nr * s->nc;
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Make ctrl->ws[parity] hold the CFL maxima of the current state.
template <class T> int refresh_wave_speeds(hf2d_sim* s, bool with_bcs) {
  if (s->ws_valid && s->ws_with_bcs == with_bcs) return HF2D_OK;
  hipLaunchKernelGGL(zero_slot_kernel, dim3(1), dim3(1), 0, s->stream, s->ctrl);
  HIPCHK(s, hipGetLastError());
  StageArgs<T> a = make_args<T>(s, s->cur, s->cur, s->cur);
  a.dt_mode = 0;
  a.dt_host = 0.0;
  a.friction = 0;
  a.rain_rate = 0.0;
  a.n_src = 0;
  a.wave_speed = 1;
  a.cfl_with_bcs = with_bcs ? 1 : 0;
  hipLaunchKernelGGL((tail_kernel<T>), dim3(tail_grid(s)), dim3(256), 0, s->stream, a, 1);
  HIPCHK(s, hipGetLastError());
  s->ws_valid = true;
  s->ws_with_bcs = with_bcs;
  return HF2D_OK;
}

template <class T> int do_max_wave_speeds(hf2d_sim* s, bool with_bcs, double* sx, double* sy) {
  int rc = refresh_wave_speeds<T>(s, with_bcs);
  if (rc) return rc;
  Ctrl c;
  HIPCHK(s, hipMemcpyAsync(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  *sx = hfm::as_f64(c.ws[c.parity][0]);
  *sy = hfm::as_f64(c.ws[c.parity][1]);
  return HF2D_OK;
}

// One step with caller dt: forward Euler or SSP-RK2 without forcing.
template <class T> int host_dt_step(hf2d_sim* s, double dt, bool rk2) {
  const int cur = s->cur, oth = 1 - s->cur;
  StageArgs<T> a = make_args<T>(s, cur, oth, cur);
  a.dt_mode = 0;
  a.dt_host = dt;
  a.host_seq = ++s->host_seq;
  int rc = launch_stage<T>(s, a, hf::MODE_STAGE, -1);
  if (rc) return rc;
  if (rk2) {
    StageArgs<T> b = make_args<T>(s, oth, cur, cur);
    b.dt_mode = 0;
    b.dt_host = dt;
    b.host_seq = a.host_seq;
    b.friction = 0;
    b.rain_rate = 0.0;
    b.n_src = 0;
    b.wave_speed = 0;
    rc = launch_stage<T>(s, b, hf::MODE_FINAL_RK2, -1);
    if (rc) return rc;
  } else {
    s->cur = oth;
  }
  state_changed(s);
  drop_graph(s);
  return HF2D_OK;
}

// Seed the active-tile list's wet flags from the current state (after the
// state was set by other means than the driver).
template <class T> int enqueue_list_init(hf2d_sim* s) {
  const int64_t nt = (int64_t)s->tiles_x * s->tiles_y;
  const unsigned g = (unsigned)((nt + 255) / 256 < 4096 ? (nt + 255) / 256 : 4096);
  {
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(g), dim3(256), 0, s->stream, s->tile_wet, nt, (unsigned char)0);
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(g), dim3(256), 0, s->stream, s->tile_run, nt, (unsigned char)0);
    hipLaunchKernelGGL((tile_wet_init_kernel<T>), dim3(tail_grid(s)), dim3(256), 0, s->stream,
                       fld<T>(s->st[s->cur][0], s->tg[s->cur][0]), s->nr, s->nc, s->tw, s->th,
                       s->tiles_x, s->tile_wet);
    HIPCHK(s, hipGetLastError());
    s->list_valid = true;
  }
  return HF2D_OK;
}

// Rebuild the active-tile list for the next step.
template <class T> int enqueue_compact(hf2d_sim* s) {
  const int64_t nt = (int64_t)s->tiles_x * s->tiles_y;
  if (!s->list_valid) {
    int rc = enqueue_list_init<T>(s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(1024), 0, s->stream, s->tile_wet, s->tile_pin, s->tile_run,
                     s->tile_act, s->tiles_x, s->tiles_y, s->tile_list, s->tile_list + nt, s->ctrl, s->cfg.cfl,
                     s->dx, s->dy, s->cfg.max_dt);
  HIPCHK(s, hipGetLastError());
  return HF2D_OK;
}

// Enqueue one device-clocked step of the configured driver.
template <class T> int enqueue_step(hf2d_sim* s) {
  const int cur = s->cur, oth = 1 - s->cur;
  StageArgs<T> a = make_args<T>(s, cur, oth, cur);
  a.dt_mode = 1;
  a.first_stage = 1;
  StageArgs<T> b = make_args<T>(s, oth, cur, cur);
  b.dt_mode = 1;
  b.first_stage = 0;
  b.friction = s->cfg.friction;
  b.friction_dry_tol = s->cfg.friction_dry_tol;
  b.rain_rate = s->cfg.rain_rate;
  b.n_src = s->n_src;
  b.wave_speed = 1;
  b.ws_part = s->ws_part;
  const bool series = s->series_host[0].n > 0 || s->series_host[1].n > 0;
  b.use_series = series ? 1 : 0;
  a.series = series ? s->series_dev : nullptr;
  if (list_usable(s)) {
    int rc = enqueue_compact<T>(s);
    if (rc) return rc;
    const int64_t nt = (int64_t)s->tiles_x * s->tiles_y;
    a.tile_list = b.tile_list = s->tile_list;
    a.n_work = b.n_work = s->tile_list + nt;
    b.tile_wet = s->tile_wet;
  } else {
    s->list_valid = false;
  }
  if (s->cfg.integrator == HF2D_SSPRK2) {
    int rc = launch_stage<T>(s, a, hf::MODE_STAGE, 0);
    if (rc) return rc;
    rc = launch_stage<T>(s, b, hf::MODE_FINAL_RK2, 1);
    if (rc) return rc;
  } else {
    // Forward Euler: stage into the other buffer, then the elementwise
    // tail copies it back with forcing and the CFL maxima.
    int rc = launch_stage<T>(s, a, hf::MODE_STAGE, 0);
    if (rc) return rc;
    StageArgs<T> t = b;
    t.ih = fld<T>(s->st[oth][0], s->tg[oth][0]);
    t.ihu = fld<T>(s->st[oth][1], s->tg[oth][1]);
    t.ihv = fld<T>(s->st[oth][2], s->tg[oth][2]);
    t.oh = fldw<T>(s->st[cur][0], s->tg[cur][0]);
    t.ohu = fldw<T>(s->st[cur][1], s->tg[cur][1]);
    t.ohv = fldw<T>(s->st[cur][2], s->tg[cur][2]);
    if (s->timing) {
      // keep the event pairing (stage, final) for the tail
      if (s->ev_used + 2 > s->ev.size()) {
        for (int k = 0; k < 64; ++k) {
          hipEvent_t e;
          HIPCHK(s, hipEventCreate(&e));
          s->ev.push_back(e);
        }
      }
      HIPCHK(s, hipEventRecord(s->ev[s->ev_used], s->stream));
    }
    hipLaunchKernelGGL((tail_kernel<T>), dim3(tail_grid(s)), dim3(256), 0, s->stream, t, 0);
    HIPCHK(s, hipGetLastError());
    if (s->timing) {
      HIPCHK(s, hipEventRecord(s->ev[s->ev_used + 1], s->stream));
      s->ev_used += 2;
    }
  }
  GaugeArgs g{};
  if (s->n_gauges > 0) {  // the value plane of the step's output depth
    g.h = s->st[s->cur][0];
    g.f32 = s->kind == HF2D_F32 ? 1 : 0;
    g.cells = s->gauge_cells;
    g.n = s->n_gauges;
    g.buf = s->gauge_buf;
    g.count = s->gauge_count;
    g.cap = s->gauge_cap;
  }
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(WS_SLOTS), 0, s->stream, s->ctrl, s->ws_part, g);
  HIPCHK(s, hipGetLastError());
  return HF2D_OK;
}

template <class T> int record_gauges_now(hf2d_sim* s) {
  hipLaunchKernelGGL((gauge_kernel<T>), dim3(1), dim3(64), 0, s->stream, s->ctrl,
                     fld<T>(s->st[s->cur][0], s->tg[s->cur][0]), s->gauge_cells, s->n_gauges, s->gauge_buf,
                     s->gauge_count, s->gauge_cap, 1);
  HIPCHK(s, hipGetLastError());
  return HF2D_OK;
}

// Enqueue n device-clocked steps: whole captured GRAPH_STEPS-step graphs
// where possible (launch-bound small grids), the rest one by one. Kernel
// timing records events per step, so it runs without the graph.
template <class T> int enqueue_n(hf2d_sim* s, int64_t n) {
  if (!s->timing && n >= GRAPH_STEPS) {
    const bool lu = list_usable(s);
    if (lu && !s->list_valid) {
      int rc = enqueue_list_init<T>(s);
      if (rc) return rc;
    }
    if (!s->gexec || s->graph_cur != s->cur || s->graph_list != lu) {
      drop_graph(s);
      HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      int rc = HF2D_OK;
      for (int64_t k = 0; k < GRAPH_STEPS && rc == HF2D_OK; ++k) rc = enqueue_step<T>(s);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s->stream, &g);
      if (rc == HF2D_OK && e != hipSuccess) rc = fail(s, HF2D_HIP_ERROR, hipGetErrorString(e));
      if (rc == HF2D_OK && hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0) != hipSuccess)
        rc = fail(s, HF2D_HIP_ERROR, "hipGraphInstantiate failed");
      if (rc != HF2D_OK) {
        if (g) (void)hipGraphDestroy(g);
        s->gexec = nullptr;
        return rc;
      }
      s->graph = g;
      s->graph_cur = s->cur;
      s->graph_list = lu;
    }
    for (; n >= GRAPH_STEPS; n -= GRAPH_STEPS) HIPCHK(s, hipGraphLaunch(s->gexec, s->stream));
  }
  for (int64_t k = 0; k < n; ++k) {
    int rc = enqueue_step<T>(s);
    if (rc) return rc;
  }
  return HF2D_OK;
}

template <class T> int prepare_driver(hf2d_sim* s) {
  return refresh_wave_speeds<T>(s, s->cfg.cfl_with_bcs != 0);
}

int read_ctrl(hf2d_sim* s, Ctrl* c) {
  HIPCHK(s, hipMemcpyAsync(c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return HF2D_OK;
}

int write_ctrl(hf2d_sim* s, const Ctrl* c) {
  HIPCHK(s, hipMemcpyAsync(s->ctrl, c, sizeof(Ctrl), hipMemcpyHostToDevice, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return HF2D_OK;
}

int ctrl_error(hf2d_sim* s, const Ctrl& c, double t_target) {
  char buf[256];
  if (c.status == 3) {
    snprintf(buf, sizeof buf, "degenerate time step dt = %g at t = %g (step %lld)", c.dt, c.t, c.steps);
    return fail(s, HF2D_DEGENERATE_DT, buf);
  }
  if (c.status == 4) {
    snprintf(buf, sizeof buf, "step budget exhausted: %lld steps, t = %g of %g s", c.steps, c.t, t_target);
    return fail(s, HF2D_STEP_BUDGET_EXHAUSTED, buf);
  }
  return HF2D_OK;
}

template <class T> int run_driver(hf2d_sim* s, bool use_target, double t_target, double dt_cap,
                                  int64_t n_steps, bool poll) {
  int rc = prepare_driver<T>(s);
  if (rc) return rc;
  Ctrl c;
  rc = read_ctrl(s, &c);
  if (rc) return rc;
  c.status = 0;
  c.done = 0;
  c.use_target = use_target ? 1 : 0;
  c.t_target = t_target;
  c.dt_cap = dt_cap;
  c.budget = use_target ? s->cfg.max_steps : (long long)n_steps;
  rc = write_ctrl(s, &c);
  if (rc) return rc;
  s->ws_valid = false;  // the slots now follow the driver
  s->driver_async = false;
  if (!use_target) {
    rc = enqueue_n<T>(s, n_steps);
    if (rc) return rc;
    if (!poll) {
      s->driver_async = true;
      return HF2D_OK;
    }
    rc = read_ctrl(s, &c);
    if (rc) return rc;
    rc = harvest_timing(s);
    if (rc) return rc;
    s->ws_valid = (c.status == 0);
    s->ws_with_bcs = s->cfg.cfl_with_bcs != 0;
    return ctrl_error(s, c, t_target);
  }
  // run_until: enqueue in chunks, poll the device clock between chunks.
  int64_t chunk = 16;
  for (;;) {
    rc = enqueue_n<T>(s, chunk);
    if (rc) return rc;
    rc = read_ctrl(s, &c);
    if (rc) return rc;
    rc = harvest_timing(s);
    if (rc) return rc;
    if (c.status != 0 || c.done != 0) break;
    if (chunk < 256) chunk *= 2;
  }
  s->ws_valid = (c.status == 0);
  s->ws_with_bcs = s->cfg.cfl_with_bcs != 0;
  return ctrl_error(s, c, t_target);
}

bool valid(const hf2d_sim* s) { return s != nullptr; }

int check_bcs_geometry(hf2d_sim* s, const int32_t kinds[4]) {
  for (int k = 0; k < 4; ++k) {
    if (kinds[k] < 0 || kinds[k] > 3) return fail(s, HF2D_INVALID_CONFIG, "unknown boundary kind");
    bool physical = kinds[k] == HF2D_BC_DISCHARGE || kinds[k] == HF2D_BC_DEPTH;
    bool xf = (k == hf::SIDE_W || k == hf::SIDE_E);
    // boundary.rs:349-364: linear bed extrapolation needs two cells across.
    if (physical && (xf ? s->nc < 2 : s->nr < 2))
      return fail(s, HF2D_INVALID_CONFIG, "physical boundary needs at least 2 cells across the face");
  }
  return HF2D_OK;
}

}  // namespace

namespace {
template <class T> int forcing_only(hf2d_sim* s, double dt, int friction, double dry_tol, double rain) {
  StageArgs<T> a = make_args<T>(s, s->cur, s->cur, s->cur);
  a.dt_mode = 0;
  a.dt_host = dt;
  a.host_seq = ++s->host_seq;
  a.friction = friction;
  a.friction_dry_tol = dry_tol;
  a.rain_rate = rain;
  a.n_src = 0;
  a.wave_speed = 0;
  hipLaunchKernelGGL((tail_kernel<T>), dim3(tail_grid(s)), dim3(256), 0, s->stream, a, 0);
  HIPCHK(s, hipGetLastError());
  state_changed(s);
  drop_graph(s);
  return HF2D_OK;
}

template <class T> int launch_point_sources(hf2d_sim* s, double dt) {
  StageArgs<T> a = make_args<T>(s, s->cur, s->cur, s->cur);
  a.n_src = s->n_src;
  hipLaunchKernelGGL((point_source_kernel<T>), dim3(1), dim3(64), 0, s->stream, a, dt);
  return HF2D_OK;
}

int upload_sources(hf2d_sim* s, const hf2d_point_source* src, int64_t k) {
  for (int64_t i = 0; i < k; ++i)
    if (src[i].row < 0 || src[i].row >= s->nr || src[i].col < 0 || src[i].col >= s->nc)
      return fail(s, HF2D_SHAPE_MISMATCH, "point source outside the grid");
  if (k > s->cap_src) {
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (s->srcs) HIPCHK(s, hipFree(s->srcs));
    s->srcs = nullptr;
    HIPCHK(s, hipMalloc(&s->srcs, (size_t)k * sizeof(PointSrc)));
    s->cap_src = (int)k;
  }
  std::vector<PointSrc> v((size_t)k);
  for (int64_t i = 0; i < k; ++i) v[(size_t)i] = PointSrc{src[i].row, src[i].col, src[i].q_mass};
  if (k > 0) {
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(s->srcs, v.data(), (size_t)k * sizeof(PointSrc), hipMemcpyHostToDevice));
  }
  s->n_src = (int)k;
  return HF2D_OK;
}
// Tiles the active-tile list never leaves out: near the domain edge, or
// holding a point source of the configured driver.
int upload_pins(hf2d_sim* s, const hf2d_point_source* src, int64_t k) {
  if (!s->tile_pin) return HF2D_OK;
  const int tw = s->tw, th = s->th;
  std::vector<unsigned char> pin((size_t)s->tiles_x * s->tiles_y, 0);
  for (int ty = 0; ty < s->tiles_y; ++ty)
    for (int tx = 0; tx < s->tiles_x; ++tx)
      pin[(size_t)ty * s->tiles_x + tx] = hf::tile_near_edge(s->nr, s->nc, tw, th, tx, ty) ? 1 : 0;
  for (int64_t i = 0; i < k; ++i) pin[(size_t)(src[i].row / th) * s->tiles_x + (size_t)(src[i].col / tw)] = 1;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipMemcpy(s->tile_pin, pin.data(), pin.size(), hipMemcpyHostToDevice));
  return HF2D_OK;
}
}  // namespace

using dual2 = hf::dualn<2>;
using dual4 = hf::dualn<4>;
using dual8 = hf::dualn<8>;
using dual16 = hf::dualn<16>;

static int tangents_of(int kind) {
  switch (kind) {
    case HF2D_DUAL: return 1;
    case HF2D_DUAL2: return 2;
    case HF2D_DUAL4: return 4;
    case HF2D_DUAL8: return 8;
    case HF2D_DUAL16: return 16;
    default: return 0;
  }
}

#define HF_DISPATCH(s, FN, ...)                                  \
  ((s)->kind == HF2D_DUAL  ? FN<dual>(__VA_ARGS__)                \
   : (s)->kind == HF2D_DUAL2 ? FN<dual2>(__VA_ARGS__)             \
   : (s)->kind == HF2D_DUAL4 ? FN<dual4>(__VA_ARGS__)             \
   : (s)->kind == HF2D_DUAL8 ? FN<dual8>(__VA_ARGS__)             \
   : (s)->kind == HF2D_DUAL16 ? FN<dual16>(__VA_ARGS__)           \
   : (s)->kind == HF2D_F32 ? FN<float>(__VA_ARGS__)               \
                           : FN<double>(__VA_ARGS__))

// ============================================================== C ABI
extern "C" {

int hf2d_create(int64_t n_rows, int64_t n_cols, double dx, double dy, int32_t scalar_kind, hf2d_sim** out) {
  if (!out) return HF2D_INVALID_ARGUMENT;
  *out = nullptr;
  if (scalar_kind != HF2D_F64 && scalar_kind != HF2D_F32 && scalar_kind != HF2D_F64_FAST &&
      tangents_of(scalar_kind) == 0)
    return HF2D_INVALID_ARGUMENT;
  if (!(dx > 0.0) || !(dy > 0.0) || n_rows <= 0 || n_cols <= 0 || n_rows >= HF2D_MAX_SIDE ||
      n_cols >= HF2D_MAX_SIDE)
    return HF2D_INVALID_CONFIG;
  hf2d_sim* s = new hf2d_sim();
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return HF2D_HIP_ERROR;
  }
  s->nr = n_rows;
  s->nc = n_cols;
  s->dx = dx;
  s->dy = dy;
  s->kind = scalar_kind;
  s->n_tan = tangents_of(scalar_kind);
  s->cfg = default_config();
  HF_DISPATCH(s, choose_tiles, s);
  const size_t n = (size_t)(n_rows * n_cols);
  const size_t bytes = n * elem_bytes(scalar_kind);
  const size_t tbytes = bytes * (size_t)s->n_tan;
  auto cleanup = [&](int code) {
    hf2d_destroy(s);
    return code;
  };
  auto zalloc = [&](void** p, size_t nb) {
    return hipMalloc(p, nb) == hipSuccess && hipMemset(*p, 0, nb) == hipSuccess;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(HF2D_HIP_ERROR);
  if (HF_DISPATCH(s, choose_persist, s) != HF2D_OK) return cleanup(HF2D_HIP_ERROR);
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < 3; ++f) {
      if (!zalloc(&s->st[b][f], bytes)) return cleanup(HF2D_HIP_ERROR);
      if (s->n_tan > 0 && !zalloc(&s->tg[b][f], tbytes)) return cleanup(HF2D_HIP_ERROR);
    }
  if (!zalloc(&s->man, bytes)) return cleanup(HF2D_HIP_ERROR);
  if (scalar_kind == HF2D_F32) {
    const int64_t rows = (n_rows + (1 << DATUM_LOG2) - 1) >> DATUM_LOG2;
    s->datum_cols = (n_cols + (1 << DATUM_LOG2) - 1) >> DATUM_LOG2;
    if (!zalloc((void**)&s->datum, (size_t)(rows * s->datum_cols) * sizeof(double))) return cleanup(HF2D_HIP_ERROR);
    if (!zalloc((void**)&s->zdev, n * sizeof(float))) return cleanup(HF2D_HIP_ERROR);
  } else {
    if (!zalloc((void**)&s->z, n * sizeof(double))) return cleanup(HF2D_HIP_ERROR);
  }
  if (s->n_tan > 0) {
    if (!zalloc((void**)&s->zt, tbytes)) return cleanup(HF2D_HIP_ERROR);
    if (!zalloc(&s->mant, tbytes)) return cleanup(HF2D_HIP_ERROR);
  }
  if (s->n_tan == 0) {
    s->tiles_x = (int)((n_cols + s->tw - 1) / s->tw);
    s->tiles_y = (int)((n_rows + s->th - 1) / s->th);
    const size_t nt = (size_t)s->tiles_x * s->tiles_y;
    if (!zalloc((void**)&s->tile_wet, nt) || !zalloc((void**)&s->tile_run, nt) ||
        !zalloc((void**)&s->tile_pin, nt) || !zalloc((void**)&s->tile_act, nt) ||
        !zalloc((void**)&s->tile_list, (nt + 1) * sizeof(int)))
      return cleanup(HF2D_HIP_ERROR);
    if (upload_pins(s, nullptr, 0) != HF2D_OK) return cleanup(HF2D_HIP_ERROR);
  }
  if (!zalloc((void**)&s->ws_part, (size_t)WS_SLOTS * WS_STRIDE * sizeof(unsigned long long)))
    return cleanup(HF2D_HIP_ERROR);
  if (hipMalloc(&s->ctrl, sizeof(Ctrl)) != hipSuccess) return cleanup(HF2D_HIP_ERROR);
  Ctrl c{};
  c.t_target = 0.0;
  c.dt_cap = __builtin_inf();
  c.compensated = scalar_kind == HF2D_F32 ? 1 : 0;
  if (hipMemcpy(s->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice) != hipSuccess) return cleanup(HF2D_HIP_ERROR);
  *out = s;
  return HF2D_OK;
}

int hf2d_destroy(hf2d_sim* s) {
  if (!s) return HF2D_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graph(s);
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < 3; ++f) {
      if (s->st[b][f]) (void)hipFree(s->st[b][f]);
      if (s->tg[b][f]) (void)hipFree(s->tg[b][f]);
    }
  void* bufs[] = {s->z,        s->zt,       s->man,      s->mant,     s->datum,    s->zdev,
                  s->ctrl,     s->srcs,     s->tile_wet, s->tile_run, s->tile_pin, s->tile_act,
                  s->tile_list, s->series_buf[0], s->series_buf[1], s->series_dev, s->gauge_cells,
                  s->gauge_buf, s->gauge_count, s->ws_part};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return HF2D_OK;
}

const char* hf2d_last_error(const hf2d_sim* s) { return s ? s->err.c_str() : "null handle"; }

int hf2d_upload_mesh(hf2d_sim* s, const double* bed, const double* bed_tangent, const double* manning,
                     const double* manning_tangent) {
  if (!valid(s) || !bed || !manning) return fail(s, HF2D_INVALID_ARGUMENT, "null mesh array");
  const size_t n = (size_t)(s->nr * s->nc);
  for (size_t k = 0; k < n; ++k)
    if (!(manning[k] >= 0.0)) return fail(s, HF2D_INVALID_CONFIG, "Manning n must be non-negative at every cell");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const size_t bytes = n * sizeof(double);
  if (s->kind == HF2D_F32) {
    // Datum per 64x64 block: the block mean in FP64; deviation rounded to FP32.
    const int64_t B = 1 << DATUM_LOG2, rows = (s->nr + B - 1) / B;
    std::vector<double> dat((size_t)(rows * s->datum_cols), 0.0);
    for (int64_t bi = 0; bi < rows; ++bi)
      for (int64_t bj = 0; bj < s->datum_cols; ++bj) {
        double sum = 0.0;
        int64_t cnt = 0;
        for (int64_t i = bi * B; i < std::min(s->nr, (bi + 1) * B); ++i)
          for (int64_t j = bj * B; j < std::min(s->nc, (bj + 1) * B); ++j, ++cnt) sum += bed[i * s->nc + j];
        dat[(size_t)(bi * s->datum_cols + bj)] = sum / (double)cnt;
      }
    std::vector<float> dev(n), man(n);
    for (int64_t i = 0; i < s->nr; ++i)
      for (int64_t j = 0; j < s->nc; ++j) {
        const size_t k = (size_t)(i * s->nc + j);
        dev[k] = (float)(bed[k] - dat[(size_t)((i >> DATUM_LOG2) * s->datum_cols + (j >> DATUM_LOG2))]);
        man[k] = (float)manning[k];
      }
    HIPCHK(s, hipMemcpy(s->datum, dat.data(), dat.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->zdev, dev.data(), n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->man, man.data(), n * sizeof(float), hipMemcpyHostToDevice));
  } else {
    HIPCHK(s, hipMemcpy(s->z, bed, bytes, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->man, manning, bytes, hipMemcpyHostToDevice));
  }
  if (s->n_tan > 0) {  // tangents: n_tan per cell, [cell][k]
    const size_t tb = bytes * (size_t)s->n_tan;
    s->has_zt = bed_tangent != nullptr;
    if (bed_tangent)
      HIPCHK(s, hipMemcpy(s->zt, bed_tangent, tb, hipMemcpyHostToDevice));
    else
      HIPCHK(s, hipMemset(s->zt, 0, tb));
    if (manning_tangent)
      HIPCHK(s, hipMemcpy(s->mant, manning_tangent, tb, hipMemcpyHostToDevice));
    else
      HIPCHK(s, hipMemset(s->mant, 0, tb));
  }
  state_changed(s);
  drop_graph(s);
  return HF2D_OK;
}

namespace {
// Copy one AoS component (stride 3) into a device plane of the grid's kind.
int put_plane(hf2d_sim* s, void* dst, const double* aos, int f) {
  const size_t n = (size_t)(s->nr * s->nc);
  if (s->kind == HF2D_F32) {
    std::vector<float> plane(n);
    for (size_t k = 0; k < n; ++k) plane[k] = aos ? (float)aos[3 * k + f] : 0.0f;
    HIPCHK(s, hipMemcpy(dst, plane.data(), n * sizeof(float), hipMemcpyHostToDevice));
  } else {
    std::vector<double> plane(n);
    for (size_t k = 0; k < n; ++k) plane[k] = aos ? aos[3 * k + f] : 0.0;
    HIPCHK(s, hipMemcpy(dst, plane.data(), n * sizeof(double), hipMemcpyHostToDevice));
  }
  return HF2D_OK;
}
// Tangents of field f between the AoS layout [cell][field][k] (k < n_tan)
// and the device buffer [cell][k].
int put_tangents(hf2d_sim* s, void* dst, const double* aos, int f) {
  const size_t n = (size_t)(s->nr * s->nc), m = (size_t)s->n_tan;
  std::vector<double> buf(n * m, 0.0);
  if (aos)
    for (size_t c = 0; c < n; ++c)
      for (size_t k = 0; k < m; ++k) buf[c * m + k] = aos[(c * 3 + (size_t)f) * m + k];
  HIPCHK(s, hipMemcpy(dst, buf.data(), n * m * sizeof(double), hipMemcpyHostToDevice));
  return HF2D_OK;
}
int get_tangents(hf2d_sim* s, const void* src, double* aos, int f) {
  const size_t n = (size_t)(s->nr * s->nc), m = (size_t)s->n_tan;
  std::vector<double> buf(n * m);
  HIPCHK(s, hipMemcpy(buf.data(), src, n * m * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t c = 0; c < n; ++c)
    for (size_t k = 0; k < m; ++k) aos[(c * 3 + (size_t)f) * m + k] = buf[c * m + k];
  return HF2D_OK;
}
int get_plane(hf2d_sim* s, const void* src, double* aos, int f) {
  const size_t n = (size_t)(s->nr * s->nc);
  if (s->kind == HF2D_F32) {
    std::vector<float> plane(n);
    HIPCHK(s, hipMemcpy(plane.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; ++k) aos[3 * k + f] = (double)plane[k];
  } else {
    std::vector<double> plane(n);
    HIPCHK(s, hipMemcpy(plane.data(), src, n * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; ++k) aos[3 * k + f] = plane[k];
  }
  return HF2D_OK;
}
}  // namespace

int hf2d_upload_state(hf2d_sim* s, const double* aos, const double* tan_aos) {
  if (!valid(s) || !aos) return fail(s, HF2D_INVALID_ARGUMENT, "null state array");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (int f = 0; f < 3; ++f) {
    int rc = put_plane(s, s->st[s->cur][f], aos, f);
    if (rc) return rc;
    if (s->n_tan > 0) {
      rc = put_tangents(s, s->tg[s->cur][f], tan_aos, f);
      if (rc) return rc;
    }
  }
  state_changed(s);
  drop_graph(s);
  return HF2D_OK;
}

int hf2d_download_state(hf2d_sim* s, double* aos, double* tan_aos) {
  if (!valid(s) || !aos) return fail(s, HF2D_INVALID_ARGUMENT, "null state array");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  for (int f = 0; f < 3; ++f) {
    int rc = get_plane(s, s->st[s->cur][f], aos, f);
    if (rc) return rc;
    if (s->n_tan > 0 && tan_aos) {
      rc = get_tangents(s, s->tg[s->cur][f], tan_aos, f);
      if (rc) return rc;
    }
  }
  return HF2D_OK;
}

namespace {
constexpr int VOL_BLOCKS = 1024;  // volume_kernel workgroups: one (sum, compensation) pair each

// Enqueue the volume reduction of the current state into part (2 *
// VOL_BLOCKS doubles on the handle's device).
hipError_t enqueue_volume(hf2d_sim* s, double* part) {
  const int64_t n = s->nr * s->nc;
  if (s->kind == HF2D_F32)
    hipLaunchKernelGGL(volume_kernel<float>, dim3(VOL_BLOCKS), dim3(256), 0, s->stream,
                       fld<float>(s->st[s->cur][0], nullptr), n, part);
  else  // f64, and the value plane alone of the Dual kinds
    hipLaunchKernelGGL(volume_kernel<double>, dim3(VOL_BLOCKS), dim3(256), 0, s->stream,
                       fld<double>(s->st[s->cur][0], nullptr), n, part);
  return hipGetLastError();
}

// Neumaier sum of the block pairs in a fixed order, times the cell area.
double fold_volume(const hf2d_sim* s, const double* hp) {
  double sum = 0.0, comp = 0.0;
  for (int k = 0; k < 2 * VOL_BLOCKS; ++k) {
    const double x = hp[k], t = sum + x;
    comp += (__builtin_fabs(sum) >= __builtin_fabs(x)) ? (sum - t) + x : (x - t) + sum;
    sum = t;
  }
  return (sum + comp) * s->dx * s->dy;
}

}  // namespace

// Scratch of hf2d_total_volume_noalloc: the per-workgroup partial sums, in a
// file-scope device array (16 KiB, whatever the grid's size) instead of a
// buffer allocated per call. Each device holds its own copy until the
// process ends. In host code `scratch` names a host-side stand-in, not the
// device array: its device address is looked up once per device
// (scratch_on) and only that address is used.
static __device__ __align__(256) double scratch[2 * 1024];

namespace {
static_assert(sizeof(scratch) == 2 * VOL_BLOCKS * sizeof(double), "one (sum, compensation) pair per workgroup");
constexpr int MAX_DEVICES = 64;
double* g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// Device address of `scratch` on the given device (looked up with that
// device current, then the caller's current device is restored).
hipError_t scratch_on(int device, double** out) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

int hf2d_total_volume(hf2d_sim* s, double* volume) {
  if (!valid(s) || !volume) return fail(s, HF2D_INVALID_ARGUMENT, "null argument");
  double* part = nullptr;
  HIPCHK(s, hipMalloc(&part, 2 * VOL_BLOCKS * sizeof(double)));
  std::vector<double> hp(2 * VOL_BLOCKS);
  hipError_t e = enqueue_volume(s, part);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hp.data(), part, 2 * VOL_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(s, HF2D_HIP_ERROR, hipGetErrorString(e));
  *volume = fold_volume(s, hp.data());
  return HF2D_OK;
}

int hf2d_total_volume_noalloc(hf2d_sim* s, double* volume) {
  if (!valid(s) || !volume) return fail(s, HF2D_INVALID_ARGUMENT, "null argument");
  double* part = nullptr;
  HIPCHK(s, scratch_on(s->device, &part));
  double hp[2 * VOL_BLOCKS];  // 16 KiB on the caller's stack, no heap
  HIPCHK(s, enqueue_volume(s, part));
  HIPCHK(s, hipMemcpyAsync(hp, part, sizeof hp, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  *volume = fold_volume(s, hp);
  return HF2D_OK;
}

int hf2d_set_boundaries(hf2d_sim* s, const int32_t kinds[4], const double params[4]) {
  if (!valid(s) || !kinds) return fail(s, HF2D_INVALID_ARGUMENT, "null boundary array");
  int rc = check_bcs_geometry(s, kinds);
  if (rc) return rc;
  for (int k = 0; k < 4; ++k) {
    s->bc_kind[k] = kinds[k];
    s->bc_param[k] = params ? params[k] : 0.0;
  }
  state_changed(s);
  drop_graph(s);
  return HF2D_OK;
}

int hf2d_max_wave_speeds(hf2d_sim* s, double* sx, double* sy) {
  if (!valid(s) || !sx || !sy) return fail(s, HF2D_INVALID_ARGUMENT, "null argument");
  return HF_DISPATCH(s, do_max_wave_speeds, s, false, sx, sy);
}

int hf2d_cfl_time_step(hf2d_sim* s, double cfl, int32_t with_bcs, double* dt) {
  if (!valid(s) || !dt) return fail(s, HF2D_INVALID_ARGUMENT, "null argument");
  double sx, sy;
  int rc = HF_DISPATCH(s, do_max_wave_speeds, s, with_bcs != 0, &sx, &sy);
  if (rc) return rc;
  const double denom = sx / s->dx + sy / s->dy;  // update.rs:150-155
  *dt = denom > 0.0 ? cfl / denom : __builtin_inf();
  return HF2D_OK;
}

int hf2d_forward_euler_step(hf2d_sim* s, double dt) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  return HF_DISPATCH(s, host_dt_step, s, dt, false);
}

int hf2d_ssprk2_step(hf2d_sim* s, double dt) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  return HF_DISPATCH(s, host_dt_step, s, dt, true);
}


int hf2d_manning_friction_step(hf2d_sim* s, double dt, double dry_tol) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  return HF_DISPATCH(s, forcing_only, s, dt, 1, dry_tol, 0.0);
}

int hf2d_apply_rain(hf2d_sim* s, double rate, double dt) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  if (rate == 0.0) return HF2D_OK;  // source.rs:89-91
  return HF_DISPATCH(s, forcing_only, s, dt, 0, 0.0, rate);
}

int hf2d_apply_point_sources(hf2d_sim* s, const hf2d_point_source* src, int64_t k, double dt) {
  if (!valid(s) || (k > 0 && !src)) return fail(s, HF2D_INVALID_ARGUMENT, "null sources");
  if (k == 0) return HF2D_OK;
  // Temporarily swap in this list; the configured driver list is restored.
  std::vector<hf2d_point_source> keep;
  int rc = HF2D_OK;
  PointSrc* saved = s->srcs;
  int saved_n = s->n_src, saved_cap = s->cap_src;
  s->srcs = nullptr;
  s->n_src = 0;
  s->cap_src = 0;
  rc = upload_sources(s, src, k);
  if (rc == HF2D_OK) {
    HF_DISPATCH(s, launch_point_sources, s, dt);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);  // the list is freed below
    if (e != hipSuccess) rc = fail(s, HF2D_HIP_ERROR, hipGetErrorString(e));
  }
  if (s->srcs) (void)hipFree(s->srcs);
  s->srcs = saved;
  s->n_src = saved_n;
  s->cap_src = saved_cap;
  state_changed(s);
  drop_graph(s);
  return rc;
}

int hf2d_configure(hf2d_sim* s, const hf2d_config* cfg, const hf2d_point_source* src, int64_t k) {
  if (!valid(s) || !cfg) return fail(s, HF2D_INVALID_ARGUMENT, "null config");
  if (!(cfg->cfl > 0.0 && cfg->cfl <= 1.0)) {
    char buf[128];
    snprintf(buf, sizeof buf, "invalid configuration: cfl must be in (0, 1], got %g", cfg->cfl);
    return fail(s, HF2D_INVALID_CONFIG, buf);
  }
  if (!(cfg->max_dt > 0.0)) {
    char buf[128];
    snprintf(buf, sizeof buf, "invalid configuration: max_dt must be positive, got %g", cfg->max_dt);
    return fail(s, HF2D_INVALID_CONFIG, buf);
  }
  if (cfg->integrator != HF2D_SSPRK2 && cfg->integrator != HF2D_FORWARD_EULER)
    return fail(s, HF2D_INVALID_CONFIG, "invalid configuration: unknown integrator");
  if (k > 0 && !src) return fail(s, HF2D_INVALID_ARGUMENT, "null sources");
  int rc = upload_sources(s, src, k);
  if (rc) return rc;
  rc = upload_pins(s, src, k);
  if (rc) return rc;
  s->cfg = *cfg;
  state_changed(s);
  drop_graph(s);
  return HF2D_OK;
}

int hf2d_set_forcing_series(hf2d_sim* s, int32_t which, int32_t n, const double* t, const double* v,
                            double time_div, double scale_mul, double scale_div) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  if (which != HF2D_SERIES_RAIN && which != HF2D_SERIES_SOURCES)
    return fail(s, HF2D_INVALID_ARGUMENT, "unknown forcing series");
  if (n < 0 || (n > 0 && (!t || !v))) return fail(s, HF2D_INVALID_ARGUMENT, "null series arrays");
  if (n > 0 && !(time_div > 0.0 && scale_div != 0.0 && std::isfinite(scale_mul)))
    return fail(s, HF2D_INVALID_CONFIG, "invalid series scaling");
  for (int32_t k = 0; k < n; ++k)
    if (!std::isfinite(t[k]) || !std::isfinite(v[k]) || (k > 0 && !(t[k] >= t[k - 1])))
      return fail(s, HF2D_INVALID_CONFIG, "series times must be finite and non-decreasing");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (s->series_buf[which]) HIPCHK(s, hipFree(s->series_buf[which]));
  s->series_buf[which] = nullptr;
  SeriesDesc d{};
  if (n > 0) {
    HIPCHK(s, hipMalloc((void**)&s->series_buf[which], 2 * (size_t)n * sizeof(double)));
    HIPCHK(s, hipMemcpy(s->series_buf[which], t, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->series_buf[which] + n, v, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    d.n = n;
    d.time_div = time_div;
    d.scale_mul = scale_mul;
    d.scale_div = scale_div;
    d.t = s->series_buf[which];
    d.v = s->series_buf[which] + n;
  }
  s->series_host[which] = d;
  if (!s->series_dev) HIPCHK(s, hipMalloc((void**)&s->series_dev, 2 * sizeof(SeriesDesc)));
  HIPCHK(s, hipMemcpy(s->series_dev, s->series_host, 2 * sizeof(SeriesDesc), hipMemcpyHostToDevice));
  s->list_valid = false;
  drop_graph(s);
  return HF2D_OK;
}

int hf2d_set_gauges(hf2d_sim* s, int32_t n, const int64_t* rows, const int64_t* cols, int64_t capacity) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  drop_graph(s);
  if (n < 0 || (n > 0 && (!rows || !cols)) || (n > 0 && capacity < 1))
    return fail(s, HF2D_INVALID_ARGUMENT, "invalid gauge arguments");
  std::vector<int64_t> cells((size_t)n);
  for (int32_t k = 0; k < n; ++k) {
    if (rows[k] < 0 || rows[k] >= s->nr || cols[k] < 0 || cols[k] >= s->nc)
      return fail(s, HF2D_SHAPE_MISMATCH, "gauge cell outside the grid");
    cells[(size_t)k] = rows[k] * s->nc + cols[k];
  }
  HIPCHK(s, hipStreamSynchronize(s->stream));
  void* old[] = {s->gauge_cells, s->gauge_buf, s->gauge_count};
  for (void* p : old)
    if (p) HIPCHK(s, hipFree(p));
  s->gauge_cells = nullptr;
  s->gauge_buf = nullptr;
  s->gauge_count = nullptr;
  s->n_gauges = 0;
  s->gauge_cap = 0;
  if (n == 0) return HF2D_OK;
  HIPCHK(s, hipMalloc((void**)&s->gauge_cells, (size_t)n * sizeof(int64_t)));
  HIPCHK(s, hipMalloc((void**)&s->gauge_buf, (size_t)capacity * (size_t)(n + 1) * sizeof(double)));
  HIPCHK(s, hipMalloc((void**)&s->gauge_count, sizeof(long long)));
  HIPCHK(s, hipMemcpy(s->gauge_cells, cells.data(), (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice));
  const long long zero = 0;
  HIPCHK(s, hipMemcpy(s->gauge_count, &zero, sizeof zero, hipMemcpyHostToDevice));
  s->n_gauges = n;
  s->gauge_cap = capacity;
  return HF_DISPATCH(s, record_gauges_now, s);
}

int hf2d_read_gauges(hf2d_sim* s, int64_t max_rows, double* out, int64_t* n_rows) {
  if (!valid(s) || !n_rows) return HF2D_INVALID_ARGUMENT;
  long long count = 0;
  if (s->n_gauges > 0) {
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(&count, s->gauge_count, sizeof count, hipMemcpyDeviceToHost));
  }
  *n_rows = count;
  const long long k = count < max_rows ? count : max_rows;
  if (out && k > 0)
    HIPCHK(s, hipMemcpy(out, s->gauge_buf, (size_t)k * (size_t)(s->n_gauges + 1) * sizeof(double),
                        hipMemcpyDeviceToHost));
  return HF2D_OK;
}

int hf2d_step(hf2d_sim* s, double dt_cap, double* dt_taken) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  int rc = HF_DISPATCH(s, run_driver, s, false, 0.0, dt_cap, 1, true);
  if (dt_taken) {
    Ctrl c;
    if (read_ctrl(s, &c) == HF2D_OK) *dt_taken = c.dt;
  }
  return rc;
}

int hf2d_run_until(hf2d_sim* s, double t_target) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  return HF_DISPATCH(s, run_driver, s, true, t_target, 0.0, 0, true);
}

int hf2d_run_steps(hf2d_sim* s, int64_t n) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  if (n <= 0) return HF2D_OK;
  return HF_DISPATCH(s, run_driver, s, false, 0.0, __builtin_inf(), n, true);
}

int hf2d_enqueue_steps(hf2d_sim* s, int64_t n) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  if (n <= 0) return HF2D_OK;
  return HF_DISPATCH(s, run_driver, s, false, 0.0, __builtin_inf(), n, false);
}

int hf2d_synchronize(hf2d_sim* s) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  Ctrl c;
  int rc = read_ctrl(s, &c);
  if (rc) return rc;
  rc = harvest_timing(s);
  if (rc) return rc;
  if (!s->driver_async) return HF2D_OK;  // nothing of the driver's to report
  s->driver_async = false;
  s->ws_valid = (c.status == 0);
  s->ws_with_bcs = s->cfg.cfl_with_bcs != 0;
  return ctrl_error(s, c, c.t_target);
}

int hf2d_get_time(hf2d_sim* s, double* t, int64_t* steps) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  Ctrl c;
  int rc = read_ctrl(s, &c);
  if (rc) return rc;
  if (t) *t = c.t;
  if (steps) *steps = c.steps;
  return HF2D_OK;
}

int hf2d_set_time(hf2d_sim* s, double t, int64_t steps) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  Ctrl c;
  int rc = read_ctrl(s, &c);
  if (rc) return rc;
  c.t = t;
  c.steps = steps;
  return write_ctrl(s, &c);
}

int hf2d_set_kernel_timing(hf2d_sim* s, int32_t enabled) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  int rc = harvest_timing(s);
  if (rc) return rc;
  s->timing = enabled != 0;
  drop_graph(s);
  s->timed_ms[0] = s->timed_ms[1] = 0.0;
  s->timed_count = 0;
  s->ev_used = 0;
  return HF2D_OK;
}

int hf2d_kernel_timing(hf2d_sim* s, double stage_ms[2], int64_t* count) {
  if (!valid(s)) return HF2D_INVALID_ARGUMENT;
  int rc = harvest_timing(s);
  if (rc) return rc;
  if (stage_ms) {
    stage_ms[0] = s->timed_ms[0];
    stage_ms[1] = s->timed_ms[1];
  }
  if (count) *count = s->timed_count;
  return HF2D_OK;
}

int hf2d_copy_probe(int64_t n_bytes, int32_t reps, double* ms_avg) {
  if (n_bytes <= 0 || reps <= 0 || !ms_avg) return HF2D_INVALID_ARGUMENT;
  const int64_t n2 = n_bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, (size_t)n2 * 16) != hipSuccess) return HF2D_HIP_ERROR;
  if (hipMalloc(&b, (size_t)n2 * 16) != hipSuccess) {
    (void)hipFree(a);
    return HF2D_HIP_ERROR;
  }
  (void)hipMemset(a, 0, (size_t)n2 * 16);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int64_t grid = (n2 + 256 * COPY_UNROLL - 1) / (256 * COPY_UNROLL);
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)grid), dim3(256), 0, 0, a, b, n2);  // warm
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(copy_kernel, dim3((unsigned)grid), dim3(256), 0, 0, a, b, n2);
  (void)hipEventRecord(e1, 0);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  *ms_avg = (double)ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return e == hipSuccess ? HF2D_OK : HF2D_HIP_ERROR;
}

int hf2d_math_probe(const double* x, const double* y, int64_t n, double* out) {
  if (!x || !y || !out || n <= 0) return HF2D_INVALID_ARGUMENT;
  double *dx_ = nullptr, *dy_ = nullptr, *dout = nullptr;
  const size_t b = (size_t)n * sizeof(double);
  hipError_t e = hipMalloc(&dx_, b);
  if (e == hipSuccess) e = hipMalloc(&dy_, b);
  if (e == hipSuccess) e = hipMalloc(&dout, 4 * b);
  if (e == hipSuccess) e = hipMemcpy(dx_, x, b, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dy_, y, b, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(math_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx_, dy_, n, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dout, 4 * b, hipMemcpyDeviceToHost);
  if (dx_) (void)hipFree(dx_);
  if (dy_) (void)hipFree(dy_);
  if (dout) (void)hipFree(dout);
  return e == hipSuccess ? HF2D_OK : HF2D_HIP_ERROR;
}

}  // extern "C"
