// Device kernels shared by the single-grid driver (hf2d.hip) and the batched
// engine (hf2d_batch.hip): the fused stage kernel, the elementwise tail
// (forcing + CFL maxima), point sources, the clock finalize, the volume
// reduction and the measurement probes.
#pragma once
#include <hip/hip_runtime.h>

#include "hf2d_tile.h"

using hf::Ctrl;
using hf::dual;
using hf::Fld;
using hf::FldW;
using hf::PointSrc;
using hf::SeriesDesc;
using hf::StageArgs;

namespace {

// ------------------------------------------------------------ tile shapes
template <class T> struct Shape;
#ifndef HF_F64_TILE_W
#define HF_F64_TILE_W 32
#endif
#ifndef HF_F64_TILE_H
#define HF_F64_TILE_H 24
#endif
#ifndef HF_F64_THREADS
#define HF_F64_THREADS 768
#endif
#ifndef HF_XCD_REMAP
#define HF_XCD_REMAP 1
#endif

// Tiles for small grids. With fewer tiles than CUs, a step lasts about one
// tile's phases, latency-bound, and a smaller tile is faster. A grid takes
// the default shape when it gives at least SMALL_GRID_TILES tiles (FP64 on
// the flood workload: the default wins from 192 tiles on, 16 x 8 at 88);
// else ShapeMid when that gives at least MID_GRID_TILES tiles; else
// ShapeSmall. Measured FP64 us/step over full runs, 16x8x512 / 32x16x1024:
// UK EA Test 4 (400 x 200) 40.3 / 36.9, Test 8A (199 x 481) 45.7 / 40.5,
// Huasco reach (200 x 67) 23.7 / 33.0 (32x24x1024, earlier: 20-40 % slower
// than the better of the two on all three).
template <class T> struct ShapeSmall {
  static constexpr bool on = false;
};
template <class T> struct ShapeMid {
  static constexpr bool on = false;
};
#ifndef HF_F64_MID_W
#define HF_F64_MID_W 32
#endif
#ifndef HF_F64_MID_H
#define HF_F64_MID_H 16
#endif
#ifndef HF_F64_MID_THREADS
#define HF_F64_MID_THREADS 1024
#endif
template <> struct ShapeMid<double> {
  static constexpr bool on = true;
  static constexpr int TW = HF_F64_MID_W, TH = HF_F64_MID_H, NT = HF_F64_MID_THREADS;
};
template <> struct ShapeSmall<double> {
  static constexpr bool on = true;
  static constexpr int TW = 16, TH = 8, NT = 512;
};
template <> struct ShapeSmall<dual> {
  static constexpr bool on = true;
  static constexpr int TW = 16, TH = 8, NT = 512;
};
template <> struct ShapeSmall<float> {
  static constexpr bool on = true;
  static constexpr int TW = 16, TH = 8, NT = 512;
};
#ifndef HF_SMALL_GRID_TILES
#define HF_SMALL_GRID_TILES 160
#endif
#ifndef HF_MID_GRID_TILES
#define HF_MID_GRID_TILES 128
#endif
constexpr long long SMALL_GRID_TILES = HF_SMALL_GRID_TILES;
constexpr long long MID_GRID_TILES = HF_MID_GRID_TILES;

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b lands on
// XCD b % 8), each with its own L2. Renumber tiles so every XCD walks one
// contiguous band of tiles: horizontally adjacent tiles, which share halo
// columns, then run on the same XCD at about the same time and the halo is
// an L2 hit instead of a second HBM fetch.
__device__ inline unsigned tile_of_block(unsigned b, unsigned nb) {
#if HF_XCD_REMAP
  const unsigned q = nb / 8u, r = nb % 8u, x = b % 8u;
  return x * q + (x < r ? x : r) + b / 8u;
#else
  (void)nb;
  return b;
#endif
}
template <> struct Shape<double> {
  static constexpr int TW = HF_F64_TILE_W, TH = HF_F64_TILE_H, NT = HF_F64_THREADS;
};
// Dual: 24 x 16 tiles with 1024 threads, 152 KB of LDS, one workgroup and
// four waves per SIMD per CU; 2.55 faces computed per cell against 3.03 on
// 16 x 8 tiles. Measured on the flood workload (G cell-updates/s, 1024^2 /
// 4096^2): 16 x 8 x 512 (two workgroups per CU) 1.97 / 2.13, 32 x 12 x 768
// 2.38 / 2.54, 32 x 12 x 1024 2.43 / 2.57, 24 x 16 x 768 2.40 / 2.54,
// 24 x 16 x 1024 2.45 / 2.59; the persistent form on 32 x 12 2.14-2.22 /
// 2.38-2.52. Grids of fewer than SMALL_GRID_TILES such tiles take 16 x 8.
#ifndef HF_DUAL_TILE_W
#define HF_DUAL_TILE_W 24
#endif
#ifndef HF_DUAL_TILE_H
#define HF_DUAL_TILE_H 16
#endif
#ifndef HF_DUAL_THREADS
#define HF_DUAL_THREADS 1024
#endif
template <> struct Shape<dual> {
  static constexpr int TW = HF_DUAL_TILE_W, TH = HF_DUAL_TILE_H, NT = HF_DUAL_THREADS;
};
#ifndef HF_DUAL2_THREADS
#define HF_DUAL2_THREADS 768
#endif
#ifndef HF_DUAL4_THREADS
#define HF_DUAL4_THREADS 512
#endif
#ifndef HF_DUAL4_TILE_W
#define HF_DUAL4_TILE_W 16
#endif
#ifndef HF_DUAL4_TILE_H
#define HF_DUAL4_TILE_H 8
#endif
// DualN<N>: (N + 1) doubles per LDS value; tiles sized to stay within one
// workgroup's LDS. For N > 4 the stage kernels run as N / 4 launches of the
// DualN<4> kernel, one per group of 4 tangent components (a DualN<8> tile
// of its own would be 8 x 4 cells at 1 workgroup of 2 waves per CU, with 4.4
// cells loaded per cell updated), so its tiles are DualN<4>'s.
template <int N> struct Shape<hf::dualn<N>> {
  static_assert(N > 4 && N % 4 == 0, "DualN<N> wider than 4 runs in groups of 4 components");
  static constexpr int TW = HF_DUAL4_TILE_W, TH = HF_DUAL4_TILE_H, NT = HF_DUAL4_THREADS;
};
// DualN<2>: 24 x 8 tiles with 768 threads (134 KB of LDS, one workgroup and
// three waves per SIMD per CU), 2.9 faces per cell against 3.03 on 16 x 8.
// Measured at 2048^2 (flood workload, G cell-updates/s): 16 x 8 x 512 1.12,
// 16 x 12 x 768 1.42, 24 x 8 x 768 1.43. Grids of fewer than
// SMALL_GRID_TILES such tiles take 16 x 8.
// DualN<4> (and the groups of 4 of wider DualN): 16 x 8 tiles with 512
// threads (159 KB of LDS; 246-256 VGPRs, so one workgroup of two waves per
// SIMD per CU whatever the tile). Against 8 x 8 tiles, which load 3.1
// cells per cell updated and keep one wave busy in the update phase:
// DualN<16> at 1024^2 9.20 -> 6.12 ms per step, DualN<4> at 256^2 0.177 ->
// 0.112 ms, 12 x 8 in between. On these tiles groups of 4 also beat pairs
// of components on the DualN<2> kernel (2048^2: DualN<4> 5.84 -> 5.00 ms,
// DualN<8> 12.5 -> 10.3 ms).
#ifndef HF_DUAL2_TILE_W
#define HF_DUAL2_TILE_W 24
#endif
#ifndef HF_DUAL2_TILE_H
#define HF_DUAL2_TILE_H 8
#endif
template <> struct Shape<hf::dualn<2>> {
  static constexpr int TW = HF_DUAL2_TILE_W, TH = HF_DUAL2_TILE_H, NT = HF_DUAL2_THREADS;
};
template <> struct ShapeSmall<hf::dualn<2>> {
  static constexpr bool on = true;
  static constexpr int TW = 16, TH = 8, NT = 512;
};
template <> struct Shape<hf::dualn<4>> {
  static constexpr int TW = HF_DUAL4_TILE_W, TH = HF_DUAL4_TILE_H, NT = HF_DUAL4_THREADS;
};
// Grids of fewer than SMALL_GRID_TILES 16 x 8 tiles: DualN<4> and the
// groups of 4 of wider DualN take 8 x 8 (64^2, one DualN<P> pass against
// the f64 primal, cost per parameter P = 4 / 8 / 16: 0.27 / 0.37 / 0.43 on
// 8 x 8, 0.29 / 0.40 / 0.46 on 16 x 8).
template <int N> struct ShapeSmall<hf::dualn<N>> {
  static constexpr bool on = N >= 4;
  static constexpr int TW = 8, TH = 8, NT = HF_DUAL4_THREADS;
};
// FP32 tiles lie inside one 64 x 64 bed-datum block (shapes that divide
// 64), so its results are the same for all of them; a 64 x 24 tile straddles
// blocks and rounds the bed differently. 32 x 16 with 512 threads (two to
// three workgroups per CU) is 1.4 % faster than 32 x 32 with 1024.
#ifndef HF_F32_TILE_W
#define HF_F32_TILE_W 32
#endif
#ifndef HF_F32_TILE_H
#define HF_F32_TILE_H 16
#endif
#ifndef HF_F32_THREADS
#define HF_F32_THREADS 512
#endif
template <> struct Shape<float> {
  static constexpr int TW = HF_F32_TILE_W, TH = HF_F32_TILE_H, NT = HF_F32_THREADS;
};

// CFL maxima slots of the driver's final stage: 64-bit atomic maxima of one
// address from every workgroup serialize (device-scope atomics are resolved
// outside the per-XCD L2s), so the workgroups of a launch spread over
// WS_SLOTS addresses, one 128-byte line apart, and finalize_kernel folds
// them. Max is associative, so the result is the single-address one.
constexpr int WS_SLOTS = 512;
constexpr int WS_STRIDE = 16;  // u64 per slot
// The batched engine's final stage spreads each scenario's atomics over
// BATCH_WS_SLOTS slots (one wavefront folds them per step), so a large
// scenario's thousands of workgroups do not serialize on one address.
constexpr int BATCH_WS_SLOTS = 64;

// max(v, v of the lane that DPP control CTRL names), in the lanes of the rows
// ROW_MASK enables (the others keep v). fmax ignores a NaN operand.
template <int CTRL, int ROW_MASK = 0xf> __device__ inline double dpp_fmax(double v) {
  const unsigned long long u = hfm::as_u64(v);
  const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
  return __builtin_fmax(v, hfm::as_f64(((unsigned long long)(unsigned)hi2 << 32) | (unsigned)lo2));
}

// Max over each row of 16 lanes, in every lane of the row: xor 1, xor 2
// (quad permutations), then the half-row and row mirrors.
__device__ inline double row_fmax(double v) {
  v = dpp_fmax<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_fmax<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_fmax<0x141>(v);  // row_half_mirror
  return dpp_fmax<0x140>(v);  // row_mirror
}

// Max over the wavefront (all 64 lanes active), uniform: row maxima, then
// row 1 and 3 take lane 15 / 47 (row_bcast:15), rows 2 and 3 take lane 31
// (row_bcast:31); lane 63 then holds the maximum.
__device__ inline double wave_fmax(double v) {
  v = row_fmax(v);
  v = dpp_fmax<0x142, 0xA>(v);
  v = dpp_fmax<0x143, 0xC>(v);
  const unsigned long long u = hfm::as_u64(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 63);
  return hfm::as_f64(((unsigned long long)hi << 32) | lo);
}

// Workgroup maxima of (sx, sy) and whether any thread saw a wet cell,
// valid in thread 0 (red: per-wave partials, 3 rows when with_wet). The
// lane folds are DPP moves, not LDS permutes.
__device__ inline void block_max(double sx, double sy, double (*red)[16], bool wet, bool with_wet, double* bx,
                                 double* by, bool* bw) {
  const int tid = threadIdx.x;
  sx = wave_fmax(sx);
  sy = wave_fmax(sy);
  const int wave = tid >> 6, lane = tid & 63;
  const bool wave_wet = with_wet ? __any(wet ? 1 : 0) != 0 : false;
  if (lane == 0) {
    red[0][wave] = sx;
    red[1][wave] = sy;
    if (with_wet) red[2][wave] = wave_wet ? 1.0 : 0.0;
  }
  __syncthreads();
  if (tid < 64) {  // wave 0 folds the (at most 16) wave partials, one per lane of row 0
    const int nw = (blockDim.x + 63) >> 6;
    const int k = lane < nw ? lane : nw - 1;  // a repeated partial leaves the max unchanged
    double x = red[0][k], y = red[1][k];
    const bool w = with_wet ? __any(red[2][k] != 0.0 ? 1 : 0) != 0 : false;
    x = row_fmax(x);
    y = row_fmax(y);
    *bx = x;
    *by = y;
    *bw = w;
  }
}

// block_max into dst (64-bit atomic max of the non-negative doubles' bit
// patterns) and, with wet_dst, the wet flag (one byte).
__device__ inline void block_max_to(double sx, double sy, double (*red)[16], unsigned long long* dst,
                                    bool wet = false, unsigned char* wet_dst = nullptr) {
  double x = 0.0, y = 0.0;
  bool w = false;
  block_max(sx, sy, red, wet, wet_dst != nullptr, &x, &y, &w);
  if (threadIdx.x == 0) {
    atomicMax(dst + 0, hfm::as_u64(x));
    atomicMax(dst + 1, hfm::as_u64(y));
    if (wet_dst) *wet_dst = w ? 1 : 0;
  }
}

// One stage of one TW x TH tile per workgroup of NT threads.
template <class T, int TW, int TH, int NT, int MODE>
__global__ __launch_bounds__(NT) void stage_kernel(StageArgs<T> a) {
  using St = hf::Stage<T, TW, TH>;
  __shared__ typename St::S s;
  double dt;
  if (!hf::stage_prologue(a, blockIdx.x == 0 && threadIdx.x == 0, &dt)) return;
  unsigned tile;
  if (a.tile_list) {  // active-tile list: positions past its length have no work
    const unsigned n = (unsigned)*a.n_work;
    if (blockIdx.x >= n) return;
    tile = (unsigned)a.tile_list[tile_of_block(blockIdx.x, n)];
  } else {
    tile = tile_of_block(blockIdx.x, gridDim.x);
  }
  const int ty = (int)(tile / (unsigned)a.tiles_x), tx = (int)(tile % (unsigned)a.tiles_x);
  const int gi0 = ty * TH, gj0 = tx * TW;
  const int tid = threadIdx.x;
  typename St::Pre pre{};
  if constexpr (St::kPrefetch) St::template prefetch<MODE>(a, gi0, gj0, tid, &pre);
  St::template phase_load<NT>(s, a, gi0, gj0, tid);
  __syncthreads();
  St::phase_diffs(s, a, gi0, gj0, tid, NT);
  __syncthreads();
  St::phase_slopes(s, a, gi0, gj0, tid, NT);
  __syncthreads();
  St::phase_faces(s, a, gi0, gj0, tid, NT);
  __syncthreads();
  St::phase_alpha(s, a, gi0, gj0, tid, NT, dt);
  __syncthreads();
  double wsx = 0.0, wsy = 0.0;
  bool wet = false;
  St::template phase_update<MODE>(s, a, gi0, gj0, tid, NT, dt, &wsx, &wsy, &wet, St::kPrefetch ? &pre : nullptr);
  if (MODE == hf::MODE_FINAL_RK2 && a.wave_speed) {
    Ctrl* c = a.ctrl;
    unsigned long long* dst =
        a.ws_part ? a.ws_part + (size_t)(blockIdx.x % WS_SLOTS) * WS_STRIDE : &c->ws[c->parity ^ 1][0];
    block_max_to(wsx, wsy, s.red, dst, wet, a.tile_wet ? a.tile_wet + tile : nullptr);
  }
}

// Stage kernels that run as a fixed grid of workgroups, each looping over
// tiles w = blockIdx.x, blockIdx.x + gridDim.x, ... in tile_of_block's
// order (a multiple of 8 workgroups keeps every workgroup on its XCD's
// band): the FP64 tile (32 x 24, 768 threads) holds 143 KB of LDS, one
// workgroup per CU, so the per-tile kernel leaves its CU idle through each
// load phase and each workgroup's start. Here the next tile's global loads
// are issued right after the current tile's are stored to LDS and overlap
// its five compute phases, and the tid-only index arithmetic is computed
// once per workgroup instead of once per tile. Measured at 4096 x 4096
// (flood workload, kernel time per stage): first stage 1.27 -> 1.16 ms,
// final stage 1.57 -> 1.43 ms, bit for bit the per-tile kernel's output.
// Per kind: on, and the phases that recompute their thread-index
// arithmetic per tile (bit 1 load, 2 diffs, 4 slopes, 8 faces, 16 alpha, 32
// update) in the first and the final stage. The masks are measured ones
// (flood workload, 4096 x 4096, kernel time per stage): bitwise FP64 final
// stage 1.57 ms per-tile, 1.43 ms with 0x0d, 1.60 ms with 0x08, 1.80 ms with
// all phases. Hoisted values that overflow the VGPR budget spill to scratch
// inside the tile loop; too many recomputed ones cost VALU issue.
// The bitwise FP64 kernels keep this form. Tiles of half the LDS (24 x 16 or
// 32 x 12 with 512 threads, two workgroups per CU) measured 2.62 ms per step
// one workgroup per tile and 2.95-3.16 ms persistent (capped at 128 VGPRs),
// against 2.57 ms here; 1024 threads on 32 x 24 (four waves per SIMD at 128
// VGPRs) spills in the final stage and measured 2.84-2.97 ms. The fast FP64
// kind takes the 32 x 12 per-tile form
// (hf2d_fast.h), and the FP32 tile (32 x 16, 512 threads, three workgroups
// per CU) runs per tile: persistent at three workgroups per CU it measured
// 1.01 / 1.13 ms per stage (0.66 / 0.74 ms with the load, slopes and faces
// phases' index arithmetic recomputed per tile) against 0.53 / 0.62 ms.
#ifndef HF_PERSIST
#define HF_PERSIST 1
#endif
#ifndef HF_PERSIST_LAUNDER_STAGE
#define HF_PERSIST_LAUNDER_STAGE 0
#endif
#ifndef HF_PERSIST_LAUNDER_FINAL
#define HF_PERSIST_LAUNDER_FINAL 0x0d
#endif
template <class T> struct Persist {
  static constexpr bool on = false;
  static constexpr unsigned launder_stage = 0, launder_final = 0;
};
template <> struct Persist<double> {
  static constexpr bool on = HF_PERSIST != 0;
  static constexpr unsigned launder_stage = HF_PERSIST_LAUNDER_STAGE, launder_final = HF_PERSIST_LAUNDER_FINAL;
};

template <class T, int TW, int TH, int NT, int MODE>
__global__ __launch_bounds__(NT) void stage_persistent_kernel(StageArgs<T> a) {
  using St = hf::Stage<T, TW, TH>;
  __shared__ typename St::S s;
  __shared__ unsigned tile_wet_flag;
  double dt;
  if (!hf::stage_prologue(a, blockIdx.x == 0 && threadIdx.x == 0, &dt)) return;
  const unsigned n = a.tile_list ? (unsigned)*a.n_work : (unsigned)((int64_t)a.tiles_x * a.tiles_y);
  unsigned w = blockIdx.x;
  if (w >= n) return;
  auto tile_at = [&](unsigned ww) {
    return a.tile_list ? (unsigned)a.tile_list[tile_of_block(ww, n)] : tile_of_block(ww, n);
  };
  const int tid = threadIdx.x;
  unsigned tile = tile_at(w);
  int gi0 = (int)(tile / (unsigned)a.tiles_x) * TH, gj0 = (int)(tile % (unsigned)a.tiles_x) * TW;
  typename St::template LoadRegs<NT> L;
  typename St::Pre pre{};
  if constexpr (St::kPrefetch && MODE == hf::MODE_STAGE) St::template prefetch<MODE>(a, gi0, gj0, tid, &pre);
  St::template load_issue<NT>(a, gi0, gj0, tid, &L);
  const bool tile_flags = MODE == hf::MODE_FINAL_RK2 && a.wave_speed && a.tile_wet;
  if (tid == 0) tile_wet_flag = 0;
  double wsx = 0.0, wsy = 0.0;
  for (;;) {
    // Phases whose bit is set in kLaunder take their thread index from tl,
    // which is opaque to the compiler: their index arithmetic is recomputed
    // per tile instead of held in registers across the loop.
    constexpr unsigned kLaunder = MODE == hf::MODE_FINAL_RK2 ? Persist<T>::launder_final : Persist<T>::launder_stage;
    int tl = tid;
    if (kLaunder) asm volatile("" : "+v"(tl));
    const int t_load = (kLaunder & 1) ? tl : tid, t_diff = (kLaunder & 2) ? tl : tid,
              t_sl = (kLaunder & 4) ? tl : tid, t_face = (kLaunder & 8) ? tl : tid,
              t_al = (kLaunder & 16) ? tl : tid, t_up = (kLaunder & 32) ? tl : tid;
    St::template load_store<NT>(s, a, gi0, gj0, t_load, L);
    typename St::Pre cur{};
    if constexpr (St::kPrefetch) {
      if (MODE == hf::MODE_STAGE)
        cur = pre;
      else
        St::template prefetch<MODE>(a, gi0, gj0, t_up, &cur);
    }
    const unsigned wn = w + gridDim.x;
    unsigned tn = 0;
    int ni0 = 0, nj0 = 0;
    __syncthreads();
    if (wn < n) {
      tn = tile_at(wn);
      ni0 = (int)(tn / (unsigned)a.tiles_x) * TH;
      nj0 = (int)(tn % (unsigned)a.tiles_x) * TW;
      St::template load_issue<NT>(a, ni0, nj0, t_load, &L);
      if constexpr (St::kPrefetch && MODE == hf::MODE_STAGE) St::template prefetch<MODE>(a, ni0, nj0, t_up, &pre);
    }
    St::phase_diffs(s, a, gi0, gj0, t_diff, NT);
    __syncthreads();
    St::phase_slopes(s, a, gi0, gj0, t_sl, NT);
    __syncthreads();
    St::phase_faces(s, a, gi0, gj0, t_face, NT);
    __syncthreads();
    St::phase_alpha(s, a, gi0, gj0, t_al, NT, dt);
    __syncthreads();
    bool wet = false;
    St::template phase_update<MODE>(s, a, gi0, gj0, t_up, NT, dt, &wsx, &wsy, &wet, St::kPrefetch ? &cur : nullptr);
    if (tile_flags && __any(wet ? 1 : 0) && (tid & 63) == 0) tile_wet_flag = 1;
    __syncthreads();  // LDS is reused by the next tile; the flag is complete
    if (tile_flags && tid == 0) {
      a.tile_wet[tile] = tile_wet_flag ? 1 : 0;
      tile_wet_flag = 0;
    }
    if (wn >= n) break;
    w = wn;
    tile = tn;
    gi0 = ni0;
    gj0 = nj0;
  }
  if (MODE == hf::MODE_FINAL_RK2 && a.wave_speed) {
    Ctrl* c = a.ctrl;
    unsigned long long* dst =
        a.ws_part ? a.ws_part + (size_t)(blockIdx.x % WS_SLOTS) * WS_STRIDE : &c->ws[c->parity ^ 1][0];
    block_max_to(wsx, wsy, s.red, dst);
  }
}

// Rebuild the active-tile list for the step about to run (one workgroup of
// 1024 threads): tiles in row-major order without the ones tile_skipped()
// leaves out; *n_work = its length. act: per-tile scratch. A step that will
// not run (hf::step_will_run on the driver's control block c with the first
// stage's cfl, dx, dy, max_dt) leaves the counts and the list as they are:
// its stage kernels return in their prologue without reading the list.
__global__ __launch_bounds__(1024) void compact_kernel(const unsigned char* __restrict__ wet,
                                                       const unsigned char* __restrict__ pinned,
                                                       unsigned char* quiet_run, unsigned char* act, int tiles_x,
                                                       int tiles_y, int* list, int* n_work, const Ctrl* c,
                                                       double cfl, double dx, double dy, double max_dt) {
  if (!hf::step_will_run(c, cfl, dx, dy, max_dt)) return;  // the same decision in every thread
  __shared__ int counts[1024];
  const int n = tiles_x * tiles_y;
  const int per = (n + 1023) / 1024;
  const int t0 = (int)threadIdx.x * per;
  const int t1 = t0 + per < n ? t0 + per : n;
  int cnt = 0;
  for (int t = t0; t < t1; ++t) {
    const bool keep = !hf::tile_skipped(wet, pinned, quiet_run, tiles_x, tiles_y, t % tiles_x, t / tiles_x);
    act[t] = keep ? 1 : 0;
    cnt += keep ? 1 : 0;
  }
  counts[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan
    const int v = (int)threadIdx.x >= off ? counts[threadIdx.x - off] : 0;
    __syncthreads();
    counts[threadIdx.x] += v;
    __syncthreads();
  }
  int pos = counts[threadIdx.x] - cnt;
  for (int t = t0; t < t1; ++t)
    if (act[t]) list[pos++] = t;
  if (threadIdx.x == 1023) *n_work = counts[1023];
}

// Per tile, whether a cell of the current state is wet (not h <= H_DRY):
// the active-tile list's starting point after the state was set by other
// means than the driver. wet must be zero on entry.
template <class T>
__global__ __launch_bounds__(256) void tile_wet_init_kernel(Fld<T> h, int64_t nr, int64_t nc, int tw, int th,
                                                            int tiles_x, unsigned char* wet) {
  using R = typename hf::ScalarOf<T>::type;
  const int64_t n = nr * nc;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    if (!(hf::val(h.ld(i)) <= R(hf::H_DRY))) {
      const int64_t gi = i / nc, gj = i % nc;
      wet[(gi / th) * tiles_x + gj / tw] = 1;  // every writer stores the same byte value
    }
  }
}

// Gauges: append (t, h at each gauge cell) as row *count of buf (rows of
// n + 1 doubles), unconditionally with force, else only when the driver is
// live. The driver uses it for the initial row; finalize_kernel appends the
// row of every step. Rows past cap are dropped.
template <class T>
__global__ __launch_bounds__(64) void gauge_kernel(const Ctrl* c, Fld<T> h, const int64_t* cells, int n,
                                                   double* buf, long long* count, long long cap, int force) {
  if (!force && (c->status != 0 || c->done != 0)) return;
  const long long row = *count;
  if (row >= cap) return;
  double* r = buf + row * (long long)(n + 1);
  for (int k = (int)threadIdx.x; k <= n; k += (int)blockDim.x)
    r[k] = k == 0 ? c->t : (double)hf::val(h.ld(cells[k - 1]));
  __syncthreads();
  if (threadIdx.x == 0) *count = row + 1;
}

__global__ __launch_bounds__(256) void fill_bytes_kernel(unsigned char* p, int64_t n, unsigned char v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

// Forward-Euler tail and the standalone source steps: elementwise over
// U_in -> U_out with the optional forcing and CFL maxima.
template <class T>
__global__ __launch_bounds__(256) void tail_kernel(StageArgs<T> a, int slot_mode) {
  __shared__ double red[2][16];
  double dt;
  if (!hf::stage_prologue(a, blockIdx.x == 0 && threadIdx.x == 0, &dt)) return;
  using St = hf::Stage<T, 32, 16>;
  double wsx = 0.0, wsy = 0.0;
  const int64_t n = (int64_t)a.nr * a.nc;
  const typename St::ForcingNow fnow = St::forcing_now(a, 0, 0, a.nr, a.nc);
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int gi = (int)(idx / a.nc), gj = (int)(idx % a.nc);
    T h = a.ih.ld(idx), hu = a.ihu.ld(idx), hv = a.ihv.ld(idx);
    St::forcing(a, fnow, gi, gj, dt, &h, &hu, &hv);
    a.oh.st(idx, h);
    a.ohu.st(idx, hu);
    a.ohv.st(idx, hv);
    if (a.wave_speed) {
      const double td = hf::tile_datum(a, gi, gj);
      T zc = hf::load_bed(a, gi, gj, idx, td);
      T zn4[4];
      zn4[hf::SIDE_N] = (a.nr >= 2 && gi == 0) ? hf::load_bed(a, gi + 1, gj, idx + a.nc, td) : zc;
      zn4[hf::SIDE_S] = (a.nr >= 2 && gi == a.nr - 1) ? hf::load_bed(a, gi - 1, gj, idx - a.nc, td) : zc;
      zn4[hf::SIDE_W] = (a.nc >= 2 && gj == 0) ? hf::load_bed(a, gi, gj + 1, idx + 1, td) : zc;
      zn4[hf::SIDE_E] = (a.nc >= 2 && gj == a.nc - 1) ? hf::load_bed(a, gi, gj - 1, idx - 1, td) : zc;
      T man = a.man.ld(idx);
      St::cell_wave_speed(a, gi, gj, h, hu, hv, zc, zn4, &man, &wsx, &wsy);
    }
  }
  if (a.wave_speed) {
    Ctrl* c = a.ctrl;
    block_max_to(wsx, wsy, red, &c->ws[slot_mode ? c->parity : (c->parity ^ 1)][0]);
  }
}

// apply_point_sources, sequential in list order (source.rs:52-69).
template <class T> __global__ void point_source_kernel(StageArgs<T> a, double dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  using R = typename hf::ScalarOf<T>::type;
  const T Z = hf::cst<T>(0.0);
  const double cell_area = a.dx * a.dy;
  for (int k = 0; k < a.n_src; ++k) {
    const int64_t idx = a.srcs[k].row * a.nc + a.srcs[k].col;
    const R dh = R(a.srcs[k].q * dt / cell_area);
    T h = hf::rmax(a.ih.ld(idx) + dh, Z);
    T hu = a.ihu.ld(idx), hv = a.ihv.ld(idx);
    if (hf::val(h) == R(0)) {
      hu = Z;
      hv = Z;
    }
    a.oh.st(idx, h);
    a.ohu.st(idx, hu);
    a.ohv.st(idx, hv);
  }
}

// Gauge cells of a driver (finalize_kernel): the depth value plane (FP64,
// or FP32 when f32 is set) and the row buffer of gauge_kernel.
struct GaugeArgs {
  const void* h;
  int f32;
  const int64_t* cells;
  int n;
  double* buf;
  long long* count;
  long long cap;
};

// max(v, v of the lane that DPP control CTRL names) as unsigned 64-bit
// integers, in the lanes of the rows ROW_MASK enables (others keep v).
template <int CTRL, int ROW_MASK = 0xf> __device__ inline unsigned long long dpp_umax(unsigned long long v) {
  const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
  const unsigned long long o = ((unsigned long long)(unsigned)hi2 << 32) | (unsigned)lo2;
  return o > v ? o : v;
}

// Unsigned 64-bit maximum over the wavefront (all lanes active), uniform;
// the same DPP sequence as wave_fmax.
__device__ inline unsigned long long wave_umax(unsigned long long v) {
  v = dpp_umax<0xB1>(v);
  v = dpp_umax<0x4E>(v);
  v = dpp_umax<0x141>(v);
  v = dpp_umax<0x140>(v);
  v = dpp_umax<0x142, 0xA>(v);
  v = dpp_umax<0x143, 0xC>(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

// End of a driver step, one workgroup of WS_SLOTS threads: fold the final
// stage's CFL slots (part, nullable) into ctrl->ws[parity ^ 1] and clear
// them, then advance the clock and flip the slot. With g.n > 0 it then
// appends the gauge row (t, h at each gauge cell) as gauge_kernel does.
// Every load that depends on no other is issued first (the slots, the
// control block, the gauge row index and depths), so the launch waits for
// one memory round trip before its stores, not one per step of the chain.
__global__ __launch_bounds__(WS_SLOTS) void finalize_kernel(Ctrl* c, unsigned long long* part, GaugeArgs g) {
  __shared__ unsigned long long red[2][WS_SLOTS / 64];
  const int t = threadIdx.x;
  unsigned long long mx = 0, my = 0;
  if (part) {
    mx = part[(size_t)t * WS_STRIDE];
    my = part[(size_t)t * WS_STRIDE + 1];
  }
  const int status = c->status, done = c->done;  // the same in every thread
  int parity = 0, compensated = 0;
  unsigned long long ws[2][2] = {};
  double ct = 0.0, cdt = 0.0, ccomp = 0.0;
  long long steps = 0, budget = 0;
  if (t == 0) {
    parity = c->parity;
    compensated = c->compensated;
    ws[0][0] = c->ws[0][0];
    ws[0][1] = c->ws[0][1];
    ws[1][0] = c->ws[1][0];
    ws[1][1] = c->ws[1][1];
    ct = c->t;
    cdt = c->dt;
    ccomp = c->t_comp;
    steps = c->steps;
    budget = c->budget;
  }
  long long row = 0;
  double gv = 0.0;
  if (g.n > 0) {
    row = *g.count;
    if (t >= 1 && t <= g.n)
      gv = g.f32 ? (double)static_cast<const float*>(g.h)[g.cells[t - 1]]
                 : static_cast<const double*>(g.h)[g.cells[t - 1]];
  }
  if (part) {
    part[(size_t)t * WS_STRIDE] = 0ULL;
    part[(size_t)t * WS_STRIDE + 1] = 0ULL;
    mx = wave_umax(mx);
    my = wave_umax(my);
    if ((t & 63) == 0) {
      red[0][t >> 6] = mx;
      red[1][t >> 6] = my;
    }
    __syncthreads();
    if (t == 0) {
      for (int k = 1; k < WS_SLOTS / 64; ++k) {
        mx = red[0][k] > mx ? red[0][k] : mx;
        my = red[1][k] > my ? red[1][k] : my;
      }
      const unsigned long long* w = ws[parity ^ 1];
      c->ws[parity ^ 1][0] = mx > w[0] ? mx : w[0];
      c->ws[parity ^ 1][1] = my > w[1] ? my : w[1];
    }
  }
  if (status != 0 || done != 0) return;
  if (t == 0) {
    if (compensated) {  // FP32 precision form: Kahan-compensated clock
      const double y = cdt - ccomp;
      const double t1 = ct + y;
      c->t_comp = (t1 - ct) - y;
      ct = t1;
    } else {
      ct = ct + cdt;  // sim.rs:225 (plain sum)
    }
    c->t = ct;
    c->steps = steps + 1;
    c->budget = budget - 1;
    c->parity = parity ^ 1;
  }
  if (g.n <= 0 || row >= g.cap) return;
  double* r = g.buf + row * (long long)(g.n + 1);
  if (t == 0) r[0] = ct;
  if (t >= 1 && t <= g.n) r[t] = gv;
  for (int k = t + WS_SLOTS; k <= g.n; k += WS_SLOTS)
    r[k] = g.f32 ? (double)static_cast<const float*>(g.h)[g.cells[k - 1]]
                 : static_cast<const double*>(g.h)[g.cells[k - 1]];
  __syncthreads();
  if (t == 0) *g.count = row + 1;
}

__global__ void zero_slot_kernel(Ctrl* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    c->ws[c->parity][0] = 0ULL;
    c->ws[c->parity][1] = 0ULL;
  }
}

// Streaming copy (dst = src), the calibration kernel for HBM counters and
// the attainable-bandwidth reference: each thread moves COPY_UNROLL 16-byte
// vectors per round, all loads issued before the stores, with nontemporal
// hints (the data is touched once). Measured on MI355X over 1 GiB: 6.25 TB/s
// with 64 Ki workgroups, against 4.5-5.6 TB/s for one vector per thread.
constexpr int COPY_UNROLL = 4;
__global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ src_, double2* __restrict__ dst_,
                                                   int64_t n2) {
  typedef double v2 __attribute__((ext_vector_type(2)));
  const v2* src = reinterpret_cast<const v2*>(src_);
  v2* dst = reinterpret_cast<v2*>(dst_);
  const int64_t stride = (int64_t)gridDim.x * 256 * COPY_UNROLL;
  for (int64_t base = (int64_t)blockIdx.x * 256 * COPY_UNROLL + threadIdx.x; base < n2; base += stride) {
    v2 v[COPY_UNROLL];
#pragma unroll
    for (int u = 0; u < COPY_UNROLL; ++u) {
      const int64_t i = base + (int64_t)u * 256;
      if (i < n2) v[u] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int u = 0; u < COPY_UNROLL; ++u) {
      const int64_t i = base + (int64_t)u * 256;
      if (i < n2) __builtin_nontemporal_store(v[u], dst + i);
    }
  }
}

__global__ void math_probe_kernel(const double* x, const double* y, int64_t n, double* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi = x[i];
  // the kernels' sqrt: the unscaled sequence from 2^-767 up, else IEEE sqrt
  out[4 * i + 0] = xi >= 0x1p-767 ? hf::sqrt_ge_tiny(xi) : __builtin_sqrt(xi);
  out[4 * i + 1] = xi / y[i];
  out[4 * i + 2] = hfm::pow_4_3(xi);
  out[4 * i + 3] = hfm::pow_1_3m(xi);
}

// Total water volume sum(h) * dx * dy in FP64: per-thread Kahan sums over a
// grid-stride range, then a fixed-order pairwise combination in LDS; one
// (sum, compensation) pair per block, combined on the host. Deterministic
// for a given grid size.
template <class T>
__global__ __launch_bounds__(256) void volume_kernel(Fld<T> h, int64_t n, double* partial) {
  __shared__ double ss[256], cc[256];
  double sum = 0.0, comp = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double y = (double)hf::val(h.ld(i)) - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  ss[threadIdx.x] = sum;
  cc[threadIdx.x] = -comp;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double a = ss[threadIdx.x], b = ss[threadIdx.x + w];
      const double t = a + b;  // TwoSum
      const double bp = t - a;
      const double err = (a - (t - bp)) + (b - bp);
      ss[threadIdx.x] = t;
      cc[threadIdx.x] = cc[threadIdx.x] + cc[threadIdx.x + w] + err;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = ss[0];
    partial[2 * blockIdx.x + 1] = cc[0];
  }
}


// ------------------------------------------------------------ host-side field views
// Field views over raw device buffers: value plane v, tangent buffer d
// (Dual: one per cell; DualN<N>: N per cell; null: tangent 0).
template <class T> struct FldOf;
template <> struct FldOf<double> {
  static Fld<double> r(const void* v, const void*) { return Fld<double>{(const double*)v}; }
  static FldW<double> w(void* v, void*) { return FldW<double>{(double*)v}; }
};
template <> struct FldOf<dual> {
  static Fld<dual> r(const void* v, const void* d) { return Fld<dual>{(const double*)v, (const double*)d}; }
  static FldW<dual> w(void* v, void* d) { return FldW<dual>{(double*)v, (double*)d}; }
};
template <> struct FldOf<float> {
  static Fld<float> r(const void* v, const void*) { return Fld<float>{(const float*)v}; }
  static FldW<float> w(void* v, void*) { return FldW<float>{(float*)v}; }
};
template <int N> struct FldOf<hf::dualn<N>> {
  static Fld<hf::dualn<N>> r(const void* v, const void* d) {
    return Fld<hf::dualn<N>>{(const double*)v, (const double*)d, N};
  }
  static FldW<hf::dualn<N>> w(void* v, void* d) { return FldW<hf::dualn<N>>{(double*)v, (double*)d, N, 1}; }
};
template <class T> Fld<T> fld(const void* v, const void* d) { return FldOf<T>::r(v, d); }
template <class T> FldW<T> fldw(void* v, void* d) { return FldOf<T>::w(v, d); }

constexpr int DATUM_LOG2 = 6;  // FP32 kind: one FP64 bed datum per 64 x 64 cells

inline size_t elem_bytes(int kind) { return kind == 2 /* HF2D_F32 */ ? sizeof(float) : sizeof(double); }

// ------------------------------------------------------------ batched engine
// One work item = one output tile of one scenario.
struct WorkItem {
  int slot;
  int tile;
};

// Stage kernel over a device work list spanning many scenarios of
// different sizes, one work item (one tile of one scenario) per workgroup:
// grid = the list's length, items mapped to workgroups in the XCD-banded
// order of single grids. Each scenario carries its own StageArgs (pointers,
// shape, boundaries, forcing) and its own Ctrl (clock, dt, CFL slots), so
// every scenario advances with its own dt; a scenario whose clock reached
// its end time (or failed) makes its tiles no-ops. Each workgroup reduces
// its CFL maxima into its scenario's slot with one atomic max.
// A persistent form (a fixed grid looping over the items) measured slower
// on the 256-scenario mix: the compiler hoists the phases' thread-index
// arithmetic and the item's descriptor out of the loop into registers (FP32
// 32 x 16: 131 VGPRs against 62; Dual 16 x 8: 233 against 123), which cuts
// the workgroups per CU. Best configurations, per step: FP32 9.36 -> 7.41
// ms, FP64 17.2 -> 16.4 ms, Dual (64 scenarios) 13.6 -> 11.3 ms.
template <class T, int TW, int TH, int NT, int MODE>
__global__ __launch_bounds__(NT) void batch_stage_kernel(const StageArgs<T>* __restrict__ descs,
                                                        const WorkItem* __restrict__ work, int n_work) {
  using St = hf::Stage<T, TW, TH>;
  __shared__ typename St::S s;
  const int tid = threadIdx.x;
  const unsigned w = blockIdx.x;
  if ((int)w >= n_work) return;
  const WorkItem it = work[tile_of_block(w, (unsigned)n_work)];
  const StageArgs<T>& a = descs[it.slot];
  double dt;
  if (!hf::stage_prologue(a, it.tile == 0 && tid == 0, &dt)) return;
  const int gi0 = (int)(it.tile / a.tiles_x) * TH, gj0 = (int)(it.tile % a.tiles_x) * TW;
  typename St::Pre pre{};
  if constexpr (St::kPrefetch) St::template prefetch<MODE>(a, gi0, gj0, tid, &pre);
  St::template phase_load<NT>(s, a, gi0, gj0, tid);
  __syncthreads();
  St::phase_diffs(s, a, gi0, gj0, tid, NT);
  __syncthreads();
  St::phase_slopes(s, a, gi0, gj0, tid, NT);
  __syncthreads();
  St::phase_faces(s, a, gi0, gj0, tid, NT);
  __syncthreads();
  St::phase_alpha(s, a, gi0, gj0, tid, NT, dt);
  __syncthreads();
  double wsx = 0.0, wsy = 0.0;
  bool wet = false;
  St::template phase_update<MODE>(s, a, gi0, gj0, tid, NT, dt, &wsx, &wsy, &wet, St::kPrefetch ? &pre : nullptr);
  if (MODE == hf::MODE_FINAL_RK2 && a.wave_speed) {
    Ctrl* c = a.ctrl;
    unsigned long long* dst =
        a.ws_part ? a.ws_part + (size_t)(blockIdx.x % BATCH_WS_SLOTS) * WS_STRIDE : &c->ws[c->parity ^ 1][0];
    block_max_to(wsx, wsy, s.red, dst);
  }
}

// Clock update of every scenario in the batch (finalize_kernel per slot),
// one wavefront per scenario: lane k folds and clears CFL slot k of the
// scenario's BATCH_WS_SLOTS (unsigned max of the bit patterns, as the
// atomics compute it), lane 0 merges the fold into the control block and
// advances the clock.
__global__ __launch_bounds__(BATCH_WS_SLOTS) void batch_finalize_kernel(Ctrl* ctrls, const int* slots, int n,
                                                                        unsigned long long* ws_part) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const int slot = slots[i];
  Ctrl* c = ctrls + slot;
  unsigned long long* p = ws_part + ((size_t)slot * BATCH_WS_SLOTS + threadIdx.x) * WS_STRIDE;
  unsigned long long x = p[0], y = p[1];
  p[0] = 0ULL;
  p[1] = 0ULL;
  for (int off = BATCH_WS_SLOTS / 2; off > 0; off >>= 1) {
    const unsigned long long xo = __shfl_xor(x, off), yo = __shfl_xor(y, off);
    x = xo > x ? xo : x;
    y = yo > y ? yo : y;
  }
  if (threadIdx.x != 0) return;
  if (c->status != 0 || c->done != 0) return;
  unsigned long long* ws = &c->ws[c->parity ^ 1][0];
  const unsigned long long w0 = ws[0], w1 = ws[1];
  ws[0] = x > w0 ? x : w0;
  ws[1] = y > w1 ? y : w1;
  if (c->compensated) {
    const double y = c->dt - c->t_comp;
    const double t1 = c->t + y;
    c->t_comp = (t1 - c->t) - y;
    c->t = t1;
  } else {
    c->t = c->t + c->dt;
  }
  c->steps = c->steps + 1;
  c->budget = c->budget - 1;
  c->parity = c->parity ^ 1;
}

}  // namespace
