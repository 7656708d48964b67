// Batched engine: many scenarios of uneven grid sizes, each with its own
// clock and dt, advanced together on one GPU.
//
// Memory: one HBM arena per batch, carved by a first-fit allocator; a
// scenario is admitted only if its footprint fits (hf2d_batch_footprint).
// Per step: one launch per RK stage over a device work list of (scenario,
// tile) items, one workgroup per item, then one finalize launch over the
// active slots;
// the step sequence is captured in a hipGraph and replayed, and re-captured
// only when the set of scenarios or the tile configuration changes. The tile
// shape / threads per workgroup is chosen by hf2d_batch_autotune, which
// times every compiled configuration on the resident mix (state snapshot and
// restore around the trials).
//
// Each scenario's arithmetic is the single-grid step's: same stage code,
// same tile decomposition for a given configuration, its own dt from its
// own CFL reduction, so a scenario's result does not depend on what else is
// in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/hydroflux_amd.h"
#include "hf2d_device.h"

namespace {

// ------------------------------------------------------------ tile configs
struct CfgInfo {
  int tw, th, nt;
};
// The autotuner's candidates per kind: the 256-cell tile shapes
// (16x16, 32x8, 64x4, 128x2) and the larger high-occupancy tiles.
const CfgInfo CFG_F32[] = {{32, 32, 1024}, {32, 16, 512}, {16, 16, 256}, {32, 8, 256}, {64, 4, 256}, {128, 2, 256}};
// 32 x 12 with 512 threads holds 79 KB of LDS: two workgroups per CU.
const CfgInfo CFG_F64[] = {{32, 24, 768}, {32, 24, 1024}, {32, 16, 1024}, {32, 12, 512},
                           {16, 16, 256}, {32, 8, 256},   {64, 4, 256},   {128, 2, 256}};
const CfgInfo CFG_DUAL[] = {{16, 8, 512}, {16, 16, 1024}, {16, 8, 256}, {24, 16, 1024}, {32, 12, 768}};
// The configuration a new batch starts on: the fastest per stage over a
// 256-scenario flood mix (98.7 M cells) on MI355X, FP32 32 x 16 x 512
// (3.40 / 3.89 ms first / final stage vs 4.57 / 5.32 ms for 32 x 32) and
// FP64 32 x 12 x 512 (7.43 / 9.07 ms vs 7.59 / 9.34 ms for 32 x 24 x 768);
// Dual 24 x 16 x 1024 on the mix's first 64 scenarios (29.2 M cells):
// 11.7 ms per step against 14.2 ms for 16 x 8 x 512.
constexpr int CFG_F32_DEFAULT = 1, CFG_F64_DEFAULT = 3, CFG_DUAL_DEFAULT = 3;

template <class T> struct Cfgs;
template <> struct Cfgs<float> {
  static const CfgInfo* list() { return CFG_F32; }
  static constexpr int n = 6;
};
template <> struct Cfgs<double> {
  static const CfgInfo* list() { return CFG_F64; }
  static constexpr int n = 8;
};
template <> struct Cfgs<dual> {
  static const CfgInfo* list() { return CFG_DUAL; }
  static constexpr int n = 5;
};

template <class T, int TW, int TH, int NT>
int launch_pair(hipStream_t st, const StageArgs<T>* d1, const StageArgs<T>* d2, const WorkItem* w, int n) {
  hipLaunchKernelGGL((batch_stage_kernel<T, TW, TH, NT, hf::MODE_STAGE>), dim3(n), dim3(NT), 0, st, d1, w, n);
  hipLaunchKernelGGL((batch_stage_kernel<T, TW, TH, NT, hf::MODE_FINAL_RK2>), dim3(n), dim3(NT), 0, st, d2, w, n);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Dispatch on the configuration index of kind T.
template <class T> int launch_cfg(int ci, hipStream_t st, const StageArgs<T>* d1, const StageArgs<T>* d2,
                                  const WorkItem* w, int n);

template <> int launch_cfg<float>(int ci, hipStream_t st, const StageArgs<float>* d1, const StageArgs<float>* d2,
                                  const WorkItem* w, int n) {
  switch (ci) {
    case 0: return launch_pair<float, 32, 32, 1024>(st, d1, d2, w, n);
    case 1: return launch_pair<float, 32, 16, 512>(st, d1, d2, w, n);
    case 2: return launch_pair<float, 16, 16, 256>(st, d1, d2, w, n);
    case 3: return launch_pair<float, 32, 8, 256>(st, d1, d2, w, n);
    case 4: return launch_pair<float, 64, 4, 256>(st, d1, d2, w, n);
    default: return launch_pair<float, 128, 2, 256>(st, d1, d2, w, n);
  }
}
template <> int launch_cfg<double>(int ci, hipStream_t st, const StageArgs<double>* d1, const StageArgs<double>* d2,
                                   const WorkItem* w, int n) {
  switch (ci) {
    case 0: return launch_pair<double, 32, 24, 768>(st, d1, d2, w, n);
    case 1: return launch_pair<double, 32, 24, 1024>(st, d1, d2, w, n);
    case 2: return launch_pair<double, 32, 16, 1024>(st, d1, d2, w, n);
    case 3: return launch_pair<double, 32, 12, 512>(st, d1, d2, w, n);
    case 4: return launch_pair<double, 16, 16, 256>(st, d1, d2, w, n);
    case 5: return launch_pair<double, 32, 8, 256>(st, d1, d2, w, n);
    case 6: return launch_pair<double, 64, 4, 256>(st, d1, d2, w, n);
    default: return launch_pair<double, 128, 2, 256>(st, d1, d2, w, n);
  }
}
template <> int launch_cfg<dual>(int ci, hipStream_t st, const StageArgs<dual>* d1, const StageArgs<dual>* d2,
                                 const WorkItem* w, int n) {
  switch (ci) {
    case 0: return launch_pair<dual, 16, 8, 512>(st, d1, d2, w, n);
    case 1: return launch_pair<dual, 16, 16, 1024>(st, d1, d2, w, n);
    case 2: return launch_pair<dual, 16, 8, 256>(st, d1, d2, w, n);
    case 3: return launch_pair<dual, 24, 16, 1024>(st, d1, d2, w, n);
    default: return launch_pair<dual, 32, 12, 768>(st, d1, d2, w, n);
  }
}

constexpr size_t ALIGN = 256;
size_t up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

struct Scen {
  bool used = false;
  int64_t nr = 0, nc = 0;
  double dx = 1.0, dy = 1.0;
  size_t off = 0, bytes = 0;
  void* st[2][3] = {{nullptr}};
  void* tg[2][3] = {{nullptr}};
  void *man = nullptr, *mant = nullptr;
  double *z = nullptr, *zt = nullptr, *datum = nullptr;
  float* zdev = nullptr;
  int64_t datum_cols = 0;
  bool has_zt = false;
  int bc_kind[4] = {1, 1, 1, 1};
  double bc_param[4] = {0, 0, 0, 0};
  hf2d_config cfg{};
  PointSrc* srcs = nullptr;
  int n_src = 0;
  double t_end = 0.0;
};

// Bytes a scenario occupies in the arena (state x2, tangents, bed, Manning, sources).
size_t footprint(int kind, int64_t nr, int64_t nc, int64_t n_src) {
  const size_t n = (size_t)(nr * nc), e = elem_bytes(kind);
  size_t b = 6 * up(n * e);
  if (kind == HF2D_DUAL) b += 6 * up(n * e) + 2 * up(n * 8) + up(n * e);  // tangents, z, zt, mant
  if (kind == HF2D_F32) {
    const int64_t B = 1 << DATUM_LOG2;
    b += up((size_t)(((nr + B - 1) / B) * ((nc + B - 1) / B)) * 8) + up(n * 4);  // datum + deviation
  } else if (kind == HF2D_F64) {
    b += up(n * 8);  // z
  }
  b += up(n * e);  // Manning
  b += up((size_t)std::max<int64_t>(n_src, 1) * sizeof(PointSrc));
  return b;
}

}  // namespace

struct hf2d_batch {
  int kind = HF2D_F32;
  int cfgi = 0;
  hipStream_t stream = nullptr;
  char* arena = nullptr;
  size_t cap = 0;
  std::map<size_t, size_t> free_blocks;  // offset -> size
  std::vector<Scen> sc;
  Ctrl* ctrls = nullptr;
  void* descs1 = nullptr;
  void* descs2 = nullptr;
  WorkItem* work = nullptr;
  int work_cap = 0, n_work = 0;
  int* slots_dev = nullptr;
  unsigned long long* ws_part = nullptr;  // BATCH_WS_SLOTS CFL slots per scenario slot
  int n_active = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  bool dirty = true;
  std::string err;
};

namespace {

int bfail(hf2d_batch* b, int code, const std::string& m) {
  if (b) b->err = m;
  return code;
}
#define BCHK(b, expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return bfail(b, HF2D_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

bool arena_alloc(hf2d_batch* b, size_t bytes, size_t* off) {
  for (auto it = b->free_blocks.begin(); it != b->free_blocks.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      b->free_blocks.erase(it);
      if (rest) b->free_blocks[at] = rest;
      return true;
    }
  }
  return false;
}

void arena_free(hf2d_batch* b, size_t off, size_t bytes) {
  auto it = b->free_blocks.emplace(off, bytes).first;
  auto nx = std::next(it);  // coalesce with the following block
  if (nx != b->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    b->free_blocks.erase(nx);
  }
  if (it != b->free_blocks.begin()) {  // and with the preceding one
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      b->free_blocks.erase(it);
    }
  }
}

template <class T> StageArgs<T> scen_args(hf2d_batch* b, int slot, int in, int out, int un, int tw, int th) {
  Scen& s = b->sc[(size_t)slot];
  StageArgs<T> a{};
  a.nr = s.nr;
  a.nc = s.nc;
  a.dx = s.dx;
  a.dy = s.dy;
  a.ih = fld<T>(s.st[in][0], s.tg[in][0]);
  a.ihu = fld<T>(s.st[in][1], s.tg[in][1]);
  a.ihv = fld<T>(s.st[in][2], s.tg[in][2]);
  a.oh = fldw<T>(s.st[out][0], s.tg[out][0]);
  a.ohu = fldw<T>(s.st[out][1], s.tg[out][1]);
  a.ohv = fldw<T>(s.st[out][2], s.tg[out][2]);
  a.nh = fld<T>(s.st[un][0], s.tg[un][0]);
  a.nhu = fld<T>(s.st[un][1], s.tg[un][1]);
  a.nhv = fld<T>(s.st[un][2], s.tg[un][2]);
  a.z = fld<T>(s.z, s.has_zt ? s.zt : nullptr);
  a.man = fld<T>(s.man, s.mant);
  a.datum = s.datum;
  a.zdev = s.zdev;
  a.datum_log2 = DATUM_LOG2;
  a.datum_cols = s.datum_cols;
  for (int k = 0; k < 4; ++k) {
    a.bc_kind[k] = s.bc_kind[k];
    a.bc_param[k] = s.bc_param[k];
  }
  a.ctrl = b->ctrls + slot;
  a.cfl = s.cfg.cfl;
  a.max_dt = s.cfg.max_dt;
  a.cfl_with_bcs = s.cfg.cfl_with_bcs;
  a.friction = s.cfg.friction;
  a.friction_dry_tol = s.cfg.friction_dry_tol;
  a.rain_rate = s.cfg.rain_rate;
  a.n_src = s.n_src;
  a.srcs = s.srcs;
  a.tiles_x = (int)((s.nc + tw - 1) / tw);
  a.tiles_y = (int)((s.nr + th - 1) / th);
  return a;
}

// Launches of the graph may still be in flight (hf2d_batch_step returns
// without waiting); the stream is drained before the executable graph is
// destroyed.
void destroy_graph(hf2d_batch* b) {
  if (b->gexec && b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->graph_steps = 0;
}

template <class T> int launch_step(hf2d_batch* b) {
  if (b->n_work > 0) {
    if (launch_cfg<T>(b->cfgi, b->stream, (const StageArgs<T>*)b->descs1, (const StageArgs<T>*)b->descs2, b->work,
                      b->n_work))
      return bfail(b, HF2D_HIP_ERROR, "stage launch failed");
  }
  if (b->n_active > 0) {
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((unsigned)b->n_active), dim3(BATCH_WS_SLOTS), 0, b->stream,
                       b->ctrls, b->slots_dev, b->n_active, b->ws_part);
    BCHK(b, hipGetLastError());
  }
  return HF2D_OK;
}

// Rebuild descriptors, work list and the step graph.
template <class T> int rebuild(hf2d_batch* b) {
  const CfgInfo ci = Cfgs<T>::list()[b->cfgi];
  std::vector<StageArgs<T>> d1(b->sc.size()), d2(b->sc.size());
  std::vector<WorkItem> w;
  std::vector<int> slots;
  for (size_t k = 0; k < b->sc.size(); ++k) {
    if (!b->sc[k].used) continue;
    StageArgs<T> a = scen_args<T>(b, (int)k, 0, 1, 0, ci.tw, ci.th);
    a.dt_mode = 1;
    a.first_stage = 1;
    StageArgs<T> c = scen_args<T>(b, (int)k, 1, 0, 0, ci.tw, ci.th);
    c.dt_mode = 1;
    c.first_stage = 0;
    c.wave_speed = 1;
    c.ws_part = b->ws_part + k * BATCH_WS_SLOTS * WS_STRIDE;
    d1[k] = a;
    d2[k] = c;
    const int nt = a.tiles_x * a.tiles_y;
    for (int t = 0; t < nt; ++t) w.push_back(WorkItem{(int)k, t});
    slots.push_back((int)k);
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  if (!d1.empty()) {
    BCHK(b, hipMemcpy(b->descs1, d1.data(), d1.size() * sizeof(StageArgs<T>), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->descs2, d2.data(), d2.size() * sizeof(StageArgs<T>), hipMemcpyHostToDevice));
  }
  if ((int)w.size() > b->work_cap) {
    if (b->work) BCHK(b, hipFree(b->work));
    b->work = nullptr;
    b->work_cap = (int)w.size() * 2;
    BCHK(b, hipMalloc(&b->work, (size_t)b->work_cap * sizeof(WorkItem)));
  }
  if (!w.empty()) BCHK(b, hipMemcpy(b->work, w.data(), w.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
  if (!slots.empty()) BCHK(b, hipMemcpy(b->slots_dev, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice));
  b->n_work = (int)w.size();
  b->n_active = (int)slots.size();
  destroy_graph(b);
  b->dirty = false;
  return HF2D_OK;
}

// Capture `steps` batch steps into a graph (no memset nodes: clears are kernels).
template <class T> int capture(hf2d_batch* b, int steps) {
  destroy_graph(b);
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = HF2D_OK;
  for (int k = 0; k < steps && rc == HF2D_OK; ++k) rc = launch_step<T>(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) return rc;
  if (e != hipSuccess) return bfail(b, HF2D_HIP_ERROR, std::string("graph capture: ") + hipGetErrorString(e));
  b->graph = g;
  BCHK(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  b->graph_steps = steps;
  return HF2D_OK;
}

template <class T> int step_n(hf2d_batch* b, int64_t n, int use_graph) {
  if (b->dirty) {
    int rc = rebuild<T>(b);
    if (rc) return rc;
  }
  if (b->n_active == 0 || n <= 0) return HF2D_OK;
  const int G = 16;
  if (use_graph && n >= G) {
    if (b->graph_steps != G) {
      int rc = capture<T>(b, G);
      if (rc) return rc;
    }
    for (int64_t k = 0; k < n / G; ++k) BCHK(b, hipGraphLaunch(b->gexec, b->stream));
    n = n % G;
  }
  for (int64_t k = 0; k < n; ++k) {
    int rc = launch_step<T>(b);
    if (rc) return rc;
  }
  return HF2D_OK;
}

// The CFL maxima of a freshly added scenario's initial state (tail kernel).
template <class T> int init_wave_speeds(hf2d_batch* b, int slot) {
  StageArgs<T> a = scen_args<T>(b, slot, 0, 0, 0, 32, 16);
  a.dt_mode = 0;
  a.dt_host = 0.0;
  a.friction = 0;
  a.rain_rate = 0.0;
  a.n_src = 0;
  a.wave_speed = 1;
  hipLaunchKernelGGL(zero_slot_kernel, dim3(1), dim3(1), 0, b->stream, a.ctrl);
  int64_t g = ((int64_t)a.nr * a.nc + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL((tail_kernel<T>), dim3((unsigned)g), dim3(256), 0, b->stream, a, 1);
  BCHK(b, hipGetLastError());
  return HF2D_OK;
}

#define HF_BDISPATCH(b, FN, ...)                                    \
  ((b)->kind == HF2D_DUAL  ? FN<dual>(__VA_ARGS__)                  \
   : (b)->kind == HF2D_F32 ? FN<float>(__VA_ARGS__)                 \
                           : FN<double>(__VA_ARGS__))

template <class T> size_t desc_size() { return sizeof(StageArgs<T>); }

template <class T> int autotune(hf2d_batch* b, int steps, int32_t* best_out, double* ms_out) {
  // Snapshot the resident states and clocks, time each configuration, restore.
  int rc = HF2D_OK;
  BCHK(b, hipStreamSynchronize(b->stream));
  std::vector<void*> snaps;
  std::vector<size_t> sizes;
  for (Scen& s : b->sc) {
    if (!s.used) continue;
    const size_t pb = (size_t)(s.nr * s.nc) * elem_bytes(b->kind);
    for (int f = 0; f < 3; ++f) {
      for (int tan = 0; tan < (b->kind == HF2D_DUAL ? 2 : 1); ++tan) {
        void* src = tan ? s.tg[0][f] : s.st[0][f];
        void* dst = nullptr;
        BCHK(b, hipMalloc(&dst, pb));
        BCHK(b, hipMemcpy(dst, src, pb, hipMemcpyDeviceToDevice));
        snaps.push_back(dst);
        sizes.push_back(pb);
      }
    }
  }
  const size_t cbytes = b->sc.size() * sizeof(Ctrl);
  std::vector<char> ctrl_host(cbytes);
  BCHK(b, hipMemcpy(ctrl_host.data(), b->ctrls, cbytes, hipMemcpyDeviceToHost));
  auto restore = [&]() -> int {
    size_t k = 0;
    for (Scen& s : b->sc) {
      if (!s.used) continue;
      for (int f = 0; f < 3; ++f)
        for (int tan = 0; tan < (b->kind == HF2D_DUAL ? 2 : 1); ++tan, ++k)
          BCHK(b, hipMemcpy(tan ? s.tg[0][f] : s.st[0][f], snaps[k], sizes[k], hipMemcpyDeviceToDevice));
    }
    BCHK(b, hipMemcpy(b->ctrls, ctrl_host.data(), cbytes, hipMemcpyHostToDevice));
    return HF2D_OK;
  };
  hipEvent_t e0, e1;
  BCHK(b, hipEventCreate(&e0));
  BCHK(b, hipEventCreate(&e1));
  int best = 0;
  double best_ms = 1e300;
  for (int c = 0; c < Cfgs<T>::n && rc == HF2D_OK; ++c) {
    b->cfgi = c;
    b->dirty = true;
    rc = step_n<T>(b, 2, 0);  // warm (also rebuilds)
    if (rc) break;
    BCHK(b, hipEventRecord(e0, b->stream));
    rc = step_n<T>(b, steps, 0);
    BCHK(b, hipEventRecord(e1, b->stream));
    BCHK(b, hipEventSynchronize(e1));
    float ms = 0.f;
    BCHK(b, hipEventElapsedTime(&ms, e0, e1));
    if (ms_out) ms_out[c] = ms / steps;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
    int r2 = restore();
    if (r2) rc = r2;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (void* p : snaps) (void)hipFree(p);
  b->cfgi = best;
  b->dirty = true;
  if (best_out) *best_out = best;
  return rc;
}

}  // namespace

extern "C" {

int hf2d_batch_create(int32_t scalar_kind, int64_t capacity_bytes, int32_t max_slots, hf2d_batch** out) {
  if (!out || capacity_bytes <= 0 || max_slots <= 0 || max_slots > HF2D_BATCH_MAX_SLOTS) return HF2D_INVALID_ARGUMENT;
  if (scalar_kind != HF2D_F64 && scalar_kind != HF2D_DUAL && scalar_kind != HF2D_F32) return HF2D_INVALID_ARGUMENT;
  *out = nullptr;
  hf2d_batch* b = new hf2d_batch();
  b->kind = scalar_kind;
  b->cfgi = scalar_kind == HF2D_F32 ? CFG_F32_DEFAULT : scalar_kind == HF2D_F64 ? CFG_F64_DEFAULT : CFG_DUAL_DEFAULT;
  b->cap = up((size_t)capacity_bytes);
  b->sc.resize((size_t)max_slots);
  auto bad = [&]() {
    hf2d_batch_destroy(b);
    return HF2D_HIP_ERROR;
  };
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return bad();
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bad();
  if (hipMalloc(&b->arena, b->cap) != hipSuccess) return bad();
  b->free_blocks[0] = b->cap;
  const size_t ds = scalar_kind == HF2D_DUAL ? desc_size<dual>() : scalar_kind == HF2D_F32 ? desc_size<float>() : desc_size<double>();
  if (hipMalloc(&b->ctrls, (size_t)max_slots * sizeof(Ctrl)) != hipSuccess) return bad();
  if (hipMemset(b->ctrls, 0, (size_t)max_slots * sizeof(Ctrl)) != hipSuccess) return bad();
  if (hipMalloc(&b->descs1, (size_t)max_slots * ds) != hipSuccess) return bad();
  if (hipMalloc(&b->descs2, (size_t)max_slots * ds) != hipSuccess) return bad();
  if (hipMalloc(&b->slots_dev, (size_t)max_slots * sizeof(int)) != hipSuccess) return bad();
  const size_t wsb = (size_t)max_slots * BATCH_WS_SLOTS * WS_STRIDE * sizeof(unsigned long long);
  if (hipMalloc(&b->ws_part, wsb) != hipSuccess) return bad();
  if (hipMemset(b->ws_part, 0, wsb) != hipSuccess) return bad();
  *out = b;
  return HF2D_OK;
}

int hf2d_batch_destroy(hf2d_batch* b) {
  if (!b) return HF2D_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  destroy_graph(b);
  void* bufs[] = {b->arena, b->ctrls, b->descs1, b->descs2, b->work, b->slots_dev, b->ws_part};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return HF2D_OK;
}

const char* hf2d_batch_last_error(const hf2d_batch* b) { return b ? b->err.c_str() : "null batch"; }

int hf2d_batch_footprint(int32_t scalar_kind, int64_t n_rows, int64_t n_cols, int64_t n_sources, int64_t* bytes) {
  if (!bytes || n_rows <= 0 || n_cols <= 0) return HF2D_INVALID_ARGUMENT;
  if (scalar_kind != HF2D_F64 && scalar_kind != HF2D_DUAL && scalar_kind != HF2D_F32) return HF2D_INVALID_ARGUMENT;
  *bytes = (int64_t)footprint(scalar_kind, n_rows, n_cols, n_sources);
  return HF2D_OK;
}

int hf2d_batch_free_bytes(const hf2d_batch* b, int64_t* largest_block, int64_t* total) {
  if (!b) return HF2D_INVALID_ARGUMENT;
  size_t lg = 0, tot = 0;
  for (auto& kv : b->free_blocks) {
    lg = std::max(lg, kv.second);
    tot += kv.second;
  }
  if (largest_block) *largest_block = (int64_t)lg;
  if (total) *total = (int64_t)tot;
  return HF2D_OK;
}

int hf2d_batch_add(hf2d_batch* b, int64_t nr, int64_t nc, double dx, double dy, const double* bed,
                   const double* manning, const double* manning_tangent, const double* state_aos,
                   const double* tangent_aos, const int32_t bc_kinds[4], const double bc_params[4],
                   const hf2d_config* cfg, const hf2d_point_source* srcs, int64_t n_src, double t_end,
                   int32_t* slot_out) {
  if (!b || !bed || !manning || !state_aos || !bc_kinds || !cfg || !slot_out) return HF2D_INVALID_ARGUMENT;
  if (!(dx > 0.0) || !(dy > 0.0) || nr <= 0 || nc <= 0 || nr >= HF2D_MAX_SIDE || nc >= HF2D_MAX_SIDE)
    return bfail(b, HF2D_INVALID_CONFIG, "bad grid");
  if (!(cfg->cfl > 0.0 && cfg->cfl <= 1.0) || !(cfg->max_dt > 0.0) || cfg->integrator != HF2D_SSPRK2)
    return bfail(b, HF2D_INVALID_CONFIG, "batch scenarios need cfl in (0,1], max_dt > 0 and SSP-RK2");
  for (int64_t k = 0; k < n_src; ++k)
    if (srcs[k].row < 0 || srcs[k].row >= nr || srcs[k].col < 0 || srcs[k].col >= nc)
      return bfail(b, HF2D_SHAPE_MISMATCH, "point source outside the grid");
  int slot = -1;
  for (size_t k = 0; k < b->sc.size(); ++k)
    if (!b->sc[k].used) {
      slot = (int)k;
      break;
    }
  if (slot < 0) return bfail(b, HF2D_CAPACITY, "no free scenario slot");
  const size_t bytes = footprint(b->kind, nr, nc, n_src);
  size_t off = 0;
  if (!arena_alloc(b, bytes, &off)) return bfail(b, HF2D_CAPACITY, "scenario does not fit in the HBM arena");
  Scen s;
  s.used = true;
  s.nr = nr;
  s.nc = nc;
  s.dx = dx;
  s.dy = dy;
  s.off = off;
  s.bytes = bytes;
  const size_t n = (size_t)(nr * nc), e = elem_bytes(b->kind);
  char* p = b->arena + off;
  auto take = [&](size_t nb) {
    void* r = p;
    p += up(nb);
    return r;
  };
  for (int bb = 0; bb < 2; ++bb)
    for (int f = 0; f < 3; ++f) s.st[bb][f] = take(n * e);
  if (b->kind == HF2D_DUAL)
    for (int bb = 0; bb < 2; ++bb)
      for (int f = 0; f < 3; ++f) s.tg[bb][f] = take(n * e);
  std::vector<char> host;
  auto put = [&](void* dst, const void* src, size_t nb) -> int {
    hipError_t err = hipMemcpy(dst, src, nb, hipMemcpyHostToDevice);
    return err == hipSuccess ? HF2D_OK : bfail(b, HF2D_HIP_ERROR, hipGetErrorString(err));
  };
  int rc = HF2D_OK;
  if (b->kind == HF2D_F32) {
    const int64_t B = 1 << DATUM_LOG2, rows = (nr + B - 1) / B;
    s.datum_cols = (nc + B - 1) / B;
    s.datum = (double*)take((size_t)(rows * s.datum_cols) * 8);
    s.zdev = (float*)take(n * 4);
    std::vector<double> dat((size_t)(rows * s.datum_cols));
    for (int64_t bi = 0; bi < rows; ++bi)
      for (int64_t bj = 0; bj < s.datum_cols; ++bj) {
        double sum = 0.0;
        int64_t cnt = 0;
        for (int64_t i = bi * B; i < std::min(nr, (bi + 1) * B); ++i)
          for (int64_t j = bj * B; j < std::min(nc, (bj + 1) * B); ++j, ++cnt) sum += bed[i * nc + j];
        dat[(size_t)(bi * s.datum_cols + bj)] = sum / (double)cnt;
      }
    std::vector<float> dev(n);
    for (int64_t i = 0; i < nr; ++i)
      for (int64_t j = 0; j < nc; ++j)
        dev[(size_t)(i * nc + j)] =
            (float)(bed[i * nc + j] - dat[(size_t)((i >> DATUM_LOG2) * s.datum_cols + (j >> DATUM_LOG2))]);
    if (!rc) rc = put(s.datum, dat.data(), dat.size() * 8);
    if (!rc) rc = put(s.zdev, dev.data(), n * 4);
  } else {
    s.z = (double*)take(n * 8);
    if (!rc) rc = put(s.z, bed, n * 8);
    if (b->kind == HF2D_DUAL) {
      s.zt = (double*)take(n * 8);
      std::vector<double> zero(n, 0.0);
      if (!rc) rc = put(s.zt, zero.data(), n * 8);
    }
  }
  s.man = take(n * e);
  if (b->kind == HF2D_DUAL) s.mant = take(n * e);
  // Manning, state planes and tangents in the grid's value type.
  auto put_plane = [&](void* dst, const double* src, size_t stride, size_t comp) -> int {
    if (b->kind == HF2D_F32) {
      std::vector<float> v(n);
      for (size_t k = 0; k < n; ++k) v[k] = src ? (float)src[stride * k + comp] : 0.0f;
      return put(dst, v.data(), n * 4);
    }
    std::vector<double> v(n);
    for (size_t k = 0; k < n; ++k) v[k] = src ? src[stride * k + comp] : 0.0;
    return put(dst, v.data(), n * 8);
  };
  if (!rc) rc = put_plane(s.man, manning, 1, 0);
  if (!rc && b->kind == HF2D_DUAL) rc = put_plane(s.mant, manning_tangent, 1, 0);
  for (int f = 0; f < 3 && !rc; ++f) {
    rc = put_plane(s.st[0][f], state_aos, 3, (size_t)f);
    if (!rc && b->kind == HF2D_DUAL) rc = put_plane(s.tg[0][f], tangent_aos, 3, (size_t)f);
  }
  s.srcs = (PointSrc*)take((size_t)std::max<int64_t>(n_src, 1) * sizeof(PointSrc));
  if (n_src > 0 && !rc) {
    std::vector<PointSrc> v((size_t)n_src);
    for (int64_t k = 0; k < n_src; ++k) v[(size_t)k] = PointSrc{srcs[k].row, srcs[k].col, srcs[k].q_mass};
    rc = put(s.srcs, v.data(), v.size() * sizeof(PointSrc));
  }
  s.n_src = (int)n_src;
  for (int k = 0; k < 4; ++k) {
    s.bc_kind[k] = bc_kinds[k];
    s.bc_param[k] = bc_params ? bc_params[k] : 0.0;
  }
  s.cfg = *cfg;
  s.t_end = t_end;
  if (rc) {
    arena_free(b, off, bytes);
    return rc;
  }
  b->sc[(size_t)slot] = s;
  Ctrl c{};
  c.t_target = t_end;
  c.use_target = 1;
  c.dt_cap = __builtin_inf();
  c.budget = cfg->max_steps;
  c.compensated = b->kind == HF2D_F32 ? 1 : 0;
  BCHK(b, hipMemcpy(b->ctrls + slot, &c, sizeof(Ctrl), hipMemcpyHostToDevice));
  rc = HF_BDISPATCH(b, init_wave_speeds, b, slot);
  if (rc) return rc;
  b->dirty = true;
  *slot_out = slot;
  return HF2D_OK;
}

int hf2d_batch_remove(hf2d_batch* b, int32_t slot) {
  if (!b || slot < 0 || slot >= (int)b->sc.size() || !b->sc[(size_t)slot].used) return HF2D_INVALID_ARGUMENT;
  BCHK(b, hipStreamSynchronize(b->stream));
  Scen& s = b->sc[(size_t)slot];
  arena_free(b, s.off, s.bytes);
  s = Scen();
  b->dirty = true;
  return HF2D_OK;
}

int hf2d_batch_step(hf2d_batch* b, int64_t n_steps, int32_t use_graph) {
  if (!b) return HF2D_INVALID_ARGUMENT;
  return HF_BDISPATCH(b, step_n, b, n_steps, use_graph);
}

int hf2d_batch_synchronize(hf2d_batch* b) {
  if (!b) return HF2D_INVALID_ARGUMENT;
  BCHK(b, hipStreamSynchronize(b->stream));
  return HF2D_OK;
}

int hf2d_batch_query(hf2d_batch* b, int32_t* idle) {
  if (!b || !idle) return HF2D_INVALID_ARGUMENT;
  const hipError_t e = hipStreamQuery(b->stream);
  if (e == hipErrorNotReady) {
    *idle = 0;
    return HF2D_OK;
  }
  BCHK(b, e);
  *idle = 1;
  return HF2D_OK;
}

// Clock and status of every slot: t, steps, status (0 ok, 3 degenerate dt,
// 4 step budget), done (reached t_end), used.
int hf2d_batch_poll(hf2d_batch* b, int32_t n_slots, double* t, int64_t* steps, int32_t* status, int32_t* done,
                    int32_t* used) {
  if (!b || n_slots < 0 || n_slots > (int)b->sc.size()) return HF2D_INVALID_ARGUMENT;
  std::vector<Ctrl> c((size_t)n_slots);
  if (n_slots > 0) {
    BCHK(b, hipMemcpyAsync(c.data(), b->ctrls, (size_t)n_slots * sizeof(Ctrl), hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
  }
  for (int k = 0; k < n_slots; ++k) {
    const bool u = b->sc[(size_t)k].used;
    if (t) t[k] = u ? c[(size_t)k].t : 0.0;
    if (steps) steps[k] = u ? c[(size_t)k].steps : 0;
    if (status) status[k] = u ? c[(size_t)k].status : 0;
    if (done) done[k] = u ? (c[(size_t)k].done || c[(size_t)k].status) : 0;
    if (used) used[k] = u ? 1 : 0;
  }
  return HF2D_OK;
}

namespace {
// Copy slot's state (and Dual tangents) into AoS host arrays. The caller
// makes sure no enqueued launch writes the slot's state.
int copy_out(hf2d_batch* b, int32_t slot, double* state_aos, double* tangent_aos) {
  Scen& s = b->sc[(size_t)slot];
  const size_t n = (size_t)(s.nr * s.nc);
  for (int f = 0; f < 3; ++f) {
    for (int tan = 0; tan < 2; ++tan) {
      double* out = tan ? tangent_aos : state_aos;
      if (!out || (tan && b->kind != HF2D_DUAL)) continue;
      const void* src = tan ? s.tg[0][f] : s.st[0][f];
      if (b->kind == HF2D_F32) {
        std::vector<float> v(n);
        BCHK(b, hipMemcpy(v.data(), src, n * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) out[3 * k + (size_t)f] = (double)v[k];
      } else {
        std::vector<double> v(n);
        BCHK(b, hipMemcpy(v.data(), src, n * 8, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) out[3 * k + (size_t)f] = v[k];
      }
    }
  }
  return HF2D_OK;
}
}  // namespace

int hf2d_batch_download(hf2d_batch* b, int32_t slot, double* state_aos, double* tangent_aos) {
  if (!b || slot < 0 || slot >= (int)b->sc.size() || !b->sc[(size_t)slot].used || !state_aos)
    return HF2D_INVALID_ARGUMENT;
  BCHK(b, hipStreamSynchronize(b->stream));
  return copy_out(b, slot, state_aos, tangent_aos);
}

int hf2d_batch_download_done(hf2d_batch* b, int32_t slot, double* state_aos, double* tangent_aos) {
  if (!b || slot < 0 || slot >= (int)b->sc.size() || !b->sc[(size_t)slot].used || !state_aos)
    return HF2D_INVALID_ARGUMENT;
  // The slot's control block, read without waiting for the batch's stream
  // (the copy runs on the null stream, which does not synchronize with it):
  // once done or failed, a slot's later steps are no-ops that write nothing.
  Ctrl c;
  BCHK(b, hipMemcpy(&c, b->ctrls + slot, sizeof(Ctrl), hipMemcpyDeviceToHost));
  if (!(c.done || c.status)) return bfail(b, HF2D_INVALID_ARGUMENT, "slot is still running");
  return copy_out(b, slot, state_aos, tangent_aos);
}

int hf2d_batch_config_count(const hf2d_batch* b, int32_t* count) {
  if (!b || !count) return HF2D_INVALID_ARGUMENT;
  *count = b->kind == HF2D_DUAL ? Cfgs<dual>::n : b->kind == HF2D_F32 ? Cfgs<float>::n : Cfgs<double>::n;
  return HF2D_OK;
}

int hf2d_batch_config_info(const hf2d_batch* b, int32_t index, int32_t* tile_w, int32_t* tile_h, int32_t* threads) {
  int32_t n = 0;
  if (hf2d_batch_config_count(b, &n) || index < 0 || index >= n) return HF2D_INVALID_ARGUMENT;
  const CfgInfo* L = b->kind == HF2D_DUAL ? CFG_DUAL : b->kind == HF2D_F32 ? CFG_F32 : CFG_F64;
  if (tile_w) *tile_w = L[index].tw;
  if (tile_h) *tile_h = L[index].th;
  if (threads) *threads = L[index].nt;
  return HF2D_OK;
}

int hf2d_batch_set_config(hf2d_batch* b, int32_t index) {
  int32_t n = 0;
  if (hf2d_batch_config_count(b, &n) || index < 0 || index >= n) return HF2D_INVALID_ARGUMENT;
  if (b->cfgi != index) {
    b->cfgi = index;
    b->dirty = true;
  }
  return HF2D_OK;
}

int hf2d_batch_autotune(hf2d_batch* b, int32_t steps, int32_t* best, double* ms_per_step) {
  if (!b || steps <= 0) return HF2D_INVALID_ARGUMENT;
  return HF_BDISPATCH(b, autotune, b, steps, best, ms_per_step);
}

}  // extern "C"
