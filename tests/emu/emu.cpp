// CPU build of the stage-kernel tile code (hydroflux_amd/csrc/hf2d_tile.h).
//
// TEST INFRASTRUCTURE ONLY. Runs the same phase functions the gfx950
// kernels run, one thread index after another per barrier-separated phase
// (a legal schedule of the kernel), with every tile-local LDS index
// bounds-checked. Used to check the kernel logic against the oracle on the
// CPU and to find indexing faults without a GPU. The device driver's
// sequencing (wave-speed refresh, first stage, final stage, finalize) is
// mirrored here on a host copy of the control block.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#define HF_BOUNDS_CHECK(cond)                                                   \
  do {                                                                          \
    if (!(cond)) {                                                              \
      std::fprintf(stderr, "LDS index out of bounds: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
      std::abort();                                                             \
    }                                                                           \
  } while (0)
#include "../../hydroflux_amd/csrc/hf2d_tile.h"

using namespace hf;

namespace {

struct Emu {
  int64_t nr, nc;
  double dx, dy;
  int dual_kind;
  int tw, th;
  std::vector<double> st[2][3], tg[2][3];
  std::vector<double> z, zt, man, mant;
  bool has_zt = false;
  int cur = 0;
  Ctrl ctrl{};
  int bc_kind[4] = {BC_WALL, BC_WALL, BC_WALL, BC_WALL};
  double bc_param[4] = {0, 0, 0, 0};
  double cfl = 0.4, max_dt = 60.0, fdt = H_DRY, rain = 0.0;
  int integrator = 1, with_bcs = 1, friction = 1;
  int64_t max_steps = 10000000;
  std::vector<PointSrc> srcs;
  double tmp_rain = 0.0, tmp_tol = 0.0;
  // active-tile list, as the device driver keeps it
  int list_on = 1;
  bool list_valid = false;
  // run_until in chunks with trailing no-op steps, as the device driver
  // (0: start no step that will not run)
  int chunked = 1;
  std::vector<unsigned char> wet, run, pin;
  std::vector<int> list;
};

// Tiles the active-tile list left out (all emulator steps).
int64_t g_tiles_skipped = 0;

template <class T> Fld<T> fld(const std::vector<double>& v, const std::vector<double>& d);
template <> Fld<double> fld<double>(const std::vector<double>& v, const std::vector<double>&) {
  return Fld<double>{v.data()};
}
template <> Fld<dual> fld<dual>(const std::vector<double>& v, const std::vector<double>& d) {
  return Fld<dual>{v.data(), d.empty() ? nullptr : d.data()};
}
template <class T> FldW<T> fldw(std::vector<double>& v, std::vector<double>& d);
template <> FldW<double> fldw<double>(std::vector<double>& v, std::vector<double>&) { return FldW<double>{v.data()}; }
template <> FldW<dual> fldw<dual>(std::vector<double>& v, std::vector<double>& d) {
  return FldW<dual>{v.data(), d.data()};
}

template <class T> StageArgs<T> args(Emu& e, int in, int out, int un, int tw, int th) {
  StageArgs<T> a{};
  a.nr = e.nr;
  a.nc = e.nc;
  a.dx = e.dx;
  a.dy = e.dy;
  a.ih = fld<T>(e.st[in][0], e.tg[in][0]);
  a.ihu = fld<T>(e.st[in][1], e.tg[in][1]);
  a.ihv = fld<T>(e.st[in][2], e.tg[in][2]);
  a.oh = fldw<T>(e.st[out][0], e.tg[out][0]);
  a.ohu = fldw<T>(e.st[out][1], e.tg[out][1]);
  a.ohv = fldw<T>(e.st[out][2], e.tg[out][2]);
  a.nh = fld<T>(e.st[un][0], e.tg[un][0]);
  a.nhu = fld<T>(e.st[un][1], e.tg[un][1]);
  a.nhv = fld<T>(e.st[un][2], e.tg[un][2]);
  static const std::vector<double> none;
  a.z = fld<T>(e.z, e.has_zt ? e.zt : none);
  a.man = fld<T>(e.man, e.mant);
  for (int k = 0; k < 4; ++k) {
    a.bc_kind[k] = e.bc_kind[k];
    a.bc_param[k] = e.bc_param[k];
  }
  a.ctrl = &e.ctrl;
  a.cfl = e.cfl;
  a.max_dt = e.max_dt;
  a.cfl_with_bcs = e.with_bcs;
  a.srcs = e.srcs.data();
  a.tiles_x = (int)((e.nc + tw - 1) / tw);
  a.tiles_y = (int)((e.nr + th - 1) / th);
  return a;
}

// One launch of the stage kernel: every tile (or every tile of the active
// list), every phase, every thread.
template <class T, int TW, int TH>
void launch(const StageArgs<T>& a, int mode, unsigned long long* dst) {
  using St = Stage<T, TW, TH>;
  std::unique_ptr<typename St::S> s(new typename St::S());
  constexpr int NT = 256;
  const int64_t nb = a.tile_list ? *a.n_work : (int64_t)a.tiles_x * a.tiles_y;
  double gx = 0.0, gy = 0.0;
  double dt;
  if (!stage_prologue(a, true, &dt)) return;  // workgroup 0's prologue
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t tile = a.tile_list ? a.tile_list[b] : b;
    const int64_t gi0 = (tile / a.tiles_x) * TH, gj0 = (tile % a.tiles_x) * TW;
    std::vector<typename St::Pre> pre(NT);
    for (int t = 0; t < NT; ++t) {
      if (mode == MODE_STAGE)
        St::template prefetch<MODE_STAGE>(a, gi0, gj0, t, &pre[t]);
      else
        St::template prefetch<MODE_FINAL_RK2>(a, gi0, gj0, t, &pre[t]);
    }
    for (int t = 0; t < NT; ++t) St::template phase_load<NT>(*s, a, gi0, gj0, t);
    for (int t = 0; t < NT; ++t) St::phase_diffs(*s, a, gi0, gj0, t, NT);
    for (int t = 0; t < NT; ++t) St::phase_slopes(*s, a, gi0, gj0, t, NT);
    for (int t = 0; t < NT; ++t) St::phase_faces(*s, a, gi0, gj0, t, NT);
    for (int t = 0; t < NT; ++t) St::phase_alpha(*s, a, gi0, gj0, t, NT, dt);
    bool wet = false;
    for (int t = 0; t < NT; ++t) {
      double wx = 0.0, wy = 0.0;
      if (mode == MODE_STAGE)
        St::template phase_update<MODE_STAGE>(*s, a, gi0, gj0, t, NT, dt, &wx, &wy, &wet, &pre[t]);
      else
        St::template phase_update<MODE_FINAL_RK2>(*s, a, gi0, gj0, t, NT, dt, &wx, &wy, &wet, &pre[t]);
      gx = rmax(gx, wx);
      gy = rmax(gy, wy);
    }
    if (a.tile_wet && mode == MODE_FINAL_RK2) a.tile_wet[tile] = wet ? 1 : 0;
  }
  if (dst && mode == MODE_FINAL_RK2 && a.wave_speed) {
    if (hfm::as_u64(gx) > dst[0]) dst[0] = hfm::as_u64(gx);
    if (hfm::as_u64(gy) > dst[1]) dst[1] = hfm::as_u64(gy);
  }
}

// Elementwise tail (forcing + CFL maxima), as tail_kernel.
template <class T> void tail(const StageArgs<T>& a, unsigned long long* dst) {
  using St = Stage<T, 32, 16>;
  double dt;
  if (!stage_prologue(a, true, &dt)) return;
  double wx = 0.0, wy = 0.0;
  const typename St::ForcingNow fnow = St::forcing_now(a, 0, 0, a.nr, a.nc);
  for (int64_t idx = 0; idx < (int64_t)a.nr * a.nc; ++idx) {
    const int gi = (int)(idx / a.nc), gj = (int)(idx % a.nc);
    T h = a.ih.ld(idx), hu = a.ihu.ld(idx), hv = a.ihv.ld(idx);
    St::forcing(a, fnow, gi, gj, dt, &h, &hu, &hv);
    a.oh.st(idx, h);
    a.ohu.st(idx, hu);
    a.ohv.st(idx, hv);
    if (a.wave_speed) {
      T zc = a.z.ld(idx);
      T zn4[4];
      zn4[SIDE_N] = (a.nr >= 2 && gi == 0) ? a.z.ld(idx + a.nc) : zc;
      zn4[SIDE_S] = (a.nr >= 2 && gi == a.nr - 1) ? a.z.ld(idx - a.nc) : zc;
      zn4[SIDE_W] = (a.nc >= 2 && gj == 0) ? a.z.ld(idx + 1) : zc;
      zn4[SIDE_E] = (a.nc >= 2 && gj == a.nc - 1) ? a.z.ld(idx - 1) : zc;
      T man = a.man.ld(idx);
      St::cell_wave_speed(a, gi, gj, h, hu, hv, zc, zn4, &man, &wx, &wy);
    }
  }
  if (a.wave_speed && dst) {
    if (hfm::as_u64(wx) > dst[0]) dst[0] = hfm::as_u64(wx);
    if (hfm::as_u64(wy) > dst[1]) dst[1] = hfm::as_u64(wy);
  }
}

template <class T, int TW, int TH> void refresh(Emu& e, bool with_bcs) {
  e.ctrl.ws[e.ctrl.parity][0] = 0;
  e.ctrl.ws[e.ctrl.parity][1] = 0;
  StageArgs<T> a = args<T>(e, e.cur, e.cur, e.cur, TW, TH);
  a.dt_mode = 0;
  a.wave_speed = 1;
  a.cfl_with_bcs = with_bcs;
  tail<T>(a, e.ctrl.ws[e.ctrl.parity]);
}

template <class T, int TW, int TH> void host_dt_step(Emu& e, double dt, bool rk2) {
  const int cur = e.cur, oth = 1 - cur;
  StageArgs<T> a = args<T>(e, cur, oth, cur, TW, TH);
  a.dt_mode = 0;
  a.dt_host = dt;
  launch<T, TW, TH>(a, MODE_STAGE, nullptr);
  if (rk2) {
    StageArgs<T> b = args<T>(e, oth, cur, cur, TW, TH);
    b.dt_mode = 0;
    b.dt_host = dt;
    launch<T, TW, TH>(b, MODE_FINAL_RK2, nullptr);
  } else {
    e.cur = oth;
  }
}

// The device driver's active-tile list (hf2d.hip enqueue_compact), on the
// host: same usability rule, same per-tile decision (tile_skipped).
template <class T, int TW, int TH> bool build_list(Emu& e, StageArgs<T>& a, StageArgs<T>& b, int64_t* n_work) {
  const bool usable = !e.dual_kind && e.integrator == 1 && e.rain == 0.0 && (!e.friction || e.fdt >= H_DRY) &&
                      e.list_on;
  if (!usable) {
    e.list_valid = false;
    return false;
  }
  const int tx = a.tiles_x, ty = a.tiles_y;
  const size_t nt = (size_t)tx * ty;
  e.pin.assign(nt, 0);
  for (int y = 0; y < ty; ++y)
    for (int x = 0; x < tx; ++x) e.pin[(size_t)y * tx + x] = tile_near_edge(e.nr, e.nc, TW, TH, x, y);
  for (const PointSrc& p : e.srcs) e.pin[(size_t)(p.row / TH) * tx + (size_t)(p.col / TW)] = 1;
  if (!e.list_valid || e.wet.size() != nt) {
    e.wet.assign(nt, 0);
    e.run.assign(nt, 0);
    for (int64_t i = 0; i < e.nr * e.nc; ++i)
      if (!(e.st[e.cur][0][(size_t)i] <= H_DRY)) e.wet[(size_t)((i / e.nc) / TH * tx + (i % e.nc) / TW)] = 1;
    e.list_valid = true;
  }
  // compact_kernel: a step that will not run leaves counts and list alone
  if (!step_will_run(&e.ctrl, e.cfl, e.dx, e.dy, e.max_dt)) {
    *n_work = (int64_t)e.list.size();
    a.tile_list = b.tile_list = e.list.data();
    b.tile_wet = e.wet.data();
    return true;
  }
  e.list.clear();
  for (int y = 0; y < ty; ++y)
    for (int x = 0; x < tx; ++x) {
      if (tile_skipped(e.wet.data(), e.pin.data(), e.run.data(), tx, ty, x, y))
        ++g_tiles_skipped;
      else
        e.list.push_back(y * tx + x);
    }
  *n_work = (int64_t)e.list.size();
  a.tile_list = b.tile_list = e.list.data();
  b.tile_wet = e.wet.data();
  return true;
}

template <class T, int TW, int TH> void driver_step(Emu& e) {
  const int cur = e.cur, oth = 1 - cur;
  StageArgs<T> a = args<T>(e, cur, oth, cur, TW, TH);
  a.dt_mode = 1;
  a.first_stage = 1;
  StageArgs<T> b = args<T>(e, oth, cur, cur, TW, TH);
  b.dt_mode = 1;
  b.first_stage = 0;
  b.friction = e.friction;
  b.friction_dry_tol = e.fdt;
  b.rain_rate = e.rain;
  b.n_src = (int)e.srcs.size();
  b.wave_speed = 1;
  int64_t n_work64 = 0;
  int n_work = 0;
  if (build_list<T, TW, TH>(e, a, b, &n_work64)) {
    n_work = (int)n_work64;
    a.n_work = b.n_work = &n_work;
  }
  launch<T, TW, TH>(a, MODE_STAGE, nullptr);
  if (e.integrator == 1) {
    launch<T, TW, TH>(b, MODE_FINAL_RK2, e.ctrl.ws[e.ctrl.parity ^ 1]);
  } else {
    StageArgs<T> t = b;
    t.ih = fld<T>(e.st[oth][0], e.tg[oth][0]);
    t.ihu = fld<T>(e.st[oth][1], e.tg[oth][1]);
    t.ihv = fld<T>(e.st[oth][2], e.tg[oth][2]);
    t.oh = fldw<T>(e.st[cur][0], e.tg[cur][0]);
    t.ohu = fldw<T>(e.st[cur][1], e.tg[cur][1]);
    t.ohv = fldw<T>(e.st[cur][2], e.tg[cur][2]);
    tail<T>(t, e.ctrl.ws[e.ctrl.parity ^ 1]);
  }
  Ctrl& c = e.ctrl;  // finalize_kernel
  if (c.status != 0 || c.done != 0) return;
  c.t = c.t + c.dt;
  c.steps = c.steps + 1;
  c.budget = c.budget - 1;
  c.parity = c.parity ^ 1;
}

template <class T, int TW, int TH> int run_driver(Emu& e, bool use_target, double t_target, int64_t n) {
  refresh<T, TW, TH>(e, e.with_bcs != 0);
  Ctrl& c = e.ctrl;
  c.status = 0;
  c.done = 0;
  c.use_target = use_target;
  c.t_target = t_target;
  c.dt_cap = __builtin_inf();
  c.budget = use_target ? e.max_steps : n;
  if (use_target) {
    // hf2d_run_until enqueues chunks of 16, 32, ..., 256 steps and polls the
    // clock between chunks: the rest of the chunk in which the clock reaches
    // its target (or the driver fails) runs as no-op steps, as here.
    // Unchunked: no step is started that will not run; the first stage's
    // prologue alone records why the driver stopped.
    int64_t chunk = 16;
    for (;;) {
      if (!e.chunked && !step_will_run(&c, e.cfl, e.dx, e.dy, e.max_dt)) {
        double dt;
        StageArgs<T> a = args<T>(e, e.cur, 1 - e.cur, e.cur, TW, TH);
        a.dt_mode = 1;
        a.first_stage = 1;
        (void)stage_prologue(a, true, &dt);
        break;
      }
      for (int64_t k = 0; k < (e.chunked ? chunk : 1); ++k) driver_step<T, TW, TH>(e);
      if (c.status != 0 || c.done != 0) break;
      if (chunk < 256) chunk *= 2;
    }
  } else {
    for (int64_t k = 0; k < n && c.status == 0; ++k) driver_step<T, TW, TH>(e);
  }
  return c.status;
}

// Tile shapes the emulator supports: the device shapes plus small odd ones
// that put tile edges at every offset from the domain boundary.
template <class T> struct Dispatch {
  template <class F> static void with_shape(int tw, int th, F f);
};

#define HF_EMU_SHAPES(X) X(64, 24) X(32, 24) X(32, 16) X(16, 16) X(8, 4) X(4, 8) X(16, 8) X(24, 16) X(24, 8)

template <class T, int TW, int TH> struct Ops {
  static void ssprk2(Emu& e, double dt) { host_dt_step<T, TW, TH>(e, dt, true); }
  static void euler(Emu& e, double dt) { host_dt_step<T, TW, TH>(e, dt, false); }
  static int run(Emu& e, bool use_target, double tt, int64_t n) { return run_driver<T, TW, TH>(e, use_target, tt, n); }
  static void ws(Emu& e, bool with_bcs) { refresh<T, TW, TH>(e, with_bcs); }
  static void forcing(Emu& e, double dt, int fric, double tol, double rain, int nsrc) {
    StageArgs<T> a = args<T>(e, e.cur, e.cur, e.cur, TW, TH);
    a.dt_mode = 0;
    a.dt_host = dt;
    a.friction = fric;
    a.friction_dry_tol = tol;
    a.rain_rate = rain;
    a.n_src = nsrc;
    a.wave_speed = 0;
    tail<T>(a, nullptr);
  }
};

template <class T> int call(Emu& e, int op, double x, int64_t n, bool flag) {
#define HF_CASE(W, H)                                   \
  if (e.tw == W && e.th == H) {                         \
    switch (op) {                                       \
      case 0: Ops<T, W, H>::ssprk2(e, x); return 0;     \
      case 1: Ops<T, W, H>::euler(e, x); return 0;      \
      case 2: return Ops<T, W, H>::run(e, flag, x, n);  \
      case 3: Ops<T, W, H>::ws(e, flag); return 0;      \
      case 4: Ops<T, W, H>::forcing(e, x, 1, e.tmp_tol, 0.0, 0); return 0; \
      case 5: Ops<T, W, H>::forcing(e, x, 0, 0.0, e.tmp_rain, 0); return 0; \
      case 6: Ops<T, W, H>::forcing(e, x, 0, 0.0, 0.0, (int)e.srcs.size()); return 0; \
    }                                                   \
  }
  HF_EMU_SHAPES(HF_CASE)
#undef HF_CASE
  std::fprintf(stderr, "emu: unsupported tile %dx%d\n", e.tw, e.th);
  std::abort();
}

int dispatch(Emu& e, int op, double x, int64_t n, bool flag) {
  if (op != 2 && op != 3) e.list_valid = false;  // the state changed outside the driver
  return e.dual_kind ? call<dual>(e, op, x, n, flag) : call<double>(e, op, x, n, flag);
}

}  // namespace

extern "C" {

void* emu_create(int64_t nr, int64_t nc, double dx, double dy, int dual_kind, int tw, int th) {
  Emu* e = new Emu();
  e->nr = nr;
  e->nc = nc;
  e->dx = dx;
  e->dy = dy;
  e->dual_kind = dual_kind;
  e->tw = tw;
  e->th = th;
  const size_t n = (size_t)(nr * nc);
  for (int b = 0; b < 2; ++b)
    for (int f = 0; f < 3; ++f) {
      e->st[b][f].assign(n, 0.0);
      if (dual_kind) e->tg[b][f].assign(n, 0.0);
    }
  e->z.assign(n, 0.0);
  e->man.assign(n, 0.0);
  if (dual_kind) {
    e->zt.assign(n, 0.0);
    e->mant.assign(n, 0.0);
  }
  e->ctrl.dt_cap = __builtin_inf();
  return e;
}

void emu_destroy(void* p) { delete static_cast<Emu*>(p); }

void emu_set_mesh(void* p, const double* bed, const double* bt, const double* man, const double* mt) {
  Emu* e = static_cast<Emu*>(p);
  e->list_valid = false;
  const size_t n = e->z.size();
  for (size_t k = 0; k < n; ++k) {
    e->z[k] = bed[k];
    e->man[k] = man[k];
  }
  if (e->dual_kind) {
    e->has_zt = bt != nullptr;
    for (size_t k = 0; k < n; ++k) {
      e->zt[k] = bt ? bt[k] : 0.0;
      e->mant[k] = mt ? mt[k] : 0.0;
    }
  }
}

void emu_set_state(void* p, const double* aos, const double* tan) {
  Emu* e = static_cast<Emu*>(p);
  e->list_valid = false;
  const size_t n = e->z.size();
  for (int f = 0; f < 3; ++f)
    for (size_t k = 0; k < n; ++k) {
      e->st[e->cur][f][k] = aos[3 * k + f];
      if (e->dual_kind) e->tg[e->cur][f][k] = tan ? tan[3 * k + f] : 0.0;
    }
}

void emu_get_state(void* p, double* aos, double* tan) {
  Emu* e = static_cast<Emu*>(p);
  const size_t n = e->z.size();
  for (int f = 0; f < 3; ++f)
    for (size_t k = 0; k < n; ++k) {
      aos[3 * k + f] = e->st[e->cur][f][k];
      if (e->dual_kind && tan) tan[3 * k + f] = e->tg[e->cur][f][k];
    }
}

void emu_set_boundaries(void* p, const int* kinds, const double* params) {
  Emu* e = static_cast<Emu*>(p);
  e->list_valid = false;
  for (int k = 0; k < 4; ++k) {
    e->bc_kind[k] = kinds[k];
    e->bc_param[k] = params[k];
  }
}

void emu_configure(void* p, double cfl, int integrator, int with_bcs, double max_dt, int friction, double fdt,
                   int64_t max_steps, double rain, const int64_t* rows, const int64_t* cols, const double* q,
                   int64_t k) {
  Emu* e = static_cast<Emu*>(p);
  e->list_valid = false;
  e->cfl = cfl;
  e->integrator = integrator;
  e->with_bcs = with_bcs;
  e->max_dt = max_dt;
  e->friction = friction;
  e->fdt = fdt;
  e->max_steps = max_steps;
  e->rain = rain;
  e->srcs.clear();
  for (int64_t s = 0; s < k; ++s) e->srcs.push_back(PointSrc{rows[s], cols[s], q[s]});
}

void emu_set_active_tiles(void* p, int on) { static_cast<Emu*>(p)->list_on = on; }
int64_t emu_tiles_skipped() { return g_tiles_skipped; }

void emu_set_chunked(void* p, int on) { static_cast<Emu*>(p)->chunked = on; }

// The active-tile list's per-tile quiet-step counts (up to cap entries);
// returns their number (0 while the list is not in use).
int64_t emu_tile_runs(void* p, unsigned char* out, int64_t cap) {
  const Emu& e = *static_cast<Emu*>(p);
  const int64_t n = e.list_valid ? (int64_t)e.run.size() : 0;
  for (int64_t k = 0; k < n && k < cap; ++k) out[k] = e.run[(size_t)k];
  return n;
}

void emu_ssprk2_step(void* p, double dt) { dispatch(*static_cast<Emu*>(p), 0, dt, 0, false); }
void emu_forward_euler_step(void* p, double dt) { dispatch(*static_cast<Emu*>(p), 1, dt, 0, false); }
int emu_run_until(void* p, double t) { return dispatch(*static_cast<Emu*>(p), 2, t, 0, true); }
int emu_run_steps(void* p, int64_t n) { return dispatch(*static_cast<Emu*>(p), 2, 0.0, n, false); }

double emu_cfl_time_step(void* p, double cfl, int with_bcs) {
  Emu* e = static_cast<Emu*>(p);
  dispatch(*e, 3, 0.0, 0, with_bcs != 0);
  double sx = hfm::as_f64(e->ctrl.ws[e->ctrl.parity][0]);
  double sy = hfm::as_f64(e->ctrl.ws[e->ctrl.parity][1]);
  double denom = sx / e->dx + sy / e->dy;
  return denom > 0.0 ? cfl / denom : __builtin_inf();
}

void emu_manning_friction_step(void* p, double dt, double tol) {
  Emu* e = static_cast<Emu*>(p);
  e->tmp_tol = tol;
  dispatch(*e, 4, dt, 0, false);
}
void emu_apply_rain(void* p, double rate, double dt) {
  Emu* e = static_cast<Emu*>(p);
  if (rate == 0.0) return;
  e->tmp_rain = rate;
  dispatch(*e, 5, dt, 0, false);
}
// Point sources through the elementwise tail (each cell applies the sources
// that hit it, in list order).
void emu_apply_point_sources(void* p, const int64_t* rows, const int64_t* cols, const double* q, int64_t k,
                             double dt) {
  Emu* e = static_cast<Emu*>(p);
  std::vector<PointSrc> keep = e->srcs;
  e->srcs.clear();
  for (int64_t s = 0; s < k; ++s) e->srcs.push_back(PointSrc{rows[s], cols[s], q[s]});
  dispatch(*e, 6, dt, 0, false);
  e->srcs = keep;
}

void emu_time(void* p, double* t, int64_t* steps) {
  Emu* e = static_cast<Emu*>(p);
  *t = e->ctrl.t;
  *steps = e->ctrl.steps;
}

// The FP32 form's stochastically rounded addition (hf2d_tile.h), host build:
// out[i] = h[i] + dh[i] rounded with key keys[i].
void emu_add_rounded_stochastically(const float* h, const float* dh, const uint32_t* keys, int64_t n, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = hf::add_rounded_stochastically(h[i], dh[i], keys[i]);
}
}
