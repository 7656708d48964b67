// One SSP-RK2 / forward-Euler stage of hydroflux solver-2d on one output tile.
//
// The body is written as barrier-separated phases over a tile-local LDS
// struct so that the identical code runs (a) on gfx950 inside the
// __global__ wrappers of hf2d.hip, with __syncthreads() between phases, and
// (b) on the CPU in the tests' emulator, which runs each phase for every
// thread index in turn (a valid schedule of a barrier-separated program).
//
// Arithmetic follows the reference expression by expression
// (solver-2d/src/update.rs:215-1211, riemann.rs:114-204, boundary.rs:193-367,
// source.rs:52-165); the build uses -ffp-contract=off so no a*b+c is fused.
//
// Dependency radius (the reason for the halo sizes): a cell's update needs
// the four faces around it, each scaled by the upstream cell's alpha
// (update.rs:992-1053). Alpha of a ring-1 neighbour needs that neighbour's
// four raw faces, whose MUSCL slopes read primitives one cell further out,
// and a slope is zeroed if any cell within +-2 along its axis is dry
// (update.rs:433-444). Hence: state (h, z, u, v) on a 3-cell halo, dry flags
// on a 4-cell halo, slopes on a 2-cell halo (along their axis; 1 across;
// computed from per-face differences), faces and alpha on a 1-cell ring.
//
// Local coordinates: tile origin (gi0, gj0); li = gi - gi0 in [0, TH),
// lj = gj - gj0 in [0, TW). x-face fj sits between local cells fj-1 and fj;
// y-face fi between local rows fi-1 and fi.
#pragma once
#include <stdint.h>

#include <type_traits>

#include "hf_scalar.h"

namespace hf {

// solver-2d/src/lib.rs:43,63,72; riemann.rs:74; boundary.rs:68,76
constexpr double GRAVITY = 9.81;
constexpr double H_DRY = 1.0e-6;
constexpr double H_VEL = 1.0e-5;
constexpr double DRY_TOL = 1.0e-12;
constexpr double DRY_INFLOW_SLOPE_THRESHOLD = 5.0e-4;
constexpr double DRY_INFLOW_H_MAX = 10.0;

enum { BC_TRANSMISSIVE = 0, BC_WALL = 1, BC_DISCHARGE = 2, BC_DEPTH = 3 };
enum { SIDE_N = 0, SIDE_S = 1, SIDE_W = 2, SIDE_E = 3 };
enum { MODE_STAGE = 0, MODE_FINAL_RK2 = 1 };

#ifndef HF_PREFETCH
#define HF_PREFETCH 1
#endif

#ifndef HF_BOUNDS_CHECK
#define HF_BOUNDS_CHECK(cond) ((void)0)
#endif

// ------------------------------------------------------------ device fields
template <class T> struct Fld;
template <> struct Fld<double> {
  const double* v;
  HF_HD double ld(int64_t k) const { return v[k]; }
};
template <> struct Fld<dual> {
  const double* v;
  const double* d;  // may be null: tangent 0 (constant bed)
  HF_HD dual ld(int64_t k) const { return dual{v[k], d ? d[k] : 0.0}; }
};
// DualN<N>: value plane, tangents of a cell at d[cell * ds + k] (null: 0).
// ds == N for a whole DualN<N> field; a view of N components out of a
// wider field (ds > N, d offset to the first component) is how the stage
// kernels of DualN<M>, M > N, run over one group of components at a time.
template <int N> struct Fld<dualn<N>> {
  const double* v;
  const double* d;
  int ds;
  HF_HD dualn<N> ld(int64_t k) const {
    dualn<N> r;
    r.v = v[k];
    for (int i = 0; i < N; ++i) r.d[i] = d ? d[k * ds + i] : 0.0;
    return r;
  }
};
template <> struct Fld<float> {
  const float* v;
  HF_HD float ld(int64_t k) const { return v[k]; }
};
template <class T> struct FldW;
template <> struct FldW<float> {
  float* v;
  HF_HD void st(int64_t k, float x) const { v[k] = x; }
};
template <> struct FldW<double> {
  double* v;
  HF_HD void st(int64_t k, double x) const { v[k] = x; }
};
template <> struct FldW<dual> {
  double* v;
  double* d;
  HF_HD void st(int64_t k, dual x) const {
    v[k] = x.v;
    d[k] = x.d;
  }
};
// wv == 0: the view stores tangents only (values are stored by the view of
// the last component group, after every group has read them).
template <int N> struct FldW<dualn<N>> {
  double* v;
  double* d;
  int ds;
  int wv;
  HF_HD void st(int64_t k, const dualn<N>& x) const {
    if (wv) v[k] = x.v;
    for (int i = 0; i < N; ++i) d[k * ds + i] = x.d[i];
  }
};

// Device-resident driver state: the clock, the dt of the step in flight and
// two ping-pong slots of CFL wave-speed maxima (bit patterns of
// non-negative doubles, so unsigned max is the float max).
struct Ctrl {
  double t;
  double dt;
  double t_target;
  double dt_cap;
  unsigned long long ws[2][2];
  long long steps;
  long long budget;
  int status;  // 0 ok, 3 degenerate dt, 4 step budget exhausted
  int done;    // run_until reached its target
  int parity;  // ws[parity] holds the maxima of the current state
  int use_target;
  double t_comp;    // Kahan compensation of t (FP32 precision form)
  int compensated;  // 1: compensated clock; 0: plain t += dt (reference)
  int pad_;
  // forcing series of the driver, evaluated at t + dt by the first stage:
  // [0] rain rate [m/s], [1] point-source rate [m^3/s]
  double series_now[2];
  int series_on[2];
};

// The clock after a step of dt, as finalize_kernel advances it.
HF_HD double clock_after(const Ctrl* c, double dt) {
  if (c->compensated) return c->t + (dt - c->t_comp);
  return c->t + dt;  // sim.rs:225
}

// A piecewise-linear forcing series (hyetograph, hydrograph) of the
// reference's operational loops: value(t_s) = (interp(t_s / time_div) *
// scale_mul) / scale_div, with interp as uk_ea_common/mod.rs (interp:
// clamp at both ends, partition_point, frac = (t - t0) / (t1 - t0),
// v0 + frac * (v1 - v0)) and the scaling as uk_ea_test4_official.rs
// (discharge_at, then / inlet cells) and uk_ea_test8a_official.rs
// (interp_min on t / 60, then * 1e-3 / 3600 for rain, * cellsize for Q).
struct SeriesDesc {
  int n;  // 0: no series
  double time_div, scale_mul, scale_div;
  const double* t;
  const double* v;
};

HF_HD double series_at(const SeriesDesc& s, double t_s) {
  const double t = t_s / s.time_div;
  double v;
  if (t <= s.t[0]) {
    v = s.v[0];
  } else if (t >= s.t[s.n - 1]) {
    v = s.v[s.n - 1];
  } else {
    int idx = 0;
    while (idx < s.n && s.t[idx] <= t) ++idx;
    const double t0 = s.t[idx - 1], t1 = s.t[idx];
    const double v0 = s.v[idx - 1], v1 = s.v[idx];
    const double frac = (t - t0) / (t1 - t0);
    v = v0 + frac * (v1 - v0);
  }
  return (v * s.scale_mul) / s.scale_div;
}

struct PointSrc {  // source.rs:36-44
  int64_t row, col;
  double q;
};

// Launch arguments of a stage: the members that do not depend on the
// scalar type, then (StageArgs<T>) the field views.
struct StageCommon {
  // Grid size. Both sides are below 2^30 (hf2d_create and the batch engine
  // reject larger ones), so cell coordinates and tile origins are 32-bit;
  // element offsets into the planes are formed in 64 bits.
  int nr, nc;
  double dx, dy;
  int bc_kind[4];  // N, S, W, E
  double bc_param[4];
  int dt_mode;  // 0: dt_host; 1: from the device clock
  int first_stage;
  double dt_host;
  // Host-clocked launches (dt_mode 0): their sequence number, the FP32
  // form's rounding key in place of the device clock's step count.
  uint32_t host_seq;
  Ctrl* ctrl;
  double cfl, max_dt;
  int cfl_with_bcs;
  int friction;
  double friction_dry_tol;
  double rain_rate;
  int n_src;
  const PointSrc* srcs;
  int wave_speed;  // reduce the output's wave speeds into ctrl->ws[parity ^ 1]
  // Final driver stage: workgroups reduce into slot blockIdx % WS_SLOTS of
  // this buffer instead of into ctrl->ws; finalize_kernel folds the slots
  // in (null: straight into ctrl->ws).
  unsigned long long* ws_part;
  int tiles_x, tiles_y;
  // Compacted active-tile list (null: every tile). Block b of a launch
  // takes list position b of the first *n_work; tile_wet (final stage,
  // nullable) receives per tile whether an output cell is wet.
  const int* tile_list;
  const int* n_work;
  unsigned char* tile_wet;
  // forcing() takes the rain rate and point-source rate from the control
  // block's series values (at t + dt) when use_series is set.
  int use_series;
  // First driver stage: the lead thread, having published dt, evaluates
  // these two series (rain, point-source discharge) at t + dt into the
  // control block for the final stage (null: none).
  const SeriesDesc* series;
  // FP32 precision form: bed = datum[block] (FP64, one per 2^datum_log2
  // square block of cells) + zdev (FP32 deviation); `z` is unused then.
  const double* datum;
  const float* zdev;
  int datum_log2;
  int64_t datum_cols;
};

template <class T> struct StageArgs : StageCommon {
  Fld<T> ih, ihu, ihv;  // stencil input
  FldW<T> oh, ohu, ohv;  // output
  Fld<T> nh, nhu, nhv;  // U^n (final SSP-RK2 stage)
  Fld<T> z, man;
};

// Bed of one cell in the tile's frame. double/dual: the stored bed. float:
// the FP64 datum of the cell's block minus the tile datum, plus the FP32
// deviation, summed in FP64 and rounded once, so a tile on 2000 m relief
// works with beds of a few metres and h + z keeps its FP32 digits.
template <class T> HF_HD T load_bed(const StageArgs<T>& a, int gi, int gj, int64_t idx, double) {
  return a.z.ld(idx);
}
template <> HF_HD float load_bed<float>(const StageArgs<float>& a, int gi, int gj, int64_t idx,
                                        double tile_datum) {
  const double d = a.datum[(int64_t)(gi >> a.datum_log2) * a.datum_cols + (gj >> a.datum_log2)];
  return (float)((d - tile_datum) + (double)a.zdev[idx]);
}
template <class T> HF_HD double tile_datum(const StageArgs<T>&, int, int) { return 0.0; }
template <> HF_HD double tile_datum<float>(const StageArgs<float>& a, int gi, int gj) {
  return a.datum[(int64_t)(gi >> a.datum_log2) * a.datum_cols + (gj >> a.datum_log2)];
}

// ------------------------------------------------------------ small pieces
// update.rs:362-370
// FP32 form: h + dh with the rounding chosen at random in proportion to
// the exact rounding error, from a hash of key (cell and step), so that the
// sum is exact in expectation. Rain of about 1e-6 m per step is a few FP32
// ulps of a deep cell; rounded to nearest, every cell of a binade rounds the
// same way on every step, which lost 4.5 % of the rain volume on a closed
// box over 5,000 steps. Deterministic for a given key; h + dh > 0.
HF_HD uint32_t hash32(uint32_t x) {  // integer finalizer: xorshift-multiply rounds
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
HF_HD float add_rounded_stochastically(float h, float dh, uint32_t key) {
  const float s = h + dh;
  const float bp = s - h;
  const float e = (h - (s - bp)) + (dh - bp);  // exact: (h + dh) - s (TwoSum)
  if (e == 0.0f || !(s > 0.0f)) return s;
  uint32_t b;
  __builtin_memcpy(&b, &s, sizeof b);
  b = e > 0.0f ? b + 1u : b - 1u;  // the other neighbour of h + dh (s > 0)
  float nb;
  __builtin_memcpy(&nb, &b, sizeof nb);
  // take nb with probability |e| / |nb - s| (nb - s is exact), without a division
  const float u = (float)(hash32(key) >> 8) * 0x1p-24f;  // [0, 1)
  return u * __builtin_fabsf(nb - s) < __builtin_fabsf(e) ? nb : s;
}

template <class T> HF_HD T minmod(T a, T b) {
  using R = typename ScalarOf<T>::type;
  if (val(a) * val(b) <= R(0)) return cst<T>(0.0);
  if (rabs(val(a)) < rabs(val(b))) return a;
  return b;
}

// update.rs:650-656
template <class T> HF_HD T interior_z_face(bool l_dry, T z_l, bool r_dry, T z_r) {
  using R = typename ScalarOf<T>::type;
  if (l_dry != r_dry) return rmax(z_l, z_r);
  return (z_l + z_r) * R(0.5);
}

// boundary.rs:335-367. zc: bed of the boundary cell, zn: bed of the next
// inner cell along the face normal.
template <class T> HF_HD T ghost_bed(int kind, T zc, T zn) {
  using R = typename ScalarOf<T>::type;
  if (kind == BC_TRANSMISSIVE || kind == BC_WALL) return zc;
  return zc * R(2.0) - zn;
}

// boundary.rs:193-219 + 274-333. `d` is dx for West/East, dy for
// North/South; `have_next` says the mesh has two cells across the normal.
template <class T>
HF_HD void ghost_state(int kind, double param, int side, T h, T hu, T hv, T zc, T zn, bool have_next,
                       double d, T man, T* gh, T* ghu, T* ghv) {
  using R = typename ScalarOf<T>::type;
  const bool xf = (side == SIDE_W || side == SIDE_E);
  *gh = h;
  *ghu = hu;
  *ghv = hv;
  if (kind == BC_WALL) {
    if (xf)
      *ghu = -hu;
    else
      *ghv = -hv;
  } else if (kind == BC_DISCHARGE) {
    if (xf)
      *ghu = cst<T>(param);
    else
      *ghv = cst<T>(param);
    if (val(h) < R(H_DRY)) {
      // bed_slope_x/y (geometry.rs:131-146), magnitude only
      double slope = have_next ? __builtin_fabs((double)val(zc - zn) / d) : 0.0;
      if (slope > DRY_INFLOW_SLOPE_THRESHOLD) {
        T hn = pow35(man * R(__builtin_fabs(param) / __builtin_sqrt(slope)));
        *gh = rmin(hn, cst<T>(DRY_INFLOW_H_MAX));
      }
    }
  } else if (kind == BC_DEPTH) {
    *gh = cst<T>(param);
  }
}

// riemann.rs:114-204. The flux is the physical flux of one side K (the
// left state when sl >= 0, the right when sr <= 0, else the side the contact
// wave s* leaves upwind), plus the star-state correction when the wave fan
// spans the face. Side K's operands are selected first and its flux is
// evaluated once: the same expressions the reference evaluates on each of
// its branches, one evaluation per face instead of one per branch.
template <class T>
HF_HD void hllc_normal_flux(T h_l, T qn_l, T qt_l, T h_r, T qn_r, T qt_r, T* fm, T* fn, T* ft) {
  using R = typename ScalarOf<T>::type;
  const bool lw = val(h_l) > R(DRY_TOL);
  const bool rw = val(h_r) > R(DRY_TOL);
  const T Z = cst<T>(0.0);
  if (!lw && !rw) {
    *fm = Z;
    *fn = Z;
    *ft = Z;
    return;
  }
  T cl = lw ? rsqrt_wet(h_l * R(GRAVITY)) : Z;  // h > DRY_TOL: g h > 2^-767
  T cr = rw ? rsqrt_wet(h_r * R(GRAVITY)) : Z;
  T un_l = lw ? div(qn_l, h_l) : Z;
  T un_r = rw ? div(qn_r, h_r) : Z;
  T ut_l = lw ? div(qt_l, h_l) : Z;
  T ut_r = rw ? div(qt_r, h_r) : Z;
  T sl, sr;
  if (lw && rw) {
    sl = rmin(un_l - cl, un_r - cr);
    sr = rmax(un_l + cl, un_r + cr);
  } else if (rw) {
    sl = un_r - cr * R(2.0);
    sr = un_r + cr;
  } else {
    sl = un_l - cl;
    sr = un_l + cl * R(2.0);
  }
  const bool exit_l = val(sl) >= R(0);
  const bool star = !exit_l && !(val(sr) <= R(0));
  T s_star = Z;
  if (star) {
    T denom = h_r * (un_r - sr) - h_l * (un_l - sl);
    s_star = (rabs(val(denom)) > R(1.0e-12)) ? div(h_r * (un_r - sr) * sl - h_l * (un_l - sl) * sr, denom)
                                             : (sl + sr) * R(0.5);
  }
  const bool left = star ? val(s_star) >= R(0) : exit_l;
  const bool wK = left ? lw : rw;
  const T hK = left ? h_l : h_r;
  const T qnK = left ? qn_l : qn_r;
  const T qtK = left ? qt_l : qt_r;
  const T unK = left ? un_l : un_r;
  const T utK = left ? ut_l : ut_r;
  T fmK = Z, fnK = Z, ftK = Z;
  if (wK) {
    fmK = qnK;
    fnK = qnK * unK + hK * hK * R(0.5 * GRAVITY);
    ftK = qnK * utK;
  }
  if (!star) {
    *fm = fmK;
    *fn = fnK;
    *ft = ftK;
    return;
  }
  // HLLC star state on side K (riemann.rs:191-203)
  const T sK = left ? sl : sr;
  T h_star = div(hK * (sK - unK), sK - s_star);
  *fm = fmK + (h_star - hK) * sK;
  *fn = fnK + (h_star * s_star - qnK) * sK;
  *ft = ftK + (h_star * utK - qtK) * sK;
}

// update.rs:256-353 with z_left = z_right = z_face. Outputs the shared
// mass and tangential fluxes and the two normal-momentum fluxes
// (minus: consumed by the left/top cell, plus: by the right/bottom cell).
template <class T>
HF_HD void well_balanced_face(T lh, T lqn, T lqt, T rh, T rqn, T rqt, T zf, T* mass, T* nm_minus,
                              T* nm_plus, T* tang) {
  using R = typename ScalarOf<T>::type;
  const T Z = cst<T>(0.0);
  T z_max = rmax(zf, zf);
  T hsl = rmax(lh + zf - z_max, Z);
  T hsr = rmax(rh + zf - z_max, Z);
  T ul = Z, vl = Z, ur = Z, vr = Z;
  if (val(lh) > R(H_VEL)) {
    ul = div(lqn, lh);
    vl = div(lqt, lh);
  }
  if (val(rh) > R(H_VEL)) {
    ur = div(rqn, rh);
    vr = div(rqt, rh);
  }
  T fm, fn, ft;
  hllc_normal_flux(hsl, hsl * ul, hsl * vl, hsr, hsr * ur, hsr * vr, &fm, &fn, &ft);
  T corr_l = (lh * lh - hsl * hsl) * R(0.5 * GRAVITY);
  T corr_r = (rh * rh - hsr * hsr) * R(0.5 * GRAVITY);
  *mass = fm;
  *nm_minus = fn + corr_l;
  *nm_plus = fn + corr_r;
  *tang = ft;
}

// Kinds held to accuracy bounds rather than to the reference's bits: the
// FP32 precision form, and the fast FP64 kind (HF_FAST_F64, a translation
// unit of its own). Their interior faces run prim_face below.
#ifndef HF_PRIM_FACE
#define HF_PRIM_FACE 1
#endif
#if defined(HF_FAST_F64)
template <class T> struct PrimFace {
  static constexpr bool on = HF_PRIM_FACE && (std::is_same<T, float>::value || std::is_same<T, double>::value);
};
#else
template <class T> struct PrimFace {
  static constexpr bool on = HF_PRIM_FACE && std::is_same<T, float>::value;
};
#endif

// well_balanced_face + hllc_normal_flux (update.rs:256-353, riemann.rs:
// 114-204) in the face's primitive states (depth h, normal and tangential
// velocity un, ut) for the kinds of PrimFace. With z_left = z_right = z_face
// the reference's hydrostatic depth max(h + z_face - z_face, 0) is h, its
// velocity (h u) / h and the HLLC's un = (h un) / h are un, and the
// correction (h^2 - hs^2) g / 2 vanishes, all in exact arithmetic; so the
// reconstructed velocities are used as they are, with the reference's
// cut-offs (velocity 0 at h <= H_VEL, a side dry at h <= DRY_TOL), and the
// eight divisions and two corrections of the conserved form drop out. Both
// normal-momentum fluxes are *nm.
template <class T>
HF_HD void prim_face(T h_l, T un_l, T ut_l, T h_r, T un_r, T ut_r, T* fm, T* fn, T* ft) {
  using R = typename ScalarOf<T>::type;
  const T Z = cst<T>(0.0);
  const bool lw = val(h_l) > R(DRY_TOL);
  const bool rw = val(h_r) > R(DRY_TOL);
  if (!lw && !rw) {
    *fm = Z;
    *fn = Z;
    *ft = Z;
    return;
  }
  const bool lv = val(h_l) > R(H_VEL), rv = val(h_r) > R(H_VEL);
  un_l = lv ? un_l : Z;
  ut_l = lv ? ut_l : Z;
  un_r = rv ? un_r : Z;
  ut_r = rv ? ut_r : Z;
  T cl = lw ? rsqrt_wet(h_l * R(GRAVITY)) : Z;
  T cr = rw ? rsqrt_wet(h_r * R(GRAVITY)) : Z;
  T sl, sr;
  if (lw && rw) {
    sl = rmin(un_l - cl, un_r - cr);
    sr = rmax(un_l + cl, un_r + cr);
  } else if (rw) {
    sl = un_r - cr * R(2.0);
    sr = un_r + cr;
  } else {
    sl = un_l - cl;
    sr = un_l + cl * R(2.0);
  }
  const bool exit_l = val(sl) >= R(0);
  const bool star = !exit_l && !(val(sr) <= R(0));
  T s_star = Z;
  if (star) {
    const T wr = h_r * (un_r - sr), wl = h_l * (un_l - sl);
    const T denom = wr - wl;
    s_star = (rabs(val(denom)) > R(1.0e-12)) ? div(wr * sl - wl * sr, denom) : (sl + sr) * R(0.5);
  }
  const bool left = star ? val(s_star) >= R(0) : exit_l;
  const bool wK = left ? lw : rw;
  const T hK = left ? h_l : h_r;
  const T unK = left ? un_l : un_r;
  const T utK = left ? ut_l : ut_r;
  const T qnK = hK * unK;
  T fmK = Z, fnK = Z, ftK = Z;
  if (wK) {
    fmK = qnK;
    fnK = qnK * unK + hK * hK * R(0.5 * GRAVITY);
    ftK = qnK * utK;
  }
  if (!star) {
    *fm = fmK;
    *fn = fnK;
    *ft = ftK;
    return;
  }
  const T sK = left ? sl : sr;
  const T h_star = div(hK * (sK - unK), sK - s_star);
  *fm = fmK + (h_star - hK) * sK;
  *fn = fnK + (h_star * s_star - qnK) * sK;
  *ft = ftK + (h_star - hK) * utK * sK;
}

// ------------------------------------------------------------ tile LDS
template <class T, int TW, int TH> struct TileSmem {
  static constexpr int W3 = TW + 6, H3 = TH + 6, N3 = W3 * H3;
  static constexpr int W4 = TW + 8, H4 = TH + 8, N4 = W4 * H4;
  // x-differences (prim_R - prim_L)/dx on x-faces fj in [-2, TW+2], rows [-1, TH]
  static constexpr int WDX = TW + 5, HDX = TH + 2, NDX = WDX * HDX;
  // y-differences on y-faces fi in [-2, TH+2], columns [-1, TW]
  static constexpr int WDY = TW + 2, HDY = TH + 5, NDY = WDY * HDY;
  // near-dry flags on rows [-2, TH+1] x columns [-2, TW+1]
  static constexpr int WND = TW + 4, HND = TH + 4, NND = WND * HND;
  static constexpr int WFX = TW + 3, HFX = TH + 2, NFX = WFX * HFX;
  static constexpr int WFY = TW + 2, HFY = TH + 3, NFY = WFY * HFY;
  static constexpr int WA = TW + 2, HA = TH + 2, NA = WA * HA;

  // minmod slopes of the cells whose faces the tile needs: along x for
  // columns [-2, TW+1] in rows [-1, TH], along y for rows [-2, TH+1] in
  // columns [-1, TW]
  static constexpr int WSX = TW + 4, HSX = TH + 2, NSX = WSX * HSX;
  static constexpr int WSY = TW + 2, HSY = TH + 4, NSY = WSY * HSY;

  T h[N3], z[N3], u[N3], v[N3];
  T sxe[NSX];
  // The u-slopes along x are last read by the face phase; alpha (tile +
  // ring, written by the alpha phase after it) takes their LDS. The tile
  // then fits in half a CU's LDS at more shapes (FP64 24 x 16: 77 KB).
  union {
    T sxu[NSX];
    T alpha[NA];
  };
  T sxv[NSX];
  T sye[NSY], syu[NSY], syv[NSY];
  // The face differences are read only to form the slopes, and the face
  // fluxes are written after that: one block of LDS holds the one, then
  // the other.
  struct Diffs {
    T dxe[NDX], dxu[NDX], dxv[NDX];
    T dye[NDY], dyu[NDY], dyv[NDY];
  };
  struct Fluxes {
    T fxm[NFX], fxa[NFX], fxb[NFX], fxt[NFX];
    T fym[NFY], fya[NFY], fyb[NFY], fyt[NFY];
  };
  union {
    Diffs d;
    Fluxes f;
  } w;
  unsigned char dry[N4];
  unsigned char nd[N3];  // bit 0: a dry cell within +-2 along x; bit 1: along y
  double red[3][16];

  static_assert(NA <= NSX, "alpha lives in the u-slope array");

  static HF_HD int i3(int li, int lj) {
    HF_BOUNDS_CHECK(li >= -3 && li < TH + 3 && lj >= -3 && lj < TW + 3);
    return (li + 3) * W3 + (lj + 3);
  }
  static HF_HD int i4(int li, int lj) {
    HF_BOUNDS_CHECK(li >= -4 && li < TH + 4 && lj >= -4 && lj < TW + 4);
    return (li + 4) * W4 + (lj + 4);
  }
  static HF_HD int idx_(int li, int fj) {
    HF_BOUNDS_CHECK(li >= -1 && li < TH + 1 && fj >= -2 && fj < TW + 3);
    return (li + 1) * WDX + (fj + 2);
  }
  static HF_HD int idy_(int fi, int lj) {
    HF_BOUNDS_CHECK(fi >= -2 && fi < TH + 3 && lj >= -1 && lj < TW + 1);
    return (fi + 2) * WDY + (lj + 1);
  }
  static HF_HD int isx(int li, int lj) {
    HF_BOUNDS_CHECK(li >= -1 && li < TH + 1 && lj >= -2 && lj < TW + 2);
    return (li + 1) * WSX + (lj + 2);
  }
  static HF_HD int isy(int li, int lj) {
    HF_BOUNDS_CHECK(li >= -2 && li < TH + 2 && lj >= -1 && lj < TW + 1);
    return (li + 2) * WSY + (lj + 1);
  }
  static HF_HD int ifx(int li, int fj) {
    HF_BOUNDS_CHECK(li >= -1 && li < TH + 1 && fj >= -1 && fj < TW + 2);
    return (li + 1) * WFX + (fj + 1);
  }
  static HF_HD int ify(int fi, int lj) {
    HF_BOUNDS_CHECK(fi >= -1 && fi < TH + 2 && lj >= -1 && lj < TW + 1);
    return (fi + 1) * WFY + (lj + 1);
  }
  static HF_HD int ia(int li, int lj) {
    HF_BOUNDS_CHECK(li >= -1 && li < TH + 1 && lj >= -1 && lj < TW + 1);
    return (li + 1) * WA + (lj + 1);
  }
};

// ------------------------------------------------------------ the stage
template <class T, int TW, int TH> struct Stage {
  using S = TileSmem<T, TW, TH>;
  using R = typename ScalarOf<T>::type;

  // FP32 precision form: depths are never recovered from an absolute water
  // surface. eta = h + z with z of metres keeps only ~1e-7 m of h in FP32,
  // which is a large fraction of a rain film; so face depths are formed as
  // h + slope * d/2 + (z_cell - z_face) with the bed term from the bed
  // difference, the face flux is evaluated in the face's own frame
  // (z_face = 0), and the bed-slope source relative to the cell's bed. In
  // exact arithmetic these are the reference's expressions; the f64 and Dual
  // paths keep the reference's own operand order.
  static constexpr bool kRel = std::is_same<T, float>::value;

  // z_L - z_face and z_R - z_face of an interior face (update.rs:650-656):
  // midpoint -> -+(z_R - z_L)/2; wet/dry face (z_face = max) -> min(., 0).
  static HF_HD void interior_dz(bool l_dry, bool r_dry, T d, T* dzl, T* dzr) {
    const T Z = cst<T>(0.0);
    if (l_dry != r_dry) {
      *dzl = rmin(d, Z);
      *dzr = rmin(-d, Z);
    } else {
      *dzl = d * R(0.5);
      *dzr = -(d * R(0.5));
    }
  }
  // z_cell - z_face on a domain-edge face: (z_cell - z_ghost) / 2 with the
  // ghost bed of boundary.rs:335-367.
  static HF_HD T boundary_dz(int kind, T zc, T zn) {
    if (kind == BC_TRANSMISSIVE || kind == BC_WALL) return cst<T>(0.0);
    return (zn - zc) * R(0.5);
  }

  // Face bed of x-face fj in local row li (update.rs:676-700).
  static HF_HD T zface_x(const S& s, const StageArgs<T>& a, int gi0, int gj0, int li, int fj) {
    const int gjf = gj0 + fj;
    if (gjf == 0) {
      T zc = s.z[S::i3(li, fj)];
      T zn = (a.nc >= 2) ? s.z[S::i3(li, fj + 1)] : zc;
      return (ghost_bed(a.bc_kind[SIDE_W], zc, zn) + zc) * R(0.5);
    }
    if (gjf == a.nc) {
      T zc = s.z[S::i3(li, fj - 1)];
      T zn = (a.nc >= 2) ? s.z[S::i3(li, fj - 2)] : zc;
      return (zc + ghost_bed(a.bc_kind[SIDE_E], zc, zn)) * R(0.5);
    }
    return interior_z_face(s.dry[S::i4(li, fj - 1)] != 0, s.z[S::i3(li, fj - 1)],
                           s.dry[S::i4(li, fj)] != 0, s.z[S::i3(li, fj)]);
  }
  // Face bed of y-face fi in local column lj (update.rs:704-728).
  static HF_HD T zface_y(const S& s, const StageArgs<T>& a, int gi0, int gj0, int fi, int lj) {
    const int gif = gi0 + fi;
    if (gif == 0) {
      T zc = s.z[S::i3(fi, lj)];
      T zn = (a.nr >= 2) ? s.z[S::i3(fi + 1, lj)] : zc;
      return (ghost_bed(a.bc_kind[SIDE_N], zc, zn) + zc) * R(0.5);
    }
    if (gif == a.nr) {
      T zc = s.z[S::i3(fi - 1, lj)];
      T zn = (a.nr >= 2) ? s.z[S::i3(fi - 2, lj)] : zc;
      return (zc + ghost_bed(a.bc_kind[SIDE_S], zc, zn)) * R(0.5);
    }
    return interior_z_face(s.dry[S::i4(fi - 1, lj)] != 0, s.z[S::i3(fi - 1, lj)],
                           s.dry[S::i4(fi, lj)] != 0, s.z[S::i3(fi, lj)]);
  }

  // Phase 1: dry flags on the 4-halo; h, z and primitives u, v on the
  // 3-halo (update.rs:405-426, 853-855). Cells outside the domain are
  // "not dry", which clips the +-2 dry windows to the domain as the
  // reference does. All global loads of a thread's NT-strided items are
  // issued before any is used (momentum too, used only where h > H_VEL),
  // so the phase waits for memory once rather than once per item and
  // again for the momentum.
  // The global values a thread loads for the load phase (NT threads: up to
  // RN items of the 4-halo each), held in registers between load_issue and
  // load_store so that the loads of one tile can be in flight while the
  // previous tile is computed.
  template <int NT> struct LoadRegs {
    static constexpr int RN = (S::N4 + NT - 1) / NT;
    T h[RN], z[RN], hu[RN], hv[RN];
  };
  template <int NT>
  static HF_HD void load_issue(const StageArgs<T>& a, int gi0, int gj0, int tid, LoadRegs<NT>* L) {
    constexpr int RN = LoadRegs<NT>::RN;
    const int ci = gi0 + TH / 2 < a.nr ? gi0 + TH / 2 : a.nr - 1;
    const int cj = gj0 + TW / 2 < a.nc ? gj0 + TW / 2 : a.nc - 1;
    const double tdat = tile_datum(a, ci, cj);
#pragma unroll
    for (int r = 0; r < RN; ++r) {
      const int k = tid + r * NT;
      const int li = k / S::W4 - 4, lj = k % S::W4 - 4;
      const int gi = gi0 + li, gj = gj0 + lj;
      if (k < S::N4 && gi >= 0 && gi < a.nr && gj >= 0 && gj < a.nc) {
        const int64_t idx = (int64_t)gi * a.nc + gj;
        L->h[r] = a.ih.ld(idx);
        if (li >= -3 && li < TH + 3 && lj >= -3 && lj < TW + 3) {
          L->z[r] = load_bed(a, gi, gj, idx, tdat);
          L->hu[r] = a.ihu.ld(idx);
          L->hv[r] = a.ihv.ld(idx);
        }
      }
    }
  }
  template <int NT>
  static HF_HD void load_store(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid, const LoadRegs<NT>& L) {
    constexpr int RN = LoadRegs<NT>::RN;
    const T Z = cst<T>(0.0);
#pragma unroll
    for (int r = 0; r < RN; ++r) {
      const int k = tid + r * NT;
      if (k >= S::N4) break;
      const int li = k / S::W4 - 4, lj = k % S::W4 - 4;
      const int gi = gi0 + li, gj = gj0 + lj;
      const bool inside = gi >= 0 && gi < a.nr && gj >= 0 && gj < a.nc;
      const bool in3 = li >= -3 && li < TH + 3 && lj >= -3 && lj < TW + 3;
      if (inside) {
        const T hh = L.h[r];
        s.dry[k] = val(hh) <= R(H_DRY);
        if (in3) {
          const int k3 = S::i3(li, lj);
          s.h[k3] = hh;
          s.z[k3] = L.z[r];
          if (val(hh) > R(H_VEL)) {
            s.u[k3] = div(L.hu[r], hh);
            s.v[k3] = div(L.hv[r], hh);
          } else {
            s.u[k3] = Z;
            s.v[k3] = Z;
          }
        }
      } else {
        s.dry[k] = 0;
        if (in3) {
          const int k3 = S::i3(li, lj);
          s.h[k3] = Z;
          s.z[k3] = Z;
          s.u[k3] = Z;
          s.v[k3] = Z;
        }
      }
    }
  }
  template <int NT>
  static HF_HD void phase_load(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid) {
    LoadRegs<NT> L;
    load_issue<NT>(a, gi0, gj0, tid, &L);
    load_store<NT>(s, a, gi0, gj0, tid, L);
  }

  // Phase 2: per-face differences of the primitives, (R - L)/dx and
  // (B - T)/dy, shared by the two cells whose minmod slopes use them
  // (update.rs:459-540 evaluates each twice; the values are identical), and
  // the near-dry flags of update.rs:433-444 (window clipped to the domain).
  static HF_HD void phase_diffs(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid, int nt) {
    const T Z = cst<T>(0.0);
    const R dx = R(a.dx), dy = R(a.dy);
    const int n_total = S::NDX + S::NDY + S::NND;
    for (int k = tid; k < n_total; k += nt) {
      if (k < S::NDX) {
        const int li = k / S::WDX - 1, fj = k % S::WDX - 2;
        const int gi = gi0 + li, gj = gj0 + fj;
        T de = Z, du = Z, dv = Z;
        if (gi >= 0 && gi < a.nr && gj >= 1 && gj < a.nc) {
          const int l = S::i3(li, fj - 1), r = S::i3(li, fj);
          if constexpr (kRel)
            de = div((s.h[r] - s.h[l]) + (s.z[r] - s.z[l]), dx);
          else
            de = div(s.h[r] + s.z[r] - (s.h[l] + s.z[l]), dx);
          du = div(s.u[r] - s.u[l], dx);
          dv = div(s.v[r] - s.v[l], dx);
        }
        s.w.d.dxe[k] = de;
        s.w.d.dxu[k] = du;
        s.w.d.dxv[k] = dv;
      } else if (k < S::NDX + S::NDY) {
        const int kk = k - S::NDX;
        const int fi = kk / S::WDY - 2, lj = kk % S::WDY - 1;
        const int gi = gi0 + fi, gj = gj0 + lj;
        T de = Z, du = Z, dv = Z;
        if (gj >= 0 && gj < a.nc && gi >= 1 && gi < a.nr) {
          const int t = S::i3(fi - 1, lj), b = S::i3(fi, lj);
          if constexpr (kRel)
            de = div((s.h[b] - s.h[t]) + (s.z[b] - s.z[t]), dy);
          else
            de = div(s.h[b] + s.z[b] - (s.h[t] + s.z[t]), dy);
          du = div(s.u[b] - s.u[t], dy);
          dv = div(s.v[b] - s.v[t], dy);
        }
        s.w.d.dye[kk] = de;
        s.w.d.dyu[kk] = du;
        s.w.d.dyv[kk] = dv;
      } else {
        const int kk = k - S::NDX - S::NDY;
        const int li = kk / S::WND - 2, lj = kk % S::WND - 2;
        // bitwise or, not ||: the ten flag reads issue together instead of
        // one branch (and LDS round trip) per read
        unsigned nx = 0, ny = 0;
        for (int d = -2; d <= 2; ++d) {
          nx |= s.dry[S::i4(li, lj + d)];
          ny |= s.dry[S::i4(li + d, lj)];
        }
        s.nd[S::i3(li, lj)] = (unsigned char)((nx != 0 ? 1 : 0) | (ny != 0 ? 2 : 0));
      }
    }
  }

  // The slope of a cell from the two face differences l, r that bracket it:
  // zero, one-sided on the first / last cell, else minmod(l, r).
  //
  // The select form reads every candidate before the choice, so the
  // difference loads do not wait behind a branch on the near-dry flag.
  // Measured on MI355X: FP32 takes it (+3.4 %); FP64 measured -1 % with it
  // and Dual's final stage spilled with it (+65 % time), so the other kinds
  // keep the branches.
  static constexpr bool kSelectSlopes = std::is_same<T, float>::value;
  static HF_HD T pick_slope(bool zero, bool first, bool last, T l, T r) {
    const T mm = minmod(l, r);
    return zero ? cst<T>(0.0) : first ? r : last ? l : mm;
  }

  // Minmod slope of cell (li, lj) along x from the face differences
  // (update.rs:459-499): zero near a dry cell or on a 1-column mesh,
  // one-sided on the first/last column, else minmod(left, right).
  static HF_HD void slope_x(const S& s, const StageArgs<T>& a, int li, int lj, int gj, T* se, T* su,
                            T* sv) {
    const T Z = cst<T>(0.0);
    if constexpr (kSelectSlopes) {
      const int l = S::idx_(li, lj), r = S::idx_(li, lj + 1);
      const bool zero = (a.nc < 2) | ((s.nd[S::i3(li, lj)] & 1) != 0);
      const bool first = gj == 0, last = gj + 1 == a.nc;
      *se = pick_slope(zero, first, last, s.w.d.dxe[l], s.w.d.dxe[r]);
      *su = pick_slope(zero, first, last, s.w.d.dxu[l], s.w.d.dxu[r]);
      *sv = pick_slope(zero, first, last, s.w.d.dxv[l], s.w.d.dxv[r]);
    } else if (a.nc < 2 || (s.nd[S::i3(li, lj)] & 1)) {
      *se = Z;
      *su = Z;
      *sv = Z;
    } else if (gj == 0) {
      const int r = S::idx_(li, lj + 1);
      *se = s.w.d.dxe[r];
      *su = s.w.d.dxu[r];
      *sv = s.w.d.dxv[r];
    } else if (gj + 1 == a.nc) {
      const int l = S::idx_(li, lj);
      *se = s.w.d.dxe[l];
      *su = s.w.d.dxu[l];
      *sv = s.w.d.dxv[l];
    } else {
      const int l = S::idx_(li, lj), r = S::idx_(li, lj + 1);
      *se = minmod(s.w.d.dxe[l], s.w.d.dxe[r]);
      *su = minmod(s.w.d.dxu[l], s.w.d.dxu[r]);
      *sv = minmod(s.w.d.dxv[l], s.w.d.dxv[r]);
    }
  }
  // update.rs:503-540
  static HF_HD void slope_y(const S& s, const StageArgs<T>& a, int li, int lj, int gi, T* se, T* su,
                            T* sv) {
    const T Z = cst<T>(0.0);
    if constexpr (kSelectSlopes) {
      const int t = S::idy_(li, lj), b = S::idy_(li + 1, lj);
      const bool zero = (a.nr < 2) | ((s.nd[S::i3(li, lj)] & 2) != 0);
      const bool first = gi == 0, last = gi + 1 == a.nr;
      *se = pick_slope(zero, first, last, s.w.d.dye[t], s.w.d.dye[b]);
      *su = pick_slope(zero, first, last, s.w.d.dyu[t], s.w.d.dyu[b]);
      *sv = pick_slope(zero, first, last, s.w.d.dyv[t], s.w.d.dyv[b]);
    } else if (a.nr < 2 || (s.nd[S::i3(li, lj)] & 2)) {
      *se = Z;
      *su = Z;
      *sv = Z;
    } else if (gi == 0) {
      const int b = S::idy_(li + 1, lj);
      *se = s.w.d.dye[b];
      *su = s.w.d.dyu[b];
      *sv = s.w.d.dyv[b];
    } else if (gi + 1 == a.nr) {
      const int t = S::idy_(li, lj);
      *se = s.w.d.dye[t];
      *su = s.w.d.dyu[t];
      *sv = s.w.d.dyv[t];
    } else {
      const int t = S::idy_(li, lj), b = S::idy_(li + 1, lj);
      *se = minmod(s.w.d.dye[t], s.w.d.dye[b]);
      *su = minmod(s.w.d.dyu[t], s.w.d.dyu[b]);
      *sv = minmod(s.w.d.dyv[t], s.w.d.dyv[b]);
    }
  }
  // Phase 3: the minmod slopes of every cell whose faces the tile needs,
  // each formed once from the two face differences that bracket it (a face
  // reads the slopes of its two cells, and each cell has two faces per
  // axis), into the slope arrays; the update phase reads the eta slopes of
  // the tile's cells from them too.
  static HF_HD void phase_slopes(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid, int nt) {
    for (int k = tid; k < S::NSX + S::NSY; k += nt) {
      T e, u, v;
      if (k < S::NSX) {
        const int li = k / S::WSX - 1, lj = k % S::WSX - 2;
        slope_x(s, a, li, lj, gj0 + lj, &e, &u, &v);
        s.sxe[k] = e;
        s.sxu[k] = u;
        s.sxv[k] = v;
      } else {
        const int kk = k - S::NSX;
        const int li = kk / S::WSY - 2, lj = kk % S::WSY - 1;
        slope_y(s, a, li, lj, gi0 + li, &e, &u, &v);
        s.sye[kk] = e;
        s.syu[kk] = u;
        s.syv[kk] = v;
      }
    }
  }

  // Raw state of a domain cell (boundary faces use it unreconstructed,
  // update.rs:892-897).
  static HF_HD void raw(const S& s, const StageArgs<T>& a, int li, int lj, int gi, int gj, T* h,
                        T* hu, T* hv) {
    const int64_t idx = (int64_t)gi * a.nc + gj;
    *h = s.h[S::i3(li, lj)];
    *hu = a.ihu.ld(idx);
    *hv = a.ihv.ld(idx);
  }

  // One x-face (local row li, face column fj) into the face arrays
  // (update.rs:560-588, 890-905).
  static HF_HD void face_x(S& s, const StageArgs<T>& a, int gi0, int gj0, int k) {
    const T Z = cst<T>(0.0);
    const int li = k / S::WFX - 1, fj = k % S::WFX - 1;
    const int gi = gi0 + li, gjf = gj0 + fj;
    T m = Z, na = Z, nb = Z, tg = Z;
    if (gi >= 0 && gi < a.nr && gjf >= 0 && gjf <= a.nc) {
      T zf = zface_x(s, a, gi0, gj0, li, fj);
      if (gjf == 0 || gjf == a.nc) {
        const bool west = (gjf == 0);
        const int lc = west ? fj : fj - 1;
        const int ln = west ? fj + 1 : fj - 2;
        const int gc = west ? 0 : a.nc - 1;
        T h, hu, hv, gh, ghu, ghv;
        raw(s, a, li, lc, gi, gc, &h, &hu, &hv);
        T zc = s.z[S::i3(li, lc)];
        T zn = (a.nc >= 2) ? s.z[S::i3(li, ln)] : zc;
        const int side = west ? SIDE_W : SIDE_E;
        ghost_state(a.bc_kind[side], a.bc_param[side], side, h, hu, hv, zc, zn, a.nc >= 2, a.dx,
                    a.man.ld((int64_t)gi * a.nc + gc), &gh, &ghu, &ghv);
        if (kRel) zf = Z;
        if (west)
          well_balanced_face(gh, ghu, ghv, h, hu, hv, zf, &m, &na, &nb, &tg);
        else
          well_balanced_face(h, hu, hv, gh, ghu, ghv, zf, &m, &na, &nb, &tg);
      } else if ((s.dry[S::i4(li, fj - 1)] & s.dry[S::i4(li, fj)]) == 0) {  // not both dry
        const R half_dx = R(0.5 * a.dx);
        const int l = S::i3(li, fj - 1), r = S::i3(li, fj);
        const int cl = S::isx(li, fj - 1), cr = S::isx(li, fj);
        const T sle = s.sxe[cl], slu = s.sxu[cl], slv = s.sxv[cl];
        const T sre = s.sxe[cr], sru = s.sxu[cr], srv = s.sxv[cr];
        T u_m = s.u[l] + slu * half_dx;
        T v_m = s.v[l] + slv * half_dx;
        T u_p = s.u[r] - sru * half_dx;
        T v_p = s.v[r] - srv * half_dx;
        T h_m, h_p;
        if constexpr (kRel) {
          T dzl, dzr;
          interior_dz(s.dry[S::i4(li, fj - 1)] != 0, s.dry[S::i4(li, fj)] != 0, s.z[l] - s.z[r], &dzl, &dzr);
          h_m = rmax(s.h[l] + sle * half_dx + dzl, Z);
          h_p = rmax(s.h[r] - sre * half_dx + dzr, Z);
          zf = Z;
        } else {
          T eta_m = s.h[l] + s.z[l] + sle * half_dx;
          T eta_p = s.h[r] + s.z[r] - sre * half_dx;
          h_m = rmax(eta_m - zf, Z);
          h_p = rmax(eta_p - zf, Z);
        }
        well_balanced_face(h_m, h_m * u_m, h_m * v_m, h_p, h_p * u_p, h_p * v_p, zf, &m, &na, &nb, &tg);
      }
    }
    s.w.f.fxm[k] = m;
    s.w.f.fxa[k] = na;
    s.w.f.fxb[k] = nb;
    s.w.f.fxt[k] = tg;
  }

  // One y-face (face row fi, local column lj) (update.rs:595-623, 907-922).
  static HF_HD void face_y(S& s, const StageArgs<T>& a, int gi0, int gj0, int k) {
    const T Z = cst<T>(0.0);
    const int fi = k / S::WFY - 1, lj = k % S::WFY - 1;
    const int gif = gi0 + fi, gj = gj0 + lj;
    T m = Z, na = Z, nb = Z, tg = Z;
    if (gj >= 0 && gj < a.nc && gif >= 0 && gif <= a.nr) {
      T zf = zface_y(s, a, gi0, gj0, fi, lj);
      if (gif == 0 || gif == a.nr) {
        const bool north = (gif == 0);
        const int lc = north ? fi : fi - 1;
        const int ln = north ? fi + 1 : fi - 2;
        const int gc = north ? 0 : a.nr - 1;
        T h, hu, hv, gh, ghu, ghv;
        raw(s, a, lc, lj, gc, gj, &h, &hu, &hv);
        T zc = s.z[S::i3(lc, lj)];
        T zn = (a.nr >= 2) ? s.z[S::i3(ln, lj)] : zc;
        const int side = north ? SIDE_N : SIDE_S;
        ghost_state(a.bc_kind[side], a.bc_param[side], side, h, hu, hv, zc, zn, a.nr >= 2, a.dy,
                    a.man.ld((int64_t)gc * a.nc + gj), &gh, &ghu, &ghv);
        // y-face: normal momentum is hv, tangential hu (riemann.rs:97-107)
        if (kRel) zf = Z;
        if (north)
          well_balanced_face(gh, ghv, ghu, h, hv, hu, zf, &m, &na, &nb, &tg);
        else
          well_balanced_face(h, hv, hu, gh, ghv, ghu, zf, &m, &na, &nb, &tg);
      } else if ((s.dry[S::i4(fi - 1, lj)] & s.dry[S::i4(fi, lj)]) == 0) {  // not both dry
        const R half_dy = R(0.5 * a.dy);
        const int t = S::i3(fi - 1, lj), b = S::i3(fi, lj);
        const int ct = S::isy(fi - 1, lj), cb = S::isy(fi, lj);
        const T ste = s.sye[ct], stu = s.syu[ct], stv = s.syv[ct];
        const T sbe = s.sye[cb], sbu = s.syu[cb], sbv = s.syv[cb];
        T u_m = s.u[t] + stu * half_dy;
        T v_m = s.v[t] + stv * half_dy;
        T u_p = s.u[b] - sbu * half_dy;
        T v_p = s.v[b] - sbv * half_dy;
        T h_m, h_p;
        if constexpr (kRel) {
          T dzt, dzb;
          interior_dz(s.dry[S::i4(fi - 1, lj)] != 0, s.dry[S::i4(fi, lj)] != 0, s.z[t] - s.z[b], &dzt, &dzb);
          h_m = rmax(s.h[t] + ste * half_dy + dzt, Z);
          h_p = rmax(s.h[b] - sbe * half_dy + dzb, Z);
          zf = Z;
        } else {
          T eta_m = s.h[t] + s.z[t] + ste * half_dy;
          T eta_p = s.h[b] + s.z[b] - sbe * half_dy;
          h_m = rmax(eta_m - zf, Z);
          h_p = rmax(eta_p - zf, Z);
        }
        well_balanced_face(h_m, h_m * v_m, h_m * u_m, h_p, h_p * v_p, h_p * u_p, zf, &m, &na, &nb, &tg);
      }
    }
    s.w.f.fym[k] = m;
    s.w.f.fya[k] = na;
    s.w.f.fyb[k] = nb;
    s.w.f.fyt[k] = tg;
  }

  // One interior face (both cells in the domain) of either axis, into the
  // face arrays (update.rs:560-623, 890-922). Faces are numbered k in
  // [0, NFX) for x and [NFX, NFX + NFY) for y; both axes run this one code
  // path, the axis entering only through which LDS arrays it reads (normal
  // and tangential velocity and slopes) and the stride between its two
  // cells, so no wavefront runs two face kinds. Faces outside the domain
  // get zero fluxes; domain-edge faces are left to boundary_faces.
  static HF_HD void face_interior(S& s, const StageArgs<T>& a, int gi0, int gj0, int k) {
    const T Z = cst<T>(0.0);
    const bool ya = k >= S::NFX;
    const int kx = k, ky = ya ? k - S::NFX : 0;
    const int rx = kx / S::WFX, ry = ky / S::WFY;
    // local cell right of an x-face / below a y-face
    const int li = ya ? ry - 1 : rx - 1;
    const int lj = ya ? ky - ry * S::WFY - 1 : kx - rx * S::WFX - 1;
    const int gi = gi0 + li, gj = gj0 + lj;
    const int g_norm = ya ? gi : gj, n_norm = ya ? a.nr : a.nc;  // along the face normal
    const int g_tan = ya ? gj : gi, n_tan = ya ? a.nc : a.nr;
    const bool inside = g_tan >= 0 && g_tan < n_tan && g_norm >= 0 && g_norm <= n_norm;
    const bool edge = g_norm == 0 || g_norm == n_norm;
    if (inside && edge) return;  // boundary_faces
    T m = Z, na = Z, nb = Z, tg = Z;
    if (inside) {
      const int r3 = S::i3(li, lj), l3 = r3 - (ya ? S::W3 : 1);
      const int r4 = S::i4(li, lj), l4 = r4 - (ya ? S::W4 : 1);
      const unsigned char dl = s.dry[l4], dr = s.dry[r4];
      T zf = interior_z_face(dl != 0, s.z[l3], dr != 0, s.z[r3]);
      if ((dl & dr) == 0) {  // not both dry
        const int cr = ya ? S::isy(li, lj) : S::isx(li, lj), cl = cr - (ya ? S::WSY : 1);
        const T* vn = ya ? s.v : s.u;  // velocity normal / tangential to the face
        const T* vt = ya ? s.u : s.v;
        const T* se = ya ? s.sye : s.sxe;
        const T* sn = ya ? s.syv : s.sxu;
        const T* st = ya ? s.syu : s.sxv;
        const R half = ya ? R(0.5 * a.dy) : R(0.5 * a.dx);
        T n_m = vn[l3] + sn[cl] * half;
        T t_m = vt[l3] + st[cl] * half;
        T n_p = vn[r3] - sn[cr] * half;
        T t_p = vt[r3] - st[cr] * half;
        T h_m, h_p;
        if constexpr (kRel) {
          T dzl, dzr;
          interior_dz(dl != 0, dr != 0, s.z[l3] - s.z[r3], &dzl, &dzr);
          h_m = rmax(s.h[l3] + se[cl] * half + dzl, Z);
          h_p = rmax(s.h[r3] - se[cr] * half + dzr, Z);
          zf = Z;
        } else {
          T eta_m = s.h[l3] + s.z[l3] + se[cl] * half;
          T eta_p = s.h[r3] + s.z[r3] - se[cr] * half;
          h_m = rmax(eta_m - zf, Z);
          h_p = rmax(eta_p - zf, Z);
        }
        if constexpr (PrimFace<T>::on) {
          prim_face(h_m, n_m, t_m, h_p, n_p, t_p, &m, &na, &tg);
          nb = na;
        } else {
          well_balanced_face(h_m, h_m * n_m, h_m * t_m, h_p, h_p * n_p, h_p * t_p, zf, &m, &na, &nb, &tg);
        }
      }
    }
    const int kk = ya ? ky : kx;
    T* om = ya ? s.w.f.fym : s.w.f.fxm;
    T* oa = ya ? s.w.f.fya : s.w.f.fxa;
    T* ob = ya ? s.w.f.fyb : s.w.f.fxb;
    T* ot = ya ? s.w.f.fyt : s.w.f.fxt;
    om[kk] = m;
    oa[kk] = na;
    ob[kk] = nb;
    ot[kk] = tg;
  }

  // The domain-edge faces in the tile's face region (ghost cells,
  // boundary.rs:193-367): x-faces in the west (column 0) and east (column
  // nc) edge columns, y-faces in the north and south edge rows.
  static HF_HD void boundary_faces(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid, int nt) {
    constexpr int NXB = 2 * (TH + 2), NYB = 2 * (TW + 2);
    for (int t = tid; t < NXB + NYB; t += nt) {
      if (t < NXB) {
        const int li = t % (TH + 2) - 1;
        const int fj = (t < TH + 2) ? -gj0 : a.nc - gj0;
        const int gi = gi0 + li;
        if (fj < -1 || fj > TW + 1 || gi < 0 || gi >= a.nr) continue;
        face_x(s, a, gi0, gj0, (li + 1) * S::WFX + (int)fj + 1);
      } else {
        const int u = t - NXB;
        const int lj = u % (TW + 2) - 1;
        const int fi = (u < TW + 2) ? -gi0 : a.nr - gi0;
        const int gj = gj0 + lj;
        if (fi < -1 || fi > TH + 1 || gj < 0 || gj >= a.nc) continue;
        face_y(s, a, gi0, gj0, ((int)fi + 1) * S::WFY + lj + 1);
      }
    }
  }

  // Phase 4: every x- and y-face of the tile and its 1-cell ring: interior
  // faces in one loop over both axes, then, in tiles that reach the domain
  // edge, the edge faces.
  static HF_HD void phase_faces(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid, int nt) {
    for (int k = tid; k < S::NFX + S::NFY; k += nt) face_interior(s, a, gi0, gj0, k);
    const bool near_edge = gi0 <= 1 || gj0 <= 1 || gi0 + TH + 1 >= a.nr || gj0 + TW + 1 >= a.nc;
    if (near_edge) boundary_faces(s, a, gi0, gj0, tid, nt);
  }

  // Phase 5: Liang-Marche outflow limiter alpha on the tile and its ring
  // (update.rs:953-982).
  static HF_HD void phase_alpha(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid, int nt,
                                double dt) {
    const R dt_dx = R(dt / a.dx), dt_dy = R(dt / a.dy);
    const T Z = cst<T>(0.0);
    for (int k = tid; k < S::NA; k += nt) {
      const int li = k / S::WA - 1, lj = k % S::WA - 1;
      const int gi = gi0 + li, gj = gj0 + lj;
      T al = cst<T>(1.0);
      if (gi >= 0 && gi < a.nr && gj >= 0 && gj < a.nc) {
        T fx_right = s.w.f.fxm[S::ifx(li, lj + 1)];
        T fx_left = s.w.f.fxm[S::ifx(li, lj)];
        T fy_bottom = s.w.f.fym[S::ify(li + 1, lj)];
        T fy_top = s.w.f.fym[S::ify(li, lj)];
        T out = Z;
        if (val(fx_right) > R(0)) out = out + fx_right * dt_dx;
        if (val(fx_left) < R(0)) out = out + (-fx_left) * dt_dx;
        if (val(fy_bottom) > R(0)) out = out + fy_bottom * dt_dy;
        if (val(fy_top) < R(0)) out = out + (-fy_top) * dt_dy;
        T avail = rmax(s.h[S::i3(li, lj)] - R(H_DRY), Z);
        if (val(out) > val(avail) && val(out) > R(0)) al = div(avail, out);
      }
      s.alpha[k] = al;
    }
  }

  // Upstream-alpha scaling of one x-face as seen from local row li
  // (update.rs:992-1025). Dry-dry interior faces are left as computed.
  static HF_HD void scaled_fx(const S& s, const StageArgs<T>& a, int gj0, int li, int fj, T* m, T* na,
                              T* nb, T* tg) {
    const int k = S::ifx(li, fj);
    *m = s.w.f.fxm[k];
    *na = s.w.f.fxa[k];
    *nb = s.w.f.fxb[k];
    *tg = s.w.f.fxt[k];
    const int gjf = gj0 + fj;
    // both upstream candidates read with the face, not after its sign is known
    const T al_l = s.alpha[S::ia(li, fj - 1)], al_r = s.alpha[S::ia(li, fj)];
    if ((s.dry[S::i4(li, fj - 1)] & s.dry[S::i4(li, fj)]) != 0 && gjf > 0 && gjf < a.nc) return;
    T al;
    if (gjf == 0)
      al = val(*m) < R(0) ? al_r : cst<T>(1.0);
    else if (gjf == a.nc)
      al = val(*m) > R(0) ? al_l : cst<T>(1.0);
    else
      al = val(*m) > R(0) ? al_l : al_r;
    *m = *m * al;
    *na = *na * al;
    *nb = *nb * al;
    *tg = *tg * al;
  }
  // update.rs:1026-1053
  static HF_HD void scaled_fy(const S& s, const StageArgs<T>& a, int gi0, int fi, int lj, T* m, T* na,
                              T* nb, T* tg) {
    const int k = S::ify(fi, lj);
    *m = s.w.f.fym[k];
    *na = s.w.f.fya[k];
    *nb = s.w.f.fyb[k];
    *tg = s.w.f.fyt[k];
    const int gif = gi0 + fi;
    const T al_t = s.alpha[S::ia(fi - 1, lj)], al_b = s.alpha[S::ia(fi, lj)];
    if ((s.dry[S::i4(fi - 1, lj)] & s.dry[S::i4(fi, lj)]) != 0 && gif > 0 && gif < a.nr) return;
    T al;
    if (gif == 0)
      al = val(*m) < R(0) ? al_b : cst<T>(1.0);
    else if (gif == a.nr)
      al = val(*m) > R(0) ? al_t : cst<T>(1.0);
    else
      al = val(*m) > R(0) ? al_t : al_b;
    *m = *m * al;
    *na = *na * al;
    *nb = *nb * al;
    *tg = *tg * al;
  }

  // Wave speeds of one cell for the CFL bound (update.rs:113-136) and, on
  // domain edges, of its ghosts (update.rs:179-201, no H_VEL cut there).
  static HF_HD void cell_wave_speed(const StageArgs<T>& a, int gi, int gj, T h, T hu, T hv, T zc,
                                    const T* zn4, const T* man, double* sx, double* sy) {
    const R hv_ = val(h);
    if (hv_ > R(H_DRY)) {
      R c = rsqrt_wet(R(GRAVITY) * hv_);
      R u = R(0), v = R(0);
      if (hv_ > R(H_VEL)) {
        u = div(val(hu), hv_);
        v = div(val(hv), hv_);
      }
      *sx = rmax(*sx, (double)(rabs(u) + c));
      *sy = rmax(*sy, (double)(rabs(v) + c));
    }
    if (!a.cfl_with_bcs) return;
    const bool edge[4] = {gi == 0, gi == a.nr - 1, gj == 0, gj == a.nc - 1};
    for (int side = 0; side < 4; ++side) {
      if (!edge[side]) continue;
      const bool xf = (side == SIDE_W || side == SIDE_E);
      const bool have_next = xf ? (a.nc >= 2) : (a.nr >= 2);
      T gh, ghu, ghv;
      ghost_state(a.bc_kind[side], a.bc_param[side], side, h, hu, hv, zc, zn4[side], have_next,
                  xf ? a.dx : a.dy, *man, &gh, &ghu, &ghv);
      const R g = val(gh);
      if (g <= R(H_DRY)) continue;
      R c = rsqrt_wet(R(GRAVITY) * g);
      R u = div(val(ghu), g);
      R v = div(val(ghv), g);
      *sx = rmax(*sx, (double)(rabs(u) + c));
      *sy = rmax(*sy, (double)(rabs(v) + c));
    }
  }

  // The step's rain rate and point-source rate (the driver series'
  // values, or the configured ones), and whether a point source lies in
  // rows [gi0, gi0 + th) x columns [gj0, gj0 + tw). Read once per thread,
  // before the launch stores anything, so the per-cell tail reloads none
  // of it.
  // The FP32 form's rounding sequence number of a launch: the device
  // clock's step count, or the host-clocked launch's own number.
  static HF_HD uint32_t rounding_seq(const StageArgs<T>& a) {
    if (a.dt_mode == 0) return a.host_seq;
    return a.ctrl ? (uint32_t)a.ctrl->steps : 0u;
  }

  // rain_key: the FP32 form's rounding key of this step.
  struct ForcingNow {
    double rain;
    double q;
    int q_series;
    int src_here;
    uint32_t rain_key;
  };
  static HF_HD ForcingNow forcing_now(const StageArgs<T>& a, int64_t gi0, int64_t gj0, int64_t th, int64_t tw) {
    ForcingNow f{a.rain_rate, 0.0, 0, 0, 0u};
    if (a.use_series) {
      const Ctrl* c = a.ctrl;
      if (c->series_on[0]) f.rain = c->series_now[0];
      if (c->series_on[1]) {
        f.q = c->series_now[1];
        f.q_series = 1;
      }
    }
    if constexpr (std::is_same<T, float>::value) {
      if (f.rain != 0.0) f.rain_key = rounding_seq(a) * 0x9E3779B9u;
    }
    for (int k = 0; k < a.n_src; ++k) {
      const int64_t r = a.srcs[k].row, cc = a.srcs[k].col;
      if (r >= gi0 && r < gi0 + th && cc >= gj0 && cc < gj0 + tw) f.src_here = 1;
    }
    return f;
  }

  // Cell-local operator-split tail in the reference callers' order
  // (sim.rs:224; uk_ea_test8a_official.rs:206-217): Manning friction
  // (source.rs:130-165), rain (source.rs:88-101), point sources
  // (source.rs:52-69).
  static HF_HD void forcing(const StageArgs<T>& a, const ForcingNow& fn, int64_t gi, int64_t gj, double dt, T* h,
                            T* hu, T* hv, const T* man_pre = nullptr) {
    const T Z = cst<T>(0.0);
    if (a.friction && val(*h) > R(a.friction_dry_tol)) {
      T n = man_pre ? *man_pre : a.man.ld(gi * a.nc + gj);
      if (val(n) != R(0)) {
        const R dt_g = R(dt * GRAVITY);
        T speed = div(rsqrt_(*hu * *hu + *hv * *hv), *h);
        T factor = div(n * n * speed, pow43(*h)) * dt_g + R(1.0);
        *hu = div(*hu, factor);
        *hv = div(*hv, factor);
      }
    }
    const double rain = fn.rain;
    if (rain != 0.0) {  // source.rs:88-101
      const R dh = R(rain * dt);
      T nh = *h + dh;
      if (val(nh) <= R(0)) {
        *h = Z;
        *hu = Z;
        *hv = Z;
      } else {
        if constexpr (std::is_same<T, float>::value)
          nh = add_rounded_stochastically(*h, dh, fn.rain_key ^ ((uint32_t)(gi * a.nc + gj) * 0x85EBCA6Bu));
        *h = nh;
      }
    }
    if (fn.src_here) {
      const double cell_area = a.dx * a.dy;
      for (int k = 0; k < a.n_src; ++k) {
        if (a.srcs[k].row != gi || a.srcs[k].col != gj) continue;
        const double q = fn.q_series ? fn.q : a.srcs[k].q;
        const R dh = R(q * dt / cell_area);
        *h = rmax(*h + dh, Z);
        if (val(*h) == R(0)) {
          *hu = Z;
          *hv = Z;
        }
      }
    }
  }

  // The global values phase_update reads for a thread's first cell
  // (k = tid), loaded before the first barrier so that their latency
  // overlaps the LDS phases instead of following them. Scalar kinds only:
  // a Dual or DualN copy would hold too many registers across the phases.
  static constexpr bool kPrefetch =
      HF_PREFETCH && (std::is_same<T, double>::value || std::is_same<T, float>::value);
  struct Pre {
    T hu, hv;        // stencil-input momentum (used where h_old > H_VEL)
    T ph, phu, phv;  // U^n (final SSP-RK2 stage)
    T man;           // Manning n (final stage: friction, Discharge ghosts)
  };
  template <int MODE>
  static HF_HD void prefetch(const StageArgs<T>& a, int gi0, int gj0, int tid, Pre* p) {
    const int gi = gi0 + tid / TW, gj = gj0 + tid % TW;
    if (tid >= TW * TH || gi >= a.nr || gj >= a.nc) return;
    const int64_t idx = (int64_t)gi * a.nc + gj;
    p->hu = a.ihu.ld(idx);
    p->hv = a.ihv.ld(idx);
    if (MODE == MODE_FINAL_RK2) {
      p->ph = a.nh.ld(idx);
      p->phu = a.nhu.ld(idx);
      p->phv = a.nhv.ld(idx);
      if (a.friction || a.cfl_with_bcs) p->man = a.man.ld(idx);
    }
  }

  // Phase 6: FV update with the explicit bed-slope source
  // (update.rs:1071-1210), then the stage epilogue: write U^(1), or the
  // SSP-RK2 combine (update.rs:1272-1286) + forcing + CFL wave speeds;
  // *wet is set when an output cell is wet (not h <= H_DRY). pre: the
  // prefetched values of cell k = tid (null: read them here).
  template <int MODE>
  static HF_HD void phase_update(S& s, const StageArgs<T>& a, int gi0, int gj0, int tid, int nt,
                                 double dt, double* wsx, double* wsy, bool* wet, const Pre* pre = nullptr) {
    const T Z = cst<T>(0.0);
    const R dt_dx = R(dt / a.dx), dt_dy = R(dt / a.dy), rdt = R(dt);
    ForcingNow fnow{};
    if (MODE == MODE_FINAL_RK2) fnow = forcing_now(a, gi0, gj0, TH, TW);
    // FP32 form: the depth update h - dh and the RK combine round
    // stochastically too (keyed by step, cell and which addition), for the
    // same reason as the rain: in slowly varying flow a cell's small
    // increment rounds the same way step after step.
    uint32_t sr_step = 0u;
    if constexpr (std::is_same<T, float>::value) sr_step = rounding_seq(a) * 0x9E3779B9u;
    for (int k = tid; k < TW * TH; k += nt) {
      const int li = k / TW, lj = k % TW;
      const int gi = gi0 + li, gj = gj0 + lj;
      if (gi >= a.nr || gj >= a.nc) continue;
      const int64_t idx = (int64_t)gi * a.nc + gj;
      const Pre* pk = (pre && k == tid) ? pre : nullptr;
      const int c3 = S::i3(li, lj);
      const T h_old = s.h[c3];
      T nh, nhu, nhv;
      // the five flags (all inside the 4-cell halo) read together, not
      // behind one short-circuit branch each
      const unsigned dry5 = s.dry[S::i4(li, lj)] & s.dry[S::i4(li - 1, lj)] & s.dry[S::i4(li + 1, lj)] &
                            s.dry[S::i4(li, lj - 1)] & s.dry[S::i4(li, lj + 1)];
      if (dry5 != 0 && gi > 0 && gi + 1 < a.nr && gj > 0 && gj + 1 < a.nc) {
        nh = h_old;
        nhu = Z;
        nhv = Z;
      } else {
        T rm, ra, rb, rt, lm, la, lb, lt, bm, ba, bb, bt, tm, ta, tb, tt;
        scaled_fx(s, a, gj0, li, lj + 1, &rm, &ra, &rb, &rt);
        scaled_fx(s, a, gj0, li, lj, &lm, &la, &lb, &lt);
        scaled_fy(s, a, gi0, li + 1, lj, &bm, &ba, &bb, &bt);
        scaled_fy(s, a, gi0, li, lj, &tm, &ta, &tb, &tt);
        // right face .minus, left face .plus, bottom face .minus, top face .plus
        T dh = (rm - lm) * dt_dx + (bm - tm) * dt_dy;
        T new_h = h_old - dh;
        if constexpr (std::is_same<T, float>::value)
          new_h = add_rounded_stochastically(h_old, -dh, sr_step ^ ((uint32_t)idx * 0x85EBCA6Bu) ^
                                                             (MODE == MODE_FINAL_RK2 ? 0x68E31DA4u : 0xB5297A4Du));
        T dhu = (ra - lb) * dt_dx + (bt - tt) * dt_dy;
        T dhv = (rt - lt) * dt_dx + (ba - tb) * dt_dy;

        const T sex = s.sxe[S::isx(li, lj)];
        const T sey = s.sye[S::isy(li, lj)];
        T s_hu, s_hv;
        if constexpr (kRel) {
          // Same source in the cell's own bed frame: e = eta - z_cell and
          // dz = z_cell - z_face, so h_face = max(e + dz, 0) and
          // z_eff - z_cell = e - h_face.
          const T zc = s.z[c3];
          T dzl, dzr, dzt, dzb, tmp;
          if (gj > 0)
            interior_dz(s.dry[S::i4(li, lj - 1)] != 0, s.dry[S::i4(li, lj)] != 0, s.z[S::i3(li, lj - 1)] - zc, &tmp,
                        &dzl);
          else
            dzl = boundary_dz(a.bc_kind[SIDE_W], zc, a.nc >= 2 ? s.z[S::i3(li, lj + 1)] : zc);
          if (gj + 1 < a.nc)
            interior_dz(s.dry[S::i4(li, lj)] != 0, s.dry[S::i4(li, lj + 1)] != 0, zc - s.z[S::i3(li, lj + 1)], &dzr,
                        &tmp);
          else
            dzr = boundary_dz(a.bc_kind[SIDE_E], zc, a.nc >= 2 ? s.z[S::i3(li, lj - 1)] : zc);
          if (gi > 0)
            interior_dz(s.dry[S::i4(li - 1, lj)] != 0, s.dry[S::i4(li, lj)] != 0, s.z[S::i3(li - 1, lj)] - zc, &tmp,
                        &dzt);
          else
            dzt = boundary_dz(a.bc_kind[SIDE_N], zc, a.nr >= 2 ? s.z[S::i3(li + 1, lj)] : zc);
          if (gi + 1 < a.nr)
            interior_dz(s.dry[S::i4(li, lj)] != 0, s.dry[S::i4(li + 1, lj)] != 0, zc - s.z[S::i3(li + 1, lj)], &dzb,
                        &tmp);
          else
            dzb = boundary_dz(a.bc_kind[SIDE_S], zc, a.nr >= 2 ? s.z[S::i3(li - 1, lj)] : zc);
          T el = (gj > 0) ? h_old - sex * R(0.5 * a.dx) : h_old;
          T er = (gj + 1 < a.nc) ? h_old + sex * R(0.5 * a.dx) : h_old;
          T hfl = rmax(el + dzl, Z), hfr = rmax(er + dzr, Z);
          s_hu = div((hfl + hfr) * ((el - hfl) - (er - hfr)) * R(0.5 * GRAVITY), R(a.dx));
          T et = (gi > 0) ? h_old - sey * R(0.5 * a.dy) : h_old;
          T eb = (gi + 1 < a.nr) ? h_old + sey * R(0.5 * a.dy) : h_old;
          T hft = rmax(et + dzt, Z), hfb = rmax(eb + dzb, Z);
          s_hv = div((hft + hfb) * ((et - hft) - (eb - hfb)) * R(0.5 * GRAVITY), R(a.dy));
        } else {
          T eta_cell = h_old + s.z[c3];
          T eta_left = (gj > 0) ? eta_cell - sex * R(0.5 * a.dx) : eta_cell;
          T eta_right = (gj + 1 < a.nc) ? eta_cell + sex * R(0.5 * a.dx) : eta_cell;
          T hfl = rmax(eta_left - zface_x(s, a, gi0, gj0, li, lj), Z);
          T hfr = rmax(eta_right - zface_x(s, a, gi0, gj0, li, lj + 1), Z);
          T zel = eta_left - hfl;
          T zer = eta_right - hfr;
          s_hu = div((hfl + hfr) * (zel - zer) * R(0.5 * GRAVITY), R(a.dx));

          T eta_top = (gi > 0) ? eta_cell - sey * R(0.5 * a.dy) : eta_cell;
          T eta_bot = (gi + 1 < a.nr) ? eta_cell + sey * R(0.5 * a.dy) : eta_cell;
          T hft = rmax(eta_top - zface_y(s, a, gi0, gj0, li, lj), Z);
          T hfb = rmax(eta_bot - zface_y(s, a, gi0, gj0, li + 1, lj), Z);
          T zet = eta_top - hft;
          T zeb = eta_bot - hfb;
          s_hv = div((hft + hfb) * (zet - zeb) * R(0.5 * GRAVITY), R(a.dy));
        }

        if (val(new_h) <= R(H_VEL)) {
          nh = rmax(new_h, Z);
          nhu = Z;
          nhv = Z;
        } else {
          T hu_old = Z, hv_old = Z;
          if (val(h_old) > R(H_VEL)) {
            hu_old = pk ? pk->hu : a.ihu.ld(idx);
            hv_old = pk ? pk->hv : a.ihv.ld(idx);
          }
          nh = new_h;
          nhu = hu_old - dhu + s_hu * rdt;
          nhv = hv_old - dhv + s_hv * rdt;
        }
      }
      if (MODE == MODE_FINAL_RK2) {
        T ph = pk ? pk->ph : a.nh.ld(idx), phu = pk ? pk->phu : a.nhu.ld(idx), phv = pk ? pk->phv : a.nhv.ld(idx);
        if constexpr (std::is_same<T, float>::value)
          nh = add_rounded_stochastically(ph, nh, sr_step ^ ((uint32_t)idx * 0x85EBCA6Bu) ^ 0x1B56C4E9u) * 0.5f;
        else
          nh = (ph + nh) * R(0.5);
        nhu = (phu + nhu) * R(0.5);
        nhv = (phv + nhv) * R(0.5);
        if (val(nh) <= R(H_VEL)) {
          nh = rmax(nh, Z);
          nhu = Z;
          nhv = Z;
        }
        forcing(a, fnow, gi, gj, dt, &nh, &nhu, &nhv, pk ? &pk->man : nullptr);
        if (!(val(nh) <= R(H_DRY))) *wet = true;
        if (a.wave_speed) {
          T zn4[4];
          zn4[SIDE_N] = (a.nr >= 2 && gi == 0) ? s.z[S::i3(li + 1, lj)] : s.z[c3];
          zn4[SIDE_S] = (a.nr >= 2 && gi == a.nr - 1) ? s.z[S::i3(li - 1, lj)] : s.z[c3];
          zn4[SIDE_W] = (a.nc >= 2 && gj == 0) ? s.z[S::i3(li, lj + 1)] : s.z[c3];
          zn4[SIDE_E] = (a.nc >= 2 && gj == a.nc - 1) ? s.z[S::i3(li, lj - 1)] : s.z[c3];
          T man = a.cfl_with_bcs ? (pk ? pk->man : a.man.ld(idx)) : Z;
          cell_wave_speed(a, gi, gj, nh, nhu, nhv, s.z[c3], zn4, &man, wsx, wsy);
        }
      }
      a.oh.st(idx, nh);
      a.ohu.st(idx, nhu);
      a.ohv.st(idx, nhv);
    }
  }
};

// ------------------------------------------------------ active-tile list
//
// The reference skips dry cells inside a step: interior dry-dry faces get a
// zero flux and an interior dry cell with four dry neighbours keeps h and
// drops its momentum (update.rs:871-900, 1060-1090). At tile granularity
// this becomes a list of the tiles a step must touch, rebuilt every step.
//
// A tile is quiet for a step when no cell of it or of its 8 neighbour tiles
// is wet (h > H_DRY, or NaN) in U^n, no cell within 2 of it lies on the
// domain edge, and it holds no point source. Then every cell of the tile
// and of its 1-cell ring is an interior isolated-dry cell in stage 1, so
// U^(1) = (h, 0, 0) there, and the same holds in stage 2; the RK combine
// returns h exactly ((h + h) * 0.5), the floor sets (max(h, 0), 0, 0), and
// without rain and with friction off for dry cells nothing else changes.
// After two quiet steps both state buffers hold that fixed point bit for
// bit (max is idempotent), so from the third quiet step on the tile is left
// out of the list and untouched; its cells add nothing to the CFL maxima.
// The driver uses the list only where this argument holds: f64 and f32
// state (a Dual tangent at h == 0 halves at every max(h, 0) tie, so it is
// not a fixed point), SSP-RK2, no rain, friction dry tolerance >= H_DRY.

// Tiles that are never quiet: within 2 cells of the domain edge.
HF_HD bool tile_near_edge(int64_t nr, int64_t nc, int tw, int th, int64_t tx, int64_t ty) {
  const int64_t gi0 = ty * th, gj0 = tx * tw;
  return gi0 < 3 || gj0 < 3 || gi0 + th > nr - 3 || gj0 + tw > nc - 3;
}

// Decide whether tile (tx, ty) is left out of this step, and advance its
// count of consecutive quiet steps (saturating at 2). wet: per tile, a wet
// cell in U^n; pinned: per tile, never quiet (edge or point source).
HF_HD bool tile_skipped(const unsigned char* wet, const unsigned char* pinned, unsigned char* quiet_run, int tiles_x,
                        int tiles_y, int tx, int ty) {
  const int t = ty * tiles_x + tx;
  bool quiet = pinned[t] == 0;
  for (int dy = -1; dy <= 1 && quiet; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int y = ty + dy, x = tx + dx;
      if (y >= 0 && y < tiles_y && x >= 0 && x < tiles_x && wet[y * tiles_x + x]) {
        quiet = false;
        break;
      }
    }
  const unsigned char r = quiet_run[t];
  quiet_run[t] = quiet ? (unsigned char)(r >= 2 ? 2 : r + 1) : (unsigned char)0;
  return quiet && r >= 2;
}

// dt of the step about to run, from the device clock: the CFL bound of
// the current state (update.rs:170-208), capped by max_dt and dt_cap
// (sim.rs:206-208).
HF_HD double dt_from_ctrl(double sx, double sy, double cap, double cfl, double dx, double dy, double max_dt) {
  const double denom = sx / dx + sy / dy;
  double dt = denom > 0.0 ? cfl / denom : __builtin_inf();
  dt = rmin(rmin(dt, max_dt), cap);
  return dt;
}

// Whether the driver step about to start runs: the decisions of the first
// stage's prologue (stage_prologue, sim.rs:209-240) on the control block
// as that prologue will read it, without its side effects. The active-tile
// compaction runs before the first stage and uses it, so that steps that
// turn out to be no-ops (the clock reached its target, the budget is spent,
// a failed or finished driver, the trailing steps of a captured chunk) do
// not advance the tiles' quiet-step counts.
HF_HD bool step_will_run(const Ctrl* c, double cfl, double dx, double dy, double max_dt) {
  if ((c->status | c->done) != 0) return false;
  if (c->use_target && !(c->t < c->t_target)) return false;
  if (c->budget == 0) return false;
  const int p = c->parity;
  const double sx = hfm::as_f64(c->ws[p][0]), sy = hfm::as_f64(c->ws[p][1]);
  const double dt = dt_from_ctrl(sx, sy, c->use_target ? (c->t_target - c->t) : c->dt_cap, cfl, dx, dy, max_dt);
  return __builtin_isfinite(dt) && dt > 0.0;
}

// Forcing-series values at the end of the step about to run (t + dt), read
// by the final stage's forcing.
HF_HD void publish_series(Ctrl* c, const SeriesDesc* s, double dt) {
  const double t_new = clock_after(c, dt);
  for (int k = 0; k < 2; ++k) {
    c->series_on[k] = s[k].n > 0 ? 1 : 0;
    if (s[k].n > 0) c->series_now[k] = series_at(s[k], t_new);
  }
}

// An empty asm statement that reads v: no instruction, but the load that
// produced v completes here, before any branch that follows, so loads
// written before it are issued together instead of one per branch.
template <class V> HF_HD void loaded_here(V v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__ __volatile__("" ::"s"(v));
#else
  (void)v;
#endif
}

// Common prologue of every stage launch. Returns false when the launch is
// a no-op (driver finished or failed); block 0 records why.
template <class T>
HF_HD bool stage_prologue(const StageArgs<T>& a, bool lead, double* dt_out) {
  if (a.dt_mode == 0) {
    *dt_out = a.dt_host;
    return true;
  }
  Ctrl* c = a.ctrl;
  // Every control-block field a decision below reads is loaded before the
  // first test (both CFL slots, selected by parity afterwards): one memory
  // round trip per launch instead of one per short-circuited condition.
  const int status = c->status, done = c->done, use_target = c->use_target, parity = c->parity;
  const double t = c->t, t_target = c->t_target, dt_cap = c->dt_cap, dt_now = c->dt;
  const long long budget = c->budget;
  const unsigned long long w00 = c->ws[0][0], w01 = c->ws[0][1], w10 = c->ws[1][0], w11 = c->ws[1][1];
  loaded_here(status | done | use_target | parity);
  loaded_here(t);
  loaded_here(t_target);
  loaded_here(dt_cap);
  loaded_here(dt_now);
  loaded_here(budget);
  loaded_here(parity ? w10 : w00);
  loaded_here(parity ? w11 : w01);
  if ((status | done) != 0) return false;
  if (!a.first_stage) {
    *dt_out = dt_now;
    return true;
  }
  if (use_target && !(t < t_target)) {  // sim.rs:233
    if (lead) c->done = 1;
    return false;
  }
  if (budget == 0) {  // sim.rs:234-240
    if (lead) c->status = 4;
    return false;
  }
  const double sx = hfm::as_f64(parity ? w10 : w00), sy = hfm::as_f64(parity ? w11 : w01);
  double dt = dt_from_ctrl(sx, sy, use_target ? (t_target - t) : dt_cap, a.cfl, a.dx, a.dy, a.max_dt);
  if (!(__builtin_isfinite(dt) && dt > 0.0)) {  // sim.rs:209-215
    if (lead) {
      c->dt = dt;
      c->status = 3;
    }
    return false;
  }
  if (lead) {
    c->dt = dt;
    c->ws[parity ^ 1][0] = 0ULL;
    c->ws[parity ^ 1][1] = 0ULL;
    if (a.series) publish_series(c, a.series, dt);
  }
  *dt_out = dt;
  return true;
}

}  // namespace hf
