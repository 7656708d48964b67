// CPU restatement of hydroflux solver-2d's second-order time step.
//
// TEST INFRASTRUCTURE ONLY. This is the oracle the HIP path is checked
// against; nothing in the product (hydroflux_amd/) links, loads or calls it.
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
//
// It follows the reference's arithmetic pass by pass and expression by
// expression (operand order, thresholds and comparisons exactly as written),
// compiled with -ffp-contract=off because Rust never contracts a*b+c.
// Citations are to the reference tree (franciscoparrao/hydroflux):
//
//   constants          solver-2d/src/lib.rs:43,63,72; riemann.rs:74;
//                      boundary.rs:68,76
//   Real / Dual        autograd/src/real.rs:21-149; autograd/src/dual.rs:12-299
//   ghost cells        solver-2d/src/boundary.rs:193-367
//   HLLC               solver-2d/src/riemann.rs:80-204
//   CFL                solver-2d/src/update.rs:113-208
//   step passes        solver-2d/src/update.rs:215-1211
//   SSP-RK2            solver-2d/src/update.rs:1243-1287
//   friction/rain/pts  solver-2d/src/source.rs:52-165
//   driver             solver-2d/src/sim.rs:36-246
//
// Parity pins (each listed in the test that checks it): the reference's
// published CSV vectors (Stoker 400x3 and 100x3, Thacker 80^2, MacDonald
// 5x50, UK EA test 6 250x50 full field, Thacker convergence step counts,
// the Stoker 400x10 step count, and the forward-mode Dual Newton trace of the
// inverse-Manning demo), plus closed forms and invariants from the
// reference's own unit tests. Row-parallel OpenMP is bit-identical to serial
// because every pass is a pure map (the same property the reference's rayon
// mode relies on).
//
// Extensions beyond the reference (marked where they occur): rain and point
// sources accept Dual states (the reference's are f64-only, source.rs:13-17);
// the forcing is a constant, so the tangent passes through and is zeroed by
// the dry clamp. `pow_mode` selects glibc pow (what the Rust reference calls)
// or the correctly rounded routine in hf_pow.h that the GPU kernels use.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../hydroflux_amd/csrc/hf_pow.h"

namespace hfo {

// solver-2d/src/lib.rs:43,63,72; riemann.rs:74; boundary.rs:68,76
constexpr double GRAVITY = 9.81;
constexpr double H_DRY = 1.0e-6;
constexpr double H_VEL = 1.0e-5;
constexpr double DRY_TOL = 1.0e-12;
constexpr double DRY_INFLOW_SLOPE_THRESHOLD = 5.0e-4;
constexpr double DRY_INFLOW_H_MAX = 10.0;

// 0: glibc pow (the reference's libm call); 1: hf_pow.h (GPU-identical).
inline int& pow_mode() {
  static int m = 0;
  return m;
}

inline double powf64(double x, double n) {
  if (pow_mode() == 1) {
    if (n == 4.0 / 3.0) return hfm::pow_4_3(x);
    if (n == 4.0 / 3.0 - 1.0) return hfm::pow_1_3m(x);
    if (n == 3.0 / 5.0) return hfm::pow_3_5(x);
    if (n == 3.0 / 5.0 - 1.0) return hfm::pow_m2_5(x);
  }
  return std::pow(x, n);
}

// ---------------------------------------------------------------- scalars
// f64 branch of `impl Real for f64` (real.rs:73-107). Rust f64::max/min
// ignore a NaN operand.
inline double value(double x) { return x; }
inline double rmax(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}
inline double rmin(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a < b ? a : b;
}
inline double rsqrt_(double a) { return std::sqrt(a); }
inline double rpowf(double a, double n) { return powf64(a, n); }
template <class T> inline T from_f64(double x);
template <> inline double from_f64<double>(double x) { return x; }

// Forward-mode dual number, autograd/src/dual.rs:12-299.
struct Dual {
  double val, dval;
};
inline double value(const Dual& d) { return d.val; }
inline Dual operator+(Dual a, Dual b) { return {a.val + b.val, a.dval + b.dval}; }
inline Dual operator-(Dual a, Dual b) { return {a.val - b.val, a.dval - b.dval}; }
inline Dual operator*(Dual a, Dual b) { return {a.val * b.val, a.dval * b.val + a.val * b.dval}; }
inline Dual operator/(Dual a, Dual b) {
  return {a.val / b.val, (a.dval * b.val - a.val * b.dval) / (b.val * b.val)};
}
inline Dual operator-(Dual a) { return {-a.val, -a.dval}; }
inline Dual operator+(Dual a, double b) { return {a.val + b, a.dval}; }
inline Dual operator-(Dual a, double b) { return {a.val - b, a.dval}; }
inline Dual operator*(Dual a, double b) { return {a.val * b, a.dval * b}; }
inline Dual operator/(Dual a, double b) { return {a.val / b, a.dval / b}; }
// dual.rs:130-155: ties average the tangents.
inline Dual rmax(Dual a, Dual b) {
  if (a.val > b.val) return a;
  if (b.val > a.val) return b;
  return {a.val, 0.5 * (a.dval + b.dval)};
}
inline Dual rmin(Dual a, Dual b) {
  if (a.val < b.val) return a;
  if (b.val < a.val) return b;
  return {a.val, 0.5 * (a.dval + b.dval)};
}
// dual.rs:46-50: tangent 0 at sqrt(0).
inline Dual rsqrt_(Dual a) {
  double s = std::sqrt(a.val);
  double d = (s == 0.0) ? 0.0 : a.dval / (2.0 * s);
  return {s, d};
}
// dual.rs:112-117
inline Dual rpowf(Dual a, double n) {
  return {powf64(a.val, n), n * powf64(a.val, n - 1.0) * a.dval};
}
template <> inline Dual from_f64<Dual>(double x) { return {x, 0.0}; }

// Vector-mode dual, autograd/src/dual_n.rs (Real impl real.rs:161-195): a
// value and N tangents. combine(o, val, a, b) sets tangent i to
// a * self_i + b * o_i; chain(val, dfdx) to self_i * dfdx; an f64 operand
// is a constant DualN (scalar_ops). Division: inv = 1 / o.val, value
// self.val * inv, factors (inv, -self.val * inv * inv).
template <int N> struct DualN {
  double val;
  double dval[N];
};
template <int N> inline DualN<N> dn_constant(double x) {
  DualN<N> r;
  r.val = x;
  for (int i = 0; i < N; ++i) r.dval[i] = 0.0;
  return r;
}
template <int N> inline DualN<N> dn_combine(const DualN<N>& s, const DualN<N>& o, double val, double a, double b) {
  DualN<N> r;
  r.val = val;
  for (int i = 0; i < N; ++i) r.dval[i] = a * s.dval[i] + b * o.dval[i];
  return r;
}
template <int N> inline DualN<N> dn_chain(const DualN<N>& s, double val, double dfdx) {
  DualN<N> r;
  r.val = val;
  for (int i = 0; i < N; ++i) r.dval[i] = s.dval[i] * dfdx;
  return r;
}
template <int N> inline double value(const DualN<N>& d) { return d.val; }
template <int N> inline DualN<N> operator+(const DualN<N>& s, const DualN<N>& o) {
  return dn_combine(s, o, s.val + o.val, 1.0, 1.0);
}
template <int N> inline DualN<N> operator-(const DualN<N>& s, const DualN<N>& o) {
  return dn_combine(s, o, s.val - o.val, 1.0, -1.0);
}
template <int N> inline DualN<N> operator*(const DualN<N>& s, const DualN<N>& o) {
  return dn_combine(s, o, s.val * o.val, o.val, s.val);
}
// Value of a DualN quotient: 0 the reference's s.val * (1 / o.val); 1 (a
// test switch, not the reference) s.val / o.val, f64's value, so that a
// DualN run follows the f64 value path exactly and its tangents can be
// compared with seeded Dual runs to rounding.
inline int& dualn_div_mode() {
  static int m = 0;
  return m;
}
template <int N> inline DualN<N> operator/(const DualN<N>& s, const DualN<N>& o) {
  const double inv = 1.0 / o.val;
  return dn_combine(s, o, dualn_div_mode() ? s.val / o.val : s.val * inv, inv, -s.val * inv * inv);
}
template <int N> inline DualN<N> operator-(const DualN<N>& s) { return dn_chain(s, -s.val, -1.0); }
template <int N> inline DualN<N> operator+(const DualN<N>& s, double x) { return s + dn_constant<N>(x); }
template <int N> inline DualN<N> operator-(const DualN<N>& s, double x) { return s - dn_constant<N>(x); }
template <int N> inline DualN<N> operator*(const DualN<N>& s, double x) { return s * dn_constant<N>(x); }
template <int N> inline DualN<N> operator/(const DualN<N>& s, double x) { return s / dn_constant<N>(x); }
template <int N> inline DualN<N> rmax(const DualN<N>& s, const DualN<N>& o) {  // dual_n.rs max
  if (s.val > o.val) return s;
  if (o.val > s.val) return o;
  return dn_combine(s, o, s.val, 0.5, 0.5);
}
template <int N> inline DualN<N> rmin(const DualN<N>& s, const DualN<N>& o) {
  if (s.val < o.val) return s;
  if (o.val < s.val) return o;
  return dn_combine(s, o, s.val, 0.5, 0.5);
}
template <int N> inline DualN<N> rsqrt_(const DualN<N>& s) {  // dual_n.rs sqrt
  const double r = std::sqrt(s.val);
  return dn_chain(s, r, r == 0.0 ? 0.0 : 0.5 / r);
}
template <int N> inline DualN<N> rpowf(const DualN<N>& s, double n) {  // dual_n.rs powf
  return dn_chain(s, powf64(s.val, n), n * powf64(s.val, n - 1.0));
}
template <> inline DualN<2> from_f64<DualN<2>>(double x) { return dn_constant<2>(x); }
template <> inline DualN<3> from_f64<DualN<3>>(double x) { return dn_constant<3>(x); }
template <> inline DualN<4> from_f64<DualN<4>>(double x) { return dn_constant<4>(x); }
template <> inline DualN<8> from_f64<DualN<8>>(double x) { return dn_constant<8>(x); }
template <> inline DualN<16> from_f64<DualN<16>>(double x) { return dn_constant<16>(x); }

template <class T> inline T zero() { return from_f64<T>(0.0); }
template <class T> inline T one() { return from_f64<T>(1.0); }

// ---------------------------------------------------------------- types
template <class T> struct Cell {  // state.rs:21-33
  T h, hu, hv;
};
template <class T> inline Cell<T> dry_cell() { return {zero<T>(), zero<T>(), zero<T>()}; }

template <class T> struct Flux {  // flux.rs:28-52 (x and y share the layout)
  T mass, x_momentum, y_momentum;
};
template <class T> inline Flux<T> zero_flux() { return {zero<T>(), zero<T>(), zero<T>()}; }

template <class T> struct FaceFlux {  // update.rs:215-250
  Flux<T> minus, plus;
};

template <class T> struct Mesh {  // geometry.rs:31-113
  int64_t nr = 0, nc = 0;
  double dx = 1.0, dy = 1.0;
  std::vector<T> bed, manning;
  const T& z(int64_t i, int64_t j) const { return bed[i * nc + j]; }
};

enum BcKind { BC_TRANSMISSIVE = 0, BC_WALL = 1, BC_DISCHARGE = 2, BC_DEPTH = 3 };
struct Bc {
  int kind = BC_WALL;
  double param = 0.0;  // q for Discharge, h for Depth
};
struct Bcs {  // boundary.rs:131-160; order N, S, W, E
  Bc north, south, west, east;
};
enum Side { NORTH = 0, SOUTH = 1, WEST = 2, EAST = 3 };

// ---------------------------------------------------------------- ghosts
// boundary.rs:274-333
template <class T> Cell<T> ghost_state(Cell<T> inner, Bc bc, Side side) {
  bool xf = (side == WEST || side == EAST);
  switch (bc.kind) {
    case BC_TRANSMISSIVE:
      return inner;
    case BC_WALL:
      if (xf) return {inner.h, -inner.hu, inner.hv};
      return {inner.h, inner.hu, -inner.hv};
    case BC_DISCHARGE:
      if (xf) return {inner.h, from_f64<T>(bc.param), inner.hv};
      return {inner.h, inner.hu, from_f64<T>(bc.param)};
    default:  // BC_DEPTH
      return {from_f64<T>(bc.param), inner.hu, inner.hv};
  }
}

// boundary.rs:335-367
template <class T> T ghost_bed(const Mesh<T>& m, Bc bc, Side side, int64_t idx) {
  const int64_t nr = m.nr, nc = m.nc;
  if (bc.kind == BC_TRANSMISSIVE || bc.kind == BC_WALL) {
    switch (side) {
      case NORTH: return m.z(0, idx);
      case SOUTH: return m.z(nr - 1, idx);
      case WEST: return m.z(idx, 0);
      default: return m.z(idx, nc - 1);
    }
  }
  switch (side) {
    case NORTH: return m.z(0, idx) * 2.0 - m.z(1, idx);
    case SOUTH: return m.z(nr - 1, idx) * 2.0 - m.z(nr - 2, idx);
    case WEST: return m.z(idx, 0) * 2.0 - m.z(idx, 1);
    default: return m.z(idx, nc - 1) * 2.0 - m.z(idx, nc - 2);
  }
}

// geometry.rs:131-146 bed_slope_x / bed_slope_y, boundary.rs:241-272
template <class T> double streamwise_slope_mag(const Mesh<T>& m, Side side, int64_t idx) {
  switch (side) {
    case WEST:
      return m.nc < 2 ? 0.0 : std::fabs(value((m.z(idx, 0) - m.z(idx, 1)) / m.dx));
    case EAST:
      return m.nc < 2 ? 0.0 : std::fabs(value((m.z(idx, m.nc - 2) - m.z(idx, m.nc - 1)) / m.dx));
    case NORTH:
      return m.nr < 2 ? 0.0 : std::fabs(value((m.z(0, idx) - m.z(1, idx)) / m.dy));
    default:
      return m.nr < 2 ? 0.0 : std::fabs(value((m.z(m.nr - 2, idx) - m.z(m.nr - 1, idx)) / m.dy));
  }
}

// boundary.rs:193-219 (Discharge over a dry inner cell: Manning normal depth).
template <class T>
void ghost_cell(const Mesh<T>& m, Cell<T> inner, Bc bc, Side side, int64_t idx, Cell<T>* g,
                T* zg) {
  Cell<T> st = ghost_state(inner, bc, side);
  T bed = ghost_bed(m, bc, side, idx);
  if (bc.kind == BC_DISCHARGE && value(inner.h) < H_DRY) {
    double slope = streamwise_slope_mag(m, side, idx);
    if (slope > DRY_INFLOW_SLOPE_THRESHOLD) {
      int64_t ci, cj;
      switch (side) {
        case WEST: ci = idx; cj = 0; break;
        case EAST: ci = idx; cj = m.nc - 1; break;
        case NORTH: ci = 0; cj = idx; break;
        default: ci = m.nr - 1; cj = idx; break;
      }
      T n = m.manning[ci * m.nc + cj];
      T hn = rpowf(n * (std::fabs(bc.param) / std::sqrt(slope)), 3.0 / 5.0);
      st.h = rmin(hn, from_f64<T>(DRY_INFLOW_H_MAX));
    }
  }
  *g = st;
  *zg = bed;
}

// ---------------------------------------------------------------- HLLC
// riemann.rs:114-204. Returns (mass, normal, tangential).
template <class T>
void hllc_normal_flux(T h_l, T qn_l, T qt_l, T h_r, T qn_r, T qt_r, T* fm, T* fn, T* ft) {
  const bool lw = value(h_l) > DRY_TOL;
  const bool rw = value(h_r) > DRY_TOL;
  if (!lw && !rw) {
    *fm = zero<T>(); *fn = zero<T>(); *ft = zero<T>();
    return;
  }
  T cl = lw ? rsqrt_(h_l * GRAVITY) : zero<T>();
  T cr = rw ? rsqrt_(h_r * GRAVITY) : zero<T>();
  T un_l = lw ? qn_l / h_l : zero<T>();
  T un_r = rw ? qn_r / h_r : zero<T>();
  T ut_l = lw ? qt_l / h_l : zero<T>();
  T ut_r = rw ? qt_r / h_r : zero<T>();
  T sl, sr;
  if (lw && rw) {
    sl = rmin(un_l - cl, un_r - cr);
    sr = rmax(un_l + cl, un_r + cr);
  } else if (!lw && rw) {
    sl = un_r - cr * 2.0;
    sr = un_r + cr;
  } else {
    sl = un_l - cl;
    sr = un_l + cl * 2.0;
  }
  T fm_l = zero<T>(), fn_l = zero<T>(), ft_l = zero<T>();
  if (lw) {
    fm_l = qn_l;
    fn_l = qn_l * un_l + h_l * h_l * (0.5 * GRAVITY);
    ft_l = qn_l * ut_l;
  }
  T fm_r = zero<T>(), fn_r = zero<T>(), ft_r = zero<T>();
  if (rw) {
    fm_r = qn_r;
    fn_r = qn_r * un_r + h_r * h_r * (0.5 * GRAVITY);
    ft_r = qn_r * ut_r;
  }
  if (value(sl) >= 0.0) {
    *fm = fm_l; *fn = fn_l; *ft = ft_l;
    return;
  }
  if (value(sr) <= 0.0) {
    *fm = fm_r; *fn = fn_r; *ft = ft_r;
    return;
  }
  T denom = h_r * (un_r - sr) - h_l * (un_l - sl);
  T s_star = (std::fabs(value(denom)) > 1.0e-12)
                 ? (h_r * (un_r - sr) * sl - h_l * (un_l - sl) * sr) / denom
                 : (sl + sr) * 0.5;
  T h_star_l = h_l * (sl - un_l) / (sl - s_star);
  T h_star_r = h_r * (sr - un_r) / (sr - s_star);
  if (value(s_star) >= 0.0) {
    *fm = fm_l + (h_star_l - h_l) * sl;
    *fn = fn_l + (h_star_l * s_star - qn_l) * sl;
    *ft = ft_l + (h_star_l * ut_l - qt_l) * sl;
  } else {
    *fm = fm_r + (h_star_r - h_r) * sr;
    *fn = fn_r + (h_star_r * s_star - qn_r) * sr;
    *ft = ft_r + (h_star_r * ut_r - qt_r) * sr;
  }
}

// riemann.rs:80-107
template <class T> Flux<T> hllc_flux_x(Cell<T> l, Cell<T> r) {
  Flux<T> f;
  hllc_normal_flux(l.h, l.hu, l.hv, r.h, r.hu, r.hv, &f.mass, &f.x_momentum, &f.y_momentum);
  return f;
}
template <class T> Flux<T> hllc_flux_y(Cell<T> l, Cell<T> r) {
  Flux<T> f;
  hllc_normal_flux(l.h, l.hv, l.hu, r.h, r.hv, r.hu, &f.mass, &f.y_momentum, &f.x_momentum);
  return f;
}

// ---------------------------------------------------------------- faces
// update.rs:256-353. `y_face` selects the y-face variant.
template <class T>
FaceFlux<T> well_balanced_face(Cell<T> left, T z_left, Cell<T> right, T z_right, bool y_face) {
  T z_max = rmax(z_left, z_right);
  T hsl = rmax(left.h + z_left - z_max, zero<T>());
  T hsr = rmax(right.h + z_right - z_max, zero<T>());
  T ul = zero<T>(), vl = zero<T>(), ur = zero<T>(), vr = zero<T>();
  if (value(left.h) > H_VEL) {
    ul = left.hu / left.h;
    vl = left.hv / left.h;
  }
  if (value(right.h) > H_VEL) {
    ur = right.hu / right.h;
    vr = right.hv / right.h;
  }
  Cell<T> rl{hsl, hsl * ul, hsl * vl};
  Cell<T> rr{hsr, hsr * ur, hsr * vr};
  Flux<T> f = y_face ? hllc_flux_y(rl, rr) : hllc_flux_x(rl, rr);
  T corr_l = (left.h * left.h - hsl * hsl) * (0.5 * GRAVITY);
  T corr_r = (right.h * right.h - hsr * hsr) * (0.5 * GRAVITY);
  FaceFlux<T> out;
  out.minus = f;
  out.plus = f;
  if (y_face) {
    out.minus.y_momentum = f.y_momentum + corr_l;
    out.plus.y_momentum = f.y_momentum + corr_r;
  } else {
    out.minus.x_momentum = f.x_momentum + corr_l;
    out.plus.x_momentum = f.x_momentum + corr_r;
  }
  return out;
}

// update.rs:362-370
template <class T> T minmod(T a, T b) {
  if (value(a) * value(b) <= 0.0) return zero<T>();
  if (std::fabs(value(a)) < std::fabs(value(b))) return a;
  return b;
}

template <class T> struct Prim {  // update.rs:394-399
  T eta, u, v;
};
template <class T> struct Slopes {  // update.rs:383-388
  T eta, u, v;
};

// update.rs:650-656
template <class T> T interior_z_face(bool l_dry, T z_l, bool r_dry, T z_r) {
  if (l_dry != r_dry) return rmax(z_l, z_r);
  return (z_l + z_r) * 0.5;
}

// update.rs:743-777
template <class T> struct Workspace {
  std::vector<Prim<T>> prim;
  std::vector<Slopes<T>> slopes_x, slopes_y;
  std::vector<T> z_face_x, z_face_y;
  std::vector<FaceFlux<T>> faces_x, faces_y;
  std::vector<T> alpha;
  std::vector<uint8_t> was_dry;
  std::vector<Cell<T>> u_n;
  void resize(int64_t nr, int64_t nc) {
    size_t n = (size_t)(nr * nc);
    prim.assign(n, Prim<T>{});
    slopes_x.assign(n, Slopes<T>{});
    slopes_y.assign(n, Slopes<T>{});
    z_face_x.assign((size_t)(nr * (nc + 1)), zero<T>());
    z_face_y.assign((size_t)((nr + 1) * nc), zero<T>());
    faces_x.assign((size_t)(nr * (nc + 1)), FaceFlux<T>{});
    faces_y.assign((size_t)((nr + 1) * nc), FaceFlux<T>{});
    alpha.assign(n, one<T>());
    was_dry.assign(n, 0);
    u_n.assign(n, dry_cell<T>());
  }
};

// ---------------------------------------------------------------- CFL
// update.rs:113-136
template <class T> void max_wave_speeds(const std::vector<Cell<T>>& st, double* sx_out, double* sy_out) {
  double sx = 0.0, sy = 0.0;
  for (const Cell<T>& s : st) {
    if (value(s.h) <= H_DRY) continue;
    double h = value(s.h);
    double c = std::sqrt(GRAVITY * h);
    double u = 0.0, v = 0.0;
    if (h > H_VEL) {
      u = value(s.hu) / h;
      v = value(s.hv) / h;
    }
    sx = rmax(sx, std::fabs(u) + c);
    sy = rmax(sy, std::fabs(v) + c);
  }
  *sx_out = sx;
  *sy_out = sy;
}

// update.rs:148-156
template <class T> double cfl_time_step(const std::vector<Cell<T>>& st, const Mesh<T>& m, double cfl) {
  double sx, sy;
  max_wave_speeds(st, &sx, &sy);
  double denom = sx / m.dx + sy / m.dy;
  return denom > 0.0 ? cfl / denom : std::numeric_limits<double>::infinity();
}

// update.rs:170-208
template <class T>
double cfl_time_step_with_bcs(const std::vector<Cell<T>>& st, const Mesh<T>& m, const Bcs& bcs,
                              double cfl) {
  double sx, sy;
  max_wave_speeds(st, &sx, &sy);
  auto fold = [&](Cell<T> g) {
    double h = value(g.h);
    if (h <= H_DRY) return;
    double c = std::sqrt(GRAVITY * h);
    double u = value(g.hu) / h;
    double v = value(g.hv) / h;
    sx = rmax(sx, std::fabs(u) + c);
    sy = rmax(sy, std::fabs(v) + c);
  };
  const int64_t nr = m.nr, nc = m.nc;
  Cell<T> g;
  T zg;
  for (int64_t i = 0; i < nr; ++i) {
    ghost_cell(m, st[i * nc + 0], bcs.west, WEST, i, &g, &zg);
    fold(g);
    ghost_cell(m, st[i * nc + nc - 1], bcs.east, EAST, i, &g, &zg);
    fold(g);
  }
  for (int64_t j = 0; j < nc; ++j) {
    ghost_cell(m, st[0 * nc + j], bcs.north, NORTH, j, &g, &zg);
    fold(g);
    ghost_cell(m, st[(nr - 1) * nc + j], bcs.south, SOUTH, j, &g, &zg);
    fold(g);
  }
  double denom = sx / m.dx + sy / m.dy;
  return denom > 0.0 ? cfl / denom : std::numeric_limits<double>::infinity();
}

// ---------------------------------------------------------------- Euler step
// update.rs:805-1211
template <class T>
void forward_euler_step_with(std::vector<Cell<T>>& st, const Mesh<T>& m, const Bcs& bcs, double dt,
                             Workspace<T>& ws) {
  const int64_t nr = m.nr, nc = m.nc;
  if (nr == 0 || nc == 0) return;
  std::vector<uint8_t>& dry = ws.was_dry;
  auto D = [&](int64_t i, int64_t j) -> bool { return dry[i * nc + j] != 0; };
  auto S = [&](int64_t i, int64_t j) -> const Cell<T>& { return st[i * nc + j]; };

  // update.rs:853-855 pre-step wet/dry snapshot
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j < nc; ++j) dry[i * nc + j] = value(st[i * nc + j].h) <= H_DRY;

  // update.rs:405-426 fill_primitives
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      const Cell<T>& s = S(i, j);
      Prim<T>& p = ws.prim[i * nc + j];
      if (value(s.h) > H_VEL) {
        p = {s.h + m.z(i, j), s.hu / s.h, s.hv / s.h};
      } else {
        p = {s.h + m.z(i, j), zero<T>(), zero<T>()};
      }
    }
  auto P = [&](int64_t i, int64_t j) -> const Prim<T>& { return ws.prim[i * nc + j]; };

  // update.rs:459-499 fill_slopes_x
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      Slopes<T>& sl = ws.slopes_x[i * nc + j];
      bool near_dry = false;
      {
        int64_t lo = j >= 2 ? j - 2 : 0, hi = std::min<int64_t>(j + 2, nc - 1);
        for (int64_t jj = lo; jj <= hi; ++jj) near_dry = near_dry || D(i, jj);
      }
      if (nc < 2 || near_dry) {
        sl = {zero<T>(), zero<T>(), zero<T>()};
      } else if (j == 0) {
        const Prim<T>&c = P(i, 0), &r = P(i, 1);
        sl = {(r.eta - c.eta) / m.dx, (r.u - c.u) / m.dx, (r.v - c.v) / m.dx};
      } else if (j + 1 == nc) {
        const Prim<T>&l = P(i, j - 1), &c = P(i, j);
        sl = {(c.eta - l.eta) / m.dx, (c.u - l.u) / m.dx, (c.v - l.v) / m.dx};
      } else {
        const Prim<T>&l = P(i, j - 1), &c = P(i, j), &r = P(i, j + 1);
        sl = {minmod((c.eta - l.eta) / m.dx, (r.eta - c.eta) / m.dx),
              minmod((c.u - l.u) / m.dx, (r.u - c.u) / m.dx),
              minmod((c.v - l.v) / m.dx, (r.v - c.v) / m.dx)};
      }
    }
  // update.rs:503-540 fill_slopes_y
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      Slopes<T>& sl = ws.slopes_y[i * nc + j];
      bool near_dry = false;
      {
        int64_t lo = i >= 2 ? i - 2 : 0, hi = std::min<int64_t>(i + 2, nr - 1);
        for (int64_t ii = lo; ii <= hi; ++ii) near_dry = near_dry || D(ii, j);
      }
      if (nr < 2 || near_dry) {
        sl = {zero<T>(), zero<T>(), zero<T>()};
      } else if (i == 0) {
        const Prim<T>&c = P(0, j), &b = P(1, j);
        sl = {(b.eta - c.eta) / m.dy, (b.u - c.u) / m.dy, (b.v - c.v) / m.dy};
      } else if (i + 1 == nr) {
        const Prim<T>&t = P(i - 1, j), &c = P(i, j);
        sl = {(c.eta - t.eta) / m.dy, (c.u - t.u) / m.dy, (c.v - t.v) / m.dy};
      } else {
        const Prim<T>&t = P(i - 1, j), &c = P(i, j), &b = P(i + 1, j);
        sl = {minmod((c.eta - t.eta) / m.dy, (b.eta - c.eta) / m.dy),
              minmod((c.u - t.u) / m.dy, (b.u - c.u) / m.dy),
              minmod((c.v - t.v) / m.dy, (b.v - c.v) / m.dy)};
      }
    }

  // update.rs:676-728 face beds
  const int64_t ncx = nc + 1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j <= nc; ++j) {
      T& z = ws.z_face_x[i * ncx + j];
      Cell<T> g;
      T zg;
      if (j == 0) {
        ghost_cell(m, S(i, 0), bcs.west, WEST, i, &g, &zg);
        z = (zg + m.z(i, 0)) * 0.5;
      } else if (j == nc) {
        ghost_cell(m, S(i, nc - 1), bcs.east, EAST, i, &g, &zg);
        z = (m.z(i, nc - 1) + zg) * 0.5;
      } else {
        z = interior_z_face(D(i, j - 1), m.z(i, j - 1), D(i, j), m.z(i, j));
      }
    }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i <= nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      T& z = ws.z_face_y[i * nc + j];
      Cell<T> g;
      T zg;
      if (i == 0) {
        ghost_cell(m, S(0, j), bcs.north, NORTH, j, &g, &zg);
        z = (zg + m.z(0, j)) * 0.5;
      } else if (i == nr) {
        ghost_cell(m, S(nr - 1, j), bcs.south, SOUTH, j, &g, &zg);
        z = (m.z(nr - 1, j) + zg) * 0.5;
      } else {
        z = interior_z_face(D(i - 1, j), m.z(i - 1, j), D(i, j), m.z(i, j));
      }
    }

  // update.rs:560-623 + 890-922 face fluxes
  const double half_dx = 0.5 * m.dx, half_dy = 0.5 * m.dy;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j <= nc; ++j) {
      FaceFlux<T>& face = ws.faces_x[i * ncx + j];
      T zf = ws.z_face_x[i * ncx + j];
      Cell<T> g;
      T zg;
      if (j == 0) {
        ghost_cell(m, S(i, 0), bcs.west, WEST, i, &g, &zg);
        face = well_balanced_face(g, zf, S(i, 0), zf, false);
      } else if (j == nc) {
        ghost_cell(m, S(i, nc - 1), bcs.east, EAST, i, &g, &zg);
        face = well_balanced_face(S(i, nc - 1), zf, g, zf, false);
      } else if (D(i, j - 1) && D(i, j)) {
        face.minus = zero_flux<T>();
        face.plus = zero_flux<T>();
      } else {
        const Prim<T>& l = P(i, j - 1);
        const Slopes<T>& sL = ws.slopes_x[i * nc + j - 1];
        T eta_m = l.eta + sL.eta * half_dx;
        T u_m = l.u + sL.u * half_dx;
        T v_m = l.v + sL.v * half_dx;
        const Prim<T>& r = P(i, j);
        const Slopes<T>& sR = ws.slopes_x[i * nc + j];
        T eta_p = r.eta - sR.eta * half_dx;
        T u_p = r.u - sR.u * half_dx;
        T v_p = r.v - sR.v * half_dx;
        T h_m = rmax(eta_m - zf, zero<T>());
        T h_p = rmax(eta_p - zf, zero<T>());
        Cell<T> rl{h_m, h_m * u_m, h_m * v_m};
        Cell<T> rr{h_p, h_p * u_p, h_p * v_p};
        face = well_balanced_face(rl, zf, rr, zf, false);
      }
    }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i <= nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      FaceFlux<T>& face = ws.faces_y[i * nc + j];
      T zf = ws.z_face_y[i * nc + j];
      Cell<T> g;
      T zg;
      if (i == 0) {
        ghost_cell(m, S(0, j), bcs.north, NORTH, j, &g, &zg);
        face = well_balanced_face(g, zf, S(0, j), zf, true);
      } else if (i == nr) {
        ghost_cell(m, S(nr - 1, j), bcs.south, SOUTH, j, &g, &zg);
        face = well_balanced_face(S(nr - 1, j), zf, g, zf, true);
      } else if (D(i - 1, j) && D(i, j)) {
        face.minus = zero_flux<T>();
        face.plus = zero_flux<T>();
      } else {
        const Prim<T>& t = P(i - 1, j);
        const Slopes<T>& sT = ws.slopes_y[(i - 1) * nc + j];
        T eta_m = t.eta + sT.eta * half_dy;
        T u_m = t.u + sT.u * half_dy;
        T v_m = t.v + sT.v * half_dy;
        const Prim<T>& b = P(i, j);
        const Slopes<T>& sB = ws.slopes_y[i * nc + j];
        T eta_p = b.eta - sB.eta * half_dy;
        T u_p = b.u - sB.u * half_dy;
        T v_p = b.v - sB.v * half_dy;
        T h_m = rmax(eta_m - zf, zero<T>());
        T h_p = rmax(eta_p - zf, zero<T>());
        Cell<T> rt{h_m, h_m * u_m, h_m * v_m};
        Cell<T> rb{h_p, h_p * u_p, h_p * v_p};
        face = well_balanced_face(rt, zf, rb, zf, true);
      }
    }

  const double dt_dx = dt / m.dx;
  const double dt_dy = dt / m.dy;
  auto FX = [&](int64_t i, int64_t j) -> FaceFlux<T>& { return ws.faces_x[i * ncx + j]; };
  auto FY = [&](int64_t i, int64_t j) -> FaceFlux<T>& { return ws.faces_y[i * nc + j]; };

  // update.rs:953-982 alpha
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      T fx_right = FX(i, j + 1).minus.mass;
      T fx_left = FX(i, j).plus.mass;
      T fy_bottom = FY(i + 1, j).minus.mass;
      T fy_top = FY(i, j).plus.mass;
      T out = zero<T>();
      if (value(fx_right) > 0.0) out = out + fx_right * dt_dx;
      if (value(fx_left) < 0.0) out = out + (-fx_left) * dt_dx;
      if (value(fy_bottom) > 0.0) out = out + fy_bottom * dt_dy;
      if (value(fy_top) < 0.0) out = out + (-fy_top) * dt_dy;
      T avail = rmax(S(i, j).h - H_DRY, zero<T>());
      ws.alpha[i * nc + j] =
          (value(out) > value(avail) && value(out) > 0.0) ? avail / out : one<T>();
    }
  auto A = [&](int64_t i, int64_t j) -> const T& { return ws.alpha[i * nc + j]; };
  auto scale = [](FaceFlux<T>& f, T a) {
    f.minus.mass = f.minus.mass * a;
    f.minus.x_momentum = f.minus.x_momentum * a;
    f.minus.y_momentum = f.minus.y_momentum * a;
    f.plus.mass = f.plus.mass * a;
    f.plus.x_momentum = f.plus.x_momentum * a;
    f.plus.y_momentum = f.plus.y_momentum * a;
  };
  // update.rs:992-1025
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j <= nc; ++j) {
      if (j > 0 && j < nc && D(i, j - 1) && D(i, j)) continue;
      FaceFlux<T>& f = FX(i, j);
      T a;
      if (j == 0) {
        a = value(f.minus.mass) < 0.0 ? A(i, 0) : one<T>();
      } else if (j == nc) {
        a = value(f.minus.mass) > 0.0 ? A(i, nc - 1) : one<T>();
      } else if (value(f.minus.mass) > 0.0) {
        a = A(i, j - 1);
      } else {
        a = A(i, j);
      }
      scale(f, a);
    }
  // update.rs:1026-1053
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i <= nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      if (i > 0 && i < nr && D(i - 1, j) && D(i, j)) continue;
      FaceFlux<T>& f = FY(i, j);
      T a;
      if (i == 0) {
        a = value(f.minus.mass) < 0.0 ? A(0, j) : one<T>();
      } else if (i == nr) {
        a = value(f.minus.mass) > 0.0 ? A(nr - 1, j) : one<T>();
      } else if (value(f.minus.mass) > 0.0) {
        a = A(i - 1, j);
      } else {
        a = A(i, j);
      }
      scale(f, a);
    }

  // update.rs:1071-1210 FV update with explicit bed-slope source
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nr; ++i)
    for (int64_t j = 0; j < nc; ++j) {
      Cell<T>& cell = st[i * nc + j];
      if (D(i, j) && i > 0 && i + 1 < nr && j > 0 && j + 1 < nc && D(i - 1, j) && D(i + 1, j) &&
          D(i, j - 1) && D(i, j + 1)) {
        cell.hu = zero<T>();
        cell.hv = zero<T>();
        continue;
      }
      const Flux<T>& fr = FX(i, j + 1).minus;
      const Flux<T>& fl = FX(i, j).plus;
      const Flux<T>& fb = FY(i + 1, j).minus;
      const Flux<T>& ft = FY(i, j).plus;
      T dh = (fr.mass - fl.mass) * dt_dx + (fb.mass - ft.mass) * dt_dy;
      T new_h = cell.h - dh;
      T dhu = (fr.x_momentum - fl.x_momentum) * dt_dx + (fb.x_momentum - ft.x_momentum) * dt_dy;
      T dhv = (fr.y_momentum - fl.y_momentum) * dt_dx + (fb.y_momentum - ft.y_momentum) * dt_dy;

      T h_old = cell.h;
      T eta_cell = h_old + m.z(i, j);
      T sex = ws.slopes_x[i * nc + j].eta;
      T eta_left = (j > 0) ? eta_cell - sex * (0.5 * m.dx) : eta_cell;
      T eta_right = (j + 1 < nc) ? eta_cell + sex * (0.5 * m.dx) : eta_cell;
      T hfl = rmax(eta_left - ws.z_face_x[i * ncx + j], zero<T>());
      T hfr = rmax(eta_right - ws.z_face_x[i * ncx + j + 1], zero<T>());
      T zel = eta_left - hfl;
      T zer = eta_right - hfr;
      T s_hu = (hfl + hfr) * (zel - zer) * (0.5 * GRAVITY) / m.dx;

      T sey = ws.slopes_y[i * nc + j].eta;
      T eta_top = (i > 0) ? eta_cell - sey * (0.5 * m.dy) : eta_cell;
      T eta_bot = (i + 1 < nr) ? eta_cell + sey * (0.5 * m.dy) : eta_cell;
      T hft = rmax(eta_top - ws.z_face_y[i * nc + j], zero<T>());
      T hfb = rmax(eta_bot - ws.z_face_y[(i + 1) * nc + j], zero<T>());
      T zet = eta_top - hft;
      T zeb = eta_bot - hfb;
      T s_hv = (hft + hfb) * (zet - zeb) * (0.5 * GRAVITY) / m.dy;

      if (value(new_h) <= H_VEL) {
        cell = {rmax(new_h, zero<T>()), zero<T>(), zero<T>()};
      } else {
        T hu_old = zero<T>(), hv_old = zero<T>();
        if (value(h_old) > H_VEL) {
          hu_old = cell.hu;
          hv_old = cell.hv;
        }
        cell.h = new_h;
        cell.hu = hu_old - dhu + s_hu * dt;
        cell.hv = hv_old - dhv + s_hv * dt;
      }
    }
}

// update.rs:1255-1287
template <class T>
void ssprk2_step_with(std::vector<Cell<T>>& st, const Mesh<T>& m, const Bcs& bcs, double dt,
                      Workspace<T>& ws) {
  ws.u_n = st;
  forward_euler_step_with(st, m, bcs, dt, ws);
  forward_euler_step_with(st, m, bcs, dt, ws);
  const int64_t n = (int64_t)st.size();
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    Cell<T>& s = st[k];
    const Cell<T>& prev = ws.u_n[k];
    s.h = (prev.h + s.h) * 0.5;
    s.hu = (prev.hu + s.hu) * 0.5;
    s.hv = (prev.hv + s.hv) * 0.5;
    if (value(s.h) <= H_VEL) {
      s.h = rmax(s.h, zero<T>());
      s.hu = zero<T>();
      s.hv = zero<T>();
    }
  }
}

// ---------------------------------------------------------------- sources
// source.rs:130-165
template <class T>
void manning_friction_step(std::vector<Cell<T>>& st, const Mesh<T>& m, double dt, double dry_tol) {
  const double dt_g = dt * GRAVITY;
  const int64_t n = (int64_t)st.size();
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    Cell<T>& s = st[k];
    const T& nn = m.manning[k];
    if (value(s.h) <= dry_tol) continue;
    if (value(nn) == 0.0) continue;
    T speed = rsqrt_(s.hu * s.hu + s.hv * s.hv) / s.h;
    T factor = nn * nn * speed / rpowf(s.h, 4.0 / 3.0) * dt_g + 1.0;
    s.hu = s.hu / factor;
    s.hv = s.hv / factor;
  }
}

// source.rs:88-101 (Dual states: extension, constant forcing)
template <class T> void apply_rain(std::vector<Cell<T>>& st, double rate, double dt) {
  if (rate == 0.0) return;
  const double dh = rate * dt;
  const int64_t n = (int64_t)st.size();
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) {
    Cell<T>& c = st[k];
    T new_h = c.h + dh;
    if (value(new_h) <= 0.0) {
      c = dry_cell<T>();
    } else {
      c.h = new_h;
    }
  }
}

struct PointSource {  // source.rs:36-44
  int64_t row, col;
  double q_mass;
};

// source.rs:52-69 (Dual states: extension, constant forcing)
template <class T>
void apply_point_sources(std::vector<Cell<T>>& st, int64_t nc, const std::vector<PointSource>& srcs,
                         double dt, double dx, double dy) {
  const double cell_area = dx * dy;
  for (const PointSource& s : srcs) {
    double dh = s.q_mass * dt / cell_area;
    Cell<T>& c = st[s.row * nc + s.col];
    c.h = rmax(c.h + dh, zero<T>());
    if (value(c.h) == 0.0) {
      c.hu = zero<T>();
      c.hv = zero<T>();
    }
  }
}

// ---------------------------------------------------------------- driver
// sim.rs:36-246, plus the caller-side forcing order of the reference's
// operational loops (uk_ea_test8a_official.rs:206-217): dt -> SSP-RK2 ->
// friction -> rain(rate, dt) -> point sources(dt). Rain and point sources
// are the caller's in the reference; here they are optional config fields.
enum Status {
  ST_OK = 0,
  ST_SHAPE_MISMATCH = 1,
  ST_INVALID_CONFIG = 2,
  ST_DEGENERATE_DT = 3,
  ST_STEP_BUDGET = 4,
};

struct Config {
  double cfl = 0.4;
  int integrator = 1;       // 0 forward Euler, 1 SSP-RK2
  int cfl_with_bcs = 1;     // Simulation::step uses cfl_time_step_with_bcs
  double max_dt = 60.0;
  int friction = 1;
  double friction_dry_tol = H_DRY;
  int64_t max_steps = 10000000;
  double rain_rate = 0.0;
};

// A piecewise-linear forcing series of the reference's operational loops:
// interp of solver-2d/examples/uk_ea_common/mod.rs (clamped at both ends,
// partition_point, frac = (t - t0) / (t1 - t0), v0 + frac * (v1 - v0)) on
// t / time_div, then (v * scale_mul) / scale_div as uk_ea_test4_official.rs
// (discharge_at, / inlet cells) and uk_ea_test8a_official.rs (interp_min,
// rain_rate_m_s, point_discharge_m3_s) scale it.
struct Series {
  std::vector<double> t, v;
  double time_div = 1.0, scale_mul = 1.0, scale_div = 1.0;
  bool on() const { return !t.empty(); }
  double at(double t_s) const {
    const double x = t_s / time_div;
    double val;
    if (x <= t.front()) {
      val = v.front();
    } else if (x >= t.back()) {
      val = v.back();
    } else {
      const size_t idx = (size_t)(std::partition_point(t.begin(), t.end(), [&](double ti) { return ti <= x; }) -
                                  t.begin());
      const double t0 = t[idx - 1], t1 = t[idx];
      const double v0 = v[idx - 1], v1 = v[idx];
      const double frac = (x - t0) / (t1 - t0);
      val = v0 + frac * (v1 - v0);
    }
    return (val * scale_mul) / scale_div;
  }
};

template <class T> struct Sim {
  Mesh<T> mesh;
  std::vector<Cell<T>> st;
  Bcs bcs;
  Config cfg;
  std::vector<PointSource> srcs;
  // the operational loops' time-varying forcing at t_new (rain rate, and
  // the rate of every point source) and their control-point records
  Series rain_series, src_series;
  std::vector<int64_t> gauges;
  std::vector<double> gauge_rows;  // rows of (t, h at each gauge)
  Workspace<T> ws;
  double t = 0.0;
  int64_t steps = 0;

  void init(int64_t nr, int64_t nc, double dx, double dy) {
    mesh.nr = nr;
    mesh.nc = nc;
    mesh.dx = dx;
    mesh.dy = dy;
    mesh.bed.assign((size_t)(nr * nc), zero<T>());
    mesh.manning.assign((size_t)(nr * nc), zero<T>());
    st.assign((size_t)(nr * nc), dry_cell<T>());
    ws.resize(nr, nc);
    bcs.north.kind = bcs.south.kind = bcs.west.kind = bcs.east.kind = BC_WALL;
  }

  // sim.rs:204-228 (+ forcing)
  int step(double dt_cap, double* dt_taken) {
    double dt = cfg.cfl_with_bcs ? cfl_time_step_with_bcs(st, mesh, bcs, cfg.cfl)
                                 : cfl_time_step(st, mesh, cfg.cfl);
    dt = std::fmin(std::fmin(dt, cfg.max_dt), dt_cap);
    if (!(std::isfinite(dt) && dt > 0.0)) {
      if (dt_taken) *dt_taken = dt;
      return ST_DEGENERATE_DT;
    }
    if (cfg.integrator == 0)
      forward_euler_step_with(st, mesh, bcs, dt, ws);
    else
      ssprk2_step_with(st, mesh, bcs, dt, ws);
    if (cfg.friction) manning_friction_step(st, mesh, dt, cfg.friction_dry_tol);
    const double t_new = t + dt;  // sim.time() after Simulation::step
    apply_rain(st, rain_series.on() ? rain_series.at(t_new) : cfg.rain_rate, dt);
    if (src_series.on()) {
      std::vector<PointSource> now = srcs;
      const double q = src_series.at(t_new);
      for (PointSource& p : now) p.q_mass = q;
      apply_point_sources(st, mesh.nc, now, dt, mesh.dx, mesh.dy);
    } else {
      apply_point_sources(st, mesh.nc, srcs, dt, mesh.dx, mesh.dy);
    }
    t += dt;
    steps += 1;
    if (dt_taken) *dt_taken = dt;
    record_gauges();
    return ST_OK;
  }

  void record_gauges() {
    if (gauges.empty()) return;
    gauge_rows.push_back(t);
    for (int64_t c : gauges) gauge_rows.push_back(value(st[(size_t)c].h));
  }

  // sim.rs:231-246
  int run_until(double t_target) {
    int64_t budget = cfg.max_steps;
    while (t < t_target) {
      if (budget == 0) return ST_STEP_BUDGET;
      int rc = step(t_target - t, nullptr);
      if (rc) return rc;
      budget -= 1;
    }
    return ST_OK;
  }

  int run_steps(int64_t n) {
    for (int64_t k = 0; k < n; ++k) {
      int rc = step(std::numeric_limits<double>::infinity(), nullptr);
      if (rc) return rc;
    }
    return ST_OK;
  }
};

}  // namespace hfo
