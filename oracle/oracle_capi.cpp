// C entry points over the CPU restatement (hf_oracle.hpp), for the tests'
// ctypes checker and bench.py's CPU-baseline leg. TEST INFRASTRUCTURE ONLY:
// the product library never links or loads this.
//
// The functions mirror the product C ABI (include/hydroflux_amd.h) so one
// test body can drive both: AoS (h, hu, hv) triples in row-major order, and
// for the dual kinds the tangents innermost: state [cell][field][k], bed and
// Manning [cell][k] (k < 1 for Dual, k < N for DualN<N>).
#include <omp.h>

#include <cstdint>
#include <cstring>

#include "hf_oracle.hpp"

using namespace hfo;

namespace {
// kind: 0 f64, 1 Dual, N (2, 4, 8, 16) DualN<N>
struct Handle {
  int kind = 0;
  Sim<double> f;
  Sim<Dual> d;
  Sim<DualN<2>> d2;
  Sim<DualN<4>> d4;
  Sim<DualN<8>> d8;
  Sim<DualN<16>> d16;
};

template <class F> auto visit(Handle* h, F&& fn) {
  switch (h->kind) {
    case 1: return fn(h->d);
    case 2: return fn(h->d2);
    case 4: return fn(h->d4);
    case 8: return fn(h->d8);
    case 16: return fn(h->d16);
    default: return fn(h->f);
  }
}

// Scalars from / to a value and its tangents (t: the k tangents, nullable).
template <class T> struct Tan;
template <> struct Tan<double> {
  static constexpr int N = 0;
  static double make(double v, const double*) { return v; }
  static double val(double x) { return x; }
  static void tangents(double, double*) {}
};
template <> struct Tan<Dual> {
  static constexpr int N = 1;
  static Dual make(double v, const double* t) { return Dual{v, t ? t[0] : 0.0}; }
  static double val(const Dual& x) { return x.val; }
  static void tangents(const Dual& x, double* t) { t[0] = x.dval; }
};
template <int M> struct Tan<DualN<M>> {
  static constexpr int N = M;
  static DualN<M> make(double v, const double* t) {
    DualN<M> r;
    r.val = v;
    for (int i = 0; i < M; ++i) r.dval[i] = t ? t[i] : 0.0;
    return r;
  }
  static double val(const DualN<M>& x) { return x.val; }
  static void tangents(const DualN<M>& x, double* t) {
    for (int i = 0; i < M; ++i) t[i] = x.dval[i];
  }
};

template <class T> void set_gauges_t(Sim<T>& s, int n, const int64_t* rows, const int64_t* cols) {
  s.gauges.clear();
  s.gauge_rows.clear();
  for (int k = 0; k < n; ++k) s.gauges.push_back(rows[k] * s.mesh.nc + cols[k]);
  s.record_gauges();
}

Bc make_bc(int kind, double p) {
  Bc b;
  b.kind = kind;
  b.param = p;
  return b;
}

std::vector<PointSource> make_srcs(const int64_t* rows, const int64_t* cols, const double* q, int64_t k) {
  std::vector<PointSource> v;
  for (int64_t s = 0; s < k; ++s) v.push_back(PointSource{rows[s], cols[s], q[s]});
  return v;
}
}  // namespace

extern "C" {

void* hfo_create(int64_t nr, int64_t nc, double dx, double dy, int kind) {
  Handle* h = new Handle();
  h->kind = kind;
  visit(h, [&](auto& s) { s.init(nr, nc, dx, dy); });
  return h;
}

void hfo_destroy(void* p) { delete static_cast<Handle*>(p); }

void hfo_set_pow_mode(int mode) { pow_mode() = mode; }
void hfo_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
void hfo_set_dualn_div_mode(int mode) { dualn_div_mode() = mode; }

// bed, manning: n cells each; tangents nullable (dual kinds only).
void hfo_set_mesh(void* p, const double* bed, const double* bed_tan, const double* manning,
                  const double* manning_tan) {
  visit(static_cast<Handle*>(p), [&](auto& s) {
    using T = typename std::decay_t<decltype(s.st)>::value_type;
    using S = decltype(T{}.h);
    constexpr int N = Tan<S>::N;
    const size_t n = s.mesh.bed.size();
    for (size_t k = 0; k < n; ++k) {
      s.mesh.bed[k] = Tan<S>::make(bed[k], bed_tan ? bed_tan + k * N : nullptr);
      s.mesh.manning[k] = Tan<S>::make(manning[k], manning_tan ? manning_tan + k * N : nullptr);
    }
  });
}

void hfo_set_state(void* p, const double* aos, const double* aos_tan) {
  visit(static_cast<Handle*>(p), [&](auto& s) {
    using T = typename std::decay_t<decltype(s.st)>::value_type;
    using S = decltype(T{}.h);
    constexpr int N = Tan<S>::N;
    const size_t n = s.st.size();
    for (size_t k = 0; k < n; ++k) {
      auto tan = [&](int f) { return aos_tan ? aos_tan + (3 * k + f) * N : nullptr; };
      s.st[k].h = Tan<S>::make(aos[3 * k], tan(0));
      s.st[k].hu = Tan<S>::make(aos[3 * k + 1], tan(1));
      s.st[k].hv = Tan<S>::make(aos[3 * k + 2], tan(2));
    }
  });
}

void hfo_get_state(void* p, double* aos, double* aos_tan) {
  visit(static_cast<Handle*>(p), [&](auto& s) {
    using T = typename std::decay_t<decltype(s.st)>::value_type;
    using S = decltype(T{}.h);
    constexpr int N = Tan<S>::N;
    const size_t n = s.st.size();
    for (size_t k = 0; k < n; ++k) {
      aos[3 * k] = Tan<S>::val(s.st[k].h);
      aos[3 * k + 1] = Tan<S>::val(s.st[k].hu);
      aos[3 * k + 2] = Tan<S>::val(s.st[k].hv);
      if (aos_tan && N > 0) {
        Tan<S>::tangents(s.st[k].h, aos_tan + (3 * k) * N);
        Tan<S>::tangents(s.st[k].hu, aos_tan + (3 * k + 1) * N);
        Tan<S>::tangents(s.st[k].hv, aos_tan + (3 * k + 2) * N);
      }
    }
  });
}

// kinds/params ordered N, S, W, E (boundary.rs:131-140).
void hfo_set_boundaries(void* p, const int* kinds, const double* params) {
  Bcs b;
  b.north = make_bc(kinds[0], params[0]);
  b.south = make_bc(kinds[1], params[1]);
  b.west = make_bc(kinds[2], params[2]);
  b.east = make_bc(kinds[3], params[3]);
  visit(static_cast<Handle*>(p), [&](auto& s) { s.bcs = b; });
}

void hfo_max_wave_speeds(void* p, double* sx, double* sy) {
  visit(static_cast<Handle*>(p), [&](auto& s) { max_wave_speeds(s.st, sx, sy); });
}

double hfo_cfl_time_step(void* p, double cfl, int with_bcs) {
  return visit(static_cast<Handle*>(p), [&](auto& s) {
    return with_bcs ? cfl_time_step_with_bcs(s.st, s.mesh, s.bcs, cfl) : cfl_time_step(s.st, s.mesh, cfl);
  });
}

void hfo_forward_euler_step(void* p, double dt) {
  visit(static_cast<Handle*>(p), [&](auto& s) { forward_euler_step_with(s.st, s.mesh, s.bcs, dt, s.ws); });
}

void hfo_ssprk2_step(void* p, double dt) {
  visit(static_cast<Handle*>(p), [&](auto& s) { ssprk2_step_with(s.st, s.mesh, s.bcs, dt, s.ws); });
}

void hfo_manning_friction_step(void* p, double dt, double dry_tol) {
  visit(static_cast<Handle*>(p), [&](auto& s) { manning_friction_step(s.st, s.mesh, dt, dry_tol); });
}

void hfo_apply_rain(void* p, double rate, double dt) {
  visit(static_cast<Handle*>(p), [&](auto& s) { apply_rain(s.st, rate, dt); });
}

void hfo_apply_point_sources(void* p, const int64_t* rows, const int64_t* cols, const double* q, int64_t k,
                             double dt) {
  std::vector<PointSource> v = make_srcs(rows, cols, q, k);
  visit(static_cast<Handle*>(p),
        [&](auto& s) { apply_point_sources(s.st, s.mesh.nc, v, dt, s.mesh.dx, s.mesh.dy); });
}

// Driver configuration (sim.rs:49-77 plus the forcing extension).
int hfo_configure(void* p, double cfl, int integrator, int cfl_with_bcs, double max_dt, int friction,
                  double friction_dry_tol, int64_t max_steps, double rain_rate, const int64_t* rows,
                  const int64_t* cols, const double* q, int64_t k) {
  if (!(cfl > 0.0 && cfl <= 1.0)) return ST_INVALID_CONFIG;  // sim.rs:141-146
  if (!(max_dt > 0.0)) return ST_INVALID_CONFIG;             // sim.rs:147-152
  Config c;
  c.cfl = cfl;
  c.integrator = integrator;
  c.cfl_with_bcs = cfl_with_bcs;
  c.max_dt = max_dt;
  c.friction = friction;
  c.friction_dry_tol = friction_dry_tol;
  c.max_steps = max_steps;
  c.rain_rate = rain_rate;
  std::vector<PointSource> v = make_srcs(rows, cols, q, k);
  visit(static_cast<Handle*>(p), [&](auto& s) {
    s.cfg = c;
    s.srcs = v;
  });
  return ST_OK;
}

int hfo_step(void* p, double dt_cap, double* dt_taken) {
  return visit(static_cast<Handle*>(p), [&](auto& s) { return s.step(dt_cap, dt_taken); });
}

int hfo_run_until(void* p, double t_target) {
  return visit(static_cast<Handle*>(p), [&](auto& s) { return s.run_until(t_target); });
}

int hfo_run_steps(void* p, int64_t n) {
  return visit(static_cast<Handle*>(p), [&](auto& s) { return s.run_steps(n); });
}

// Driver forcing series (which: 0 rain, 1 point sources; n = 0 removes).
void hfo_set_series(void* p, int which, int n, const double* t, const double* v, double time_div,
                    double scale_mul, double scale_div) {
  Series sr;
  sr.t.assign(t, t + n);
  sr.v.assign(v, v + n);
  sr.time_div = time_div;
  sr.scale_mul = scale_mul;
  sr.scale_div = scale_div;
  visit(static_cast<Handle*>(p), [&](auto& s) { (which == 0 ? s.rain_series : s.src_series) = sr; });
}

// Gauges: record (t, h) at the cells after every driver step; the current
// state is the first row.
void hfo_set_gauges(void* p, int n, const int64_t* rows, const int64_t* cols) {
  visit(static_cast<Handle*>(p), [&](auto& s) { set_gauges_t(s, n, rows, cols); });
}
int64_t hfo_gauge_values(void* p, double* out, int64_t max_values) {
  return visit(static_cast<Handle*>(p), [&](auto& s) {
    const int64_t n = (int64_t)s.gauge_rows.size();
    if (out) std::memcpy(out, s.gauge_rows.data(), (size_t)(n < max_values ? n : max_values) * sizeof(double));
    return n;
  });
}

void hfo_time(void* p, double* t, int64_t* steps) {
  visit(static_cast<Handle*>(p), [&](auto& s) {
    *t = s.t;
    *steps = s.steps;
  });
}

void hfo_set_time(void* p, double t, int64_t steps) {
  visit(static_cast<Handle*>(p), [&](auto& s) {
    s.t = t;
    s.steps = steps;
  });
}

// Correctly rounded powers shared with the kernels, exposed for tests.
double hfo_pow_4_3(double x) { return hfm::pow_4_3(x); }
double hfo_pow_1_3m(double x) { return hfm::pow_1_3m(x); }
double hfo_pow_3_5(double x) { return hfm::pow_3_5(x); }
double hfo_pow_m2_5(double x) { return hfm::pow_m2_5(x); }
double hfo_libm_pow(double x, double n) { return std::pow(x, n); }

// The reference's DualN unit-test function (autograd/src/dual_n.rs tests,
// components_match_independent_scalar_passes) in DualN<3> and in three
// seeded Dual passes: out = [vec.val, vec.dval[0..3], scalar val k, scalar dval k for k < 3].
void hfo_dualn_probe(double av, double bv, double cv, double* out) {
  using D3 = DualN<3>;
  auto seeded = [](double v, int k) {
    D3 r = dn_constant<3>(v);
    r.dval[k] = 1.0;
    return r;
  };
  auto absn = [](const D3& x) {
    const double sign = x.val > 0.0 ? 1.0 : (x.val < 0.0 ? -1.0 : 0.0);
    return dn_chain(x, std::fabs(x.val), sign);
  };
  D3 a = seeded(av, 0), b = seeded(bv, 1), c = seeded(cv, 2);
  D3 wave = rsqrt_(a * 9.81);
  D3 fric = (b * b) / rpowf(a, 4.0 / 3.0);
  D3 vec = rmax(wave * fric + absn(c), dn_constant<3>(1e-9)) - c / (a + 1.0);
  out[0] = vec.val;
  for (int i = 0; i < 3; ++i) out[1 + i] = vec.dval[i];
  for (int k = 0; k < 3; ++k) {
    Dual a1{av, k == 0 ? 1.0 : 0.0}, b1{bv, k == 1 ? 1.0 : 0.0}, c1{cv, k == 2 ? 1.0 : 0.0};
    // Dual abs (dual.rs:88-100): tangent sign * dval, 0 at 0
    const double sgn = c1.val > 0.0 ? 1.0 : (c1.val < 0.0 ? -1.0 : 0.0);
    Dual absc{std::fabs(c1.val), sgn * c1.dval};
    Dual w1 = rsqrt_(a1 * 9.81);
    Dual f1 = (b1 * b1) / rpowf(a1, 4.0 / 3.0);
    Dual sc = rmax(w1 * f1 + absc, Dual{1e-9, 0.0}) - c1 / (a1 + 1.0);
    out[4 + 2 * k] = sc.val;
    out[5 + 2 * k] = sc.dval;
  }
}
}
