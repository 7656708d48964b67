// Scalar types for the stage kernels: double, a forward-mode dual, and
// float for the FP32 precision form.
//
// The dual follows the reference's rules exactly (autograd/src/dual.rs:12-299,
// real.rs:115-149): quotient rule with (a'b - ab')/(b*b), sqrt(0) has tangent
// 0, max/min ties average the tangents, powf(v, n) has tangent
// n * v^(n-1) * v'. Every branch decision is taken on the value alone.
//
// Each value type T has a scalar type R = ScalarOf<T>::type (double for
// double and dual, float for float); constants, thresholds, dt and the grid
// spacings enter T arithmetic as R. For double and dual this is the same
// IEEE double arithmetic as writing the literals directly.
//
// Shared by the device code (hipcc) and the CPU build of the same tile code
// (g++), so the double and dual paths use only IEEE operations and the
// deterministic powers of hf_pow.h.
#pragma once
#include <math.h>

#include "hf_pow.h"

namespace hf {

struct dual {
  double v, d;
};

template <class T> struct ScalarOf {
  using type = double;
};
template <> struct ScalarOf<float> {
  using type = float;
};

HF_HD double val(double x) { return x; }
HF_HD double val(dual x) { return x.v; }
HF_HD float val(float x) { return x; }

HF_HD double rabs(double x) { return __builtin_fabs(x); }
HF_HD float rabs(float x) { return __builtin_fabsf(x); }

// cst<T>(x): the constant x as a T (tangents 0).
template <class T> struct CstOf;
template <> struct CstOf<double> {
  static HF_HD double make(double x) { return x; }
};
template <> struct CstOf<dual> {
  static HF_HD dual make(double x) { return dual{x, 0.0}; }
};
template <> struct CstOf<float> {
  static HF_HD float make(double x) { return (float)x; }
};
template <class T> HF_HD T cst(double x) { return CstOf<T>::make(x); }

HF_HD dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
HF_HD dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
HF_HD dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
HF_HD dual operator/(dual a, dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
HF_HD dual operator-(dual a) { return {-a.v, -a.d}; }
HF_HD dual operator+(dual a, double b) { return {a.v + b, a.d}; }
HF_HD dual operator-(dual a, double b) { return {a.v - b, a.d}; }
HF_HD dual operator*(dual a, double b) { return {a.v * b, a.d * b}; }
HF_HD dual operator/(dual a, double b) { return {a.v / b, a.d / b}; }

// Division. double and dual: IEEE division, as the reference. float (the
// precision form): reciprocal then multiply, ~1 ulp, a fraction of the
// correctly rounded sequence's cost.
#if defined(HF_FAST_F64) && defined(__HIP_DEVICE_COMPILE__)
// The fast FP64 kind (its own translation unit, built with FMA
// contraction): 1/b from v_rcp_f64 (~2^-26) and two Newton steps (~1 ulp),
// then a multiply. The reciprocal is a pure function of b, so divisions by
// one depth share it (hu/h, hv/h). Not correctly rounded; b = 0 gives NaN.
HF_HD double rcp_fast(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  return __builtin_fma(r, e, r);
}
HF_HD double div(double a, double b) { return a * rcp_fast(b); }
#else
HF_HD double div(double a, double b) { return a / b; }
#endif
HF_HD dual div(dual a, dual b) { return a / b; }
HF_HD dual div(dual a, double b) { return a / b; }
#if defined(__HIP_DEVICE_COMPILE__) && !defined(HF_F32_IEEE)
HF_HD float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
HF_HD float fsqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
#else
HF_HD float div(float a, float b) { return a / b; }
HF_HD float fsqrt_(float a) { return sqrtf(a); }
#endif

// f64::max / f64::min ignore a NaN operand; IEEE maxNum/minNum do the same.
HF_HD double rmax(double a, double b) { return __builtin_fmax(a, b); }
HF_HD double rmin(double a, double b) { return __builtin_fmin(a, b); }
HF_HD float rmax(float a, float b) { return __builtin_fmaxf(a, b); }
HF_HD float rmin(float a, float b) { return __builtin_fminf(a, b); }
HF_HD dual rmax(dual a, dual b) {
  if (a.v > b.v) return a;
  if (b.v > a.v) return b;
  return {a.v, 0.5 * (a.d + b.d)};
}
HF_HD dual rmin(dual a, dual b) {
  if (a.v < b.v) return a;
  if (b.v < a.v) return b;
  return {a.v, 0.5 * (a.d + b.d)};
}

#if defined(HF_FAST_F64) && defined(__HIP_DEVICE_COMPILE__)
// Fast FP64 kind: sqrt from v_rsq_f64 and two Newton steps on (sqrt x,
// 1 / (2 sqrt x)) (~1 ulp), for x > 0; 0 stays 0.
HF_HD double sqrt_fast(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  double g = x * r, h = r * 0.5;
  const double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  const double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return x > 0.0 ? g : x;
}
HF_HD double rsqrt_(double a) { return sqrt_fast(a); }
#else
HF_HD double rsqrt_(double a) { return __builtin_sqrt(a); }
#endif

// sqrt of x >= 2^-767 (or +inf): the device's correctly rounded sqrt
// sequence without its range scaling. The compiler's IEEE sqrt scales
// x < 2^-767 up before v_rsq_f64, refines by these Newton steps, scales
// back and passes 0 and +inf through; for x >= 2^-767 the scalings are by
// 2^0 and the steps below are the same operations on the same operands, so
// the result is the same bits (hf2d_math_probe checks it on the device).
// Host builds use the library sqrt, which is correctly rounded too.
HF_HD double sqrt_ge_tiny(double x) {
#if defined(HF_FAST_F64) && defined(__HIP_DEVICE_COMPILE__)
  return sqrt_fast(x);
#elif defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rsq(x);
  double g = x * r, h = r * 0.5;
  const double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return x < __builtin_inf() ? g : x;
#else
  return __builtin_sqrt(x);
#endif
}
HF_HD float rsqrt_(float a) { return fsqrt_(a); }
HF_HD dual rsqrt_(dual a) {
  double s = __builtin_sqrt(a.v);
  double d = (s == 0.0) ? 0.0 : a.d / (2.0 * s);
  return {s, d};
}

// sqrt of a value known to be at least 2^-767 (a wet depth times g: h >
// 1e-12): the same result as rsqrt_, without the small-argument scaling.
HF_HD double rsqrt_wet(double a) { return sqrt_ge_tiny(a); }
HF_HD float rsqrt_wet(float a) { return fsqrt_(a); }
HF_HD dual rsqrt_wet(dual a) {
  double s = sqrt_ge_tiny(a.v);
  return {s, a.d / (2.0 * s)};
}

// x.powf(4.0 / 3.0) (source.rs:161) with the deterministic power routine.
#if defined(HF_FAST_F64) && defined(__HIP_DEVICE_COMPILE__)
// Fast FP64 kind: x^(4/3) = x cbrt(x) for x > 0, cbrt from an FP32
// exp2(log2(x) / 3) seed (~2^-22) and two Newton steps in FP64
// z <- (2 z + x / z^2) / 3.
HF_HD double pow43(double x) {
  double z = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (1.0f / 3.0f));
  z = (2.0 * z + x * rcp_fast(z * z)) * (1.0 / 3.0);
  z = (2.0 * z + x * rcp_fast(z * z)) * (1.0 / 3.0);
  return x * z;
}
#else
HF_HD double pow43(double x) { return hfm::pow_4_3(x); }
#endif
HF_HD dual pow43(dual x) {
  const double n = 4.0 / 3.0;
  return {hfm::pow_4_3(x.v), n * hfm::pow_1_3m(x.v) * x.d};
}
HF_HD float pow43(float x) { return x * cbrtf(x); }

// x.powf(3.0 / 5.0) for the Discharge-on-dry ghost depth (boundary.rs:212),
// the only other power on the path, with the deterministic routines.
HF_HD double pow35(double x) { return hfm::pow_3_5(x); }
HF_HD dual pow35(dual x) {
  const double n = 3.0 / 5.0;
  return {hfm::pow_3_5(x.v), n * hfm::pow_m2_5(x.v) * x.d};
}
HF_HD float pow35(float x) { return powf(x, 0.6f); }

// ------------------------------------------------------------ DualN<N>
//
// The reference's vector-mode dual (autograd/src/dual_n.rs, Real impl at
// real.rs:161-195): a value and N tangents. Its rules are not Dual's:
// every binary operation is combine(a, b, val, ca, cb) with tangents
// ca * a_i + cb * b_i, every unary one chain(val, dfdx) with a_i * dfdx;
// division takes inv = 1 / b and the value a * inv; sqrt's factor is
// 0.5 / s (0 at s == 0); max/min ties average as 0.5 * a_i + 0.5 * b_i;
// an f64 operand is promoted to a constant (tangents 0) first
// (dual_n.rs scalar_ops). Branches read the value alone.
template <int N> struct dualn {
  double v;
  double d[N];
};

template <int N> HF_HD dualn<N> dn_const(double x) {
  dualn<N> r;
  r.v = x;
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N> HF_HD dualn<N> dn_combine(const dualn<N>& a, const dualn<N>& b, double val, double ca, double cb) {
  dualn<N> r;
  r.v = val;
  for (int i = 0; i < N; ++i) r.d[i] = ca * a.d[i] + cb * b.d[i];
  return r;
}
template <int N> HF_HD dualn<N> dn_chain(const dualn<N>& a, double val, double dfdx) {
  dualn<N> r;
  r.v = val;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * dfdx;
  return r;
}

template <int N> struct CstOf<dualn<N>> {
  static HF_HD dualn<N> make(double x) { return dn_const<N>(x); }
};

template <int N> HF_HD double val(const dualn<N>& x) { return x.v; }
template <int N> HF_HD dualn<N> operator+(const dualn<N>& a, const dualn<N>& b) {
  return dn_combine(a, b, a.v + b.v, 1.0, 1.0);
}
template <int N> HF_HD dualn<N> operator-(const dualn<N>& a, const dualn<N>& b) {
  return dn_combine(a, b, a.v - b.v, 1.0, -1.0);
}
template <int N> HF_HD dualn<N> operator*(const dualn<N>& a, const dualn<N>& b) {
  return dn_combine(a, b, a.v * b.v, b.v, a.v);
}
template <int N> HF_HD dualn<N> operator/(const dualn<N>& a, const dualn<N>& b) {
  const double inv = 1.0 / b.v;
  return dn_combine(a, b, a.v * inv, inv, -a.v * inv * inv);
}
template <int N> HF_HD dualn<N> operator-(const dualn<N>& a) { return dn_chain(a, -a.v, -1.0); }
template <int N> HF_HD dualn<N> operator+(const dualn<N>& a, double b) { return a + dn_const<N>(b); }
template <int N> HF_HD dualn<N> operator-(const dualn<N>& a, double b) { return a - dn_const<N>(b); }
template <int N> HF_HD dualn<N> operator*(const dualn<N>& a, double b) { return a * dn_const<N>(b); }
template <int N> HF_HD dualn<N> operator/(const dualn<N>& a, double b) { return a / dn_const<N>(b); }
template <int N> HF_HD dualn<N> div(const dualn<N>& a, const dualn<N>& b) { return a / b; }
template <int N> HF_HD dualn<N> div(const dualn<N>& a, double b) { return a / dn_const<N>(b); }
template <int N> HF_HD dualn<N> rmax(const dualn<N>& a, const dualn<N>& b) {
  if (a.v > b.v) return a;
  if (b.v > a.v) return b;
  return dn_combine(a, b, a.v, 0.5, 0.5);
}
template <int N> HF_HD dualn<N> rmin(const dualn<N>& a, const dualn<N>& b) {
  if (a.v < b.v) return a;
  if (b.v < a.v) return b;
  return dn_combine(a, b, a.v, 0.5, 0.5);
}
template <int N> HF_HD dualn<N> rsqrt_(const dualn<N>& a) {
  const double s = __builtin_sqrt(a.v);
  return dn_chain(a, s, s == 0.0 ? 0.0 : 0.5 / s);
}
template <int N> HF_HD dualn<N> rsqrt_wet(const dualn<N>& a) {
  const double s = sqrt_ge_tiny(a.v);
  return dn_chain(a, s, 0.5 / s);
}
// powf(4/3) and powf(3/5) with the deterministic powers: chain(v^n, n v^(n-1))
template <int N> HF_HD dualn<N> pow43(const dualn<N>& x) {
  const double n = 4.0 / 3.0;
  return dn_chain(x, hfm::pow_4_3(x.v), n * hfm::pow_1_3m(x.v));
}
template <int N> HF_HD dualn<N> pow35(const dualn<N>& x) {
  const double n = 3.0 / 5.0;
  return dn_chain(x, hfm::pow_3_5(x.v), n * hfm::pow_m2_5(x.v));
}

}  // namespace hf
