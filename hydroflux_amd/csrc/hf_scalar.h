// Scalar types for the stage kernels: double and a forward-mode dual.
//
// The dual follows the reference's rules exactly (autograd/src/dual.rs:12-299,
// real.rs:115-149): quotient rule with (a'b - ab')/(b*b), sqrt(0) has tangent
// 0, max/min ties average the tangents, powf(v, n) has tangent
// n * v^(n-1) * v'. Every branch decision is taken on the value alone.
//
// Shared by the device code (hipcc) and the CPU build of the same tile code
// (g++), so it uses only IEEE operations and the deterministic powers of
// hf_pow.h.
#pragma once
#include <math.h>

#include "hf_pow.h"

namespace hf {

struct dual {
  double v, d;
};

HF_HD double val(double x) { return x; }
HF_HD double val(dual x) { return x.v; }

template <class T> HF_HD T cst(double x);
template <> HF_HD double cst<double>(double x) { return x; }
template <> HF_HD dual cst<dual>(double x) { return dual{x, 0.0}; }

HF_HD dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
HF_HD dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
HF_HD dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
HF_HD dual operator/(dual a, dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
HF_HD dual operator-(dual a) { return {-a.v, -a.d}; }
HF_HD dual operator+(dual a, double b) { return {a.v + b, a.d}; }
HF_HD dual operator-(dual a, double b) { return {a.v - b, a.d}; }
HF_HD dual operator*(dual a, double b) { return {a.v * b, a.d * b}; }
HF_HD dual operator/(dual a, double b) { return {a.v / b, a.d / b}; }

// f64::max / f64::min ignore a NaN operand; IEEE maxNum/minNum do the same.
HF_HD double rmax(double a, double b) { return __builtin_fmax(a, b); }
HF_HD double rmin(double a, double b) { return __builtin_fmin(a, b); }
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

HF_HD double rsqrt_(double a) { return __builtin_sqrt(a); }
HF_HD dual rsqrt_(dual a) {
  double s = __builtin_sqrt(a.v);
  double d = (s == 0.0) ? 0.0 : a.d / (2.0 * s);
  return {s, d};
}

// x.powf(4.0 / 3.0) (source.rs:161) with the deterministic power routine.
HF_HD double pow43(double x) { return hfm::pow_4_3(x); }
HF_HD dual pow43(dual x) {
  const double n = 4.0 / 3.0;
  return {hfm::pow_4_3(x.v), n * hfm::pow_1_3m(x.v) * x.d};
}

// General x.powf(n) for the Discharge-on-dry ghost depth (boundary.rs:212),
// the only other power on the path.
HF_HD double powf_gen(double x, double n) { return pow(x, n); }
HF_HD dual powf_gen(dual x, double n) { return {pow(x.v, n), n * pow(x.v, n - 1.0) * x.d}; }

}  // namespace hf
