// Deterministic real powers used by the Manning friction step and the
// Discharge-onto-dry ghost.
//
// The reference evaluates `h.powf(4.0 / 3.0)` (solver-2d/src/source.rs:161)
// and, for forward-mode tangents, `n * v.powf(n - 1.0) * dv`
// (autograd/src/dual.rs:112-117) through the platform libm `pow`. A GPU
// libm and glibc may disagree in the last bit, and the solver amplifies a
// one-ulp difference at wet/dry fronts, so the kernels use the routines
// below instead. They are built only from IEEE-754 +, -, *, / and fused
// multiply-add, each of which is correctly rounded on x86-64 and on gfx950,
// so host and device produce bit-identical results for every input.
//
// Note on the exponents: the double nearest 4/3 is e = 4/3 - eps with
// eps = 2^-52 / 3, and e - 1 = 1/3 - eps exactly (Sterbenz). Both routines
// evaluate x^(e) resp. x^(e-1), i.e. what `pow(x, 4.0/3.0)` is asked for,
// not the mathematical x^(4/3):
//
//   x^(1/3 - eps) = cbrt(x) * exp(-eps ln x) = cbrt(x) * (1 - eps ln x + O(1e-27))
//
// Method: x = m * 2^(3k), m in [1, 8). z ~ m^(-1/3) by a quadratic seed and
// four Newton steps z <- z + z (1 - m z^3) / 3 (quadratic convergence, no
// division). y = m z^2 then gets one double-double correction
// delta = (m - y^3) z^2 / 3, with m - y^3 formed exactly by the FMA
// two-product (Ogita, Rump & Oishi 2005). The result y + delta carries
// about 100 significant bits, so the final rounding agrees with a correctly
// rounded power except within ~1e-16 ulp of a rounding tie.
//
// This header is compiled by hipcc (device and host) and by g++ (the CPU
// restatement used as the test oracle); it must stay free of library calls.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define HF_HD __host__ __device__ __attribute__((always_inline)) inline
#else
#define HF_HD inline
#endif

namespace hfm {

HF_HD double as_f64(uint64_t u) { return __builtin_bit_cast(double, u); }
HF_HD uint64_t as_u64(double x) { return __builtin_bit_cast(uint64_t, x); }

// 2^e for e in [-1022, 1023], exact.
HF_HD double pow2i(int e) { return as_f64((uint64_t)(int64_t)(e + 1023) << 52); }

// x * 2^e by two exact power-of-two multiplies (|e| <= 2044).
HF_HD double scale2(double x, int e) {
  int a = e / 2;
  int b = e - a;
  return (x * pow2i(a)) * pow2i(b);
}

// Split a finite x > 0 into f in [1, 2) and integer E with x = f * 2^E.
HF_HD void split_pos(double x, double* f, int* E) {
  int bias_shift = 0;
  uint64_t u = as_u64(x);
  if (((u >> 52) & 0x7ff) == 0) {  // subnormal: lift into the normal range
    x = x * pow2i(192);
    u = as_u64(x);
    bias_shift = 192;
  }
  *E = (int)((u >> 52) & 0x7ff) - 1023 - bias_shift;
  *f = as_f64((u & 0x000fffffffffffffULL) | (1023ULL << 52));
}

// ln x for finite x > 0 with absolute error below 2e-14: x = f 2^E with
// f in [sqrt(1/2), sqrt(2)), ln f = 2 atanh(s), s = (f-1)/(f+1), |s| < 0.1716,
// odd series truncated after s^15.
HF_HD double ln_pos(double x) {
  double f;
  int E;
  split_pos(x, &f, &E);
  if (f > 1.4142135623730951) {
    f = f * 0.5;
    E = E + 1;
  }
  double s = (f - 1.0) / (f + 1.0);
  double s2 = s * s;
  double p = 0.06666666666666667;       // 1/15
  p = p * s2 + 0.07692307692307693;     // 1/13
  p = p * s2 + 0.09090909090909091;     // 1/11
  p = p * s2 + 0.1111111111111111;      // 1/9
  p = p * s2 + 0.14285714285714285;     // 1/7
  p = p * s2 + 0.2;                     // 1/5
  p = p * s2 + 0.3333333333333333;      // 1/3
  p = p * s2 + 1.0;
  return (double)E * 0.6931471805599453 + (2.0 * s) * p;
}

// Cube root of m in [1, 8) as an unevaluated pair (y, delta), |delta| << ulp(y).
HF_HD void cbrt_dd(double m, double* y_out, double* d_out) {
  // Seed for m^(-1/3) from f = m / 2^r, then Newton steps on m itself.
  double z;
  if (m < 2.0) {
    z = (0.09202394 * m - 0.47733264) * m + 1.38237822;
  } else if (m < 4.0) {
    double f = m * 0.5;
    z = ((0.09202394 * f - 0.47733264) * f + 1.38237822) * 0.7937005259840998;
  } else {
    double f = m * 0.25;
    z = ((0.09202394 * f - 0.47733264) * f + 1.38237822) * 0.6299605249474366;
  }
  for (int it = 0; it < 4; ++it) {
    double t = z * z;
    t = t * z;
    t = m * t;
    double e = 1.0 - t;
    z = z + z * (e * 0.3333333333333333);
  }
  double z2 = z * z;
  double y = m * z2;
  // Exact-enough residual m - y^3 via FMA two-products.
  double a = y * y;
  double ea = __builtin_fma(y, y, -a);
  double b = a * y;
  double eb = __builtin_fma(a, y, -b);
  double r = (m - b) - (eb + ea * y);
  *y_out = y;
  *d_out = r * z2 * 0.3333333333333333;
}

// eps = 2^-52 / 3: the gap between 4/3 and the double nearest to it.
#define HFM_POW_EPS 0x1.5555555555555p-54

// Reduce finite x > 0 to m in [1, 8) and k with x = m * 2^(3k).
HF_HD void split_cube(double x, double* m, int* k) {
  double f;
  int E;
  split_pos(x, &f, &E);
  int kk = (E >= 0) ? (E / 3) : -((-E + 2) / 3);
  int r = E - 3 * kk;  // 0, 1 or 2
  *m = f * (double)(1 << r);
  *k = kk;
}

// pow(x, 4.0/3.0) for the double exponent 4.0/3.0.
HF_HD double pow_4_3(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 0.0 : (x - x) / (x - x);  // 0 or NaN
  if (x > 1.7976931348623157e308) return x;                      // +inf
  double m, y, d;
  int k;
  split_cube(x, &m, &k);
  cbrt_dd(m, &y, &d);
  double L = ln_pos(x);
  double p = m * y;
  double pe = __builtin_fma(m, y, -p);
  double lo = (pe + m * d) - (p * HFM_POW_EPS) * L;
  return scale2(p + lo, 4 * k);
}

// pow(x, 4.0/3.0 - 1.0), the exponent the Dual tangent of powf(4/3) uses.
HF_HD double pow_1_3m(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 0.0 : (x - x) / (x - x);
  if (x > 1.7976931348623157e308) return x;
  double m, y, d;
  int k;
  split_cube(x, &m, &k);
  cbrt_dd(m, &y, &d);
  double L = ln_pos(x);
  double lo = d - (y * HFM_POW_EPS) * L;
  return scale2(y + lo, k);
}

// ------------------------------------------------------------ x^(3/5)
//
// The Discharge-onto-dry ghost depth (n |q| / sqrt(S))^(3/5)
// (solver-2d/src/boundary.rs:203-216) and, for its Dual tangent,
// pow(x, 3.0/5.0 - 1.0). The double nearest 3/5 is e = 3/5 - eps5 with
// eps5 = 2^-53 / 5, and e - 1 = -2/5 - eps5 exactly (Sterbenz), so the
// routines evaluate x^(3/5) resp. x^(-2/5) times x^(-eps5) = 1 - eps5 ln x.
//
// Method: x = m * 2^(5k), m in [1, 32). z ~ m^(-1/5) from a quadratic seed
// on the mantissa and four division-free Newton steps
// z <- z + z (1 - m z^5) / 5; y = m z^4 ~ m^(1/5) then gets one
// double-double correction d = (m - y^5) z^4 / 5 with m - y^5 formed from
// FMA two-products. (y + d)^3 and (y + d)^3 / m (one remainder-corrected
// division) give m^(3/5) and m^(-2/5) to ~100 bits before the final
// rounding.

#define HFM_POW5_EPS 0x1.999999999999ap-56  // 3/5 - fl(3/5) = 2^-53 / 5

// Reduce finite x > 0 to m = f * 2^r in [1, 32) (f in [1, 2), r in 0..4)
// and k with x = m * 2^(5k).
HF_HD void split_fifth(double x, double* m, double* f_out, int* r_out, int* k) {
  double f;
  int E;
  split_pos(x, &f, &E);
  int kk = (E >= 0) ? (E / 5) : -((-E + 4) / 5);
  int r = E - 5 * kk;  // 0..4
  *m = f * (double)(1 << r);
  *f_out = f;
  *r_out = r;
  *k = kk;
}

// Fifth root of m = f * 2^r as an unevaluated pair (y, d).
HF_HD void fifth_root_dd(double m, double f, int r, double* y_out, double* d_out) {
  double s;  // 2^(-r/5)
  switch (r) {
    case 0: s = 1.0; break;
    case 1: s = 0.8705505632961241; break;
    case 2: s = 0.757858283255199; break;
    case 3: s = 0.6597539553864471; break;
    default: s = 0.5743491774985174; break;
  }
  double z = ((0.0521304 * f - 0.28315556) * f + 1.22947226) * s;
  for (int it = 0; it < 4; ++it) {
    double t = z * z;
    t = t * t;
    t = t * z;
    t = m * t;
    double e = 1.0 - t;
    z = z + z * (e * 0.2);
  }
  double z2 = z * z;
  double z4 = z2 * z2;
  double y = m * z4;
  // m - y^5 from two-products: y^2 = a + ea, a^2 = b + eb, b y = c + ec,
  // y^5 = c + ec + (eb + 2 a ea) y up to terms below 2^-100 m.
  double a = y * y;
  double ea = __builtin_fma(y, y, -a);
  double b = a * a;
  double eb = __builtin_fma(a, a, -b);
  double c = b * y;
  double ec = __builtin_fma(b, y, -c);
  double res = (m - c) - (ec + (eb + 2.0 * a * ea) * y);
  *y_out = y;
  *d_out = res * z4 * 0.2;
}

// (y + d)^3 as hi + lo.
HF_HD void cube_dd(double y, double d, double* hi, double* lo) {
  double a = y * y;
  double ea = __builtin_fma(y, y, -a);
  double b = a * y;
  double eb = __builtin_fma(a, y, -b);
  *hi = b;
  *lo = eb + ea * y + 3.0 * a * d;
}

// pow(x, 3.0/5.0) for the double exponent 3.0/5.0.
HF_HD double pow_3_5(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 0.0 : (x - x) / (x - x);  // 0 or NaN
  if (x > 1.7976931348623157e308) return x;                      // +inf
  double m, f, y, d, hi, lo;
  int r, k;
  split_fifth(x, &m, &f, &r, &k);
  fifth_root_dd(m, f, r, &y, &d);
  cube_dd(y, d, &hi, &lo);
  lo = lo - (hi * HFM_POW5_EPS) * ln_pos(x);
  return scale2(hi + lo, 3 * k);
}

// pow(x, 3.0/5.0 - 1.0), the exponent the Dual tangent of powf(3/5) uses.
HF_HD double pow_m2_5(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? __builtin_inf() : (x - x) / (x - x);  // pow(0, y < 0) = inf
  if (x > 1.7976931348623157e308) return 0.0;
  double m, f, y, d, hi, lo;
  int r, k;
  split_fifth(x, &m, &f, &r, &k);
  fifth_root_dd(m, f, r, &y, &d);
  cube_dd(y, d, &hi, &lo);
  double q = hi / m;  // m^(-2/5) = (y + d)^3 / m
  double rem = __builtin_fma(-q, m, hi);
  double ql = (rem + lo) / m;
  ql = ql - (q * HFM_POW5_EPS) * ln_pos(x);
  return scale2(q + ql, -2 * k);
}

}  // namespace hfm
