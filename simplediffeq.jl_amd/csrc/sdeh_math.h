// Scalar math shared by the HIP integrators and by host code.
//
// The adaptive controller of the reference raises the error estimate to the
// powers beta1 = 7/50 and beta2 = 2/25 with `@fastmath ^`
// (src/tsit5/gpuatsit5.jl:279,283; src/verner/gpuvern7.jl:406,410;
// src/verner/gpuvern9.jl:575,579), i.e. whatever libm pow the host links.
// Device libraries and host libm do not round pow identically, so the
// integrators use this pow instead: a plain log/exp composition built only from
// IEEE-754 operations (+, -, *, /, fma, round-to-nearest-even, bit moves) that
// every conforming target rounds identically.  The error is below 0.6 ulp over
// the controller's range (so it equals a correctly rounded pow in all but
// near-halfway cases); what it buys is that a CPU build and a gfx950 build of
// the same integrator produce bit-identical step sequences.
//
// Also here: Julia's NaN-propagating `min`/`max` for IEEE floats, which the
// reference uses outside `@fastmath` (gpuatsit5.jl:287,290,293), and the
// `@fastmath` variants used at gpuatsit5.jl:289.
//
// Compile with floating-point contraction disabled (-ffp-contract=off): every
// fused multiply-add below is written as an explicit fma().
#pragma once

#if defined(__HIP__)
#define SDEH_HD __host__ __device__
#else
#define SDEH_HD
#endif

#if !defined(__HIPCC_RTC__)
#include <math.h>
#include <stdint.h>
#else
// hipRTC declares the fixed-width integer types in __hip_internal only.
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace sdeh_math {

SDEH_HD inline uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
SDEH_HD inline double double_of(uint64_t b) { return __builtin_bit_cast(double, b); }

// Julia Base `min`/`max` on IEEE floats: NaN in either argument gives NaN, and
// -0.0 orders below +0.0.
template <class T>
SDEH_HD inline T jl_min(T x, T y) {
  if (x != x) return x;
  if (y != y) return y;
  if (x == y) return __builtin_signbit(x) ? x : y;
  return y < x ? y : x;
}
template <class T>
SDEH_HD inline T jl_max(T x, T y) {
  if (x != x) return x;
  if (y != y) return y;
  if (x == y) return __builtin_signbit(x) ? y : x;
  return y > x ? y : x;
}
// Julia `@fastmath` min/max (Base.FastMath.min_fast/max_fast): a single
// comparison select; the result with a NaN argument is not specified by the
// reference.  Callers reach these only with non-NaN arguments.
template <class T>
SDEH_HD inline T fast_min(T x, T y) { return y < x ? y : x; }
template <class T>
SDEH_HD inline T fast_max(T x, T y) { return y > x ? y : x; }

// Natural log of a finite x > 0 (subnormals included) as an unevaluated sum
// hi + lo.  x = 2^k * m with m in [sqrt(1/2), sqrt(2)); f = m - 1 (exact);
// s = f / (2 + f); log(1 + f) = 2 atanh(s) = f + c with
// c = s (f^2/2 + R(s^2)) - f^2/2 and R(z) = sum_{j>=1} 2 z^j / (2j + 1)
// (Taylor series, 11 terms, |z| <= 0.0295).  k ln 2 + f is summed exactly.
SDEH_HD inline void log_pos_dd(double x, double* hi, double* lo) {
  int k = 0;
  if (x < 0x1p-1022) {  // subnormal: scale into the normal range
    x *= 0x1p54;
    k = -54;
  }
  uint64_t b = bits_of(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = double_of((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 0x1.6a09e667f3bcdp+0) {  // sqrt(2)
    m *= 0.5;
    e += 1;
  }
  k += e;
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double r = 0x1.642c8590b2164p-4;          // 2/23
  r = fma(r, z, 0x1.8618618618618p-4);      // 2/21
  r = fma(r, z, 0x1.af286bca1af28p-4);      // 2/19
  r = fma(r, z, 0x1.e1e1e1e1e1e1ep-4);      // 2/17
  r = fma(r, z, 0x1.1111111111111p-3);      // 2/15
  r = fma(r, z, 0x1.3b13b13b13b14p-3);      // 2/13
  r = fma(r, z, 0x1.745d1745d1746p-3);      // 2/11
  r = fma(r, z, 0x1.c71c71c71c71cp-3);      // 2/9
  r = fma(r, z, 0x1.2492492492492p-2);      // 2/7
  r = fma(r, z, 0x1.999999999999ap-2);      // 2/5
  r = fma(r, z, 0x1.5555555555555p-1);      // 2/3
  r = r * z;
  const double hfsq = 0.5 * f * f;
  const double c = fma(s, hfsq + r, -hfsq);
  const double dk = (double)k;
  const double ln2_hi = 0x1.62e4200000000p-1;  // ln 2 with the low 32 bits zero
  const double ln2_lo = 0x1.fdf473de6af28p-22;
  const double a = dk * ln2_hi;  // exact: |k| < 2^11, ln2_hi has 21 significant bits
  const double sum = a + f;      // two-sum: a + f == sum + err exactly
  const double bb = sum - a;
  const double err = (a - (sum - bb)) + (f - bb);
  const double tail = err + fma(dk, ln2_lo, c);
  const double h = sum + tail;
  *hi = h;
  *lo = tail - (h - sum);
}

// e^(zh + zl) for finite zh, |zl| <= ulp(zh).  n = round(zh / ln 2);
// r = zh - n ln 2 (exact for the high part of ln 2); e^r by its Taylor
// polynomial of degree 13 (|r| <= 0.347), with 1 + r summed exactly; the
// result is scaled by 2^n through the exponent bits in two factors, so a
// subnormal result is formed by one final rounding multiply.
SDEH_HD inline double exp_dd(double zh, double zl) {
  if (zh > 709.782712893384) return __builtin_inf();
  if (zh < -745.2) return 0.0;
  const double n = __builtin_rint(zh * 0x1.71547652b82fep+0);  // 1/ln 2
  const double r0 = fma(-n, 0x1.62e4200000000p-1, zh);
  const double rl0 = fma(-n, 0x1.fdf473de6af28p-22, zl);
  const double rh = r0 + rl0;
  const double rl = (r0 - rh) + rl0;
  double q = 0x1.6124613a86d09p-33;         // 1/13!
  q = fma(q, rh, 0x1.1eed8eff8d898p-29);    // 1/12!
  q = fma(q, rh, 0x1.ae64567f544e4p-26);    // 1/11!
  q = fma(q, rh, 0x1.27e4fb7789f5cp-22);    // 1/10!
  q = fma(q, rh, 0x1.71de3a556c734p-19);    // 1/9!
  q = fma(q, rh, 0x1.a01a01a01a01ap-16);    // 1/8!
  q = fma(q, rh, 0x1.a01a01a01a01ap-13);    // 1/7!
  q = fma(q, rh, 0x1.6c16c16c16c17p-10);    // 1/6!
  q = fma(q, rh, 0x1.1111111111111p-7);     // 1/5!
  q = fma(q, rh, 0x1.5555555555555p-5);     // 1/4!
  q = fma(q, rh, 0x1.5555555555555p-3);     // 1/3!
  q = fma(q, rh, 0.5);
  const double t = fma(rh * rh, q, fma(rl, rh, rl));  // e^r - 1 - rh
  const double s1 = 1.0 + rh;                         // fast two-sum, |rh| < 1
  const double e1 = rh - (s1 - 1.0);
  const double p = s1 + (e1 + t);
  const int ni = (int)n;
  // 2^ni = 2^h1 * 2^h2 with both halves normal doubles (ni in [-1075, 1024]).
  const int h1 = ni / 2;
  const int h2 = ni - h1;
  const double s2a = double_of((uint64_t)(h1 + 1023) << 52);
  const double s2b = double_of((uint64_t)(h2 + 1023) << 52);
  return (p * s2a) * s2b;
}

// x^y for the controller: x >= 0 (an RMS error norm), y > 0 a fixed exponent,
// split as log then exp so one logarithm can serve several exponents (the
// controller raises the same error estimate to beta1 now and to beta2 for the
// next step).  pow_from(pow_log(x), y) == pow_ctrl(x, y) bit for bit.
// Special values: NaN x gives NaN, 0 gives 0, +Inf gives +Inf, negative x
// gives NaN, 1 gives 1.
struct PowLog {
  double hi, lo;
  double special;  // result for every y when is_special
  bool is_special;
};

SDEH_HD inline PowLog pow_log(double x) {
  PowLog r;
  r.hi = r.lo = 0.0;
  r.is_special = true;
  if (x != x) {
    r.special = x;
  } else if (x == 0.0) {
    r.special = 0.0;
  } else if (x < 0.0) {
    r.special = __builtin_nan("");
  } else if (x == __builtin_inf()) {
    r.special = x;
  } else if (x == 1.0) {
    r.special = 1.0;
  } else {
    r.is_special = false;
    r.special = 0.0;
    log_pos_dd(x, &r.hi, &r.lo);
  }
  return r;
}

SDEH_HD inline double pow_from(const PowLog& l, double y) {
  if (l.is_special) return l.special;
  const double ph = y * l.hi;
  const double pl = fma(y, l.hi, -ph) + y * l.lo;
  return exp_dd(ph, pl);
}

SDEH_HD inline double pow_ctrl(double x, double y) { return pow_from(pow_log(x), y); }

}  // namespace sdeh_math
