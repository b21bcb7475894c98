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

#include "sdeh_powtab.h"

namespace sdeh_math {

SDEH_HD inline uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
SDEH_HD inline double double_of(uint64_t b) { return __builtin_bit_cast(double, b); }

// Julia Base `min`/`max` on IEEE floats: NaN in either argument gives NaN, and
// -0.0 orders below +0.0.
// Written as a chain of selects (in order of precedence from the last one up:
// NaN x, NaN y, equal values, the comparison) so that device code evaluates it
// without branches.
template <class T>
SDEH_HD inline T jl_min(T x, T y) {
  T r = y < x ? y : x;
  r = x == y ? (__builtin_signbit(x) ? x : y) : r;
  r = y != y ? y : r;
  return x != x ? x : r;
}
template <class T>
SDEH_HD inline T jl_max(T x, T y) {
  T r = y > x ? y : x;
  r = x == y ? (__builtin_signbit(x) ? y : x) : r;
  r = y != y ? y : r;
  return x != x ? x : r;
}
// jl_max / jl_min for arguments that are never negative (absolute values): the
// -0.0/+0.0 rule cannot apply, so a comparison select plus NaN propagation gives
// the same result (x's NaN first, as above).
template <class T>
SDEH_HD inline T jl_max_nonneg(T x, T y) {
  T r = x < y ? y : x;
  r = y != y ? y : r;
  return x != x ? x : r;
}
template <class T>
SDEH_HD inline T jl_min_nonneg(T x, T y) {
  T r = y < x ? y : x;
  r = y != y ? y : r;
  return x != x ? x : r;
}
// Julia `@fastmath` min/max (Base.FastMath.min_fast/max_fast): a single
// comparison select; the result with a NaN argument is not specified by the
// reference.  Callers reach these only with non-NaN arguments.
template <class T>
SDEH_HD inline T fast_min(T x, T y) { return y < x ? y : x; }
template <class T>
SDEH_HD inline T fast_max(T x, T y) { return y > x ? y : x; }

// Natural log of a finite x > 0 (normal or subnormal) as an unevaluated sum
// hi + lo, |error| < 2^-66 |log x| + 2^-90 (table-driven, Tang's construction):
//   x = 2^k z, z in [0x1.6ap-1, 0x1.6ap0), the subinterval index i taken from
//   the top 7 bits of the bit pattern of z relative to 0x1.6ap-1;
//   r = z * invc_i - 1 is exact (invc_i has 8 significant bits, |r| < 2^-7);
//   log x = k ln2 + lc_i + log1p(r), lc_i = -log(invc_i) = lh_i + ll_i;
//   log1p(r) = r - r^2/2 + r^3 P(r), P the Taylor series 1/3 - r/4 + ... - r^6/9
//   (truncation below |r|^10/10 < 2^-70).
// The large terms are added with error-free transformations (Knuth's two-sum;
// the fast form where the first operand dominates or is zero): k ln2_hi is exact
// (ln2_hi has 42 significant bits, |k| < 2^11).
// log_norm_dd is the core for a normal x scaled by 2^-kadj (log of x * 2^kadj);
// log_pos_dd scales subnormal inputs into the normal range first.
SDEH_HD inline void log_norm_dd(double xs, int kadj, double* hi, double* lo) {
  using namespace sdeh_powtab;
  const uint64_t ix = bits_of(xs);
  const uint64_t tmp = ix - LOG_OFF;
  const int i = (int)((tmp >> LOG_SHIFT) & (uint64_t)LOG_MASK);
  const int k = (int)((int64_t)tmp >> 52) - kadj;
  const double z = double_of(ix - (tmp & 0xfff0000000000000ULL));
  const double invc = LOG_TAB[i][0], lh = LOG_TAB[i][1], ll = LOG_TAB[i][2];
  const double r = fma(z, invc, -1.0);  // exact
  const double dk = (double)k;
  const double a = dk * LN2_HI;  // exact
  // s1 + e1 = a + lh (|a| >= ln 2 > |lh| unless k == 0)
  const double s1 = a + lh;
  const double e1 = (a - s1) + lh;
  // s2 + e2 = s1 + r (two-sum)
  const double s2 = s1 + r;
  const double bb = s2 - s1;
  const double e2 = (s1 - (s2 - bb)) + (r - bb);
  // -r^2/2 = ar2 + e3 exactly
  const double ar = -0.5 * r;
  const double ar2 = r * ar;
  const double e3 = fma(ar, r, -ar2);
  // h + e4 = s2 + ar2 (|s2| >= |ar2|: s2 = r when k = 0 and invc = 1,
  // otherwise |s2| > 2^-9 > r^2/2)
  const double h = s2 + ar2;
  const double e4 = (s2 - h) + ar2;
  double p = -0x1.c71c71c71c71cp-4;      // -1/9
  p = fma(p, r, 0x1p-3);                 //  1/8
  p = fma(p, r, -0x1.2492492492492p-3);  // -1/7
  p = fma(p, r, 0x1.5555555555555p-3);   //  1/6
  p = fma(p, r, -0x1.999999999999ap-3);  // -1/5
  p = fma(p, r, 0x1p-2);                 //  1/4
  p = fma(p, r, -0x1.5555555555555p-2);  // -1/3
  const double r3 = (r * r) * r;
  const double t = fma(-r3, p, fma(dk, LN2_LO, ll)) + ((e1 + e2) + (e3 + e4));
  const double y = h + t;
  *hi = y;
  *lo = (h - y) + t;
}

SDEH_HD inline void log_pos_dd(double x, double* hi, double* lo) {
  const bool sub = x < 0x1p-1022;
  log_norm_dd(sub ? x * 0x1p54 : x, sub ? 54 : 0, hi, lo);
}

// e^(zh + zl) for |zl| <= 2^-40 |zh| + 2^-60, zh in [-745.2, 709.78] (outside:
// 0 or +Inf); table-driven (Tang):  n = round(zh * 128/ln2) = 128 m + j,
// u = (zh - n ln2/128) + zl with the first difference exact (Sterbenz),
// e^u - 1 = u + u^2 Q(u), Q the Taylor series 1/2 + u/6 + ... + u^4/720
// (|u| < 2^-8.4, truncation below 2^-70); e^z = 2^m (th_j + (th_j (e^u - 1) + tl_j)).
// exp_core returns res and m with e^(zh + zl) = res * 2^m; exp_dd applies 2^m as
// two exact power-of-two factors so that a subnormal result is rounded by the
// last multiply only.
SDEH_HD inline double exp_core(double zc, double zl, int* m_out) {
  using namespace sdeh_powtab;
  const double nd = __builtin_rint(zc * INV_LN2N);
  const int n = (int)nd;
  const double r0 = fma(-nd, LN2N_HI, zc);  // exact
  const double u = r0 + fma(-nd, LN2N_LO, zl);
  double q = 0x1.6c16c16c16c17p-10;   // 1/720
  q = fma(q, u, 0x1.1111111111111p-7);  // 1/120
  q = fma(q, u, 0x1.5555555555555p-5);  // 1/24
  q = fma(q, u, 0x1.5555555555555p-3);  // 1/6
  q = fma(q, u, 0.5);
  const double em1 = fma(u * u, q, u);
  const int j = n & EXP_MASK;
  const int m = n >> EXP_BITS;  // arithmetic shift: floor(n / 128)
  const double th = EXP_TAB[j][0], tl = EXP_TAB[j][1];
  *m_out = m;
  return th + fma(th, em1, tl);
}

SDEH_HD inline double exp_dd(double zh, double zl) {
  const bool over = zh > 709.782712893384;
  const bool under = zh < -745.2;
  int m;
  const double res = exp_core((over || under) ? 0.0 : zh, zl, &m);
  // 2^m = 2^m1 * 2^m2 with both halves normal (m in [-1075, 1024]).
  const int m1 = m / 2;
  const int m2 = m - m1;
  const double s2a = double_of((uint64_t)(m1 + 1023) << 52);
  const double s2b = double_of((uint64_t)(m2 + 1023) << 52);
  const double v = (res * s2a) * s2b;
  return over ? __builtin_inf() : (under ? 0.0 : v);
}

// x^y for the controller: x >= 0 (an RMS error norm), 0 < y <= 1 a fixed
// exponent, split as log then exp so one logarithm can serve several exponents
// (the controller raises the same error estimate to beta1 now and to beta2 for
// the next step).  pow_from(pow_log(x), y) == pow_ctrl(x, y) bit for bit.
// Error below 0.52 ulp for results in the normal range.  Special values: NaN x
// gives x, 0 gives 0, +Inf gives +Inf, negative x gives NaN, 1 gives 1 (the
// table path returns exactly 1 there).  Every path is evaluated with selects.
struct PowLog {
  double hi, lo;
  double special;  // result for every y when is_special
  bool is_special;
};

SDEH_HD inline PowLog pow_log(double x) {
  const bool nan_or_neg = !(x >= 0.0);
  const bool zero = x == 0.0;
  const bool inf = x == __builtin_inf();
  PowLog r;
  r.is_special = nan_or_neg || zero || inf;
  r.special = zero ? 0.0 : ((x != x || inf) ? x : __builtin_nan(""));
  log_pos_dd(r.is_special ? 1.0 : x, &r.hi, &r.lo);
  return r;
}

SDEH_HD inline double pow_from(const PowLog& l, double y) {
  const double ph = y * l.hi;
  const double pl = fma(y, l.hi, -ph) + y * l.lo;
  const double v = exp_dd(ph, pl);
  return l.is_special ? l.special : v;
}

SDEH_HD inline double pow_ctrl(double x, double y) { return pow_from(pow_log(x), y); }

// The same pow on the controller's own arguments, with less work: x is an RMS
// error estimate or qold (0, +Inf, NaN, or a normal number: the square root of a
// binary64 is never subnormal) and 0 < y <= 0.9, so |y log x| < 700 and x^y is a
// normal number.  No subnormal input scaling, no range saturation, and 2^m applied
// as one factor (exact for a normal result).  On that domain the results are
// pow_ctrl's bit for bit.
SDEH_HD inline PowLog pow_log_ctrl(double x) {
  const bool nan_or_neg = !(x >= 0.0);
  const bool zero = x == 0.0;
  const bool inf = x == __builtin_inf();
  PowLog r;
  r.is_special = nan_or_neg || zero || inf;
  r.special = zero ? 0.0 : ((x != x || inf) ? x : __builtin_nan(""));
  log_norm_dd(r.is_special ? 1.0 : x, 0, &r.hi, &r.lo);
  return r;
}

SDEH_HD inline double pow_from_ctrl(const PowLog& l, double y) {
  const double ph = y * l.hi;
  const double pl = fma(y, l.hi, -ph) + y * l.lo;
  int m;
  const double res = exp_core(ph, pl, &m);
  const double v = res * double_of((uint64_t)(m + 1023) << 52);
  return l.is_special ? l.special : v;
}

}  // namespace sdeh_math
