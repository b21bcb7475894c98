// Fixed-step time grid: Julia's `tspan[1]:dt:tspan[2]` for Float64
// (the grid the fixed-step solvers iterate over, gpuatsit5.jl:92,
// gpuvern7.jl:100, gpuvern9.jl:100).
//
// Julia Base (base/twiceprecision.jl, `(:)(start::T, step::T, stop::T) where
// T<:IEEEFloat`) first looks for exact rationals start = a/d1, step = b/d2,
// stop = c/d3 with denominators found by a continued fraction (Base.rat,
// bounded by 2^24); if all three round-trip and the common denominator keeps
// start*den and step*den exactly representable, the length is
// (den*stop_n - stop_d*start_n + step_n*stop_d) / (step_n*stop_d) (integer
// division) and element i is the rational (start_n + i*step_n)/den to 2x
// precision; otherwise it takes start and step literally,
// len = round((stop-start)/step) + 1 less an overshoot correction, and
// element i is start + i*step to 2x precision.  Here both element forms are
// produced as the correctly rounded binary64 value (integer quotient of exact
// integers, or an exact two-product/two-sum).
#pragma once

#include <math.h>
#include <stdint.h>
#include <stdlib.h>

namespace sdeh_range {

struct Grid {
  int rational = 0;
  int64_t start_n = 0, step_n = 0, den = 1;
  double start = 0.0, step = 0.0;
  int64_t len = 0;
};

inline void rat(double x, int64_t* num, int64_t* den) {
  double y = x;
  int64_t a = 1, d = 1, b = 0, c = 0;
  const double m = 16777216.0;  // maxintfloat(Float32)
  while (fabs(y) <= m) {
    const int64_t f = (int64_t)trunc(y);
    y -= (double)f;
    int64_t t = f * a + c;
    c = a;
    a = t;
    t = f * b + d;
    d = b;
    b = t;
    const int64_t mx = llabs(a) > llabs(b) ? llabs(a) : llabs(b);
    if (mx > (int64_t)m) {
      *num = c;
      *den = d;
      return;
    }
    if ((double)a / (double)b == x) break;
    y = 1.0 / y;
  }
  *num = a;
  *den = b;
}

inline int64_t gcd64(int64_t a, int64_t b) {
  a = llabs(a);
  b = llabs(b);
  while (b != 0) {
    const int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

inline bool between(double a, double x, double b) {
  return (a <= x && x <= b) || (b <= x && x <= a);
}

inline Grid make(double start, double step, double stop) {
  Grid g;
  g.start = start;
  g.step = step;
  int64_t sn, sd;
  rat(step, &sn, &sd);
  if (sd != 0 && (double)sn / (double)sd == step) {
    int64_t an, ad, cn, cd;
    rat(start, &an, &ad);
    rat(stop, &cn, &cd);
    if (ad != 0 && cd != 0 && (double)an / (double)ad == start && (double)cn / (double)cd == stop) {
      const int64_t den = ad / gcd64(ad, sd) * sd;
      const double m = 9007199254740992.0;  // maxintfloat(Float64)
      if (den != 0 && fabs(start * (double)den) <= m && fabs(step * (double)den) <= m &&
          den % ad == 0 && den % sd == 0) {
        const int64_t a0 = (int64_t)nearbyint(start * (double)den);
        const int64_t s0 = (int64_t)nearbyint(step * (double)den);
        int64_t len = (den * cn - cd * a0 + s0 * cd) / (s0 * cd);
        if (len < 0) len = 0;
        if (between(start, start + (double)(len - 1) * step, stop + step / 2) &&
            !between(start, start + (double)len * step, stop)) {
          g.rational = 1;
          g.start_n = a0;
          g.step_n = s0;
          g.den = den;
          g.len = len;
          return g;
        }
      }
    }
  }
  const double lf = (stop - start) / step;
  int64_t len;
  if (lf < 0) {
    len = 0;
  } else if (lf == 0) {
    len = 1;
  } else {
    len = (int64_t)nearbyint(lf) + 1;
    const double stop2 = start + (double)(len - 1) * step;
    len -= (int64_t)((start < stop && stop < stop2) + (start > stop && stop > stop2));
  }
  g.len = len;
  return g;
}

inline double elem(const Grid& g, int64_t i) {
  if (g.rational) return (double)(g.start_n + i * g.step_n) / (double)g.den;
  const double di = (double)i;
  const double hi = di * g.step;
  const double lo = fma(di, g.step, -hi);
  const double s = g.start + hi;
  const double bb = s - g.start;
  const double e = (g.start - (s - bb)) + (hi - bb);
  return s + (e + lo);
}

}  // namespace sdeh_range
