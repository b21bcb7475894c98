// Device integrators: the six GPU-compatible solvers of SimpleDiffEq.jl as
// HIP kernels for gfx950, one trajectory per lane.
//
//   GPUSimpleTsit5   src/tsit5/gpuatsit5.jl:55-147    sdeh_fixed<..., ALG_TSIT5>
//   GPUSimpleATsit5  src/tsit5/gpuatsit5.jl:205-336   sdeh_adaptive<..., ALG_ATSIT5>
//   GPUSimpleVern7   src/verner/gpuvern7.jl:55-242    sdeh_fixed<..., ALG_VERN7>
//   GPUSimpleAVern7  src/verner/gpuvern7.jl:300-536   sdeh_adaptive<..., ALG_AVERN7>
//   GPUSimpleVern9   src/verner/gpuvern9.jl:55-353    sdeh_fixed<..., ALG_VERN9>
//   GPUSimpleAVern9  src/verner/gpuvern9.jl:411-779   sdeh_adaptive<..., ALG_AVERN9>
//   GPUSimpleRK4     src/rk4/gpurk4.jl:53-98          sdeh_fixed<..., ALG_RK4>
//   GPUSimpleEuler   src/euler/gpueuler.jl:53-90      sdeh_fixed<..., ALG_EULER>
//
// Execution model.  The reference runs one sequential `while t < tf / while
// EEst > 1` nest per trajectory.  Here a lane owns a trajectory and the nest is
// flattened: every iteration of the kernel loop is one step attempt for every
// live lane, accepted or rejected, so a lane that accepts never waits for a
// neighbour that rejects.  Lanes whose trajectory ended either idle until the
// wave drains (one-launch schedule) or take the next trajectory from a device
// atomic queue (persistent schedule): the wave claims `chunk` queue slots with
// one atomic and hands them to its free lanes in lane order (ballot + mbcnt).
// The queue hands out slots < n only, so every wave reaches its exit once all
// slots are claimed and its lanes have finished; every trajectory ends after
// at most `maxiters` attempts.
//
// Numerics.  State, stages and the controller scalars are in T (double or
// float); time t, tf and saveat are binary64.  Arithmetic follows one fixed
// order, the fused reading of the reference's @muladd bodies:
//   uprev + dt*(a1*k1 + ... + am*km):  acc = a1*k1; acc = fma(aj, kj, acc);
//                                      out = fma(dt, acc, uprev)
//   uprev + dt*a21*k1 (one term):     out = fma(dt*a21, k1, uprev)
//   stage time t + c*dt:               fma(c, dt, t) in binary64
//   abstol + max(|uprev|,|u|)*reltol:  fma(max, reltol, abstol)
//   ODE_DEFAULT_NORM:                  sqrt(((e1*e1 + e2*e2) + ...) / N)
//   bθs polynomials:                   Horner with fma (Base.evalpoly)
// and the library is compiled with -ffp-contract=off, so the compiler adds no
// other fusion.  In binary64 every operation used is correctly rounded on
// gfx950 (fma, +, *, /, sqrt) and the controller's pow is sdeh_math::pow_ctrl,
// so a lane reproduces the CPU restatement (oracle, fused mode) bit for bit.
// The binary32 mode uses hardware reciprocal, square root, log2 and exp2 in the
// controller and error scaling: it is a precision mode judged by its error
// against binary64 results, not by bits.
#pragma once

#include "sdeh_kargs.h"
#include "sdeh_math.h"
#include "sdeh_tableaus.h"

namespace sdeh {

enum { ALG_TSIT5 = 0, ALG_ATSIT5 = 1, ALG_VERN7 = 2, ALG_AVERN7 = 3, ALG_VERN9 = 4,
       ALG_AVERN9 = 5, ALG_RK4 = 6, ALG_EULER = 7 };
enum { SAVE_END = 0, SAVE_AT = 1, SAVE_STEPS = 2 };
enum { RET_DEFAULT = 0, RET_DTMIN = 1, RET_NAN = 2, RET_MAXITERS = 3, RET_SAVE_OVERFLOW = 4 };
// A staged solve's first launch parks trajectories still running after
// stage_cap attempts; the second launch resumes them and overwrites the code.
enum { RET_PARKED = -2 };

template <class T>
struct Id {
  typedef T type;
};

// ---------------------------------------------------------------------------
// Scalar operations by state type.
// ---------------------------------------------------------------------------
template <class T>
struct Ops;

// Ops<T>::Log / pw_from: x^y through one logarithm that several exponents can
// share; pw_from(plog(x), y) == pw(x, y).
template <>
struct Ops<double> {
  typedef sdeh_math::PowLog Log;
  __device__ static inline double div(double a, double b) { return a / b; }
  __device__ static inline double root(double x) { return __builtin_sqrt(x); }
  // The controller's arguments stay in pow_log_ctrl's domain (normal, 0, Inf or
  // NaN; exponents 7/50 and 2/25), where it returns pow_ctrl's bits.
  __device__ static inline double pw(double x, double y) { return sdeh_math::pow_ctrl(x, y); }
  __device__ static inline Log plog(double x) { return sdeh_math::pow_log_ctrl(x); }
  __device__ static inline double pw_from(const Log& l, double y) {
    return sdeh_math::pow_from_ctrl(l, y);
  }

  // Division and square root with the range handling of the IEEE sequences left
  // out, for operands the caller has checked.  gfx950 divides binary64 as
  // v_div_scale (twice), Newton's iteration on v_rcp_f64, v_div_fmas and
  // v_div_fixup; for |a| in [2^-700, 2^700] and |b| in [2^-60, 2^60] (so b and 1/b
  // are normal, exponent(a) - exponent(b) < 768, the biased exponent of a exceeds 53
  // and a/b is normal) both scalings are by 2^0, v_div_fmas is a plain fma and the
  // fix-up returns the quotient unchanged, so div_r below is the same operations on
  // the same values.  Likewise sqrt is a scaling for x < 2^-767,
  // Newton's iteration on v_rsq_f64, the unscaling and a fix-up for 0 and +Inf:
  // for x in [2^-767, 2^1023] only the iteration remains.  Same bits as `/` and
  // `__builtin_sqrt` on that domain.
  static constexpr bool kFast = true;
  __device__ static inline double rcp_nr(double b) {  // b's refined reciprocal
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
  }
  __device__ static inline double div_r(double a, double b, double r) {  // r = rcp_nr(b)
    const double q = a * r;
    const double rem = fma(-b, q, a);
    return fma(rem, r, q);
  }
  __device__ static inline double div_in(double a, double b) { return div_r(a, b, rcp_nr(b)); }
  __device__ static inline double root_in(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
  }
};

template <>
struct Ops<float> {
  typedef float Log;
  static constexpr bool kFast = false;  // binary32 divides through v_rcp_f32 already
  __device__ static inline float rcp_nr(float b) { return b; }
  __device__ static inline float div_r(float a, float b, float) { return div(a, b); }
  __device__ static inline float div_in(float a, float b) { return div(a, b); }
  __device__ static inline float root_in(float x) { return root(x); }
  __device__ static inline float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  __device__ static inline float root(float x) { return __builtin_amdgcn_sqrtf(x); }
  __device__ static inline float pw(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
  }
  __device__ static inline Log plog(float x) { return __builtin_amdgcn_logf(x); }
  __device__ static inline float pw_from(Log l, float y) { return __builtin_amdgcn_exp2f(y * l); }
};

// ---------------------------------------------------------------------------
// Lane layout.  A right-hand side type R declares G, its lanes per trajectory.
// G = 1: a lane holds a whole trajectory (R::N components).  G > 1 (a power of
// two dividing 64): the G adjacent lanes of a group hold one trajectory, lane c
// holding component c (R::N == 1 per lane, G components in all), and R::f
// exchanges the state across the group through LDS.  The stage, error and
// dense-output arithmetic is the same code in both layouts; only loads, stores,
// the error norm and work distribution look at G.
//
// Wave layout (R::WAVE, for right-hand sides evaluated on the matrix cores): the
// G lanes of a trajectory are 64/G apart, so a wave holds 64/G trajectories, lane
// l serving trajectory slot l % (64/G) of the wave at group position q = l / (64/G),
// and lane q holds components q, q + G, q + 2G, ... (R::N of them).  R::f is then a
// wave-wide operation: the integrator calls it in uniform control flow only
// (adaptive_body_wave).
// ---------------------------------------------------------------------------
template <class R, class = int>
struct IsWave {
  static constexpr bool value = false;
};
template <class R>
struct IsWave<R, decltype((void)R::WAVE, 0)> {
  static constexpr bool value = R::WAVE;
};

template <class R>
__device__ __forceinline__ int lane_comp() {
  if constexpr (IsWave<R>::value)
    return (int)((threadIdx.x & 63u) / (unsigned)(64 / R::G));
  else
    return R::G == 1 ? 0 : (int)(threadIdx.x & (unsigned)(R::G - 1));
}

// Component index of this lane's register c.
template <class R>
__device__ __forceinline__ int comp_idx(int c) {
  if constexpr (IsWave<R>::value) return lane_comp<R>() + R::G * c;
  else return lane_comp<R>() * R::N + c;
}

// all[j] = v of lane j of this lane's group (j = 0..G-1), for lanes in uniform
// control flow within their group.  One LDS slot per thread of the workgroup
// (blocks of at most 256 threads); a wave's LDS accesses complete in program
// order, and the wave-scope fences keep the compiler from moving them.
template <class T, int G>
__device__ __forceinline__ void group_gather(T v, T (&all)[G]) {
  __shared__ T xbuf[256];
  const unsigned base = threadIdx.x & ~(unsigned)(G - 1);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  xbuf[threadIdx.x] = v;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int j = 0; j < G; ++j) all[j] = xbuf[base + j];
}

// Wave layout: this wave's share (NW values) of a block's LDS exchange buffer
// (blocks of at most 256 threads).  One buffer per NW; a kernel uses one NW for all
// its exchanges, which never overlap.
template <class T, int NW>
__device__ __forceinline__ T* wave_lds() {
  __shared__ T buf[4 * NW];
  return buf + (threadIdx.x >> 6) * NW;
}

// Wave layout: all[c] = component c (c = 0 .. N*G-1) of this lane's trajectory,
// for a wave in uniform control flow.  LDS [slot][component] in the wave's share,
// slots WAVE_SLOT_STRIDE<T, R> values apart: one more than the components, so that
// the slots of a wave start in different banks (at a stride of 16 binary64 values,
// 128 B, every slot's component c fell in the same bank pair: 8-way conflicts on
// each exchange).
template <class T, class R>
constexpr int wave_slot_stride() {
  return R::N * R::G + 1;
}

template <class T, class R, int NW>
__device__ __forceinline__ void wave_gather(const T (&v)[R::N], T (&all)[R::N * R::G]) {
  constexpr int FULL = R::N * R::G, SLOTS = 64 / R::G, XS = wave_slot_stride<T, R>();
  static_assert(XS * SLOTS <= NW, "one wave's share of the exchange buffer");
  T* w = wave_lds<T, NW>() + (threadIdx.x & (unsigned)(SLOTS - 1)) * XS;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int c = 0; c < R::N; ++c) w[comp_idx<R>(c)] = v[c];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int c = 0; c < FULL; ++c) all[c] = w[c];
}

// ---------------------------------------------------------------------------
// Linear combinations over stage vectors (arrays held in registers).
// ---------------------------------------------------------------------------
template <class T, int N>
__device__ __forceinline__ void acc_more(T (&)[N]) {}

template <class T, int N, class... R>
__device__ __forceinline__ void acc_more(T (&acc)[N], const typename Id<T>::type a,
                                         const T (&k)[N], const R&... rest) {
#pragma unroll
  for (int c = 0; c < N; ++c) acc[c] = fma(a, k[c], acc[c]);
  acc_more(acc, rest...);
}

// acc = a0*k0 + a1*k1 + ... (fma chain, left to right)
template <class T, int N, class... R>
__device__ __forceinline__ void lsum(T (&acc)[N], const typename Id<T>::type a0,
                                     const T (&k0)[N], const R&... rest) {
#pragma unroll
  for (int c = 0; c < N; ++c) acc[c] = a0 * k0[c];
  acc_more(acc, rest...);
}

// out = base + h * (a0*k0 + a1*k1 + ...)
template <class T, int N, class... R>
__device__ __forceinline__ void comb(T (&out)[N], const T (&base)[N],
                                     const typename Id<T>::type h, const R&... terms) {
  T acc[N];
  lsum(acc, terms...);
#pragma unroll
  for (int c = 0; c < N; ++c) out[c] = fma(h, acc[c], base[c]);
}

// out = base + s*k (first stage, s = dt * a21)
template <class T, int N>
__device__ __forceinline__ void comb1(T (&out)[N], const T (&base)[N],
                                      const typename Id<T>::type s, const T (&k)[N]) {
#pragma unroll
  for (int c = 0; c < N; ++c) out[c] = fma(s, k[c], base[c]);
}

// Compensated form of out = base + h*acc for the binary32 precision mode:
// y = h*acc - comp; out = base + y; comp' = (out - base) - y.
template <class T, int N>
__device__ __forceinline__ void kahan_update(T (&out)[N], T (&comp_out)[N], const T (&base)[N],
                                             const T (&comp)[N], T h, const T (&acc)[N]) {
#pragma unroll
  for (int c = 0; c < N; ++c) {
    const T y = fma(h, acc[c], -comp[c]);
    const T s = base[c] + y;
    comp_out[c] = (s - base[c]) - y;
    out[c] = s;
  }
}

template <class T, int N>
__device__ __forceinline__ void copyv(T (&d)[N], const T (&s)[N]) {
#pragma unroll
  for (int c = 0; c < N; ++c) d[c] = s[c];
}

__device__ __forceinline__ double stage_t(double t, double c, double dt) {
  return fma(c, dt, t);
}

// tmp <- tmp / (abstol + max(|uprev|, |u|) * reltol); return the sum of squares of
// ODE_DEFAULT_NORM(tmp) over all components of the trajectory (gathered from the
// group when G > 1; the sum runs over components 0, 1, ... in that order either
// way).  The norm is sqrt(sum / NormLen<R>).
//
// Julia's max(|uprev|, |u|) is the larger value, or NaN if either is NaN.  Here u
// = uprev + (terms), so a NaN in uprev is a NaN in u, and one comparison select
// that keeps |u| unless |u| < |uprev| gives the same value (a NaN only through u,
// and a NaN makes the norm NaN whatever its payload).
// NW: the wave layout's exchange buffer (values per wave; 0 = the gather's own size).
template <class T, class R, int NW = 0>
__device__ __forceinline__ T err_sumsq(T (&tmp)[R::N], const T (&uprev)[R::N],
                                       const T (&u)[R::N], T abstol, T reltol) {
  constexpr int N = R::N;
#pragma unroll
  for (int c = 0; c < N; ++c) {
    const T au = fabs(u[c]), ap = fabs(uprev[c]);
    const T m = au < ap ? ap : au;
    tmp[c] = Ops<T>::div(tmp[c], fma(m, reltol, abstol));
  }
  if constexpr (IsWave<R>::value) {
    constexpr int FULL = N * R::G;
    constexpr int NWX = NW > 0 ? NW : (64 / R::G) * wave_slot_stride<T, R>();
    T all[FULL];
    wave_gather<T, R, NWX>(tmp, all);
    T s = all[0] * all[0];
#pragma unroll
    for (int c = 1; c < FULL; ++c) s = s + all[c] * all[c];
    return s;
  } else if constexpr (R::G > 1) {
    static_assert(N == 1, "a lane of a group holds one component");
    T all[R::G];
    group_gather<T, R::G>(tmp[0], all);
    T s = all[0] * all[0];
#pragma unroll
    for (int c = 1; c < R::G; ++c) s = s + all[c] * all[c];
    return s;
  } else {
    T s = tmp[0] * tmp[0];
#pragma unroll
    for (int c = 1; c < N; ++c) s = s + tmp[c] * tmp[c];
    return s;
  }
}

// err_sumsq of one lane's trajectory with the scaled errors divided by bare Newton
// iterations (Ops<T>::div_in), err left unchanged.  Valid when every scale b =
// m * reltol + abstol is a normal number below 2^1022 (1/b normal too) and at least
// 2^-420 (sdeh_kargs::err_bare) and the sum it returns lies in [2^-700, 2^700]: then
// it is err_sumsq's sum, bit for bit.  Each term t = e / b falls in one of three cases.
// |t| in [2^-540, 2^350] (the sum bounds every term by 2^350): the numerator is
// at least 2^-960 and exceeds the divisor by fewer than 768 binades, the quotient is
// normal, and the IEEE sequence's scalings and fix-up do nothing, so div_in runs
// the same operations.  |t| < 2^-540 (or e = 0): both quotients square to +0.
// |t| > 2^350 (or e Inf or NaN): the bare quotient is at least 2^349, Inf or NaN, so
// the returned sum leaves the range and the caller computes err_sumsq instead.
template <class T, class R>
__device__ __forceinline__ T err_sumsq_in(const T (&e)[R::N], const T (&uprev)[R::N],
                                          const T (&u)[R::N], T abstol, T reltol) {
  static_assert(R::G == 1, "one lane per trajectory");
  T s = T(0);
#pragma unroll
  for (int c = 0; c < R::N; ++c) {
    const T au = fabs(u[c]), ap = fabs(uprev[c]);
    const T m = au < ap ? ap : au;
    const T t = Ops<T>::div_in(e[c], fma(m, reltol, abstol));
    s = c == 0 ? t * t : s + t * t;
  }
  return s;
}

template <class R>
struct NormLen {
  static constexpr int value = R::N * R::G;
  static constexpr bool pow2 = (value & (value - 1)) == 0;
};

// The step-size controller from the error sum of squares s:
// EEst = sqrt(s / len), q11 = EEst^beta1, and the proposed dt / q, where
// q = clamp((q11 / qold^beta2) / gamma) on accept and min(1/qmin, q11 / gamma) on
// reject (gpuatsit5.jl:279-299, gpuvern7.jl:406-426, gpuvern9.jl:575-595).  It
// is evaluated for every lane without branching on accept/reject: the reject
// step dt / min(inv_qmin, q11/gamma) and the accept step
// dt / clamp((q11/qold^beta2)/gamma) share their three divisions, a reject
// dividing q11 by 1 (exact), and EEst == 0 dividing inv_qmax by 1.  EEst^beta1 now
// and EEst^beta2 (the next step's qold^beta2 when qold becomes EEst) share one
// logarithm, returned in `lg`: same bits as two pow calls.
//
// IN: every division and the root go through Ops<T>::div_r / div_in / root_in,
// valid when s is in [2^-700, 2^700] and dt <= 2^700.  Then EEst is in
// [2^-351, 2^350], q11 in [2^-50, 2^49], qold^beta2 in [0.47, 1] (qold =
// max(EEst, 1e-4) of an accepted step, EEst <= 1), q in [1/qmax, 1/qmin] and dt >=
// the solver's dtmin (>= 1e-14): every numerator lies in [2^-700, 2^700] and every
// divisor (len, qold^beta2 or 1, gamma, q) in [2^-60, 2^60], the range those need.
// For other s (0, tiny, huge, Inf, NaN) the IEEE operations run instead and the
// IN results are discarded; computing them is plain arithmetic (the logarithm's and
// exponential's table indices are masked to the tables) on any input.
template <class T>
struct CtrlK {
  T beta1, inv_qmax, inv_qmin, gamma, r_len, r_gamma;
};
template <class T>
struct CtrlOut {
  T EEst, dtq;
  typename Ops<T>::Log lg;
};
template <class T>
__device__ __forceinline__ bool ctrl_in_range(T s, T dt) {
  if constexpr (!Ops<T>::kFast) return false;
  else return s >= T(0x1p-700) && s <= T(0x1p700) && dt <= T(0x1p700);
}
template <class T, class R, bool IN>
__device__ __forceinline__ CtrlOut<T> controller(T s, T dt, T qold_pow, const CtrlK<T>& k) {
  constexpr int L = NormLen<R>::value;
  CtrlOut<T> o;
  T x;
  if constexpr (!IN || NormLen<R>::pow2) x = Ops<T>::div(s, T(L));  // a power of 2: exact
  else x = Ops<T>::div_r(s, T(L), k.r_len);
  o.EEst = IN ? Ops<T>::root_in(x) : Ops<T>::root(x);
  o.lg = Ops<T>::plog(o.EEst);
  const T q11 = Ops<T>::pw_from(o.lg, k.beta1);
  const bool accept = !(o.EEst > T(1));
  const bool zero = (o.EEst == T(0));
  const T num = (accept && zero) ? k.inv_qmax : q11;
  const T den = (accept && !zero) ? qold_pow : T(1);
  const T qg = IN ? Ops<T>::div_r(Ops<T>::div_in(num, den), k.gamma, k.r_gamma)
                  : Ops<T>::div(Ops<T>::div(num, den), k.gamma);
  // With EEst not NaN, q11 and qold^beta2 are finite or +Inf and qg is not NaN:
  // Julia's min (gpuatsit5.jl:287) reduces to one comparison select.
  const T q = accept ? sdeh_math::fast_max(k.inv_qmax, sdeh_math::fast_min(k.inv_qmin, qg))
                     : sdeh_math::fast_min(k.inv_qmin, qg);
  o.dtq = IN ? Ops<T>::div_in(dt, q) : Ops<T>::div(dt, q);
  return o;
}

// Horner evaluation of evalpoly(θ, (0, r...)) / evalpoly(θ, (0, 0, r...)).
template <class T, int NR>
__device__ __forceinline__ T poly1(T th, const T (&r)[NR]) {
  T h = r[NR - 1];
#pragma unroll
  for (int i = NR - 2; i >= 0; --i) h = fma(th, h, r[i]);
  return fma(th, h, T(0));
}
template <class T, int NR>
__device__ __forceinline__ T poly2(T th, const T (&r)[NR]) {
  T h = r[NR - 1];
#pragma unroll
  for (int i = NR - 2; i >= 0; --i) h = fma(th, h, r[i]);
  h = fma(th, h, T(0));
  return fma(th, h, T(0));
}

// ---------------------------------------------------------------------------
// Stage blocks.  Each computes the trial state u from uprev over (t, dt).
// `kahan` variants return the compensation of the u update.
// ---------------------------------------------------------------------------

#define TS(x) T(sdeh_tab::tsit5::x)

// Tsit5 stages k2..k7 from k1 (FSAL); gpuatsit5.jl:99-110 / 258-270.
template <class T, class R, bool KAHAN>
__device__ __forceinline__ void tsit5_stages(const T (&uprev)[R::N], const T (&k1)[R::N],
                                             T (&k2)[R::N], T (&k3)[R::N], T (&k4)[R::N],
                                             T (&k5)[R::N], T (&k6)[R::N], T (&k7)[R::N],
                                             T (&u)[R::N], const T (&comp)[R::N],
                                             T (&comp_new)[R::N], const T* p, const T* ps,
                                             double t, T dt) {
  constexpr int N = R::N;
  T tmp[N];
  comb1(tmp, uprev, dt * TS(a21), k1);
  R::f(k2, tmp, p, ps, stage_t(t, (double)TS(c1), (double)dt));
  comb(tmp, uprev, dt, TS(a31), k1, TS(a32), k2);
  R::f(k3, tmp, p, ps, stage_t(t, (double)TS(c2), (double)dt));
  comb(tmp, uprev, dt, TS(a41), k1, TS(a42), k2, TS(a43), k3);
  R::f(k4, tmp, p, ps, stage_t(t, (double)TS(c3), (double)dt));
  comb(tmp, uprev, dt, TS(a51), k1, TS(a52), k2, TS(a53), k3, TS(a54), k4);
  R::f(k5, tmp, p, ps, stage_t(t, (double)TS(c4), (double)dt));
  comb(tmp, uprev, dt, TS(a61), k1, TS(a62), k2, TS(a63), k3, TS(a64), k4, TS(a65), k5);
  R::f(k6, tmp, p, ps, t + (double)dt);
  if constexpr (KAHAN) {
    T acc[N];
    lsum(acc, TS(a71), k1, TS(a72), k2, TS(a73), k3, TS(a74), k4, TS(a75), k5, TS(a76), k6);
    kahan_update(u, comp_new, uprev, comp, dt, acc);
  } else {
    comb(u, uprev, dt, TS(a71), k1, TS(a72), k2, TS(a73), k3, TS(a74), k4, TS(a75), k5,
         TS(a76), k6);
  }
  R::f(k7, u, p, ps, t + (double)dt);
}

// Tsit5 dense output (tsit5.jl:385-399 weights) at θ over (uprev, h).
template <class T, int N>
__device__ __forceinline__ void tsit5_dense(T (&out)[N], const T (&uprev)[N], T h, T th,
                                            const T (&k1)[N], const T (&k2)[N],
                                            const T (&k3)[N], const T (&k4)[N],
                                            const T (&k5)[N], const T (&k6)[N],
                                            const T (&k7)[N]) {
  const T r1[4] = {TS(r11), TS(r12), TS(r13), TS(r14)};
  const T r2[3] = {TS(r22), TS(r23), TS(r24)};
  const T r3[3] = {TS(r32), TS(r33), TS(r34)};
  const T r4[3] = {TS(r42), TS(r43), TS(r44)};
  const T r5[3] = {TS(r52), TS(r53), TS(r54)};
  const T r6[3] = {TS(r62), TS(r63), TS(r64)};
  const T r7[3] = {TS(r72), TS(r73), TS(r74)};
  comb(out, uprev, h, poly1(th, r1), k1, poly2(th, r2), k2, poly2(th, r3), k3, poly2(th, r4),
       k4, poly2(th, r5), k5, poly2(th, r6), k6, poly2(th, r7), k7);
}
#undef TS

#define V7(x) T(sdeh_tab::vern7::x)
#define V7X(x) T(sdeh_tab::vern7x::x)
#define V7I(x) T(sdeh_tab::vern7i::x)

template <class T, int N>
struct V7K {
  T k1[N], k4[N], k5[N], k6[N], k7[N], k8[N], k9[N];
};

// Verner 7 stages k1..k9 (+k10, error) and u; gpuvern7.jl:104-136 / 360-402.
// Returns the embedded error accumulator in err (adaptive only).
template <class T, class R, bool ADAPT, bool KAHAN>
__device__ __forceinline__ void vern7_stages(const T (&uprev)[R::N], V7K<T, R::N>& K,
                                             T (&u)[R::N], T (&err)[R::N],
                                             const T (&comp)[R::N], T (&comp_new)[R::N],
                                             const T* p, const T* ps, double t, T dt) {
  constexpr int N = R::N;
  T tmp[N], k2[N], k3[N];
  R::f(K.k1, uprev, p, ps, t);
  comb1(tmp, uprev, dt * V7(a021), K.k1);
  R::f(k2, tmp, p, ps, stage_t(t, (double)V7(c2), (double)dt));
  comb(tmp, uprev, dt, V7(a031), K.k1, V7(a032), k2);
  R::f(k3, tmp, p, ps, stage_t(t, (double)V7(c3), (double)dt));
  comb(tmp, uprev, dt, V7(a041), K.k1, V7(a043), k3);
  R::f(K.k4, tmp, p, ps, stage_t(t, (double)V7(c4), (double)dt));
  comb(tmp, uprev, dt, V7(a051), K.k1, V7(a053), k3, V7(a054), K.k4);
  R::f(K.k5, tmp, p, ps, stage_t(t, (double)V7(c5), (double)dt));
  comb(tmp, uprev, dt, V7(a061), K.k1, V7(a063), k3, V7(a064), K.k4, V7(a065), K.k5);
  R::f(K.k6, tmp, p, ps, stage_t(t, (double)V7(c6), (double)dt));
  comb(tmp, uprev, dt, V7(a071), K.k1, V7(a073), k3, V7(a074), K.k4, V7(a075), K.k5, V7(a076),
       K.k6);
  R::f(K.k7, tmp, p, ps, stage_t(t, (double)V7(c7), (double)dt));
  comb(tmp, uprev, dt, V7(a081), K.k1, V7(a083), k3, V7(a084), K.k4, V7(a085), K.k5, V7(a086),
       K.k6, V7(a087), K.k7);
  R::f(K.k8, tmp, p, ps, stage_t(t, (double)V7(c8), (double)dt));
  comb(tmp, uprev, dt, V7(a091), K.k1, V7(a093), k3, V7(a094), K.k4, V7(a095), K.k5, V7(a096),
       K.k6, V7(a097), K.k7, V7(a098), K.k8);
  T k10[N];
  if constexpr (ADAPT) {
    T g10[N];
    comb(g10, uprev, dt, V7(a101), K.k1, V7(a103), k3, V7(a104), K.k4, V7(a105), K.k5,
         V7(a106), K.k6, V7(a107), K.k7);
    R::f(K.k9, tmp, p, ps, t + (double)dt);
    R::f(k10, g10, p, ps, t + (double)dt);
  } else {
    R::f(K.k9, tmp, p, ps, t + (double)dt);
  }
  if constexpr (KAHAN) {
    T acc[N];
    lsum(acc, V7(b1), K.k1, V7(b4), K.k4, V7(b5), K.k5, V7(b6), K.k6, V7(b7), K.k7, V7(b8),
         K.k8, V7(b9), K.k9);
    kahan_update(u, comp_new, uprev, comp, dt, acc);
  } else {
    comb(u, uprev, dt, V7(b1), K.k1, V7(b4), K.k4, V7(b5), K.k5, V7(b6), K.k6, V7(b7), K.k7,
         V7(b8), K.k8, V7(b9), K.k9);
  }
  if constexpr (ADAPT) {
    lsum(err, V7(btilde1), K.k1, V7(btilde4), K.k4, V7(btilde5), K.k5, V7(btilde6), K.k6,
         V7(btilde7), K.k7, V7(btilde8), K.k8, V7(btilde9), K.k9, V7(btilde10), k10);
  }
}

// Vern7 extra stages k11..k16 at times ts + c*h (gpuvern7.jl:147-190 / 436-508).
template <class T, class R>
__device__ __forceinline__ void vern7_extra(const T (&uprev)[R::N], const V7K<T, R::N>& K,
                                            T (&x)[6][R::N], const T* p, const T* ps,
                                            double ts, T h) {
  constexpr int N = R::N;
  T tmp[N];
  comb(tmp, uprev, h, V7X(a1101), K.k1, V7X(a1104), K.k4, V7X(a1105), K.k5, V7X(a1106), K.k6,
       V7X(a1107), K.k7, V7X(a1108), K.k8, V7X(a1109), K.k9);
  R::f(x[0], tmp, p, ps, stage_t(ts, (double)V7X(c11), (double)h));
  comb(tmp, uprev, h, V7X(a1201), K.k1, V7X(a1204), K.k4, V7X(a1205), K.k5, V7X(a1206), K.k6,
       V7X(a1207), K.k7, V7X(a1208), K.k8, V7X(a1209), K.k9, V7X(a1211), x[0]);
  R::f(x[1], tmp, p, ps, stage_t(ts, (double)V7X(c12), (double)h));
  comb(tmp, uprev, h, V7X(a1301), K.k1, V7X(a1304), K.k4, V7X(a1305), K.k5, V7X(a1306), K.k6,
       V7X(a1307), K.k7, V7X(a1308), K.k8, V7X(a1309), K.k9, V7X(a1311), x[0], V7X(a1312),
       x[1]);
  R::f(x[2], tmp, p, ps, stage_t(ts, (double)V7X(c13), (double)h));
  comb(tmp, uprev, h, V7X(a1401), K.k1, V7X(a1404), K.k4, V7X(a1405), K.k5, V7X(a1406), K.k6,
       V7X(a1407), K.k7, V7X(a1408), K.k8, V7X(a1409), K.k9, V7X(a1411), x[0], V7X(a1412),
       x[1], V7X(a1413), x[2]);
  R::f(x[3], tmp, p, ps, stage_t(ts, (double)V7X(c14), (double)h));
  comb(tmp, uprev, h, V7X(a1501), K.k1, V7X(a1504), K.k4, V7X(a1505), K.k5, V7X(a1506), K.k6,
       V7X(a1507), K.k7, V7X(a1508), K.k8, V7X(a1509), K.k9, V7X(a1511), x[0], V7X(a1512),
       x[1], V7X(a1513), x[2]);
  R::f(x[4], tmp, p, ps, stage_t(ts, (double)V7X(c15), (double)h));
  comb(tmp, uprev, h, V7X(a1601), K.k1, V7X(a1604), K.k4, V7X(a1605), K.k5, V7X(a1606), K.k6,
       V7X(a1607), K.k7, V7X(a1608), K.k8, V7X(a1609), K.k9, V7X(a1611), x[0], V7X(a1612),
       x[1], V7X(a1613), x[2]);
  R::f(x[5], tmp, p, ps, stage_t(ts, (double)V7X(c16), (double)h));
}

// Vern7 dense output (bθs at verner_tableaus.jl:1337-1360).
template <class T, int N>
__device__ __forceinline__ void vern7_dense(T (&out)[N], const T (&uprev)[N], T h, T th,
                                            const V7K<T, N>& K, const T (&x)[6][N]) {
  const T r1[7] = {V7I(r011), V7I(r012), V7I(r013), V7I(r014), V7I(r015), V7I(r016),
                   V7I(r017)};
#define SDEH_R7(j) \
  const T r##j[6] = {V7I(r##j##2), V7I(r##j##3), V7I(r##j##4), V7I(r##j##5), V7I(r##j##6), \
                     V7I(r##j##7)};
  SDEH_R7(04) SDEH_R7(05) SDEH_R7(06) SDEH_R7(07) SDEH_R7(08) SDEH_R7(09)
  SDEH_R7(11) SDEH_R7(12) SDEH_R7(13) SDEH_R7(14) SDEH_R7(15) SDEH_R7(16)
#undef SDEH_R7
  comb(out, uprev, h, poly1(th, r1), K.k1, poly2(th, r04), K.k4, poly2(th, r05), K.k5,
       poly2(th, r06), K.k6, poly2(th, r07), K.k7, poly2(th, r08), K.k8, poly2(th, r09), K.k9,
       poly2(th, r11), x[0], poly2(th, r12), x[1], poly2(th, r13), x[2], poly2(th, r14), x[3],
       poly2(th, r15), x[4], poly2(th, r16), x[5]);
}
#undef V7
#undef V7X
#undef V7I

#define V9(x) T(sdeh_tab::vern9::x)
#define V9X(x) T(sdeh_tab::vern9x::x)
#define V9I(x) T(sdeh_tab::vern9i::x)

// The nine stage vectors the Vern9 dense output reads (slots of b1Θ, b8Θ..b15Θ).
template <class T, int N>
struct V9S {
  T s[9][N];
};

// Verner 9 stages k1..k15 (+k16, error) and u; gpuvern9.jl:104-174 / 468-561.
// On return `S` holds the dense-output slots: k1, k8..k15 (renamed as at
// gpuvern9.jl:564-573) or, when `alias_quirk`, k1..k9 (the fixed solver's
// unrenamed variables, gpuvern9.jl:216-331).
template <class T, class R, bool ADAPT, bool KAHAN, bool KEEP>
__device__ __forceinline__ void vern9_stages(const T (&uprev)[R::N], V9S<T, R::N>& S,
                                             T (&u)[R::N], T (&err)[R::N],
                                             const T (&comp)[R::N], T (&comp_new)[R::N],
                                             const T* p, const T* ps, double t, T dt,
                                             bool alias_quirk) {
  constexpr int N = R::N;
  T tmp[N], k1[N], k2[N], k3[N], k4[N], k5[N], k6[N], k7[N], k8[N], k9[N], k10[N], k11[N],
      k12[N], k13[N], k14[N], k15[N];
  R::f(k1, uprev, p, ps, t);
  comb1(tmp, uprev, dt * V9(a0201), k1);
  R::f(k2, tmp, p, ps, stage_t(t, (double)V9(c1), (double)dt));
  comb(tmp, uprev, dt, V9(a0301), k1, V9(a0302), k2);
  R::f(k3, tmp, p, ps, stage_t(t, (double)V9(c2), (double)dt));
  comb(tmp, uprev, dt, V9(a0401), k1, V9(a0403), k3);
  R::f(k4, tmp, p, ps, stage_t(t, (double)V9(c3), (double)dt));
  comb(tmp, uprev, dt, V9(a0501), k1, V9(a0503), k3, V9(a0504), k4);
  R::f(k5, tmp, p, ps, stage_t(t, (double)V9(c4), (double)dt));
  comb(tmp, uprev, dt, V9(a0601), k1, V9(a0604), k4, V9(a0605), k5);
  R::f(k6, tmp, p, ps, stage_t(t, (double)V9(c5), (double)dt));
  comb(tmp, uprev, dt, V9(a0701), k1, V9(a0704), k4, V9(a0705), k5, V9(a0706), k6);
  R::f(k7, tmp, p, ps, stage_t(t, (double)V9(c6), (double)dt));
  comb(tmp, uprev, dt, V9(a0801), k1, V9(a0806), k6, V9(a0807), k7);
  R::f(k8, tmp, p, ps, stage_t(t, (double)V9(c7), (double)dt));
  comb(tmp, uprev, dt, V9(a0901), k1, V9(a0906), k6, V9(a0907), k7, V9(a0908), k8);
  R::f(k9, tmp, p, ps, stage_t(t, (double)V9(c8), (double)dt));
  if constexpr (KEEP) {
    if (alias_quirk) {
      copyv(S.s[1], k2); copyv(S.s[2], k3); copyv(S.s[3], k4); copyv(S.s[4], k5);
      copyv(S.s[5], k6); copyv(S.s[6], k7); copyv(S.s[7], k8); copyv(S.s[8], k9);
    }
  }
  comb(tmp, uprev, dt, V9(a1001), k1, V9(a1006), k6, V9(a1007), k7, V9(a1008), k8, V9(a1009),
       k9);
  R::f(k10, tmp, p, ps, stage_t(t, (double)V9(c9), (double)dt));
  comb(tmp, uprev, dt, V9(a1101), k1, V9(a1106), k6, V9(a1107), k7, V9(a1108), k8, V9(a1109),
       k9, V9(a1110), k10);
  R::f(k11, tmp, p, ps, stage_t(t, (double)V9(c10), (double)dt));
  comb(tmp, uprev, dt, V9(a1201), k1, V9(a1206), k6, V9(a1207), k7, V9(a1208), k8, V9(a1209),
       k9, V9(a1210), k10, V9(a1211), k11);
  R::f(k12, tmp, p, ps, stage_t(t, (double)V9(c11), (double)dt));
  comb(tmp, uprev, dt, V9(a1301), k1, V9(a1306), k6, V9(a1307), k7, V9(a1308), k8, V9(a1309),
       k9, V9(a1310), k10, V9(a1311), k11, V9(a1312), k12);
  R::f(k13, tmp, p, ps, stage_t(t, (double)V9(c12), (double)dt));
  comb(tmp, uprev, dt, V9(a1401), k1, V9(a1406), k6, V9(a1407), k7, V9(a1408), k8, V9(a1409),
       k9, V9(a1410), k10, V9(a1411), k11, V9(a1412), k12, V9(a1413), k13);
  R::f(k14, tmp, p, ps, stage_t(t, (double)V9(c13), (double)dt));
  comb(tmp, uprev, dt, V9(a1501), k1, V9(a1506), k6, V9(a1507), k7, V9(a1508), k8, V9(a1509),
       k9, V9(a1510), k10, V9(a1511), k11, V9(a1512), k12, V9(a1513), k13, V9(a1514), k14);
  T k16[N];
  if constexpr (ADAPT) {
    T g16[N];
    comb(g16, uprev, dt, V9(a1601), k1, V9(a1606), k6, V9(a1607), k7, V9(a1608), k8,
         V9(a1609), k9, V9(a1610), k10, V9(a1611), k11, V9(a1612), k12, V9(a1613), k13);
    R::f(k15, tmp, p, ps, t + (double)dt);
    R::f(k16, g16, p, ps, t + (double)dt);
  } else {
    R::f(k15, tmp, p, ps, t + (double)dt);
  }
  if constexpr (KAHAN) {
    T acc[N];
    lsum(acc, V9(b1), k1, V9(b8), k8, V9(b9), k9, V9(b10), k10, V9(b11), k11, V9(b12), k12,
         V9(b13), k13, V9(b14), k14, V9(b15), k15);
    kahan_update(u, comp_new, uprev, comp, dt, acc);
  } else {
    comb(u, uprev, dt, V9(b1), k1, V9(b8), k8, V9(b9), k9, V9(b10), k10, V9(b11), k11, V9(b12),
         k12, V9(b13), k13, V9(b14), k14, V9(b15), k15);
  }
  if constexpr (ADAPT) {
    lsum(err, V9(btilde1), k1, V9(btilde8), k8, V9(btilde9), k9, V9(btilde10), k10,
         V9(btilde11), k11, V9(btilde12), k12, V9(btilde13), k13, V9(btilde14), k14,
         V9(btilde15), k15, V9(btilde16), k16);
  }
  if constexpr (KEEP) {
    copyv(S.s[0], k1);
    if (!alias_quirk) {
      copyv(S.s[1], k8); copyv(S.s[2], k9); copyv(S.s[3], k10); copyv(S.s[4], k11);
      copyv(S.s[5], k12); copyv(S.s[6], k13); copyv(S.s[7], k14); copyv(S.s[8], k15);
    }
  }
}

// Vern9 extra stages (k11..k20 of the reference) at ts + c*h
// (gpuvern9.jl:200-317 / 608-745).
template <class T, class R>
__device__ __forceinline__ void vern9_extra(const T (&uprev)[R::N], const V9S<T, R::N>& S,
                                            T (&x)[10][R::N], const T* p, const T* ps,
                                            double ts, T h) {
  constexpr int N = R::N;
  T tmp[N];
#define SDEH_M9(r)                                                                         \
  V9X(a##r##01), S.s[0], V9X(a##r##08), S.s[1], V9X(a##r##09), S.s[2], V9X(a##r##10),      \
      S.s[3], V9X(a##r##11), S.s[4], V9X(a##r##12), S.s[5], V9X(a##r##13), S.s[6],         \
      V9X(a##r##14), S.s[7], V9X(a##r##15), S.s[8]
  comb(tmp, uprev, h, SDEH_M9(17));
  R::f(x[0], tmp, p, ps, stage_t(ts, (double)V9X(c17), (double)h));
  comb(tmp, uprev, h, SDEH_M9(18), V9X(a1817), x[0]);
  R::f(x[1], tmp, p, ps, stage_t(ts, (double)V9X(c18), (double)h));
  comb(tmp, uprev, h, SDEH_M9(19), V9X(a1917), x[0], V9X(a1918), x[1]);
  R::f(x[2], tmp, p, ps, stage_t(ts, (double)V9X(c19), (double)h));
  comb(tmp, uprev, h, SDEH_M9(20), V9X(a2017), x[0], V9X(a2018), x[1], V9X(a2019), x[2]);
  R::f(x[3], tmp, p, ps, stage_t(ts, (double)V9X(c20), (double)h));
  comb(tmp, uprev, h, SDEH_M9(21), V9X(a2117), x[0], V9X(a2118), x[1], V9X(a2119), x[2],
       V9X(a2120), x[3]);
  R::f(x[4], tmp, p, ps, stage_t(ts, (double)V9X(c21), (double)h));
  comb(tmp, uprev, h, SDEH_M9(22), V9X(a2217), x[0], V9X(a2218), x[1], V9X(a2219), x[2],
       V9X(a2220), x[3], V9X(a2221), x[4]);
  R::f(x[5], tmp, p, ps, stage_t(ts, (double)V9X(c22), (double)h));
  comb(tmp, uprev, h, SDEH_M9(23), V9X(a2317), x[0], V9X(a2318), x[1], V9X(a2319), x[2],
       V9X(a2320), x[3], V9X(a2321), x[4]);
  R::f(x[6], tmp, p, ps, stage_t(ts, (double)V9X(c23), (double)h));
  comb(tmp, uprev, h, SDEH_M9(24), V9X(a2417), x[0], V9X(a2418), x[1], V9X(a2419), x[2],
       V9X(a2420), x[3], V9X(a2421), x[4]);
  R::f(x[7], tmp, p, ps, stage_t(ts, (double)V9X(c24), (double)h));
  comb(tmp, uprev, h, SDEH_M9(25), V9X(a2517), x[0], V9X(a2518), x[1], V9X(a2519), x[2],
       V9X(a2520), x[3], V9X(a2521), x[4]);
  R::f(x[8], tmp, p, ps, stage_t(ts, (double)V9X(c25), (double)h));
  comb(tmp, uprev, h, SDEH_M9(26), V9X(a2617), x[0], V9X(a2618), x[1], V9X(a2619), x[2],
       V9X(a2620), x[3], V9X(a2621), x[4]);
  R::f(x[9], tmp, p, ps, stage_t(ts, (double)V9X(c26), (double)h));
#undef SDEH_M9
}

// Vern9 dense output (bθs at verner_tableaus.jl:1362-1398): the 19 weights of θ,
// then out = uprev + h * (w_1 k1 + w_8 k8 + ... + w_26 k26).
constexpr int V9W = 19;
template <class T>
__device__ __forceinline__ void vern9_weights(T th, T (&w)[V9W]) {
  const T r1[9] = {V9I(r011), V9I(r012), V9I(r013), V9I(r014), V9I(r015),
                   V9I(r016), V9I(r017), V9I(r018), V9I(r019)};
#define SDEH_R9(j)                                                                      \
  const T r##j[8] = {V9I(r##j##2), V9I(r##j##3), V9I(r##j##4), V9I(r##j##5),           \
                     V9I(r##j##6), V9I(r##j##7), V9I(r##j##8), V9I(r##j##9)};
  SDEH_R9(08) SDEH_R9(09) SDEH_R9(10) SDEH_R9(11) SDEH_R9(12) SDEH_R9(13) SDEH_R9(14)
  SDEH_R9(15) SDEH_R9(17) SDEH_R9(18) SDEH_R9(19) SDEH_R9(20) SDEH_R9(21) SDEH_R9(22)
  SDEH_R9(23) SDEH_R9(24) SDEH_R9(25) SDEH_R9(26)
#undef SDEH_R9
  w[0] = poly1(th, r1);
  w[1] = poly2(th, r08); w[2] = poly2(th, r09); w[3] = poly2(th, r10); w[4] = poly2(th, r11);
  w[5] = poly2(th, r12); w[6] = poly2(th, r13); w[7] = poly2(th, r14); w[8] = poly2(th, r15);
  w[9] = poly2(th, r17); w[10] = poly2(th, r18); w[11] = poly2(th, r19);
  w[12] = poly2(th, r20); w[13] = poly2(th, r21); w[14] = poly2(th, r22);
  w[15] = poly2(th, r23); w[16] = poly2(th, r24); w[17] = poly2(th, r25);
  w[18] = poly2(th, r26);
}

template <class T, int N>
__device__ __forceinline__ void vern9_combine(T (&out)[N], const T (&uprev)[N], T h,
                                              const T (&w)[V9W], const V9S<T, N>& S,
                                              const T (&x)[10][N]) {
  comb(out, uprev, h, w[0], S.s[0], w[1], S.s[1], w[2], S.s[2], w[3], S.s[3], w[4], S.s[4],
       w[5], S.s[5], w[6], S.s[6], w[7], S.s[7], w[8], S.s[8], w[9], x[0], w[10], x[1],
       w[11], x[2], w[12], x[3], w[13], x[4], w[14], x[5], w[15], x[6], w[16], x[7],
       w[17], x[8], w[18], x[9]);
}

template <class T, int N>
__device__ __forceinline__ void vern9_dense(T (&out)[N], const T (&uprev)[N], T h, T th,
                                            const V9S<T, N>& S, const T (&x)[10][N]) {
  T w[V9W];
  vern9_weights(th, w);
  vern9_combine(out, uprev, h, w, S, x);
}
#undef V9
#undef V9X
#undef V9I

// ---------------------------------------------------------------------------
// Output helpers (structure-of-arrays, component major).
// ---------------------------------------------------------------------------
// Row `row` of a [rows][components][n] output: this lane's components of `traj`.
template <class T, class R>
__device__ __forceinline__ void store_state(void* base, int64_t row, int64_t n, int64_t traj,
                                            const T (&u)[R::N]) {
  constexpr int FULL = R::N * R::G;
  T* o = (T*)base;
#pragma unroll
  for (int c = 0; c < R::N; ++c) o[(row * FULL + comp_idx<R>(c)) * n + traj] = u[c];
}

// Row `row` of a saved-state output with `rows` rows, in the layout the kernel
// arguments select: point-major [rows][components][n] or trajectory-major
// [n][rows][components] (a lane then writes its components contiguously).
template <class T, class R>
__device__ __forceinline__ void store_save(const sdeh_kargs& a, void* base, int64_t row,
                                           int64_t rows, int64_t traj, const T (&u)[R::N]) {
  constexpr int FULL = R::N * R::G;
  if (!a.traj_major) {
    store_state<T, R>(base, row, a.n, traj, u);
    return;
  }
  T* o = (T*)base + (traj * rows + row) * FULL;
#pragma unroll
  for (int c = 0; c < R::N; ++c) o[comp_idx<R>(c)] = u[c];
}

__device__ __forceinline__ void store_save_time(const sdeh_kargs& a, int64_t row, int64_t traj,
                                                double t) {
  if (a.traj_major) a.t_steps[traj * a.max_saves + row] = t;
  else a.t_steps[row * a.n + traj] = t;
}

// save_everystep output of one trajectory.  Fixed rows: point k goes to row k of
// the trajectory's max_saves rows.  Packed (a.packed): point k goes to row
// base + k of the shared output while k < lim, where base = steps_off[traj] and lim
// = n_steps[traj] come from the counting pass and the offsets scan.  Packed rows
// are held back in registers and written PK_ROWS at a time (consecutive rows:
// contiguous bytes, so L2 sees whole runs of a trajectory's lines instead of one
// 24-byte row per step, lines it would otherwise write back partly filled); the
// trajectory's end writes the rest.
#ifndef SDEH_PK_ROWS
#define SDEH_PK_ROWS 4
#endif
constexpr int PK_ROWS = SDEH_PK_ROWS;

template <class T, int N>
struct StepSink {
  int64_t base;  // packed: first row of the trajectory
  int32_t lim;   // rows available (packed) or max_saves
  int32_t n;     // points offered
  int32_t kept;  // points stored (written or held back)
  int32_t held;  // packed: rows held back, the last `held` of bt/bu
  double bt[PK_ROWS];
  T bu[PK_ROWS][N];
};

template <class R, class T, int N>
__device__ __forceinline__ void sink_start(const sdeh_kargs& a, StepSink<T, N>& s,
                                           int64_t traj) {
  s.n = s.kept = s.held = 0;
  if (a.packed) {
    s.base = a.steps_off[traj];
    s.lim = a.n_steps[traj];
  } else {
    s.base = 0;
    s.lim = a.max_saves;
  }
}

// Packed: write the rows held back (rows kept - held .. kept - 1).
template <class T, class R>
__device__ __forceinline__ void flush_steps(const sdeh_kargs& a, StepSink<T, R::N>& s) {
  if (!a.packed || s.held == 0) return;
  constexpr int FULL = R::N * R::G;
  const bool lead = lane_comp<R>() == 0;
#pragma unroll
  for (int k = 0; k < PK_ROWS; ++k) {
    if (k >= PK_ROWS - s.held) {
      const int64_t row = s.base + s.kept - (PK_ROWS - k);
      T* o = (T*)a.u_steps + row * FULL;
#pragma unroll
      for (int c = 0; c < R::N; ++c) o[comp_idx<R>(c)] = s.bu[k][c];
      if (lead) a.t_steps[row] = s.bt[k];
    }
  }
  s.held = 0;
}

template <class T, class R>
__device__ __forceinline__ void save_step(const sdeh_kargs& a, StepSink<T, R::N>& s,
                                          int64_t traj, double t, const T (&u)[R::N]) {
  const bool lead = lane_comp<R>() == 0;
  ++s.n;
  if (s.kept >= s.lim) return;
  if (a.packed) {
#pragma unroll
    for (int k = 0; k + 1 < PK_ROWS; ++k) {  // shift: static register indices
      s.bt[k] = s.bt[k + 1];
#pragma unroll
      for (int c = 0; c < R::N; ++c) s.bu[k][c] = s.bu[k + 1][c];
    }
    s.bt[PK_ROWS - 1] = t;
#pragma unroll
    for (int c = 0; c < R::N; ++c) s.bu[PK_ROWS - 1][c] = u[c];
    ++s.kept;
    if (++s.held == PK_ROWS) flush_steps<T, R>(a, s);
    return;
  }
  store_save<T, R>(a, a.u_steps, s.kept, a.max_saves, traj, u);
  if (lead) store_save_time(a, s.kept, traj, t);
  ++s.kept;
}

// Leader's end-of-trajectory write of the save_everystep count (after
// flush_steps on every lane of the trajectory).
template <class T, int N>
__device__ __forceinline__ void finish_steps(const sdeh_kargs& a, const StepSink<T, N>& s,
                                             int64_t traj) {
  a.n_steps[traj] = s.kept;
}

template <class T, class R>
__device__ __forceinline__ void load_state(T (&u)[R::N], const void* base, int64_t n,
                                           int64_t traj) {
#pragma unroll
  for (int c = 0; c < R::N; ++c) u[c] = ((const T*)base)[comp_idx<R>(c) * n + traj];
}

// Per-lane parameters: p[c][traj] for c < R::P, or the right-hand side's own
// lane loader in a group layout.
template <class T, class R, int P>
__device__ __forceinline__ void load_params(T (&pl)[P], const sdeh_kargs& a, int64_t traj) {
  if constexpr (R::G > 1) {
    R::template load_lane<T>(pl, (const T*)a.p, (const T*)a.ps, a.n, traj, lane_comp<R>());
  } else {
#pragma unroll
    for (int c = 0; c < R::P; ++c) pl[c] = ((const T*)a.p)[c * a.n + traj];
  }
}

template <class T>
__device__ __forceinline__ T nan_of() {
  return (T)__builtin_nan("");
}

// ---------------------------------------------------------------------------
// Wave-level work queue for the persistent schedule.
// ---------------------------------------------------------------------------
// A loop-invariant value held in vector registers.  The one-lane kernels run out
// of scalar registers first; a scalar the compiler spills is reloaded with a
// v_readlane (a vector instruction) on every attempt that uses it.
template <class X>
__device__ __forceinline__ X in_vgpr(X x) {
  asm volatile("" : "+v"(x));
  return x;
}
// in_vgpr where it costs no occupancy: in the binary64 endpoint kernels (the
// binary32 ones have scalar registers to spare and would cross 64 vector registers,
// and the save modes' kernels would lose a wave per SIMD).
template <bool ON, class X>
__device__ __forceinline__ X in_vgpr_if(X x) {
  if constexpr (ON) return in_vgpr(x);
  return x;
}

__device__ __forceinline__ unsigned lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Claim queue slots for the lanes flagged in `need` (called in uniform control
// flow by the whole wave).  Returns this lane's slot or -1.  `base`/`lim` is
// the wave's current claimed range of `chunk` slots; `drained` becomes true once
// the queue is exhausted.  (Claims that shrink as the queue empties balance the
// end of the launch better on paper but measured slower: each extra claim is a
// device-scope atomic round trip the wave waits for.)
__device__ __forceinline__ int64_t wave_claim(const sdeh_kargs& a, bool need_me, int64_t& base,
                                              int64_t& lim, bool& drained) {
  int64_t mine = -1;
  unsigned long long need = __ballot(need_me);
  while (need != 0ull && !drained) {
    const int cnt = __popcll(need);
    if (base >= lim) {
      unsigned long long old = 0;
      if (__lane_id() == 0) old = atomicAdd(a.queue, (unsigned long long)a.chunk);
      old = (unsigned long long)__shfl((long long)old, 0);
      base = (int64_t)old;
      lim = base + a.chunk < a.n ? base + a.chunk : a.n;
      if (base >= a.n) {
        drained = true;
        break;
      }
    }
    const int64_t avail = lim - base;
    const bool in_need = ((need >> __lane_id()) & 1ull) != 0ull;
    const int64_t rank = (int64_t)lane_rank(need);
    if (in_need && rank < avail) mine = base + rank;
    const int64_t taken = (int64_t)cnt < avail ? (int64_t)cnt : avail;
    base += taken;
    need = __ballot(in_need && mine < 0);
  }
  return mine;
}

// wave_claim that also names the claimed slot's trajectory, as slot_traj does.  The
// wave caches the hand-out order's entries for 64 consecutive slots, win .. win + 63,
// one per lane in `word`, and reloads them with one coalesced load when the slots it
// hands out leave that window; a starting lane then takes its entry from the cache
// with a lane shuffle instead of loading it after the claim, a memory round trip the
// whole wave would wait for.  Without an order the cache holds the slots themselves.
// `win` starts below -64, so that the first hand-out loads the window.
__device__ __forceinline__ int64_t wave_claim_win(const sdeh_kargs& a, bool need_me,
                                                  int64_t& base, int64_t& lim, bool& drained,
                                                  int64_t& win, int64_t& word, int64_t& traj) {
  int64_t mine = -1;
  traj = -1;
  unsigned long long need = __ballot(need_me);
  while (need != 0ull && !drained) {
    const int cnt = __popcll(need);
    if (base >= lim) {
      unsigned long long old = 0;
      if (__lane_id() == 0) old = atomicAdd(a.queue, (unsigned long long)a.chunk);
      old = (unsigned long long)__shfl((long long)old, 0);
      base = (int64_t)old;
      lim = base + a.chunk < a.n ? base + a.chunk : a.n;
      if (base >= a.n) {
        drained = true;
        break;
      }
    }
    const int64_t avail = lim - base;
    const bool in_need = ((need >> __lane_id()) & 1ull) != 0ull;
    const int64_t rank = (int64_t)lane_rank(need);
    const int64_t taken = (int64_t)cnt < avail ? (int64_t)cnt : avail;
    if (base < win || base + taken > win + 64) {  // the hand-out leaves the window
      win = base;
      const int64_t e = win + (int64_t)__lane_id();
      word = !a.order ? e : (e < a.n ? a.order[e] : int64_t(-1));
    }
    const bool gets = in_need && rank < avail;
    const int src = gets ? (int)(base + rank - win) : (int)__lane_id();
    const int64_t v = (int64_t)__shfl((long long)word, src);
    if (gets) {
      mine = base + rank;
      traj = (uint64_t)v < (uint64_t)a.n ? v : int64_t(-1);
    }
    base += taken;
    need = __ballot(in_need && mine < 0);
  }
  return mine;
}

// The trajectory handed out at `slot`: order[slot], or slot itself without an order.
// An entry outside [0, n) names no trajectory: it gives -1, which no lane takes up, so
// that a hand-out order is never dereferenced out of bounds (the host checks caller
// orders to be permutations; the library's own orders are permutations by construction).
__device__ __forceinline__ int64_t slot_traj(const sdeh_kargs& a, int64_t slot) {
  if (!a.order) return slot;
  const int64_t v = a.order[slot];
  return (uint64_t)v < (uint64_t)a.n ? v : int64_t(-1);
}

// The debug visit counter of trajectory `traj` (sdeh_kargs::visits), one per take-up.
__device__ __forceinline__ void note_visit(const sdeh_kargs& a, int64_t traj) {
  if (a.visits) atomicAdd(a.visits + traj, 1);
}

// Save time i, or 0 past the end (callers test i < n_saveat first).  The next save
// time is kept in a register between steps and the one after it is loaded while
// the current point's dense output is computed, so the save test on each step
// waits on no global load.
__device__ __forceinline__ double save_time(const sdeh_kargs& a, int64_t i) {
  return i < a.n_saveat ? a.saveat[i] : 0.0;
}

// ---------------------------------------------------------------------------
// Adaptive solvers: ATsit5, AVern7, AVern9.
// ---------------------------------------------------------------------------
// Adaptive solvers in the wave layout.  The same per-trajectory arithmetic as
// adaptive_body, arranged so that every right-hand-side evaluation runs with the
// whole wave active (R::f is one wave-wide matrix operation over all the wave's
// trajectories): the stages are computed by every lane each iteration, lanes
// without a running trajectory on stale values whose results are dropped; the
// FSAL seed of new trajectories and the extra stages of dense output run for the
// whole wave when any of its lanes needs them.  Decisions and outputs are per
// lane, as in adaptive_body.
template <class T, class R, int ALG, int SAVE, bool KAHAN>
__device__ __forceinline__ void adaptive_body_wave(const sdeh_kargs& a) {
  constexpr int N = R::N;
  constexpr int P = R::P;
  constexpr int SLOTS = 64 / R::G;  // trajectories per wave
  constexpr bool V9 = (ALG == ALG_AVERN9);
  // LDS exchange per wave: the error norm's components, and (AVern9 with saveat) the
  // dense-output weights of R::G points per trajectory.  Each trajectory's share is
  // padded to an odd number of values (DWS) so that the wave's trajectories read
  // their weights from different banks.
  constexpr int DWS = R::G * V9W + 1;
  constexpr int NWG = SLOTS * wave_slot_stride<T, R>();
  constexpr int NWL = (V9 && SAVE == SAVE_AT && SLOTS * DWS > NWG) ? SLOTS * DWS : NWG;
  constexpr double DTMIN = V9 ? (double)1.0e-7f : 1.0e-14;  // gpuvern9.jl:466
  constexpr double SNAP = DTMIN;                            // gpuvern9.jl:591
  // build_adaptive_controller_cache(T), src/SimpleDiffEq.jl:67-77
  const T beta1 = T(7.0 / 50), beta2 = T(2.0 / 25), qmax = T(10.0), qmin = T(1.0 / 5),
          gamma = T(9.0 / 10), qoldinit = T(1.0e-4);
  const T inv_qmax = T(1) / qmax, inv_qmin = T(1) / qmin;
  const T qoldinit_pow = Ops<T>::pw(qoldinit, beta2);
  const T abstol = (T)a.abstol, reltol = (T)a.reltol;
  const CtrlK<T> ck{beta1, inv_qmax, inv_qmin, gamma, Ops<T>::rcp_nr(T(NormLen<R>::value)),
                    Ops<T>::rcp_nr(gamma)};
  const T* ps = (const T*)a.ps;
  const int64_t n = a.n;
  const int lane = (int)__lane_id();

  // Lane state; lanes without a trajectory keep finite values.
  T uprev[N], comp[N], fsal[N], pl[P];
#pragma unroll
  for (int c = 0; c < N; ++c) uprev[c] = comp[c] = fsal[c] = T(0);
#pragma unroll
  for (int c = 0; c < P; ++c) pl[c] = T(0);
  R::template init_lane<T>(pl, ps);  // the wave-wide operands, for every lane
  double t = 0.0, tf = 0.0;
  T dt = T(1), qold = qoldinit, qold_pow = qoldinit_pow;
  int32_t nacc = 0, nrej = 0, cur = 0;
  StepSink<T, N> sk{};
  double ts_next = 0.0;  // save_time(cur)
  int64_t attempts = 0, traj = -1;
  bool active = false;

  int64_t qbase = 0, qlim = 0;
  bool drained = !a.persistent;
  const bool lead = lane_comp<R>() == 0;  // writes the per-trajectory scalars
  if (!a.persistent) {
    const int64_t slot =
        (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * SLOTS + (lane & (SLOTS - 1));
    if (slot < n) {
      traj = slot_traj(a, slot);
      active = traj >= 0;
    }
  }
  bool fresh = active;

  for (;;) {
    if (a.persistent) {
      bool want = !active && lead;
      if (a.refill_min > 1) {  // a batch of free slots, as in adaptive_body
        const int kmin = a.refill_min < SLOTS ? a.refill_min : SLOTS;
        if (__popcll(__ballot(want)) < kmin) want = false;
      }
      int64_t slot = wave_claim(a, want, qbase, qlim, drained);
      slot = (int64_t)__shfl((long long)slot, lane & (SLOTS - 1));  // from the slot's lead lane
      if (slot >= 0) {
        traj = slot_traj(a, slot);
        active = fresh = traj >= 0;
      }
    }
    if (__ballot(active) == 0ull) break;

    if (fresh) {  // start trajectory `traj` (gpuatsit5.jl:221-241)
      if (lead) note_visit(a, traj);
      load_state<T, R>(uprev, a.u0, n, traj);
#pragma unroll
      for (int c = 0; c < N; ++c) comp[c] = T(0);
      R::template load_lane<T>(pl, (const T*)a.p, ps, n, traj, lane_comp<R>());
      t = a.tspan ? a.tspan[traj] : a.t0;
      tf = a.tspan ? a.tspan[n + traj] : a.tf;
      dt = (T)(a.dt0 ? a.dt0[traj] : a.dt);
      qold = qoldinit;
      qold_pow = qoldinit_pow;
      nacc = nrej = 0;
      attempts = 0;
      cur = 0;
      if constexpr (SAVE == SAVE_AT) {
        if (a.n_saveat > 0 && t == a.saveat[0]) {
          store_save<T, R>(a, a.u_saveat, 0, a.n_saveat, traj, uprev);
          cur = 1;
        }
        ts_next = save_time(a, cur);
      }
      if constexpr (SAVE == SAVE_STEPS) {
        sink_start<R>(a, sk, traj);
        save_step<T, R>(a, sk, traj, t, uprev);
      }
    }
    if constexpr (ALG == ALG_ATSIT5) {  // FSAL seed, :241
      if (__ballot(fresh) != 0ull) {
        T k1[N];
        R::f(k1, uprev, pl, ps, t);
        if (fresh) copyv(fsal, k1);
      }
    }
    fresh = false;

    int ret = -1;  // -1: running
    bool run = false;
    if (active) {
      if (!(t < tf)) ret = RET_DEFAULT;
      else if ((double)dt < DTMIN) ret = RET_DTMIN;  // :256
      else if (attempts >= a.maxiters) ret = RET_MAXITERS;
      else run = true;
    }

    // ---- one attempt for every lane: stages, trial state u and the embedded error ----
    T u[N], err[N], comp_new[N];
    T k2[N], k3[N], k4[N], k5[N], k6[N], k7[N];  // ATsit5 stages (FSAL k1 = fsal)
    V7K<T, N> K7;                                 // AVern7 stages
    V9S<T, N> S9;                                 // AVern9 stages
    if constexpr (ALG == ALG_ATSIT5) {
      tsit5_stages<T, R, KAHAN>(uprev, fsal, k2, k3, k4, k5, k6, k7, u, comp, comp_new, pl, ps,
                                t, dt);
      // dt * (btilde1*k1 + ... + btilde7*k7), gpuatsit5.jl:272-275
      lsum(err, T(sdeh_tab::tsit5::btilde1), fsal, T(sdeh_tab::tsit5::btilde2), k2,
           T(sdeh_tab::tsit5::btilde3), k3, T(sdeh_tab::tsit5::btilde4), k4,
           T(sdeh_tab::tsit5::btilde5), k5, T(sdeh_tab::tsit5::btilde6), k6,
           T(sdeh_tab::tsit5::btilde7), k7);
    } else if constexpr (V9) {
      vern9_stages<T, R, true, KAHAN, SAVE == SAVE_AT>(uprev, S9, u, err, comp, comp_new, pl,
                                                        ps, t, dt, false);
    } else {
      vern7_stages<T, R, true, KAHAN>(uprev, K7, u, err, comp, comp_new, pl, ps, t, dt);
    }
#pragma unroll
    for (int c = 0; c < N; ++c) err[c] = dt * err[c];
    const T ssq = err_sumsq<T, R, NWL>(err, uprev, u, abstol, reltol);
    CtrlOut<T> co;
    if constexpr (Ops<T>::kFast) {
      co = controller<T, R, true>(ssq, dt, qold_pow, ck);
      if (run && !ctrl_in_range(ssq, dt)) co = controller<T, R, false>(ssq, dt, qold_pow, ck);
    } else {
      co = controller<T, R, false>(ssq, dt, qold_pow, ck);
    }
    const T EEst = co.EEst;

    // ---- PI controller: accept/reject (gpuatsit5.jl:279-299, gpuvern7.jl:406-426,
    // gpuvern9.jl:575-595) ----
    bool stepped = false;  // accepted: t has advanced, uprev not yet
    double told = 0.0;
    T dtold = T(0);
    if (run) {
      ++attempts;
      if (EEst != EEst) {
        ret = RET_NAN;
      } else if (EEst > T(1)) {
        dt = co.dtq;
        ++nrej;
      } else {
        qold = sdeh_math::fast_max(EEst, qoldinit);  // jl_max: EEst >= 0, not NaN
        qold_pow = (qold == EEst) ? Ops<T>::pw_from(co.lg, beta2) : qoldinit_pow;
        dtold = dt;
        const double rem = tf - t - (double)dtold;
        dt = (T)sdeh_math::fast_min((double)fabs(co.dtq), fabs(rem));
        told = t;
        t = (rem < SNAP) ? tf : t + (double)dtold;
        ++nacc;
        stepped = true;
      }
    }
    if constexpr (SAVE == SAVE_AT) {
      if constexpr (ALG == ALG_ATSIT5) {  // gpuatsit5.jl:305-316
        if (stepped) {
          while (cur < a.n_saveat && ts_next <= t) {
            const double ts_after = save_time(a, cur + 1);
            const T th = (T)((ts_next - told) / (double)dtold);
            T o[N];
            tsit5_dense(o, uprev, dtold, th, fsal, k2, k3, k4, k5, k6, k7);
            store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
            ++cur;
            ts_next = ts_after;
          }
        }
      } else {
        // Extra stages depend only on the step, not on θ: computed once, for the
        // whole wave when one of its lanes has a save point in the step.
        const bool saves = stepped && cur < a.n_saveat && ts_next <= t;
        if (__ballot(saves) != 0ull) {
          if constexpr (V9) {
            T x[10][N];
            vern9_extra<T, R>(uprev, S9, x, pl, ps, told, dtold);  // told + c*dtold
            if (saves) {
              // The weights depend on the point only: the G lanes of a trajectory take
              // G of its points at a time, lane q the weights of point cur + q, and
              // share them through LDS; every lane then forms its components of each
              // point in order.  Same values as vern9_dense point by point.
              constexpr int PB = R::G;
              T* wl = wave_lds<T, NWL>() + (lane & (SLOTS - 1)) * DWS;
              const int q = lane_comp<R>();
              while (cur < a.n_saveat && ts_next <= t) {
                const int64_t pq = cur + q;
                const double tq = pq < a.n_saveat ? a.saveat[pq] : ts_next;
                T w[V9W];
                vern9_weights((T)((tq - told) / (double)dtold), w);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();  // the previous batch's reads are done
#pragma unroll
                for (int i = 0; i < V9W; ++i) wl[q * V9W + i] = w[i];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                for (int m = 0; m < PB && cur < a.n_saveat && ts_next <= t; ++m) {
                  // the weights are read from LDS where the combination uses them
                  const T(&wm)[V9W] = *reinterpret_cast<const T(*)[V9W]>(wl + m * V9W);
                  T o[N];
                  vern9_combine(o, uprev, dtold, wm, S9, x);
                  store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
                  ++cur;
                  ts_next = save_time(a, cur);
                }
              }
            }
          } else {
            T x[6][N];
            // Quirk (gpuvern7.jl:449): stage times from the post-step t.
            vern7_extra<T, R>(uprev, K7, x, pl, ps, a.quirks ? t : told, dtold);
            if (saves) {
              while (cur < a.n_saveat && ts_next <= t) {
                const double ts_after = save_time(a, cur + 1);
                const T th = (T)((ts_next - told) / (double)dtold);
                T o[N];
                vern7_dense(o, uprev, dtold, th, K7, x);
                store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
                ++cur;
                ts_next = ts_after;
              }
            }
          }
        }
      }
    }
    if (stepped) {
      if constexpr (SAVE == SAVE_STEPS) save_step<T, R>(a, sk, traj, t, u);
      copyv(uprev, u);
      if constexpr (ALG == ALG_ATSIT5) copyv(fsal, k7);
      if constexpr (KAHAN) copyv(comp, comp_new);
      if (!(t < tf)) ret = RET_DEFAULT;
    }

    if (ret >= 0) {  // trajectory ends: write outputs (gpuatsit5.jl:322-329)
      if (SAVE == SAVE_STEPS && ret == RET_DEFAULT && sk.kept < sk.n) ret = RET_SAVE_OVERFLOW;
      store_state<T, R>(a.u_end, 0, n, traj, uprev);
      if constexpr (SAVE == SAVE_STEPS) flush_steps<T, R>(a, sk);
      if (lead) {
        a.t_end[traj] = t;
        a.retcode[traj] = ret;
        if (a.naccept) a.naccept[traj] = nacc;
        if (a.nreject) a.nreject[traj] = nrej;
        if constexpr (SAVE == SAVE_STEPS) finish_steps(a, sk, traj);
      }
      active = false;
    }
  }
}

// A staged solve's parked trajectory, one record per trajectory (layout and size
// SDEH_PARK_BYTES in sdeh_kargs.h): t, tf, counts and flag, then u, fsal, comp,
// p, dt and qold^beta2.
template <class T, class R, bool F, bool K>
struct ParkRec {
  static constexpr int N = R::N, P = R::P > 0 ? R::P : 1, NP = R::P;
  __device__ static char* at(const sdeh_kargs& a, int64_t traj) {
    return (char*)a.st_rec + traj * a.st_stride;
  }
  __device__ static int32_t flag(const sdeh_kargs& a, int64_t traj) {
    return ((const int32_t*)at(a, traj))[6];  // byte 24
  }
  __device__ static void set_flag(const sdeh_kargs& a, int64_t traj, int32_t f) {
    ((int32_t*)at(a, traj))[6] = f;
  }
  __device__ static void store(const sdeh_kargs& a, int64_t traj, const T (&u)[N],
                               const T (&fsal)[N], const T (&comp)[N], const T (&pl)[P],
                               double t, double tf, T dt, T qpow, int32_t nacc, int32_t nrej) {
    char* r = at(a, traj);
    ((double*)r)[0] = t;
    ((double*)r)[1] = tf;
    ((int32_t*)r)[4] = nacc;
    ((int32_t*)r)[5] = nrej;
    ((int32_t*)r)[6] = 1;
    T* v = (T*)(r + 32);
#pragma unroll
    for (int c = 0; c < N; ++c) *v++ = u[c];
    if constexpr (F) {
#pragma unroll
      for (int c = 0; c < N; ++c) *v++ = fsal[c];
    }
    if constexpr (K) {
#pragma unroll
      for (int c = 0; c < N; ++c) *v++ = comp[c];
    }
#pragma unroll
    for (int c = 0; c < NP; ++c) *v++ = pl[c];
    v[0] = dt;
    v[1] = qpow;
  }
  __device__ static void load(const sdeh_kargs& a, int64_t traj, T (&u)[N], T (&fsal)[N],
                              T (&comp)[N], T (&pl)[P], double& t, double& tf, T& dt, T& qpow,
                              int32_t& nacc, int32_t& nrej) {
    const char* r = at(a, traj);
    t = ((const double*)r)[0];
    tf = ((const double*)r)[1];
    nacc = ((const int32_t*)r)[4];
    nrej = ((const int32_t*)r)[5];
    const T* v = (const T*)(r + 32);
#pragma unroll
    for (int c = 0; c < N; ++c) u[c] = *v++;
    if constexpr (F) {
#pragma unroll
      for (int c = 0; c < N; ++c) fsal[c] = *v++;
    }
#pragma unroll
    for (int c = 0; c < N; ++c) comp[c] = T(0);
    if constexpr (K) {
#pragma unroll
      for (int c = 0; c < N; ++c) comp[c] = *v++;
    }
#pragma unroll
    for (int c = 0; c < NP; ++c) pl[c] = *v++;
    dt = v[0];
    qpow = v[1];
  }
};

template <class T, class R, int ALG, int SAVE, bool KAHAN>
__device__ __forceinline__ void adaptive_body(const sdeh_kargs& a) {
  if constexpr (IsWave<R>::value) {
    adaptive_body_wave<T, R, ALG, SAVE, KAHAN>(a);
    return;
  }
  constexpr int N = R::N;
  constexpr int P = R::P > 0 ? R::P : 1;
  constexpr double DTMIN = (ALG == ALG_AVERN9) ? (double)1.0e-7f : 1.0e-14;  // gpuvern9.jl:466
  constexpr double SNAP = DTMIN;                                              // gpuvern9.jl:591
  // build_adaptive_controller_cache(T), src/SimpleDiffEq.jl:67-77
  const T beta1 = T(7.0 / 50), beta2 = T(2.0 / 25), qmax = T(10.0), qmin = T(1.0 / 5),
          gamma = T(9.0 / 10), qoldinit = T(1.0e-4);
  const T inv_qmax = T(1) / qmax, inv_qmin = T(1) / qmin;
  const T qoldinit_pow = Ops<T>::pw(qoldinit, beta2);
  constexpr bool PIN = sizeof(T) == 8 && SAVE == SAVE_END;
  const T abstol = in_vgpr_if<PIN>((T)a.abstol), reltol = in_vgpr_if<PIN>((T)a.reltol);
  const CtrlK<T> ck{beta1, inv_qmax, inv_qmin, gamma, Ops<T>::rcp_nr(T(NormLen<R>::value)),
                    Ops<T>::rcp_nr(gamma)};
  const T* ps = (const T*)a.ps;
  const int64_t n = a.n;

  // Lane state.
  T uprev[N], comp[N], fsal[N], pl[P];
  double t = 0.0, tf = 0.0;
  T dt = T(0), qold = T(0), qold_pow = T(0);
  int32_t nacc = 0, nrej = 0, cur = 0;
  StepSink<T, N> sk{};
  double ts_next = 0.0;  // save_time(cur)
  int64_t attempts = 0, traj = -1;
  bool active = false;

  int64_t qbase = 0, qlim = 0;
  bool drained = !a.persistent;
  // the persistent queue's cache of hand-out order entries (wave_claim_win): in the
  // binary32 one-lane endpoint kernels, whose occupancy is set by scalar registers, so
  // the vector registers it takes cost no wave (the binary64 and save modes' kernels
  // would lose one)
  constexpr bool ORDER_WIN = sizeof(T) == 4 && R::G == 1 && SAVE == SAVE_END;
  int64_t qwin = -128, qword = -1;
  const bool lead = lane_comp<R>() == 0;  // writes the per-trajectory scalars
  if (!a.persistent) {
    const int64_t slot = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / R::G;
    if (slot < n) {
      traj = slot_traj(a, slot);
      active = traj >= 0;
    }
  }
  bool fresh = active;
  // Staging (stage_cap / resume) applies to one lane per trajectory with endpoint
  // output.
  constexpr bool STAGED = SAVE == SAVE_END && R::G == 1;
  using Park = ParkRec<T, R, ALG == ALG_ATSIT5, KAHAN>;
  const int64_t att_lim = in_vgpr_if<PIN>((STAGED && a.stage_cap > 0 && a.stage_cap < a.maxiters)
                                              ? (int64_t)a.stage_cap
                                              : a.maxiters);

  for (;;) {
    if (a.persistent) {
      bool want = !active && lead;
      // a batch of at least refill_min free slots (at most the wave's slot count, so
      // that a wave whose lanes are all free always claims)
      if (a.refill_min > 1) {
        const int kmin = a.refill_min < 64 / R::G ? a.refill_min : 64 / R::G;
        if (__popcll(__ballot(want)) < kmin) want = false;
      }
      for (;;) {
        int64_t slot, claimed = -1;
        if constexpr (ORDER_WIN)
          slot = wave_claim_win(a, want, qbase, qlim, drained, qwin, qword, claimed);
        else
          slot = wave_claim(a, want, qbase, qlim, drained);
        if constexpr (R::G > 1)  // the group leader claimed; its group follows
          slot = (int64_t)__shfl((long long)slot, (int)(__lane_id() & ~(unsigned)(R::G - 1)));
        if (slot >= 0) {
          traj = ORDER_WIN ? claimed : slot_traj(a, slot);
          active = fresh = traj >= 0;
          // a resumed launch skips the trajectories that ended in the first stage
          if (STAGED && a.resume && active && Park::flag(a, traj) == 0) active = fresh = false;
        }
        want = !active && lead;
        if (!STAGED || !a.resume || drained || __ballot(want && slot >= 0) == 0ull) break;
      }
    }
    if (__ballot(active) == 0ull) break;
    if (!active) continue;

    if (fresh && STAGED && a.resume) {  // continue a parked trajectory
      fresh = false;
      note_visit(a, traj);
      Park::load(a, traj, uprev, fsal, comp, pl, t, tf, dt, qold_pow, nacc, nrej);
      qold = qoldinit;  // written before it is read again
      attempts = (int64_t)nacc + nrej;
      cur = 0;
    }
    if (fresh) {  // start trajectory `traj` (gpuatsit5.jl:221-241)
      fresh = false;
      if (lead) note_visit(a, traj);
      load_state<T, R>(uprev, a.u0, n, traj);
#pragma unroll
      for (int c = 0; c < N; ++c) comp[c] = T(0);
      load_params<T, R>(pl, a, traj);
      t = a.tspan ? a.tspan[traj] : a.t0;
      tf = a.tspan ? a.tspan[n + traj] : a.tf;
      dt = (T)(a.dt0 ? a.dt0[traj] : a.dt);
      qold = qoldinit;
      qold_pow = qoldinit_pow;
      nacc = nrej = 0;
      attempts = 0;
      cur = 0;
      if constexpr (SAVE == SAVE_AT) {
        if (a.n_saveat > 0 && t == a.saveat[0]) {
          store_save<T, R>(a, a.u_saveat, 0, a.n_saveat, traj, uprev);
          cur = 1;
        }
        ts_next = save_time(a, cur);
      }
      if constexpr (SAVE == SAVE_STEPS) {
        sink_start<R>(a, sk, traj);
        save_step<T, R>(a, sk, traj, t, uprev);
      }
      if constexpr (ALG == ALG_ATSIT5) R::f(fsal, uprev, pl, ps, t);  // FSAL seed, :241
    }

    int ret = -1;  // -1: running
    if (!(t < tf)) {
      ret = RET_DEFAULT;
    } else if ((double)dt < DTMIN) {  // :256
      ret = RET_DTMIN;
    } else if (attempts >= att_lim) {  // maxiters, or the first stage's cap (parked)
      ret = attempts >= a.maxiters ? RET_MAXITERS : RET_PARKED;
    } else {
      ++attempts;
      T u[N], err[N], comp_new[N];
      // ---- one attempt: stages, trial state u and the embedded error ----
      constexpr bool V9 = (ALG == ALG_AVERN9);
      T k2[N], k3[N], k4[N], k5[N], k6[N], k7[N];  // ATsit5 stages (FSAL k1 = fsal)
      V7K<T, N> K7;                                 // AVern7 stages
      V9S<T, N> S9;                                 // AVern9 stages
      if constexpr (ALG == ALG_ATSIT5) {
        tsit5_stages<T, R, KAHAN>(uprev, fsal, k2, k3, k4, k5, k6, k7, u, comp, comp_new, pl,
                                  ps, t, dt);
        // dt * (btilde1*k1 + ... + btilde7*k7), gpuatsit5.jl:272-275
        lsum(err, T(sdeh_tab::tsit5::btilde1), fsal, T(sdeh_tab::tsit5::btilde2), k2,
             T(sdeh_tab::tsit5::btilde3), k3, T(sdeh_tab::tsit5::btilde4), k4,
             T(sdeh_tab::tsit5::btilde5), k5, T(sdeh_tab::tsit5::btilde6), k6,
             T(sdeh_tab::tsit5::btilde7), k7);
      } else if constexpr (V9) {
        // Verner: k1 is evaluated at the start of every attempt (gpuvern7.jl:360,
        // gpuvern9.jl:468).
        vern9_stages<T, R, true, KAHAN, SAVE == SAVE_AT>(uprev, S9, u, err, comp, comp_new, pl,
                                                          ps, t, dt, false);
      } else {
        vern7_stages<T, R, true, KAHAN>(uprev, K7, u, err, comp, comp_new, pl, ps, t, dt);
      }
#pragma unroll
      for (int c = 0; c < N; ++c) err[c] = dt * err[c];
      CtrlOut<T> co;
      // The range-checked controller runs for every lane; the IEEE one again only
      // where the check fails (a branch the wave skips when no lane needs it).
      // Measured on the 2^20-trajectory Lorenz ATsit5 solve: 1.135 ms, against 1.27
      // ms for an if/else on the check and 1.19 ms with IEEE operations only.
      // One lane per trajectory: under tolerances that allow it (a.err_bare, the
      // same for every lane) the error norm's divisions are range-checked too, by the
      // same check on the sum they give; a bare sum out of range is recomputed with
      // IEEE divisions.
      if constexpr (Ops<T>::kFast && R::G == 1) {
        T ssq;
        if (a.err_bare) ssq = err_sumsq_in<T, R>(err, uprev, u, abstol, reltol);
        else ssq = err_sumsq<T, R>(err, uprev, u, abstol, reltol);
        const bool in = ctrl_in_range(ssq, dt);
        if (!in && a.err_bare) ssq = err_sumsq<T, R>(err, uprev, u, abstol, reltol);
        co = controller<T, R, true>(ssq, dt, qold_pow, ck);
        if (!in) co = controller<T, R, false>(ssq, dt, qold_pow, ck);
      } else if constexpr (Ops<T>::kFast) {
        const T ssq = err_sumsq<T, R>(err, uprev, u, abstol, reltol);
        co = controller<T, R, true>(ssq, dt, qold_pow, ck);
        if (!ctrl_in_range(ssq, dt)) co = controller<T, R, false>(ssq, dt, qold_pow, ck);
      } else {
        const T ssq = err_sumsq<T, R>(err, uprev, u, abstol, reltol);
        co = controller<T, R, false>(ssq, dt, qold_pow, ck);
      }
      const T EEst = co.EEst;
      if (EEst != EEst) {
        ret = RET_NAN;
      } else {
        // ---- PI controller: accept/reject (gpuatsit5.jl:279-299,
        // gpuvern7.jl:406-426, gpuvern9.jl:575-595) ----
        const bool accept = !(EEst > T(1));
        const T dtq = co.dtq;
        if (!accept) {
          dt = dtq;
          ++nrej;
        } else {
          qold = sdeh_math::fast_max(EEst, qoldinit);  // jl_max: EEst >= 0, not NaN
          qold_pow = (qold == EEst) ? Ops<T>::pw_from(co.lg, beta2) : qoldinit_pow;
          const T dtold = dt;
          const double rem = tf - t - (double)dtold;
          // Julia's min: neither argument is NaN here (dtq is dt / q with q in
          // [1/qmax, 1/qmin]; t and tf are finite or tf = +Inf), so one select.
          dt = (T)sdeh_math::fast_min((double)fabs(dtq), fabs(rem));
          const double told = t;
          t = (rem < SNAP) ? tf : t + (double)dtold;
          ++nacc;
          if constexpr (SAVE == SAVE_AT) {
            if constexpr (ALG == ALG_ATSIT5) {  // gpuatsit5.jl:305-316
              while (cur < a.n_saveat && ts_next <= t) {
                const double ts_after = save_time(a, cur + 1);
                const T th = (T)((ts_next - told) / (double)dtold);
                T o[N];
                tsit5_dense(o, uprev, dtold, th, fsal, k2, k3, k4, k5, k6, k7);
                store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
                ++cur;
                ts_next = ts_after;
              }
            } else if (cur < a.n_saveat && ts_next <= t) {
              // Extra stages depend only on the step, not on θ: computed once.
              if constexpr (V9) {
                T x[10][N];
                vern9_extra<T, R>(uprev, S9, x, pl, ps, told, dtold);  // told + c*dtold
                while (cur < a.n_saveat && ts_next <= t) {
                  const double ts_after = save_time(a, cur + 1);
                  const T th = (T)((ts_next - told) / (double)dtold);
                  T o[N];
                  vern9_dense(o, uprev, dtold, th, S9, x);
                  store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
                  ++cur;
                  ts_next = ts_after;
                }
              } else {
                T x[6][N];
                // Quirk (gpuvern7.jl:449): stage times from the post-step t.
                vern7_extra<T, R>(uprev, K7, x, pl, ps, a.quirks ? t : told, dtold);
                while (cur < a.n_saveat && ts_next <= t) {
                  const double ts_after = save_time(a, cur + 1);
                  const T th = (T)((ts_next - told) / (double)dtold);
                  T o[N];
                  vern7_dense(o, uprev, dtold, th, K7, x);
                  store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
                  ++cur;
                  ts_next = ts_after;
                }
              }
            }
          }
          if constexpr (SAVE == SAVE_STEPS) save_step<T, R>(a, sk, traj, t, u);
          copyv(uprev, u);
          if constexpr (ALG == ALG_ATSIT5) copyv(fsal, k7);
          if constexpr (KAHAN) copyv(comp, comp_new);
          if (!(t < tf)) ret = RET_DEFAULT;
        }
      }
    }

    if (ret != -1) {  // trajectory ends or parks: write outputs (gpuatsit5.jl:322-329)
      if (SAVE == SAVE_STEPS && ret == RET_DEFAULT && sk.kept < sk.n) ret = RET_SAVE_OVERFLOW;
      const bool parked = STAGED && ret == RET_PARKED;
      if (!parked) {
        store_state<T, R>(a.u_end, 0, n, traj, uprev);
        if constexpr (SAVE == SAVE_STEPS) flush_steps<T, R>(a, sk);
        if (lead) {
          a.t_end[traj] = t;
          a.retcode[traj] = ret;
          if (a.naccept) a.naccept[traj] = nacc;
          if (a.nreject) a.nreject[traj] = nrej;
          if constexpr (SAVE == SAVE_STEPS) finish_steps(a, sk, traj);
        }
      }
      if (STAGED && a.stage_cap > 0) {
        double key = 0.0;
        if (parked) {
          Park::store(a, traj, uprev, fsal, comp, pl, t, tf, dt, qold_pow, nacc, nrej);
          // remaining attempts extrapolated from the first stage's progress
          const double t0 = a.tspan ? a.tspan[traj] : a.t0;
          key = (double)a.stage_cap * ((tf - t) / (t - t0));
        } else {
          Park::set_flag(a, traj, 0);
        }
        a.st_key[traj] = key;
      }
      active = false;
    }
  }
}

// ---------------------------------------------------------------------------
// Several trajectories per lane (TPL slots; one lane per trajectory's state,
// endpoint output, ATsit5 and AVern7).  Every iteration runs one attempt of every
// slot: the slots' stage and controller arithmetic sits in one straight-line block,
// so the compiler interleaves the independent chains (instruction-level parallelism
// in place of waves), and then each slot that attempted takes its own decision and
// bookkeeping exactly as adaptive_body does, with the same arithmetic: a
// trajectory's results do not depend on its slot or neighbour.  A slot without a
// running trajectory computes on its last values (all finite or not, results
// dropped) until it is given a new one; only attempting slots can send the wave to
// the IEEE controller.  One launch: lane g takes trajectories TPL g .. TPL g + TPL-1
// (of the hand-out order); persistent: each empty slot claims from the queue.
// ---------------------------------------------------------------------------
template <class T, class R, int ALG, bool KAHAN, int TPL>
__device__ __forceinline__ void adaptive_body_multi(const sdeh_kargs& a) {
  static_assert(R::G == 1 && !IsWave<R>::value, "one lane per trajectory state");
  static_assert(ALG == ALG_ATSIT5 || ALG == ALG_AVERN7, "ATsit5 and AVern7");
  constexpr int N = R::N;
  constexpr int P = R::P > 0 ? R::P : 1;
  constexpr double DTMIN = 1.0e-14, SNAP = 1.0e-14;  // gpuatsit5.jl:256,294; gpuvern7.jl:383,421
  // build_adaptive_controller_cache(T), src/SimpleDiffEq.jl:67-77
  const T beta1 = T(7.0 / 50), beta2 = T(2.0 / 25), qmax = T(10.0), qmin = T(1.0 / 5),
          gamma = T(9.0 / 10), qoldinit = T(1.0e-4);
  const T inv_qmax = T(1) / qmax, inv_qmin = T(1) / qmin;
  const T qoldinit_pow = Ops<T>::pw(qoldinit, beta2);
  const T abstol = (T)a.abstol, reltol = (T)a.reltol;
  const CtrlK<T> ck{beta1, inv_qmax, inv_qmin, gamma, Ops<T>::rcp_nr(T(NormLen<R>::value)),
                    Ops<T>::rcp_nr(gamma)};
  const T* ps = (const T*)a.ps;
  const int64_t n = a.n;

  T uprev[TPL][N], comp[TPL][N], fsal[TPL][N], pl[TPL][P];
  double t[TPL], tf[TPL];
  T dt[TPL], qold_pow[TPL];
  int32_t nacc[TPL], nrej[TPL];
  int64_t attempts[TPL], traj[TPL];
  bool active[TPL], fresh[TPL];
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
#pragma unroll
    for (int c = 0; c < N; ++c) uprev[k][c] = comp[k][c] = fsal[k][c] = T(0);
#pragma unroll
    for (int c = 0; c < P; ++c) pl[k][c] = T(0);
    t[k] = tf[k] = 0.0;
    dt[k] = T(1);
    qold_pow[k] = qoldinit_pow;
    nacc[k] = nrej[k] = 0;
    attempts[k] = 0;
    traj[k] = -1;
    active[k] = fresh[k] = false;
  }
  int64_t qbase = 0, qlim = 0;
  bool drained = !a.persistent;
  if (!a.persistent) {
    const int64_t base = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * TPL;
#pragma unroll
    for (int k = 0; k < TPL; ++k)
      if (base + k < n) {
        traj[k] = slot_traj(a, base + k);
        active[k] = fresh[k] = traj[k] >= 0;
      }
  }

  for (;;) {
    if (a.persistent) {
#pragma unroll
      for (int k = 0; k < TPL; ++k) {
        const int64_t slot = wave_claim(a, !active[k], qbase, qlim, drained);
        if (slot >= 0) {
          traj[k] = slot_traj(a, slot);
          active[k] = fresh[k] = traj[k] >= 0;
        }
      }
    }
    bool any = false;
#pragma unroll
    for (int k = 0; k < TPL; ++k) any = any || active[k];
    if (__ballot(any) == 0ull) break;

#pragma unroll
    for (int k = 0; k < TPL; ++k)
      if (fresh[k]) {  // start trajectory traj[k] (gpuatsit5.jl:221-241)
        fresh[k] = false;
        note_visit(a, traj[k]);
        load_state<T, R>(uprev[k], a.u0, n, traj[k]);
#pragma unroll
        for (int c = 0; c < N; ++c) comp[k][c] = T(0);
        load_params<T, R>(pl[k], a, traj[k]);
        t[k] = a.tspan ? a.tspan[traj[k]] : a.t0;
        tf[k] = a.tspan ? a.tspan[n + traj[k]] : a.tf;
        dt[k] = (T)(a.dt0 ? a.dt0[traj[k]] : a.dt);
        qold_pow[k] = qoldinit_pow;
        nacc[k] = nrej[k] = 0;
        attempts[k] = 0;
        if constexpr (ALG == ALG_ATSIT5) R::f(fsal[k], uprev[k], pl[k], ps, t[k]);  // :241
      }

    int ret[TPL];
    bool go[TPL];
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
      ret[k] = -1;
      go[k] = false;
      if (active[k]) {
        if (!(t[k] < tf[k])) ret[k] = RET_DEFAULT;
        else if ((double)dt[k] < DTMIN) ret[k] = RET_DTMIN;  // :256
        else if (attempts[k] >= a.maxiters) ret[k] = RET_MAXITERS;
        else {
          go[k] = true;
          ++attempts[k];
        }
      }
    }
    // ---- one attempt of every slot: stages, trial state and embedded error ----
    T u[TPL][N], err[TPL][N], comp_new[TPL][N];
    T k2[TPL][N], k3[TPL][N], k4[TPL][N], k5[TPL][N], k6[TPL][N], k7[TPL][N];
    V7K<T, N> K7[TPL];
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
      if constexpr (ALG == ALG_ATSIT5) {
        tsit5_stages<T, R, KAHAN>(uprev[k], fsal[k], k2[k], k3[k], k4[k], k5[k], k6[k], k7[k],
                                  u[k], comp[k], comp_new[k], pl[k], ps, t[k], dt[k]);
        lsum(err[k], T(sdeh_tab::tsit5::btilde1), fsal[k], T(sdeh_tab::tsit5::btilde2), k2[k],
             T(sdeh_tab::tsit5::btilde3), k3[k], T(sdeh_tab::tsit5::btilde4), k4[k],
             T(sdeh_tab::tsit5::btilde5), k5[k], T(sdeh_tab::tsit5::btilde6), k6[k],
             T(sdeh_tab::tsit5::btilde7), k7[k]);
      } else {
        vern7_stages<T, R, true, KAHAN>(uprev[k], K7[k], u[k], err[k], comp[k], comp_new[k],
                                        pl[k], ps, t[k], dt[k]);
      }
    }
    CtrlOut<T> co[TPL];
    T ssq[TPL];
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
#pragma unroll
      for (int c = 0; c < N; ++c) err[k][c] = dt[k] * err[k][c];
      ssq[k] = err_sumsq<T, R>(err[k], uprev[k], u[k], abstol, reltol);
      if constexpr (Ops<T>::kFast) co[k] = controller<T, R, true>(ssq[k], dt[k], qold_pow[k], ck);
      else co[k] = controller<T, R, false>(ssq[k], dt[k], qold_pow[k], ck);
    }
    if constexpr (Ops<T>::kFast) {
#pragma unroll
      for (int k = 0; k < TPL; ++k)  // the IEEE controller again where the check fails
        if (go[k] && !ctrl_in_range(ssq[k], dt[k]))
          co[k] = controller<T, R, false>(ssq[k], dt[k], qold_pow[k], ck);
    }
    // ---- each attempting slot's decision (gpuatsit5.jl:279-299, gpuvern7.jl:406-426) ----
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
      if (!go[k]) continue;
      const T EEst = co[k].EEst;
      if (EEst != EEst) {
        ret[k] = RET_NAN;
        continue;
      }
      if (EEst > T(1)) {
        dt[k] = co[k].dtq;
        ++nrej[k];
        continue;
      }
      const T qold = sdeh_math::fast_max(EEst, qoldinit);
      qold_pow[k] = (qold == EEst) ? Ops<T>::pw_from(co[k].lg, beta2) : qoldinit_pow;
      const T dtold = dt[k];
      const double rem = tf[k] - t[k] - (double)dtold;
      dt[k] = (T)sdeh_math::fast_min((double)fabs(co[k].dtq), fabs(rem));
      t[k] = (rem < SNAP) ? tf[k] : t[k] + (double)dtold;
      ++nacc[k];
      copyv(uprev[k], u[k]);
      if constexpr (ALG == ALG_ATSIT5) copyv(fsal[k], k7[k]);
      if constexpr (KAHAN) copyv(comp[k], comp_new[k]);
      if (!(t[k] < tf[k])) ret[k] = RET_DEFAULT;
    }
#pragma unroll
    for (int k = 0; k < TPL; ++k)
      if (ret[k] != -1) {  // the trajectory ends: outputs (gpuatsit5.jl:322-329)
        store_state<T, R>(a.u_end, 0, n, traj[k], uprev[k]);
        a.t_end[traj[k]] = t[k];
        a.retcode[traj[k]] = ret[k];
        if (a.naccept) a.naccept[traj[k]] = nacc[k];
        if (a.nreject) a.nreject[traj[k]] = nrej[k];
        active[k] = false;
      }
  }
}

// ---------------------------------------------------------------------------
// Fixed-step solvers: Tsit5, Vern7, Vern9 over t0:dt:tf.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double grid_elem(const sdeh_kargs& a, int64_t i) {
  if (a.g_rational) return (double)(a.g_start_n + i * a.g_step_n) / (double)a.g_den;
  const double di = (double)i;
  const double hi = di * a.g_step;
  const double lo = fma(di, a.g_step, -hi);
  const double s = a.g_start + hi;
  const double bb = s - a.g_start;
  const double e = (a.g_start - (s - bb)) + (hi - bb);
  return s + (e + lo);
}

template <class T, class R, int ALG, int SAVE>
__device__ __forceinline__ void fixed_body(const sdeh_kargs& a) {
  static_assert(!IsWave<R>::value, "fixed-step solvers run the lane and group layouts");
  constexpr int N = R::N;
  constexpr int P = R::P > 0 ? R::P : 1;
  const int64_t slot = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / R::G;
  if (slot >= a.n) return;  // whole groups leave together
  const bool lead = lane_comp<R>() == 0;
  const int64_t n = a.n;
  const int64_t traj = slot_traj(a, slot);
  if (traj < 0) return;  // the group's slot names no trajectory
  if (lead) note_visit(a, traj);
  const T* ps = (const T*)a.ps;
  T u[N], uprev[N], pl[P], zero[N], dummy[N], err[N];
  load_state<T, R>(u, a.u0, n, traj);
#pragma unroll
  for (int c = 0; c < N; ++c) zero[c] = T(0);
  load_params<T, R>(pl, a, traj);
  const T dt = (T)a.dt;
  double t = a.t0;
  int32_t cur = 0;
  StepSink<T, N> sk{};
  if constexpr (SAVE == SAVE_STEPS) sink_start<R>(a, sk, traj);
  double ts_next = 0.0;  // save_time(cur)
  if constexpr (SAVE == SAVE_AT) {
    if (a.n_saveat > 0 && t == a.saveat[0]) {
      store_save<T, R>(a, a.u_saveat, 0, a.n_saveat, traj, u);
      cur = 1;
    }
    ts_next = save_time(a, cur);
  }
  if constexpr (SAVE == SAVE_STEPS) save_step<T, R>(a, sk, traj, t, u);
  T k7[N];
  if constexpr (ALG == ALG_TSIT5) R::f(k7, u, pl, ps, t);  // gpuatsit5.jl:86
  for (int64_t i = 1; i < a.g_len; ++i) {
    copyv(uprev, u);
    t = grid_elem(a, i - 1);  // gpuatsit5.jl:98
    if constexpr (ALG == ALG_RK4 || ALG == ALG_EULER) {
      // gpurk4.jl:75-82 / gpueuler.jl:73-75 evaluate f from t = ts[i], the END of
      // the step (quirk); with quirks off, from its start ts[i-1].
      const double t_next = grid_elem(a, i);
      const double te = a.quirks ? t_next : t;
      T k1[N];
      R::f(k1, u, pl, ps, te);
      if constexpr (ALG == ALG_RK4) {
        const T half = T(1) / T(2), sixth = T(1) / T(6);  // gpurk4.jl:70-71
        T tmp[N], k2[N], k3[N], k4[N];
        comb1(tmp, uprev, dt * half, k1);
        R::f(k2, tmp, pl, ps, stage_t(te, (double)half, (double)dt));
        comb1(tmp, uprev, dt * half, k2);
        R::f(k3, tmp, pl, ps, stage_t(te, (double)half, (double)dt));
        comb1(tmp, uprev, dt, k3);
        R::f(k4, tmp, pl, ps, te + (double)dt);
        // u = uprev + dt*sixth*(k1 + 2k2 + 2k3 + k4), the sum left to right
        // (the doublings are exact)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          const T s4 = ((k1[c] + T(2) * k2[c]) + T(2) * k3[c]) + k4[c];
          u[c] = fma(dt * sixth, s4, uprev[c]);
        }
      } else {
        comb1(u, uprev, dt, k1);
      }
      t = t_next;  // us[i] belongs to ts[i]
    } else if constexpr (ALG == ALG_TSIT5) {
      T k1[N], k2[N], k3[N], k4[N], k5[N], k6[N];
      copyv(k1, k7);
      tsit5_stages<T, R, false>(uprev, k1, k2, k3, k4, k5, k6, k7, u, zero, dummy, pl, ps, t, dt);
      t += (double)dt;
      if constexpr (SAVE == SAVE_AT) {  // :116-127
        while (cur < a.n_saveat && ts_next <= t) {
          const double ts_after = save_time(a, cur + 1);
          const T th = (T)((ts_next - (t - (double)dt)) / (double)dt);
          T o[N];
          tsit5_dense(o, uprev, dt, th, k1, k2, k3, k4, k5, k6, k7);
          store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
          ++cur;
          ts_next = ts_after;
        }
      }
    } else if constexpr (ALG == ALG_VERN7) {
      V7K<T, N> K;
      vern7_stages<T, R, false, false>(uprev, K, u, err, zero, dummy, pl, ps, t, dt);
      const double told = t;
      t += (double)dt;  // gpuvern7.jl:138
      if constexpr (SAVE == SAVE_AT) {
        if (cur < a.n_saveat && ts_next <= t) {
          T x[6][N];
          vern7_extra<T, R>(uprev, K, x, pl, ps, a.quirks ? t : told, dt);  // quirk :160
          while (cur < a.n_saveat && ts_next <= t) {
            const double ts_after = save_time(a, cur + 1);
            const T th = (T)((ts_next - (t - (double)dt)) / (double)dt);
            T o[N];
            vern7_dense(o, uprev, dt, th, K, x);
            store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
            ++cur;
            ts_next = ts_after;
          }
        }
      }
    } else {
      V9S<T, N> S;
      vern9_stages<T, R, false, false, SAVE == SAVE_AT>(uprev, S, u, err, zero, dummy, pl, ps, t,
                                                         dt, a.quirks != 0);
      const double told = t;
      t += (double)dt;  // gpuvern9.jl:185
      if constexpr (SAVE == SAVE_AT) {
        if (cur < a.n_saveat && ts_next <= t) {
          T x[10][N];
          vern9_extra<T, R>(uprev, S, x, pl, ps, a.quirks ? t : told, dt);  // quirk :222
          while (cur < a.n_saveat && ts_next <= t) {
            const double ts_after = save_time(a, cur + 1);
            const T th = (T)((ts_next - (t - (double)dt)) / (double)dt);
            T o[N];
            vern9_dense(o, uprev, dt, th, S, x);
            store_save<T, R>(a, a.u_saveat, cur, a.n_saveat, traj, o);
            ++cur;
            ts_next = ts_after;
          }
        }
      }
    }
    if constexpr (SAVE == SAVE_STEPS) save_step<T, R>(a, sk, traj, t, u);
  }
  store_state<T, R>(a.u_end, 0, n, traj, u);
  if constexpr (SAVE == SAVE_STEPS) flush_steps<T, R>(a, sk);
  if (!lead) return;
  a.t_end[traj] = t;
  int ret = RET_DEFAULT;
  if (SAVE == SAVE_STEPS && sk.kept < sk.n) ret = RET_SAVE_OVERFLOW;
  a.retcode[traj] = ret;
  if (a.naccept) a.naccept[traj] = (int32_t)(a.g_len > 0 ? a.g_len - 1 : 0);
  if (a.nreject) a.nreject[traj] = 0;
  if constexpr (SAVE == SAVE_STEPS) finish_steps(a, sk, traj);
}

}  // namespace sdeh

namespace sdeh {
// Waves per SIMD the compiler is asked to keep resident.  The binary64 adaptive
// kernels of small states with endpoint output need a few VGPRs more than the 96
// that five waves allow when left alone (ATsit5/Lorenz: 101-103); held to 96 they
// lose nothing and run 6% faster (five waves instead of four).  The wave-layout
// (matrix-core) kernels with saveat output take 328 registers left alone (one wave
// per SIMD); held to two waves they spill 31 registers to scratch and still run 22%
// faster (GLV-16 AVern9, 2^17 trajectories, 256 points: 10.8 against 13.8 ms).
// Other kernels keep the compiler's choice (1 = no minimum).
template <class T, class R, int ALG, int SAVE>
struct WavesHint {
  static constexpr int value =
      IsWave<R>::value ? (SAVE == SAVE_AT ? 2 : 1)
                       : ((sizeof(T) == 8 && R::N * R::G <= 4 && SAVE == SAVE_END &&
                           ALG == ALG_ATSIT5) ? 5 : 1);
};
}  // namespace sdeh

// Kernel entry templates.  Instantiated with extern "C" names by the
// ahead-of-time translation unit and by the hipRTC wrapper.  WAVES > 0 asks the
// compiler for that many waves per SIMD instead of the WavesHint default (at most
// 512 / WAVES VGPRs per lane, in steps of 8): the autotuner's register-cap
// variants of the same arithmetic.
template <class T, class R, int ALG, int SAVE, bool KAHAN, int WAVES = 0>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES
                                                 : sdeh::WavesHint<T, R, ALG, SAVE>::value)))
    sdeh_adaptive_kernel(sdeh_kargs a) {
  sdeh::adaptive_body<T, R, ALG, SAVE, KAHAN>(a);
}

// TPL trajectories per lane (adaptive_body_multi).
template <class T, class R, int ALG, bool KAHAN, int TPL>
__global__ void __launch_bounds__(256) sdeh_adaptive_multi_kernel(sdeh_kargs a) {
  sdeh::adaptive_body_multi<T, R, ALG, KAHAN, TPL>(a);
}

template <class T, class R, int ALG, int SAVE>
__global__ void __launch_bounds__(256) sdeh_fixed_kernel(sdeh_kargs a) {
  sdeh::fixed_body<T, R, ALG, SAVE>(a);
}
