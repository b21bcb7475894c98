// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// C entry points of the CPU restatement, loaded with ctypes by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline.  The ensemble loop
// splits [0, n_traj) into contiguous per-thread chunks, the batching
// SciMLBase's EnsembleThreads uses, and runs one reference-style solve per
// trajectory.  Buffers follow the library's layout (simplediffeq_hip.h),
// host memory only.
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "oracle.hpp"

using namespace oracle;

namespace {

template <class T, class F>
int run_ensemble(const sdeh_solve_desc* d, const sdeh_io* io, const Mode& md, int nthreads) {
  constexpr int N = F::N, P = F::P;
  const int64_t n = d->n_traj;
  const T* u0 = (const T*)io->u0;
  const T* pp = (const T*)io->p;
  F f;
  f.ps = (const T*)io->p_shared;
  std::atomic<int> any_dtmin{0}, any_overflow{0};
  const bool saveat = d->saveat != nullptr && d->n_saveat > 0;
  const bool everystep = !saveat && d->save_everystep;

  auto work = [&](int64_t lo, int64_t hi) {
    T u0l[N > 0 ? N : 1], pl[P > 0 ? P : 1], uend[N > 0 ? N : 1];
    std::vector<T> us(saveat ? (size_t)(d->n_saveat * N) : 0);
    std::vector<double> st;
    std::vector<T> su;
    for (int64_t i = lo; i < hi; ++i) {
      for (int c = 0; c < N; ++c) u0l[c] = u0[c * n + i];
      for (int c = 0; c < P; ++c) pl[c] = pp[c * n + i];
      Traj<T> tr;
      tr.u0 = u0l;
      tr.p = pl;
      tr.t0 = io->tspan ? io->tspan[i] : d->t0;
      tr.tf = io->tspan ? io->tspan[n + i] : d->tf;
      tr.dt = io->dt0 ? io->dt0[i] : d->dt;
      tr.abstol = (T)d->abstol;
      tr.reltol = (T)d->reltol;
      tr.saveat = saveat ? d->saveat : nullptr;
      tr.n_saveat = saveat ? d->n_saveat : 0;
      tr.save_everystep = d->save_everystep;
      tr.u_end = uend;
      if (saveat) {
        std::fill(us.begin(), us.end(), std::numeric_limits<T>::quiet_NaN());
        tr.us = us.data();
      }
      if (everystep) {
        st.clear();
        su.clear();
        tr.steps_t = &st;
        tr.steps_u = &su;
      }
      solve_one<T, F>(d->alg, f, tr, md);
      for (int c = 0; c < N; ++c) ((T*)io->u_end)[c * n + i] = uend[c];
      if (io->t_end) io->t_end[i] = tr.t_end;
      int rc = tr.retcode;
      if (rc == SDEH_RET_DTLESSTHANMIN) any_dtmin = 1;
      if (io->naccept) io->naccept[i] = (int32_t)tr.naccept;
      if (io->nreject) io->nreject[i] = (int32_t)tr.nreject;
      const bool tm = d->save_layout == SDEH_SAVE_TRAJ_MAJOR;
      if (saveat && io->u_saveat) {
        T* o = (T*)io->u_saveat;
        const int64_t ns = d->n_saveat;
        for (int64_t j = 0; j < ns; ++j)
          for (int c = 0; c < N; ++c)
            o[tm ? (i * ns + j) * N + c : (j * N + c) * n + i] = us[j * N + c];
      }
      if (everystep) {
        const int64_t cnt = (int64_t)st.size();
        const int64_t keep = std::min<int64_t>(cnt, d->max_saves);
        const int64_t ms = d->max_saves;
        for (int64_t s = 0; s < keep; ++s) {
          if (io->t_steps) io->t_steps[tm ? i * ms + s : s * n + i] = st[s];
          if (io->u_steps)
            for (int c = 0; c < N; ++c)
              ((T*)io->u_steps)[tm ? (i * ms + s) * N + c : (s * N + c) * n + i] = su[s * N + c];
        }
        if (io->n_steps) io->n_steps[i] = (int32_t)keep;
        if (cnt > d->max_saves) {
          any_overflow = 1;
          if (rc == SDEH_RET_DEFAULT) rc = SDEH_RET_SAVE_OVERFLOW;
        }
      }
      if (io->retcode) io->retcode[i] = rc;
    }
  };

  if (nthreads <= 1 || n < 2) {
    work(0, n);
  } else {
    std::vector<std::thread> th;
    const int64_t per = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; ++k) {
      const int64_t lo = std::min<int64_t>(n, k * per), hi = std::min<int64_t>(n, lo + per);
      if (lo >= hi) break;
      th.emplace_back(work, lo, hi);
    }
    for (auto& x : th) x.join();
  }
  if (any_dtmin) return SDEH_E_DTMIN;
  if (any_overflow) return SDEH_E_SAVE_OVERFLOW;
  return SDEH_OK;
}

template <class T>
int dispatch_rhs(const sdeh_solve_desc* d, const sdeh_io* io, const Mode& md, int nthreads) {
  switch (d->rhs) {
    case SDEH_RHS_LORENZ: return run_ensemble<T, Lorenz<T>>(d, io, md, nthreads);
    case SDEH_RHS_GLV16: return run_ensemble<T, GLV16<T>>(d, io, md, nthreads);
    case SDEH_RHS_EXPDECAY3: return run_ensemble<T, ExpDecay3<T>>(d, io, md, nthreads);
    case SDEH_RHS_AFFINE1: return run_ensemble<T, Affine1<T>>(d, io, md, nthreads);
    case SDEH_RHS_LINEAR2: return run_ensemble<T, Linear2<T>>(d, io, md, nthreads);
    default: return SDEH_E_UNSUPPORTED;
  }
}

}  // namespace

extern "C" {

// Solve an ensemble on the CPU.  arith: ARITH_FUSED (0) or ARITH_PLAIN (1);
// pow_kind: POW_PORTABLE (0) or POW_LIBM (1); nthreads: worker threads.
int oracle_solve(const sdeh_solve_desc* d, const sdeh_io* io, int arith, int pow_kind,
                 int nthreads) {
  if (!d || !io || d->n_traj < 0 || d->alg < 0 || d->alg > SDEH_ALG_EULER) return SDEH_E_INVALID;
  if (d->rhs == SDEH_RHS_GLV16 && !io->p_shared) return SDEH_E_INVALID;
  Mode md;
  md.arith = arith;
  md.pow = pow_kind;
  md.quirks = d->quirks;
  md.maxiters = d->maxiters > 0 ? d->maxiters : 10000000;
  if (d->dtype == SDEH_F64) return dispatch_rhs<double>(d, io, md, nthreads);
  if (d->dtype == SDEH_F32) return dispatch_rhs<float>(d, io, md, nthreads);
  return SDEH_E_INVALID;
}

double oracle_pow_ctrl(double x, double y) { return sdeh_math::pow_ctrl(x, y); }

// The pow's two halves, for their own accuracy tests.
void oracle_log_dd(double x, double* hi_lo) { sdeh_math::log_pos_dd(x, &hi_lo[0], &hi_lo[1]); }
double oracle_exp_dd(double zh, double zl) { return sdeh_math::exp_dd(zh, zl); }
// Julia min/max and the kernels' reduced forms: op 0 jl_max, 1 jl_min,
// 2 jl_max_nonneg, 3 jl_min_nonneg, 4 fast_max, 5 fast_min.
double oracle_minmax(int op, double x, double y) {
  switch (op) {
    case 0: return sdeh_math::jl_max(x, y);
    case 1: return sdeh_math::jl_min(x, y);
    case 2: return sdeh_math::jl_max_nonneg(x, y);
    case 3: return sdeh_math::jl_min_nonneg(x, y);
    case 4: return sdeh_math::fast_max(x, y);
    case 5: return sdeh_math::fast_min(x, y);
  }
  return 0.0;
}
// The kernels' controller-domain pow (pow_log_ctrl + pow_from_ctrl).
double oracle_pow_ctrl_domain(double x, double y) {
  return sdeh_math::pow_from_ctrl(sdeh_math::pow_log_ctrl(x), y);
}

// Julia range restatement: length and the first min(len, cap) elements.
int64_t oracle_range(double start, double step, double stop, double* elems, int64_t cap,
                     int32_t* rational) {
  const JRange r = jl_range(start, step, stop);
  for (int64_t i = 0; i < std::min(cap, r.len); ++i) elems[i] = r.elem(i);
  if (rational) *rational = r.rational;
  return r.len;
}

// Dense-output weights at θ (binary64): Tsit5 (7), Vern7 (13), Vern9 (19).
int oracle_btheta(int which, double th, double* out) {
  if (which == 0) {
    namespace tb = sdeh_tab::tsit5;
    const double rs[22] = {tb::r11, tb::r12, tb::r13, tb::r14, tb::r22, tb::r23, tb::r24, tb::r32,
                           tb::r33, tb::r34, tb::r42, tb::r43, tb::r44, tb::r52, tb::r53, tb::r54,
                           tb::r62, tb::r63, tb::r64, tb::r72, tb::r73, tb::r74};
    out[0] = evalpoly_z(th, rs, 4, 1);
    for (int j = 0; j < 6; ++j) out[j + 1] = evalpoly_z(th, rs + 4 + 3 * j, 3, 2);
    return 7;
  }
  if (which == 1) {
    // Unit vectors through vern7_dense give the weights: out = 0 + 1 * (sum b_j e_j).
    constexpr int N = 13;
    V7Stages<double, N> K{};
    double x[6][N] = {};
    double* slots[13] = {K.k1, K.k4, K.k5, K.k6, K.k7, K.k8, K.k9,
                         x[0], x[1], x[2], x[3], x[4], x[5]};
    for (int j = 0; j < 13; ++j) slots[j][j] = 1.0;
    double zero[N] = {};
    vern7_dense<double, N>(ARITH_FUSED, out, zero, 1.0, th, K, x);
    return 13;
  }
  if (which == 2) {
    constexpr int N = 19;
    double s[9][N] = {}, x[10][N] = {};
    const double* sl[9];
    for (int j = 0; j < 9; ++j) {
      s[j][j] = 1.0;
      sl[j] = s[j];
    }
    for (int j = 0; j < 10; ++j) x[j][9 + j] = 1.0;
    double zero[N] = {};
    vern9_dense<double, N>(ARITH_FUSED, out, zero, 1.0, th, sl, x);
    return 19;
  }
  return 0;
}

}  // extern "C"
