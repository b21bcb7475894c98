// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// A test program for an AddressSanitizer + UndefinedBehaviorSanitizer build of the CPU
// restatement (make -C oracle sanitize): every solver, both state types, every
// built-in right-hand side, endpoint / saveat (both layouts) / every-step output
// (with and without overflow), one and several threads, and the edge cases the
// tests use (NaN states, dt below dtmin, maxiters, empty and reversed spans,
// per-trajectory spans and dt0, zero trajectories).  It checks the outcome codes
// and that every reached output is finite where the problem is finite; the
// sanitizers check every access and every arithmetic conversion.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/simplediffeq_hip.h"

extern "C" int oracle_solve(const sdeh_solve_desc* d, const sdeh_io* io, int arith, int pow_kind,
                            int nthreads);

namespace {

struct Rhs {
  int64_t id;
  int n, p, s;
};
const Rhs kRhs[] = {{SDEH_RHS_LORENZ, 3, 3, 0},
                    {SDEH_RHS_GLV16, 16, 16, 256},
                    {SDEH_RHS_EXPDECAY3, 3, 1, 0},
                    {SDEH_RHS_AFFINE1, 1, 2, 0},
                    {SDEH_RHS_LINEAR2, 2, 4, 0}};

int failures = 0;

void check(bool ok, const char* what, int alg, int dtype, int64_t rhs, int mode) {
  if (!ok) {
    std::fprintf(stderr, "FAIL %s: alg %d dtype %d rhs %lld mode %d\n", what, alg, dtype,
                 (long long)rhs, mode);
    ++failures;
  }
}

double lcg(uint64_t* s) {  // uniform [0, 1)
  *s = *s * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(*s >> 11) * (1.0 / 9007199254740992.0);
}

template <class T>
void run(const Rhs& r, int alg, int mode, int threads, int64_t n) {
  const bool adaptive = alg == SDEH_ALG_ATSIT5 || alg == SDEH_ALG_AVERN7 || alg == SDEH_ALG_AVERN9;
  const bool grid_only = alg == SDEH_ALG_RK4 || alg == SDEH_ALG_EULER;
  const int dtype = sizeof(T) == 8 ? SDEH_F64 : SDEH_F32;
  // modes: 0 endpoint, 1 saveat point-major, 2 saveat trajectory-major,
  // 3 every step (room), 4 every step (overflow), 5 edge cases (endpoint)
  if (grid_only && (mode == 1 || mode == 2)) return;
  sdeh_solve_desc d;
  std::memset(&d, 0, sizeof d);
  d.quirks = 1;
  d.alg = alg;
  d.dtype = dtype;
  d.rhs = r.id;
  d.n_traj = n;
  d.t0 = 0.0;
  d.tf = r.id == SDEH_RHS_GLV16 ? 2.0 : 1.0;
  d.dt = adaptive ? 0.1 : 0.05;
  d.abstol = 1e-6;
  d.reltol = 1e-4;
  d.save_everystep = mode == 3 || mode == 4;
  d.max_saves = mode == 3 ? 400 : (mode == 4 ? 3 : 0);
  d.save_layout = mode == 2 ? SDEH_SAVE_TRAJ_MAJOR : SDEH_SAVE_POINT_MAJOR;
  std::vector<double> saveat = {0.0, 0.25, 0.5, 0.75, d.tf};
  if (mode == 1 || mode == 2) {
    d.saveat = saveat.data();
    d.n_saveat = (int64_t)saveat.size();
  }
  uint64_t seed = 12345 + (uint64_t)alg * 7 + (uint64_t)mode * 131 + (uint64_t)r.id * 17;
  std::vector<T> u0((size_t)(r.n * n)), p((size_t)(r.p * n)), ps((size_t)r.s);
  for (auto& x : u0) x = (T)(r.id == SDEH_RHS_GLV16 ? 0.1 + 0.9 * lcg(&seed) : 2 * lcg(&seed) - 1);
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < r.p; ++c) {
      double v = lcg(&seed);
      if (r.id == SDEH_RHS_LORENZ) v = c == 0 ? 8 + 4 * v : (c == 1 ? 21 * v : 2 + v);
      else if (r.id == SDEH_RHS_GLV16) v = 0.5 + v;
      else v = -2 + 2.5 * v;
      p[(size_t)(c * n + i)] = (T)v;
    }
  for (int i = 0; i < 16 && r.s; ++i)
    for (int j = 0; j < 16; ++j) ps[(size_t)(i * 16 + j)] = (T)(i == j ? -1.0 : 0.1 * (lcg(&seed) - 0.5));
  std::vector<double> tspan, dt0;
  if (mode == 5 && adaptive && n >= 6) {  // edge cases in the first trajectories
    u0[1] = (T)NAN;                                    // NaN state
    tspan.assign((size_t)(2 * n), 0.0);
    for (int64_t i = 0; i < n; ++i) tspan[(size_t)(n + i)] = d.tf * (0.5 + lcg(&seed));
    tspan[(size_t)(n + 2)] = 0.0;                      // empty span
    tspan[(size_t)(n + 3)] = -1.0;                     // reversed span
    dt0.assign((size_t)n, 0.1);
    dt0[4] = 1e-16;                                    // below dtmin
    d.maxiters = 50;                                   // some reach the cap
  }
  std::vector<T> u_end((size_t)(r.n * n)), u_saveat, u_steps;
  std::vector<double> t_end((size_t)n), t_steps;
  std::vector<int32_t> ret((size_t)n), na((size_t)n), nr((size_t)n), ns((size_t)n);
  if (d.n_saveat) u_saveat.assign((size_t)(d.n_saveat * r.n * n), (T)0);
  if (d.save_everystep) {
    u_steps.assign((size_t)(d.max_saves * r.n * n), (T)0);
    t_steps.assign((size_t)(d.max_saves * n), 0.0);
  }
  sdeh_io io{};
  io.u0 = u0.data();
  io.p = p.data();
  io.p_shared = r.s ? ps.data() : nullptr;
  io.n_shared = r.s;
  io.tspan = tspan.empty() ? nullptr : tspan.data();
  io.dt0 = dt0.empty() ? nullptr : dt0.data();
  io.u_end = u_end.data();
  io.t_end = t_end.data();
  io.retcode = ret.data();
  io.naccept = na.data();
  io.nreject = nr.data();
  io.u_saveat = u_saveat.empty() ? nullptr : u_saveat.data();
  io.u_steps = u_steps.empty() ? nullptr : u_steps.data();
  io.t_steps = t_steps.empty() ? nullptr : t_steps.data();
  io.n_steps = d.save_everystep ? ns.data() : nullptr;
  const int st = oracle_solve(&d, &io, 0, 0, threads);
  check(st == SDEH_OK || st == SDEH_E_DTMIN || st == SDEH_E_SAVE_OVERFLOW, "status", alg, dtype,
        r.id, mode);
  bool any_overflow = false;
  for (int64_t i = 0; i < n; ++i) any_overflow |= ret[(size_t)i] == SDEH_RET_SAVE_OVERFLOW;
  if (mode == 3 || mode == 4)  // overflow status exactly when a trajectory lost points
    check((st == SDEH_E_SAVE_OVERFLOW) == any_overflow, "overflow status", alg, dtype, r.id, mode);
  for (int64_t i = 0; i < n; ++i) {
    const int rc = ret[(size_t)i];
    check(rc >= SDEH_RET_DEFAULT && rc <= SDEH_RET_SAVE_OVERFLOW, "retcode", alg, dtype, r.id, mode);
    if (mode != 5 && !grid_only) check(rc == SDEH_RET_DEFAULT || rc == SDEH_RET_SAVE_OVERFLOW,
                                       "plain run ends normally", alg, dtype, r.id, mode);
    if (rc == SDEH_RET_DEFAULT)
      for (int c = 0; c < r.n; ++c)
        check(std::isfinite((double)u_end[(size_t)(c * n + i)]), "finite end state", alg, dtype,
              r.id, mode);
  }
}

}  // namespace

int main() {
  int runs = 0;
  for (const Rhs& r : kRhs)
    for (int alg = SDEH_ALG_TSIT5; alg <= SDEH_ALG_EULER; ++alg)
      for (int mode = 0; mode <= 5; ++mode)
        for (int threads : {1, 3}) {
          const int64_t n = mode == 5 ? 9 : (threads == 1 ? 7 : 29);
          run<double>(r, alg, mode, threads, n);
          run<float>(r, alg, mode, threads, n);
          runs += 2;
        }
  // zero trajectories
  run<double>(kRhs[0], SDEH_ALG_ATSIT5, 0, 2, 0);
  std::printf("sanitize_check: %d ensemble solves, %d failures\n", runs + 1, failures);
  return failures ? 1 : 0;
}
