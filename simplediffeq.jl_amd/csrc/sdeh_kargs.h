// Kernel argument block shared by the host runtime and the device integrators
// (ahead-of-time and hipRTC-compiled alike).  Plain data only; passed by value
// as the single kernel argument.
#pragma once

#if !defined(__HIPCC_RTC__)
#include <stdint.h>
#else
// hipRTC declares the fixed-width integer types in __hip_internal only.
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

struct sdeh_kargs {
  const void* u0;          // [N][n] state type
  const void* p;           // [P][n] state type
  const void* ps;          // shared RHS block, state type
  const double* tspan;     // [2][n] or null (then t0, tf)
  const double* dt0;       // [n] or null (then dt)
  const int64_t* order;    // [n] lane hand-out order or null (identity)
  void* u_end;             // [N][n]
  double* t_end;           // [n]
  int32_t* retcode;        // [n]
  int32_t* naccept;        // [n] or null
  int32_t* nreject;        // [n] or null
  const double* saveat;    // [n_saveat] device copy
  void* u_saveat;          // [n_saveat][N][n]
  void* u_steps;           // [max_saves][N][n]
  double* t_steps;         // [max_saves][n]
  int32_t* n_steps;        // [n]
  unsigned long long* queue;  // persistent schedule: next unclaimed slot
  int64_t n;
  int64_t n_saveat;
  int64_t maxiters;
  double t0, tf, dt, abstol, reltol;
  // Fixed-step time grid t0:dt:tf (Julia range semantics), element i:
  //   rational: (g_start_n + i * g_step_n) / g_den, else g_start + i * g_step.
  int64_t g_len, g_start_n, g_step_n, g_den;
  double g_start, g_step;
  int32_t g_rational;
  int32_t max_saves;
  int32_t quirks;
  int32_t persistent;      // 1: lanes refill from *queue in chunks of `chunk`
  int32_t chunk;
  int32_t traj_major;      // saved-state outputs [n][rows][N] instead of [rows][N][n]
  // Packed save_everystep (packed != 0): point k of trajectory i goes to row
  // steps_off[i] + k of u_steps ([row][N]) / t_steps ([row]) while k < n_steps[i],
  // the row count a first pass and the offsets scan left there.
  int32_t packed;
  // Staged endpoint solves (one lane per trajectory, adaptive, endpoint output).
  // stage_cap > 0: a trajectory still running after stage_cap attempts parks: its
  // whole state goes to one record of st_stride bytes at st_rec + i * st_stride
  // (layout below) instead of the outputs; every trajectory writes its record's
  // flag (1 parked, 0 ended) and st_key[i]: the attempts it is expected still to
  // need, stage_cap * (tf - t) / (t - t0), or 0 when it ended.  resume != 0: a
  // trajectory handed out continues from its record if the flag is set and is
  // skipped otherwise.  One record per trajectory keeps a resume to one or two
  // cache lines however the hand-out order scatters the trajectories.
  int32_t stage_cap;
  const int64_t* steps_off;
  void* st_rec;
  double* st_key;
  int64_t st_stride;
  int32_t resume;
  // Persistent schedule: a wave claims new trajectories only once at least
  // refill_min of its trajectory slots are free (0 or 1: as soon as one is), so that
  // the loads that start them are waited for once per batch.
  int32_t refill_min;
  // 1 when abstol is in [2^-420, 2^1020] and reltol in [0, 2^-3], so every scale
  // m * reltol + abstol of the error norm (m = max(|uprev|, |u|) finite) is a normal
  // number below 2^1022 with a normal reciprocal: the binary64 one-lane kernels may
  // then divide the scaled errors by bare Newton iterations (the error sum of
  // squares is the IEEE one whenever it lies in the controller's checked range).
  int32_t err_bare;
  // Optional [n] debug counters: a lane that takes trajectory i up in a launch
  // (a fresh start, or a staged solve resuming it) adds 1 to visits[i].
  int32_t* visits;
};

// Parked-trajectory record: double t, tf; int32 naccept, nreject, flag, pad; then
// state-type arrays u[N], fsal[N] (F = 1 for ATsit5), comp[N] (K = 1 with Kahan
// sums), p[P], dt, qold^beta2.  Bytes for element size ES:
#define SDEH_PARK_BYTES(N, P, ES, F, K) (32 + (ES) * ((N) * (1 + (F) + (K)) + (P) + 2))
