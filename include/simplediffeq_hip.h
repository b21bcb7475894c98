/*
 * libsimplediffeq_hip — C ABI of the MI355X (gfx950) ensemble integrator for
 * SimpleDiffEq.jl's GPU-compatible out-of-place solvers.
 *
 * One call of sdeh_solve() replaces an ensemble of per-trajectory calls of
 *
 *   DiffEqBase.solve(prob::ODEProblem, alg::GPUSimpleTsit5;  saveat, save_everystep, dt)
 *       src/tsit5/gpuatsit5.jl:55-147
 *   DiffEqBase.solve(prob::ODEProblem, alg::GPUSimpleATsit5; dt, saveat, save_everystep,
 *                    abstol, reltol)                         src/tsit5/gpuatsit5.jl:205-336
 *   DiffEqBase.solve(prob::ODEProblem, alg::GPUSimpleVern7;  ...) src/verner/gpuvern7.jl:55-242
 *   DiffEqBase.solve(prob::ODEProblem, alg::GPUSimpleAVern7; ...) src/verner/gpuvern7.jl:300-536
 *   DiffEqBase.solve(prob::ODEProblem, alg::GPUSimpleVern9;  ...) src/verner/gpuvern9.jl:55-353
 *   DiffEqBase.solve(prob::ODEProblem, alg::GPUSimpleAVern9; ...) src/verner/gpuvern9.jl:411-779
 *
 * that an EnsembleProblem under EnsembleThreads would make, one trajectory per
 * GPU lane.  Keyword defaults mirror those signatures (sdeh_desc_init()).
 *
 * Data layout: every per-trajectory array is structure-of-arrays, component
 * major: element (component c, trajectory i) of an array with n_traj
 * trajectories is at index c * n_traj + i.  saveat output is
 * [n_saveat][n_state][n_traj].  Time values are always binary64.
 *
 * Errors: every function returns an sdeh_status (0 = success).  A failing call
 * leaves a message retrievable with sdeh_last_error().  Per-trajectory outcomes
 * go to the retcode array (sdeh_retcode); the reference throws
 * error("dt<dtmin") where this library records SDEH_RET_DTLESSTHANMIN for that
 * trajectory, and sdeh_solve() then returns SDEH_E_DTMIN after writing all
 * outputs.
 *
 * Threading: a context may be used by one host thread at a time.  Work is
 * ordered on the context's stream (or the caller's stream, see sdeh_solve()).
 * A device-memory call may return before its kernels finish; they keep using
 * the context's device scratch (work queue, staged saveat times, counters), so
 * a later call on the same context that names a different stream first waits
 * for the earlier stream.  Calls on one stream simply queue behind each other.
 * For concurrent ensembles on several streams, use one context per stream.
 */
#ifndef SIMPLEDIFFEQ_HIP_H
#define SIMPLEDIFFEQ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDEH_ABI_VERSION 5

typedef enum sdeh_status {
  SDEH_OK = 0,
  SDEH_E_INVALID = -1,      /* bad argument (message says which) */
  SDEH_E_HIP = -2,          /* HIP runtime error */
  SDEH_E_COMPILE = -3,      /* hipRTC compilation failed (log in sdeh_last_error) */
  SDEH_E_UNSUPPORTED = -4,  /* valid request the library does not implement */
  SDEH_E_NOMEM = -5,        /* device allocation failed */
  SDEH_E_DTMIN = -6,        /* solve finished; >= 1 trajectory stopped at dt < dtmin */
  SDEH_E_SAVE_OVERFLOW = -7 /* solve finished; >= 1 trajectory had more steps than max_saves */
} sdeh_status;

/* Algorithms: the six reference algorithm structs. */
typedef enum sdeh_alg {
  SDEH_ALG_TSIT5 = 0,   /* GPUSimpleTsit5   gpuatsit5.jl:52  */
  SDEH_ALG_ATSIT5 = 1,  /* GPUSimpleATsit5  gpuatsit5.jl:200 */
  SDEH_ALG_VERN7 = 2,   /* GPUSimpleVern7   gpuvern7.jl:52   */
  SDEH_ALG_AVERN7 = 3,  /* GPUSimpleAVern7  gpuvern7.jl:295  */
  SDEH_ALG_VERN9 = 4,   /* GPUSimpleVern9   gpuvern9.jl:52   */
  SDEH_ALG_AVERN9 = 5,  /* GPUSimpleAVern9  gpuvern9.jl:406  */
  SDEH_ALG_RK4 = 6,     /* GPUSimpleRK4     gpurk4.jl:50     (fixed step, no saveat) */
  SDEH_ALG_EULER = 7    /* GPUSimpleEuler   gpueuler.jl:50   (fixed step, no saveat) */
} sdeh_alg;

/* Element type of the state u (the reference's eltype(u0)). */
typedef enum sdeh_dtype { SDEH_F64 = 0, SDEH_F32 = 1 } sdeh_dtype;

/* Per-trajectory outcome (the reference's sol.retcode plus the failures it throws). */
typedef enum sdeh_retcode {
  SDEH_RET_DEFAULT = 0,        /* ReturnCode.Default: integrated to tf */
  SDEH_RET_DTLESSTHANMIN = 1,  /* the reference's error("dt<dtmin") */
  SDEH_RET_NAN = 2,            /* error estimate or time became NaN (reference: unspecified) */
  SDEH_RET_MAXITERS = 3,       /* attempt cap reached (the reference has no cap) */
  SDEH_RET_SAVE_OVERFLOW = 4   /* save_everystep output exceeded max_saves (or, packed,
                                  steps_capacity); integration completed */
} sdeh_retcode;

/* Built-in right-hand sides (compiled ahead of time into the library).
 * Arbitrary RHS functions come from sdeh_compile_rhs(). */
typedef enum sdeh_builtin_rhs {
  SDEH_RHS_LORENZ = 1,    /* n_state 3, n_param 3 (sigma, rho, beta) */
  SDEH_RHS_GLV16 = 2,     /* n_state 16, n_param 16 (r_i); p_shared: A, 16x16 row-major */
  SDEH_RHS_EXPDECAY3 = 3, /* n_state 3, n_param 1: du = p0 * u */
  SDEH_RHS_AFFINE1 = 4,   /* n_state 1, n_param 2: du = p0 * u + p1 * t (non-autonomous) */
  SDEH_RHS_LINEAR2 = 5    /* n_state 2, n_param 4: du = [p0 p1; p2 p3] * u */
} sdeh_builtin_rhs;

/* Layout of the saved-state outputs (u_saveat, u_steps, t_steps).  u_end and the
 * other per-trajectory outputs are always [component][trajectory]. */
typedef enum sdeh_save_layout {
  SDEH_SAVE_POINT_MAJOR = 0, /* u_saveat [n_saveat][n_state][n_traj], u_steps
                                [max_saves][n_state][n_traj], t_steps [max_saves][n_traj] */
  SDEH_SAVE_TRAJ_MAJOR = 1,  /* u_saveat [n_traj][n_saveat][n_state], u_steps
                                [n_traj][max_saves][n_state], t_steps [n_traj][max_saves]:
                                each trajectory's saves contiguous, as the reference's
                                per-trajectory solution holds them; a lane writes whole
                                rows, which the GPU stores far faster */
  SDEH_SAVE_PACKED = 2       /* save_everystep only, no per-trajectory cap (max_saves is
                                not used): trajectory i's points are rows
                                steps_offset[i] .. steps_offset[i+1]-1 of u_steps
                                [steps_capacity][n_state] and t_steps [steps_capacity];
                                n_steps[i] = their count.  The reference's per-trajectory
                                `us`/`ts` vectors grow with push! (gpuatsit5.jl:113-114,
                                302-303); here the adaptive solvers first run an
                                endpoint pass that counts each trajectory's steps, a
                                scan turns the counts into offsets, and the saving pass
                                writes every row in place (fixed-step solvers know
                                their counts and skip the first pass).  When the rows
                                do not fit steps_capacity, trajectories are laid out in
                                index order until it is full: the later ones keep a
                                prefix of their rows (retcode SAVE_OVERFLOW), and
                                stats.steps_needed gives the capacity that fits. */
} sdeh_save_layout;

/* Scheduling of trajectories onto lanes. */
typedef enum sdeh_schedule {
  SDEH_SCHED_AUTO = 0,       /* library choice (autotuned) */
  SDEH_SCHED_ONE_LAUNCH = 1, /* one lane per trajectory, grid covers the ensemble */
  SDEH_SCHED_PERSISTENT = 2, /* resident grid; lanes refill from a device atomic queue */
  SDEH_SCHED_STAGED = 3      /* adaptive solvers with endpoint output, one lane per
                                trajectory, no caller order: a one-launch pass of 32
                                attempts per trajectory, then the trajectories still
                                running, largest expected remainder first, on the
                                persistent grid (same results as the other schedules);
                                elsewhere (other outputs or layouts, io->order, graph
                                capture, chunked host-memory solves) it runs as
                                SDEH_SCHED_PERSISTENT */
} sdeh_schedule;

/* Precision mode (north-star precision work). */
typedef enum sdeh_precision {
  SDEH_PREC_PLAIN = 0,  /* stages and update in the state type */
  SDEH_PREC_KAHAN = 1   /* FP32 state with a compensated u update; t stays FP64 */
} sdeh_precision;

typedef enum sdeh_memory { SDEH_MEM_HOST = 0, SDEH_MEM_DEVICE = 1 } sdeh_memory;

/* Solve descriptor: the keyword arguments of the reference solve methods plus
 * ensemble and scheduling fields.  Initialise with sdeh_desc_init(). */
typedef struct sdeh_solve_desc {
  int32_t alg;              /* sdeh_alg */
  int32_t dtype;            /* sdeh_dtype */
  int64_t rhs;              /* sdeh_builtin_rhs or a handle from sdeh_compile_rhs */
  int64_t n_traj;           /* ensemble size (trajectories) */
  double t0, tf;            /* common tspan, used when io->tspan is NULL */
  double dt;                /* fixed step, or initial step of the adaptive solvers;
                               reference default 0.1f0 = 0.10000000149011612 */
  double abstol, reltol;    /* reference defaults 1f-6 and 1f-3 (as binary64) */
  int32_t save_everystep;   /* reference default true; with saveat == NULL,
                               every accepted step is written (capacity max_saves) */
  int32_t max_saves;        /* per-trajectory capacity of the save_everystep output */
  const double* saveat;     /* host array of n_saveat save times, ascending, or NULL */
  int64_t n_saveat;
  int64_t maxiters;         /* attempt cap per trajectory (0 = library default 10^7) */
  int32_t quirks;           /* 1: reproduce the reference's time quirks (fixed Vern9
                               saveat stage aliasing, post-step saveat stage times,
                               RK4/Euler stages from the END of each step);
                               0: corrected */
  int32_t schedule;         /* sdeh_schedule */
  int32_t precision;        /* sdeh_precision */
  int32_t block;            /* 0 = tuned; else workgroup size (multiple of 64, <= 1024) */
  int32_t grid;             /* 0 = tuned; persistent schedule: number of workgroups */
  int32_t chunk;            /* 0 = tuned; persistent schedule: trajectories per queue
                               grab (the last grab takes what is left) */
  int32_t lanes;            /* lanes per trajectory: 0 = for binary64, the right-hand
                               side's matrix-core layout where it has a kernel for the
                               solver and output (GLV16: 4 lanes 16 apart, 4 components
                               each, A u on v_mfma_f64_16x16x4f64; adaptive solvers,
                               endpoint and saveat), else its group layout (GLV16: 16
                               adjacent lanes, one component each), else 1; 1 = one
                               lane holds a whole trajectory; or either layout's size
                               (4, 16) itself (4 also runs binary32 ATsit5, on
                               v_mfma_f32_16x16x4f32).  Results are identical. */
  int32_t save_layout;      /* sdeh_save_layout of u_saveat, u_steps and t_steps */
  int32_t span_order;       /* 0: with per-trajectory tspans (io->tspan) and no io->order,
                               hand trajectories to lanes longest span first (fewer
                               long trajectories start late; results are unchanged);
                               1: index order */
  int32_t waves;            /* register cap: 0 = tuned (else the kernel's default);
                               4, 6 or 8 = the variant compiled for that many waves per
                               SIMD (at most 512 / waves VGPRs per lane), built for the
                               Lorenz right-hand side's adaptive ATsit5 and AVern7
                               endpoint kernels (SDEH_E_UNSUPPORTED elsewhere).  Same
                               arithmetic, same results.  (ABI 4; was reserved, 0.) */
  int32_t per_lane;         /* trajectories per lane: 0 = tuned (else 1); 2 = the variant
                               whose lanes each integrate two trajectories, their
                               stages interleaved (instruction-level parallelism in
                               place of waves), built for the Lorenz right-hand side's
                               adaptive ATsit5 (binary64, binary32, compensated) and
                               AVern7 (binary64) endpoint kernels and for registered
                               right-hand sides; one-launch and persistent schedules (a
                               staged shape runs persistent).  Same results.  (ABI 4) */
  int32_t refill;           /* persistent and staged schedules: a wave takes new
                               trajectories only once at least `refill` of its
                               trajectory slots are free, so that the loads that start
                               them are waited for once per batch.  0 = tuned (else 4
                               with one lane per trajectory, 1 with several); values
                               above the wave's slot count (64 / lanes) act as that
                               count; the two-per-lane kernels refill as soon as a slot
                               is free.  Same results.  (ABI 5; was reserved, 0.) */
} sdeh_solve_desc;

/* Solve buffers.  With memory == SDEH_MEM_HOST these are host pointers and the
 * library stages them through device memory; with SDEH_MEM_DEVICE they are
 * device pointers used in place.  Optional inputs may be NULL. */
typedef struct sdeh_io {
  const void* u0;          /* [n_state][n_traj] of dtype */
  const void* p;           /* [n_param][n_traj] of dtype (may be NULL if n_param == 0) */
  const void* p_shared;    /* RHS-specific block common to all trajectories, of dtype */
  int64_t n_shared;        /* element count of p_shared */
  const double* tspan;     /* optional [2][n_traj]: row 0 = t0, row 1 = tf */
  const double* dt0;       /* optional [n_traj] initial dt per trajectory (adaptive) */
  const int64_t* order;    /* optional [n_traj] permutation: the order in which
                              trajectories are handed to lanes; results stay indexed
                              by trajectory.  Checked to be a permutation before the
                              solve (SDEH_E_INVALID otherwise), except the order that
                              sdeh_plan_order last wrote on this context; the
                              kernels never read an entry outside [0, n_traj) */
  void* u_end;             /* [n_state][n_traj] state at the final time */
  double* t_end;           /* [n_traj] final time */
  int32_t* retcode;        /* [n_traj] sdeh_retcode */
  int32_t* naccept;        /* [n_traj] accepted steps (optional) */
  int32_t* nreject;        /* [n_traj] rejected attempts (optional) */
  void* u_saveat;          /* [n_saveat][n_state][n_traj] (or trajectory-major, see
                              save_layout); unreached points are NaN */
  void* u_steps;           /* [max_saves][n_state][n_traj] save_everystep states */
  double* t_steps;         /* [max_saves][n_traj] save_everystep times */
  int32_t* n_steps;        /* [n_traj] number of saved steps (u0 included) */
  int64_t* steps_offset;   /* SDEH_SAVE_PACKED: [n_traj + 1] first row of each
                              trajectory; steps_offset[n_traj] = rows used */
  int64_t steps_capacity;  /* SDEH_SAVE_PACKED: rows of u_steps and t_steps */
  int32_t* visits;         /* optional [n_traj] debug counters, device memory only (ABI 4):
                              each time a lane takes trajectory i up in a launch (its
                              start, or a staged solve's resume) visits[i] gains 1, so a
                              caller can check that every trajectory is handed out
                              exactly once per launch.  NULL: no counting */
} sdeh_io;

/* Timing and work counters of the last sdeh_solve() on a context. */
typedef struct sdeh_stats {
  double kernel_ms;        /* device time of the integration kernel(s) (HIP events; the
                              packed layout: both passes and the offsets scan) */
  double total_ms;         /* host wall time of the call, staging copies included */
  int64_t attempts;        /* sum over trajectories of naccept + nreject */
  int64_t accepts;
  int32_t block, grid, chunk, schedule; /* launch shape actually used */
  int32_t lanes;           /* lanes per trajectory used */
  int32_t waves;           /* register-cap variant used (0 = the default kernel) */
  int64_t steps_needed;    /* SDEH_SAVE_PACKED: total points offered by all
                              trajectories (the steps_capacity that holds them all) */
  int32_t per_lane;        /* trajectories per lane used */
  int32_t refill;          /* refill batch used (0: one launch) (ABI 5) */
} sdeh_stats;

typedef struct sdeh_ctx sdeh_ctx;

/* Fill d with the reference solve defaults (alg ATSIT5, F64, dt 0.1f0,
 * abstol 1f-6, reltol 1f-3, save_everystep true, quirks on, auto schedule). */
void sdeh_desc_init(sdeh_solve_desc* d);

int sdeh_abi_version(void);

/* Create a context on HIP device `device`.  The caller owns *out. */
int sdeh_create(sdeh_ctx** out, int device);
void sdeh_destroy(sdeh_ctx* ctx);
const char* sdeh_last_error(const sdeh_ctx* ctx);

/* Shape of an RHS (built-in id or compiled handle). */
int sdeh_rhs_info(sdeh_ctx* ctx, int64_t rhs, int32_t* n_state, int32_t* n_param,
                  int64_t* n_shared);

/* Register an out-of-place RHS given as HIP C source (the output of the RHS
 * tracer).  `src` must define
 *   template <class T> __device__ inline void <fn_name>(T* du, const T* u,
 *       const T* p, const T* ps, double t);
 * Integrator kernels for it are compiled with hipRTC for gfx950 on first use
 * and cached (in memory, and on disk under $SDEH_CACHE_DIR when set).
 * Replaces the Julia-side specialisation of solve() on the user's f. */
int sdeh_compile_rhs(sdeh_ctx* ctx, const char* src, const char* fn_name, int32_t n_state,
                     int32_t n_param, int64_t n_shared, int64_t* handle);

/* Integrate an ensemble.  `stream` is a hipStream_t (NULL = the context's stream).
 * With SDEH_MEM_HOST the call returns after results are back in host memory;
 * with SDEH_MEM_DEVICE it returns once the work is queued, unless `sync` != 0. */
int sdeh_solve(sdeh_ctx* ctx, const sdeh_solve_desc* d, const sdeh_io* io, int32_t memory,
               void* stream, int32_t sync);

/* sdeh_solve without any device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Device buffers only (as
 * SDEH_MEM_DEVICE); the caller passes every input and output, naccept and nreject
 * included.  Output: endpoint, or saveat with at most 1024 times (more:
 * SDEH_E_INVALID); save_everystep is SDEH_E_UNSUPPORTED, and so is io->order.
 * Trajectories are handed to lanes in index order (no longest-span-first order),
 * and a staged launch shape runs as the persistent one; results are the same bits as
 * sdeh_solve's.  The library's scratch for it is one static device array per GPU
 * (8,256 bytes: the work-queue counter, the counts, the saveat times), shared by
 * every sdeh_solve_static on that GPU: calls that overlap on one GPU (other contexts,
 * other streams, other threads) race on it and must be serialised by the caller.
 * `stream` and `sync` as in sdeh_solve; stats are filled when sync != 0. */
int sdeh_solve_static(sdeh_ctx* ctx, const sdeh_solve_desc* d, const sdeh_io* io, void* stream,
                      int32_t sync);

int sdeh_get_stats(const sdeh_ctx* ctx, sdeh_stats* out);

/* Repeated ensembles: capture one device-memory solve into a HIP graph and
 * replay it.  The graph re-arms its own work queue on each replay. */
int sdeh_graph_capture(sdeh_ctx* ctx, const sdeh_solve_desc* d, const sdeh_io* io,
                       int64_t* graph_handle);
int sdeh_graph_launch(sdeh_ctx* ctx, int64_t graph_handle, void* stream);
int sdeh_graph_destroy(sdeh_ctx* ctx, int64_t graph_handle);

/* Launch-shape autotuner: time candidate (block, grid, chunk, schedule) shapes
 * for the given problem on device buffers, then the register-cap variants (waves)
 * of the best one and the two-trajectories-per-lane variant (one launch and
 * persistent) where they are built, and store the best for (alg, rhs, dtype, save
 * mode, precision, lanes).  Later solves with tuned fields left 0 use it. */
int sdeh_autotune(sdeh_ctx* ctx, const sdeh_solve_desc* d, const sdeh_io* io,
                  sdeh_stats* best);

/* Synchronise the context stream and the stream of the context's last call. */
int sdeh_synchronize(sdeh_ctx* ctx);

/* The fixed-step time grid t0:dt:tf with Julia's Float64 range rules (the
 * `_ts = tspan[1]:dt:tspan[2]` of gpuatsit5.jl:92, gpuvern7.jl:100,
 * gpuvern9.jl:100): returns length(t0:dt:tf) (the solvers take length - 1
 * steps) and writes the first min(length, cap) elements to elems (may be
 * NULL).  Returns -1 when dt == 0. */
int64_t sdeh_fixed_grid(double t0, double dt, double tf, double* elems, int64_t cap);

/* ---- Learned cost model: 6 -> 64 -> 64 -> 2 MLP (FP32), trained on device. ----
 * Layout of the parameter vector (FP32, n = 4738): W1[64][6], b1[64],
 * W2[64][64], b2[64], W3[2][64], b3[2].  Hidden activations are max(x, 0) (ReLU). */
#define SDEH_MLP_NIN 6
#define SDEH_MLP_NH 64
#define SDEH_MLP_NOUT 2
#define SDEH_MLP_NPARAM 4738

/* Forward pass on device buffers: x [n][6] -> y [n][2] (row-major). */
int sdeh_mlp_forward(sdeh_ctx* ctx, const float* params, const float* x, float* y, int64_t n,
                     void* stream);
/* The binning step of the Lorenz-ensemble cost model in one pass over device
 * buffers (no counterpart in the reference; it plans the hand-out order that
 * sdeh_io.order takes).  Inputs binary64, component-major as in sdeh_io: p[3][n],
 * u0[3][n], tspan[2][n].  Features in FP32 from the inputs rounded to FP32:
 * (p_c - feat[2c]) * (1 / feat[2c+1]) for c = 0, 1, 2; (log(tf - t0) - feat[6]) *
 * (1 / feat[7]) (the difference in binary64); sqrt((u0_0^2 + u0_1^2) + u0_2^2) /
 * sqrt(3) * 2 - 1; u0_2.  The forward pass runs on the f16 matrix cores: weights,
 * layer 1's biases, features and activations rounded to f16, products accumulated
 * in FP32, layer 2's and 3's biases added in FP32; its y agrees with
 * sdeh_mlp_forward's FP32 pass to a few 1e-3.  Then pred =
 * y * y_norm[2..3] + y_norm[0..1] (FP32: std, then mean), key[i] = -pred_0 (an
 * ascending sort of key is the hand-out order, largest predicted cost first) and
 * dt0[i] = exp(pred_1) in FP32, widened.  pred [n][2] may be NULL. */
int sdeh_mlp_plan_lorenz(sdeh_ctx* ctx, const float* params, const double* p, const double* u0,
                         const double* tspan, int64_t n, const double* feat, const float* y_norm,
                         float* key, double* dt0, float* pred, void* stream);
/* The hand-out order from a plan's keys (no counterpart in the reference; it turns
 * sdeh_mlp_plan_lorenz's `key` into the permutation sdeh_io.order takes, with no
 * sort outside the library): order[0..n) lists the trajectories by increasing key,
 * i.e. by decreasing predicted cost exp(-key), counting-sorted on a grid of 1/16
 * octave of that cost (16 buckets per doubling, evenly spaced in log2, covering
 * 2^-30 .. 2^34; costs outside go to the end buckets).  Trajectories that share a
 * bucket keep an unspecified relative order; NaN keys come last.  Device buffers;
 * key [n] FP32, order [n] int64, n < 2^32; ordered on `stream` (NULL = the
 * context's), using the context's device scratch like sdeh_solve.  The context
 * remembers the last order buffer written here and its n: a device solve given that
 * buffer as io->order, for the same n_traj, skips the permutation check (one pass
 * over n and a host synchronisation), so the buffer must still hold what this call
 * wrote. */
int sdeh_plan_order(sdeh_ctx* ctx, const float* key, int64_t n, int64_t* order, void* stream);
/* Loss (mean squared error over n x 2 outputs) and its gradient with respect
 * to the parameters, on device buffers.  grad has SDEH_MLP_NPARAM floats;
 * loss is one float. */
int sdeh_mlp_loss_grad(sdeh_ctx* ctx, const float* params, const float* x, const float* y,
                       int64_t n, float* loss, float* grad, void* stream);
/* One Adam update (in place) on device buffers; t is the 1-based step number. */
int sdeh_mlp_adam(sdeh_ctx* ctx, float* params, const float* grad, float* m, float* v,
                  float lr, float beta1, float beta2, float eps, int32_t t, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SIMPLEDIFFEQ_HIP_H */
