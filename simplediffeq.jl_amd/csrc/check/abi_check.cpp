// Test program for the host-sanitized build of libsimplediffeq_hip (make asan): the
// library's host code (descriptor validation, planning, arena sizing and reuse,
// the chunked host pipeline, the staged schedule's scratch, the packed layout's
// capacity protocol, graphs, hipRTC registration, the cost-model entry points and
// every refusal path) runs under AddressSanitizer and UndefinedBehaviorSanitizer
// while the kernels run on the GPU.  Device results are checked against each other
// (host- and device-memory solves, every schedule, the allocation-free solve):
// they must be the same bits.  Prints one line per step; exit 0 when all pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../../include/simplediffeq_hip.h"

namespace {

int failures = 0;

#define EXPECT(cond, what)                                                 \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::printf("FAIL %s (line %d)\n", what, __LINE__);                  \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

#define HIPOK(expr)                                                        \
  do {                                                                     \
    if ((expr) != hipSuccess) {                                            \
      std::printf("HIP call failed: %s (line %d)\n", #expr, __LINE__);     \
      return 2;                                                            \
    }                                                                      \
  } while (0)

double lcg(uint64_t* s) {
  *s = *s * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(*s >> 11) * (1.0 / 9007199254740992.0);
}

// Lorenz inputs, component-major: u0 [3][n], p [3][n], tspan [2][n].
void lorenz(int64_t n, uint64_t seed, std::vector<double>& u0, std::vector<double>& p,
            std::vector<double>& ts) {
  u0.resize(3 * n);
  p.resize(3 * n);
  ts.resize(2 * n);
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) u0[c * n + i] = 2 * lcg(&seed) - 1;
    p[i] = 8 + 4 * lcg(&seed);
    p[n + i] = 28 * lcg(&seed);
    p[2 * n + i] = 2 + lcg(&seed);
    ts[i] = 0.0;
    ts[n + i] = 0.5 + 3 * lcg(&seed);
  }
}

struct Host {
  std::vector<double> u_end, t_end, u_saveat;
  std::vector<int32_t> ret, na, nr;
  void size(int64_t n, int64_t ns) {
    u_end.assign(3 * n, 0);
    t_end.assign(n, 0);
    u_saveat.assign(ns * 3 * n, 0);
    ret.assign(n, -9);
    na.assign(n, 0);
    nr.assign(n, 0);
  }
  void io(sdeh_io* o) {
    o->u_end = u_end.data();
    o->t_end = t_end.data();
    o->retcode = ret.data();
    o->naccept = na.data();
    o->nreject = nr.data();
    o->u_saveat = u_saveat.empty() ? nullptr : u_saveat.data();
  }
};

bool same(const void* a, const void* b, size_t bytes) { return std::memcmp(a, b, bytes) == 0; }

template <class T>
T* dev(size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, count * sizeof(T) + 256) != hipSuccess) return nullptr;
  return (T*)q;
}

}  // namespace

int main() {
  sdeh_ctx* c = nullptr;
  EXPECT(sdeh_create(&c, 0) == SDEH_OK, "create");
  if (!c) return 2;
  EXPECT(sdeh_abi_version() == SDEH_ABI_VERSION, "abi version");

  // ---- host-memory solves: small (one launch) and pipelined (>= 2^17) ----
  for (int64_t n : {1000, (1 << 17) + 77}) {
    std::vector<double> u0, p, ts;
    lorenz(n, 7 + n, u0, p, ts);
    std::vector<double> saveat = {0.0, 0.3, 0.9, 1.7};
    sdeh_solve_desc d;
    sdeh_desc_init(&d);
    d.alg = SDEH_ALG_ATSIT5;
    d.n_traj = n;
    d.save_everystep = 0;
    d.saveat = saveat.data();
    d.n_saveat = (int64_t)saveat.size();
    sdeh_io io{};
    io.u0 = u0.data();
    io.p = p.data();
    io.tspan = ts.data();
    Host h;
    h.size(n, d.n_saveat);
    h.io(&io);
    EXPECT(sdeh_solve(c, &d, &io, SDEH_MEM_HOST, nullptr, 1) == SDEH_OK, "host solve");
    // the same on device buffers
    double *du0 = dev<double>(3 * n), *dp = dev<double>(3 * n), *dts = dev<double>(2 * n),
           *due = dev<double>(3 * n), *dte = dev<double>(n), *dus = dev<double>(4 * 3 * n);
    int32_t *drc = dev<int32_t>(n), *dna = dev<int32_t>(n), *dnr = dev<int32_t>(n);
    HIPOK(hipMemcpy(du0, u0.data(), 24 * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dp, p.data(), 24 * n, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dts, ts.data(), 16 * n, hipMemcpyHostToDevice));
    sdeh_io dio{};
    dio.u0 = du0;
    dio.p = dp;
    dio.tspan = dts;
    dio.u_end = due;
    dio.t_end = dte;
    dio.retcode = drc;
    dio.naccept = dna;
    dio.nreject = dnr;
    dio.u_saveat = dus;
    for (int sched : {SDEH_SCHED_ONE_LAUNCH, SDEH_SCHED_PERSISTENT}) {
      d.schedule = sched;
      EXPECT(sdeh_solve(c, &d, &dio, SDEH_MEM_DEVICE, nullptr, 1) == SDEH_OK, "device solve");
      Host g;
      g.size(n, d.n_saveat);
      HIPOK(hipMemcpy(g.u_end.data(), due, 24 * n, hipMemcpyDeviceToHost));
      HIPOK(hipMemcpy(g.u_saveat.data(), dus, 4 * 24 * n, hipMemcpyDeviceToHost));
      HIPOK(hipMemcpy(g.na.data(), dna, 4 * n, hipMemcpyDeviceToHost));
      EXPECT(same(g.u_end.data(), h.u_end.data(), 24 * n), "host = device u_end");
      EXPECT(same(g.u_saveat.data(), h.u_saveat.data(), 4 * 24 * n), "host = device saveat");
      EXPECT(same(g.na.data(), h.na.data(), 4 * n), "host = device naccept");
    }
    // endpoint only, staged schedule (parked records in the arena), then the
    // allocation-free solve of the same: same bits
    d.saveat = nullptr;
    d.n_saveat = 0;
    d.schedule = SDEH_SCHED_STAGED;
    dio.u_saveat = nullptr;
    EXPECT(sdeh_solve(c, &d, &dio, SDEH_MEM_DEVICE, nullptr, 1) == SDEH_OK, "staged solve");
    std::vector<double> a(3 * n), b(3 * n);
    HIPOK(hipMemcpy(a.data(), due, 24 * n, hipMemcpyDeviceToHost));
    d.schedule = SDEH_SCHED_PERSISTENT;
    EXPECT(sdeh_solve_static(c, &d, &dio, nullptr, 1) == SDEH_OK, "static solve");
    HIPOK(hipMemcpy(b.data(), due, 24 * n, hipMemcpyDeviceToHost));
    EXPECT(same(a.data(), b.data(), 24 * n), "staged = static");
    // plan order on keys, and the autotuner with its register-cap stage
    std::vector<float> key(n);
    uint64_t s = 99;
    for (auto& k : key) k = (float)(-4 - 2 * lcg(&s));
    float* dkey = dev<float>(n);
    int64_t* dord = dev<int64_t>(n);
    HIPOK(hipMemcpy(dkey, key.data(), 4 * n, hipMemcpyHostToDevice));
    EXPECT(sdeh_plan_order(c, dkey, n, dord, nullptr) == SDEH_OK, "plan order");
    EXPECT(sdeh_synchronize(c) == SDEH_OK, "synchronize");
    std::vector<int64_t> ord(n);
    HIPOK(hipMemcpy(ord.data(), dord, 8 * n, hipMemcpyDeviceToHost));
    std::vector<char> seen(n, 0);
    bool perm = true;
    for (int64_t v : ord) perm = perm && v >= 0 && v < n && !seen[v]++;
    EXPECT(perm, "plan order is a permutation");
    d.schedule = SDEH_SCHED_AUTO;
    sdeh_stats best{};
    EXPECT(sdeh_autotune(c, &d, &dio, &best) == SDEH_OK, "autotune");
    EXPECT(best.kernel_ms > 0, "autotune timing");
    // a graph of the solve, replayed
    int64_t g = 0;
    EXPECT(sdeh_graph_capture(c, &d, &dio, &g) == SDEH_OK, "graph capture");
    for (int k = 0; k < 3; ++k) EXPECT(sdeh_graph_launch(c, g, nullptr) == SDEH_OK, "graph launch");
    EXPECT(sdeh_synchronize(c) == SDEH_OK, "synchronize");
    HIPOK(hipMemcpy(b.data(), due, 24 * n, hipMemcpyDeviceToHost));
    EXPECT(same(a.data(), b.data(), 24 * n), "graph replay = eager");
    EXPECT(sdeh_graph_destroy(c, g) == SDEH_OK, "graph destroy");
    for (void* q : {(void*)du0, (void*)dp, (void*)dts, (void*)due, (void*)dte, (void*)dus,
                    (void*)drc, (void*)dna, (void*)dnr, (void*)dkey, (void*)dord})
      HIPOK(hipFree(q));
    std::printf("step: ensemble of %lld trajectories done\n", (long long)n);
  }

  // ---- packed every-step output: the capacity protocol ----
  {
    const int64_t n = 3001;
    std::vector<double> u0, p, ts;
    lorenz(n, 3, u0, p, ts);
    sdeh_solve_desc d;
    sdeh_desc_init(&d);
    d.n_traj = n;
    d.tf = 2.0;
    d.save_layout = SDEH_SAVE_PACKED;
    sdeh_io io{};
    io.u0 = u0.data();
    io.p = p.data();
    Host h;
    h.size(n, 0);
    h.io(&io);
    std::vector<int32_t> ns(n);
    std::vector<int64_t> off(n + 1);
    io.n_steps = ns.data();
    io.steps_offset = off.data();
    io.steps_capacity = 0;  // ask for the size
    EXPECT(sdeh_solve(c, &d, &io, SDEH_MEM_HOST, nullptr, 1) == SDEH_E_SAVE_OVERFLOW,
           "packed: size query");
    sdeh_stats st{};
    sdeh_get_stats(c, &st);
    std::vector<double> us(3 * st.steps_needed), tsv(st.steps_needed);
    io.u_steps = us.data();
    io.t_steps = tsv.data();
    io.steps_capacity = st.steps_needed;
    EXPECT(sdeh_solve(c, &d, &io, SDEH_MEM_HOST, nullptr, 1) == SDEH_OK, "packed solve");
    EXPECT(off[n] == st.steps_needed, "packed rows");
    std::printf("step: packed every-step done\n");
  }

  // ---- hipRTC registration: a valid and an invalid source ----
  {
    const char* src =
        "template <class T> __device__ inline void decay(T* du, const T* u, const T* p,\n"
        "    const T* ps, double t) { du[0] = p[0] * u[0]; }\n";
    int64_t h = 0;
    EXPECT(sdeh_compile_rhs(c, src, "decay", 1, 1, 0, &h) == SDEH_OK, "compile rhs");
    EXPECT(sdeh_compile_rhs(c, "this is not C", "decay", 1, 1, 0, &h) == SDEH_E_COMPILE,
           "compile error reported");
    EXPECT(std::strlen(sdeh_last_error(c)) > 0, "compile log");
    EXPECT(sdeh_compile_rhs(c, src, "1bad", 1, 1, 0, &h) == SDEH_E_INVALID, "bad name");
    std::printf("step: hipRTC done\n");
  }

  // ---- the cost-model entry points ----
  {
    const int64_t n = 4099;
    std::vector<float> w(SDEH_MLP_NPARAM), x(n * 6), y(n * 2);
    uint64_t s = 5;
    for (auto& v : w) v = (float)(0.2 * (lcg(&s) - 0.5));
    for (auto& v : x) v = (float)(2 * lcg(&s) - 1);
    for (auto& v : y) v = (float)(lcg(&s));
    float *dw = dev<float>(w.size()), *dx = dev<float>(x.size()), *dy = dev<float>(y.size()),
          *dout = dev<float>(y.size()), *dg = dev<float>(w.size()), *dm = dev<float>(w.size()),
          *dv = dev<float>(w.size()), *dl = dev<float>(1);
    HIPOK(hipMemcpy(dw, w.data(), 4 * w.size(), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dx, x.data(), 4 * x.size(), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dy, y.data(), 4 * y.size(), hipMemcpyHostToDevice));
    HIPOK(hipMemset(dm, 0, 4 * w.size()));
    HIPOK(hipMemset(dv, 0, 4 * w.size()));
    EXPECT(sdeh_mlp_forward(c, dw, dx, dout, n, nullptr) == SDEH_OK, "mlp forward");
    EXPECT(sdeh_mlp_loss_grad(c, dw, dx, dy, n, dl, dg, nullptr) == SDEH_OK, "mlp grad");
    EXPECT(sdeh_mlp_adam(c, dw, dg, dm, dv, 1e-3f, 0.9f, 0.999f, 1e-8f, 1, nullptr) == SDEH_OK,
           "mlp adam");
    EXPECT(sdeh_synchronize(c) == SDEH_OK, "synchronize");
    float loss = NAN;
    HIPOK(hipMemcpy(&loss, dl, 4, hipMemcpyDeviceToHost));
    EXPECT(std::isfinite(loss) && loss > 0, "mlp loss");
    for (float* q : {dw, dx, dy, dout, dg, dm, dv, dl}) HIPOK(hipFree(q));
    std::printf("step: cost model done\n");
  }

  // ---- refusals: each a clean status, no fault ----
  {
    sdeh_solve_desc d;
    sdeh_desc_init(&d);
    sdeh_io io{};
    d.n_traj = 10;
    d.alg = 99;
    EXPECT(sdeh_solve(c, &d, &io, SDEH_MEM_HOST, nullptr, 1) == SDEH_E_INVALID, "bad alg");
    d.alg = SDEH_ALG_ATSIT5;
    EXPECT(sdeh_solve(c, &d, &io, SDEH_MEM_HOST, nullptr, 1) == SDEH_E_INVALID, "missing io");
    d.waves = 5;
    EXPECT(sdeh_solve(c, &d, &io, SDEH_MEM_HOST, nullptr, 1) == SDEH_E_INVALID, "bad waves");
    d.waves = 0;
    std::vector<double> hu0(30, 0.5), hp(30, 1.0), hue(30), hte(10);
    std::vector<int32_t> hrc(10);
    io.u0 = hu0.data();  // host pointers passed as device memory
    io.p = hp.data();
    io.u_end = hue.data();
    io.t_end = hte.data();
    io.retcode = hrc.data();
    EXPECT(sdeh_solve(c, &d, &io, SDEH_MEM_DEVICE, nullptr, 1) == SDEH_E_INVALID,
           "host pointers refused as device memory");
    EXPECT(sdeh_solve_static(c, &d, &io, nullptr, 1) == SDEH_E_INVALID,
           "static: counts required / host pointers");
    EXPECT(sdeh_plan_order(c, nullptr, 4, nullptr, nullptr) == SDEH_E_INVALID, "plan order args");
    EXPECT(sdeh_graph_launch(c, 12345, nullptr) == SDEH_E_INVALID, "unknown graph");
    std::printf("step: refusals done\n");
  }

  sdeh_destroy(c);
  std::printf("abi_check: %d failures\n", failures);
  return failures ? 1 : 0;
}
