// Built-in right-hand sides for the device integrators.  Each is the
// out-of-place f(u, p, t) of a benchmark or test problem, written as a struct
// with compile-time state/parameter counts and one static device function:
//   N: state components per lane, P: per-lane parameters, S: shared elements,
//   G: lanes per trajectory (1, or a group size; see the lane layout in
//      sdeh_integrators.h).
//   f(du, u, p, ps, t): du = f(u, p, t); p = this lane's parameters (registers),
//   ps = the block shared by all trajectories (uniform loads).
// Evaluation order is exactly as written (the library is built without
// floating-point contraction), matching the CPU restatement.
// hipRTC-compiled user right-hand sides get the same struct shape from the
// generated wrapper in the runtime.
#pragma once

namespace sdeh_rhs {

// Lorenz (test/gpusimpleatsit5_tests.jl:3-13): p = (sigma, rho, beta).
struct Lorenz {
  static constexpr int N = 3, P = 3, S = 0, G = 1;
  template <class T>
  __device__ static inline void f(T* du, const T* u, const T* p, const T* __restrict__ ps,
                                  double t) {
    const T s = p[0], r = p[1], b = p[2];
    du[0] = s * (u[1] - u[0]);
    du[1] = u[0] * (r - u[2]) - u[1];
    du[2] = u[0] * u[1] - b * u[2];
  }
};

// Generalised Lotka-Volterra, 16 species: du_i = u_i (r_i + sum_j A_ij u_j),
// the sum accumulated from r_i as fma(A_ij, u_j, acc) for j = 0..15.
// p = r (16 per trajectory); ps = A, 16x16 row-major, shared.
// A is read through the constant address space (the same matrix for every lane,
// at compile-time offsets), so it arrives by scalar loads into SGPRs and feeds
// the FMAs directly.  The pointer passes through an empty asm on each evaluation:
// the compiler can then no longer prove the loads loop-invariant, so it keeps them
// inside the evaluation (batched s_load_dwordx16) instead of loading all of A (256
// values) once and parking it in VGPR lanes, from which every FMA needed a
// v_readlane first.
struct GLV16 {
  static constexpr int N = 16, P = 16, S = 256, G = 1;
  template <class T>
  __device__ static inline void f(T* du, const T* u, const T* p, const T* __restrict__ ps,
                                  double t) {
    const T* pa = ps;
    asm volatile("" : "+s"(pa));
    const __attribute__((address_space(4))) T* A =
        (const __attribute__((address_space(4))) T*)pa;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      T acc = p[i];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc = fma(A[i * 16 + j], u[j], acc);
      du[i] = u[i] * acc;
    }
  }
};

// GLV-16 in the group layout: 16 adjacent lanes hold one trajectory, lane c
// holding component c.  The lane's parameters are its growth rate r_c and its row
// of A (pl = r_c, A_c0 .. A_c15, loaded when the trajectory starts), so the
// product needs no loads inside the integration loop; the state is exchanged
// through LDS once per evaluation.  Arithmetic per component is GLV16's:
// acc = r_c, then fma(A_cj, u_j, acc) for j = 0..15, du_c = u_c * acc.
struct GLV16G {
  static constexpr int N = 1, P = 17, S = 256, G = 16;
  template <class T>
  __device__ static inline void load_lane(T* pl, const T* p, const T* __restrict__ ps, int64_t n,
                                          int64_t traj, int c) {
    pl[0] = p[c * n + traj];
#pragma unroll
    for (int j = 0; j < 16; ++j) pl[1 + j] = ps[c * 16 + j];
  }
  template <class T>
  __device__ static inline void f(T* du, const T* u, const T* p, const T* __restrict__ ps,
                                  double t) {
    T all[16];
    sdeh::group_gather<T, 16>(u[0], all);
    T acc = p[0];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc = fma(p[1 + j], all[j], acc);
    du[0] = u[0] * acc;
  }
};

// u' = p0 u, three components (test/gpu_ode_regression.jl:2-9 with p0 = -1).
struct ExpDecay3 {
  static constexpr int N = 3, P = 1, S = 0, G = 1;
  template <class T>
  __device__ static inline void f(T* du, const T* u, const T* p, const T* __restrict__ ps,
                                  double t) {
#pragma unroll
    for (int c = 0; c < 3; ++c) du[c] = p[0] * u[c];
  }
};

// Non-autonomous scalar problem u' = p0 u + p1 t.
struct Affine1 {
  static constexpr int N = 1, P = 2, S = 0, G = 1;
  template <class T>
  __device__ static inline void f(T* du, const T* u, const T* p, const T* __restrict__ ps,
                                  double t) {
    du[0] = p[0] * u[0] + p[1] * (T)t;
  }
};

// u' = M u, M = [p0 p1; p2 p3].
struct Linear2 {
  static constexpr int N = 2, P = 4, S = 0, G = 1;
  template <class T>
  __device__ static inline void f(T* du, const T* u, const T* p, const T* __restrict__ ps,
                                  double t) {
    du[0] = p[0] * u[0] + p[1] * u[1];
    du[1] = p[2] * u[0] + p[3] * u[1];
  }
};

}  // namespace sdeh_rhs
