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

// GLV-16 on the matrix cores (binary64; binary32 ATsit5 below), in the wave layout: 4 lanes per
// trajectory and 16 trajectories per wave.  Trajectory slot j of a wave is on lanes
// j, j+16, j+32, j+48; lane q (= lane / 16) holds components q, q+4, q+8, q+12.
// A u for the wave's 16 trajectories is four v_mfma_f64_16x16x4f64 (D = A B + C,
// A 16x4, B 4x16, C and D 16x16): product kb covers components 4kb .. 4kb+3.  Its B
// operand is the lane's own component q + 4kb (register kb: no data movement), its A
// operand the lane's fixed entry A[lane % 16][4kb + q], C the running sums, whose
// row for lane q, register r is component q + 4r (the lane's own).  The instruction
// forms each output as fma(a3, b3, fma(a2, b2, fma(a1, b1, fma(a0, b0, c)))) over
// k = 0..3 in order (lane layout and rounding checked on the device, 2^20 outputs
// bit for bit), so acc_i = r_i, then fma(A_ij, u_j, acc) for j = 0..15:
// GLV16's arithmetic, with no exchange of the state between lanes.
// Binary32 runs v_mfma_f32_16x16x4f32, which forms each output the same way (fma
// chain over k = 0..3, checked on the device: 38,400 random outputs bit for bit) but
// puts output rows 4(lane / 16) + r in lane `lane`, register r.  Its A operand's row
// i' therefore carries component i' / 4 + 4 (i' % 4), the component that lane i' / 4
// holds in register i' % 4, and the layout above is unchanged.
// pl[0..3] = r of the lane's components (per trajectory), pl[4..7] = the lane's A
// operands (the same for every trajectory: loaded once for every lane).
struct GLV16M {
  static constexpr int N = 4, P = 8, S = 256, G = 4;
  static constexpr bool WAVE = true;
  template <class T>
  __device__ static inline void init_lane(T* pl, const T* __restrict__ ps) {
    const int l = (int)__lane_id();
    const int row = sizeof(T) == 8 ? (l & 15) : ((l & 15) >> 2) + 4 * (l & 3);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) pl[4 + kb] = ps[row * 16 + 4 * kb + (l >> 4)];
  }
  template <class T>
  __device__ static inline void load_lane(T* pl, const T* p, const T* __restrict__ ps, int64_t n,
                                          int64_t traj, int q) {
#pragma unroll
    for (int r = 0; r < 4; ++r) pl[r] = p[(q + 4 * r) * n + traj];
  }
  template <class T>
  __device__ static inline void f(T* du, const T* u, const T* p, const T* __restrict__ ps,
                                  double t) {
    if constexpr (sizeof(T) == 8) {
      typedef double d4 __attribute__((ext_vector_type(4)));
      d4 acc = {p[0], p[1], p[2], p[3]};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(p[4 + kb], u[kb], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) du[r] = u[r] * acc[r];
    } else {
      typedef float f4 __attribute__((ext_vector_type(4)));
      f4 acc = {p[0], p[1], p[2], p[3]};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(p[4 + kb], u[kb], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) du[r] = u[r] * acc[r];
    }
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
