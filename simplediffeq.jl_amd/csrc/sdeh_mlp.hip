// Learned cost model kernels: a 6 -> 64 -> 64 -> 2 MLP (ReLU hidden layers,
// linear output, FP32) trained on device.  It predicts, from a trajectory's
// parameters, initial state and horizon, the log of its accepted-step count
// and the log of a good initial dt; the runtime's caller uses the predictions
// to bin trajectories into wavefronts of similar cost and to seed dt0.  The
// network has 4738 parameters; its two 64 x 64 products (and the backward pass's)
// run on the matrix cores, and every reduction has a fixed order (deterministic).
//
// Parameter layout (FP32): W1[64][6], b1[64], W2[64][64], b2[64], W3[2][64], b3[2].
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SDEH_HIDDEN __attribute__((visibility("hidden")))

namespace {

constexpr int NIN = 6, NH = 64, NOUT = 2;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NOUT * NH;
constexpr int NPARAM = OFF_B3 + NOUT;
static_assert(NPARAM == 4738, "parameter count");

__device__ inline void load_params(float* sw, const float* __restrict__ params) {
  for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) sw[i] = params[i];
  __syncthreads();
}

// ---- inference on the matrix cores ------------------------------------------
// Each wave takes tiles of 32 samples.  Layer 1 (6 -> 64) is vector work: lane
// (row i = lane % 32, parity q = lane / 32) computes the 32 hidden units k = 2m + q
// of sample i, exactly the A operand v_mfma_f32_32x32x2f32 wants (A[i][k] with
// k = lane / 32 for K-step m).  Layer 2 (64 x 64) is 2 x 32 MFMAs per tile: D tile t
// (units 32t .. 32t+31) = sum over m of A_m x B_m with B_m[k][j] = W2[32t + j][2m + k],
// held in 64 registers per lane.  D register r of lane l holds sample row
// 8(r/4) + 4(l/32) + r%4 and unit l % 32 (+32 t); after the ReLU the tile goes
// through LDS (rows padded to 65 floats) so that layer 3 (64 -> 2) reads each
// sample's row.
typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int H2P = NH + 1;

__device__ __forceinline__ float relu(float x) { return fmaxf(x, 0.0f); }

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Feature sources and output sinks of the forward pass.
struct XFromBuffer {  // x [n][6] FP32
  const float* __restrict__ x;
  __device__ void operator()(int64_t smp, bool valid, float (&xi)[NIN]) const {
#pragma unroll
    for (int j = 0; j < NIN; ++j) xi[j] = valid ? x[smp * NIN + j] : 0.f;
  }
};

struct YToBuffer {  // y [n][2] FP32
  float* y;
  __device__ void operator()(int64_t smp, float o0, float o1) const {
    y[smp * NOUT] = o0;
    y[smp * NOUT + 1] = o1;
  }
};

// Binning step: pred = y * std + mean (FP32, two roundings), key = -pred_0 (an
// ascending stable sort of the keys hands out the largest predicted costs first),
// dt0 = exp(pred_1) in FP32 widened to binary64; pred is optional.
struct YPlan {
  float* key;
  double* dt0;
  float* pred;
  float m0, m1, s0, s1;
  __device__ void operator()(int64_t smp, float o0, float o1) const {
    const float a0 = o0 * s0, a1 = o1 * s1;
    const float p0 = a0 + m0, p1 = a1 + m1;
    key[smp] = -p0;
    dt0[smp] = (double)expf(p1);
    if (pred) {
      pred[smp * NOUT] = p0;
      pred[smp * NOUT + 1] = p1;
    }
  }
};

// ---- the binning step on the f16 matrix cores -------------------------------
// sdeh_mlp_plan_lorenz's forward pass: the same network with its weights and
// hidden activations rounded to f16 and every product accumulated in FP32 on
// v_mfma_f32_32x32x16_f16 (K = 16 per instruction).  Each layer is computed
// transposed, D^T = W X^T (M = units, N = 32 samples), because the MFMA output
// layout then hands the next layer its B operand without moving data: in output
// tile t, register r of lane l holds sample l % 32 and unit
// 32 t + 8 (r / 4) + 4 (l / 32) + r % 4.  The next layer's K step m takes registers
// 8 (m % 2) .. 8 (m % 2) + 7 of tile m / 2, so element e of lane half q stands for
// unit 16 m + 8 (e / 4) + 4 q + e % 4 (unit_of_k).  The weights, the A operand, are
// loaded under the same labelling of K; the product only needs both operands to
// agree, since the hardware pairs element e of lane half q of A with the same
// element of B.  Layer 1 has one K step: lane half 0 holds the parameter and horizon
// features (elements 0-3), lane half 1 the initial-state ones (elements 0, 1) and a
// constant 1 (element 2) against the biases b1.  Layer 2's biases are the
// accumulator's initial value, in the output layout.  Layer 3 (two outputs) is rows
// 0 and 1 of a 32-row tile, its biases added in FP32.  The activations are max(x, 0)
// of the f16-rounded sums.  The predictions only order and seed the solves;
// sdeh_mlp_forward keeps the FP32 forward pass.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));

// Eight floats as f16 (rounded toward zero, two per v_cvt_pkrtz_f16_f32).
__device__ __forceinline__ half8 pack8(const float (&v)[8]) {
  half8 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const fp16x2 h = __builtin_amdgcn_cvt_pkrtz(v[2 * i], v[2 * i + 1]);
    r[2 * i] = (_Float16)h[0];
    r[2 * i + 1] = (_Float16)h[1];
  }
  return r;
}

// max(x, 0) of registers b .. b + 7 of an output tile, as a B operand: the max is
// taken after the rounding, two values per v_pk_max_f16 (the same result as
// rounding max(x, 0): rounding toward zero keeps the sign and 0 stays 0).
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ half8 relu_pack8(const floatx16& d, int b) {
  const half2v z = {(_Float16)0.f, (_Float16)0.f};
  half8 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const fp16x2 h = __builtin_amdgcn_cvt_pkrtz(d[b + 2 * i], d[b + 2 * i + 1]);
    const half2v m = __builtin_elementwise_max(half2v{(_Float16)h[0], (_Float16)h[1]}, z);
    r[2 * i] = m[0];
    r[2 * i + 1] = m[1];
  }
  return r;
}

// The hidden unit that element e of lane half q stands for in K step m.
__device__ __forceinline__ constexpr int unit_of_k(int m, int q, int e) {
  return 16 * m + 8 * (e >> 2) + 4 * q + (e & 3);
}

// Features of a sample for lane half q.  load() reads the binary64 inputs (q = 0:
// p_0, p_1, p_2, t0, tf; q = 1: u0_0, u0_1, u0_2) of sample min(smp, n - 1), so a
// tile's tail and the prefetches past the last tile read in bounds; operator() makes
// the FP32 features and the f16 K step: q = 0: (p_c - off[c]) * inv[c] for c = 0, 1, 2
// and (log(tf - t0) - off[3]) * inv[3]; q = 1: |u0| / sqrt(3) * 2 - 1, u0_2, 1.
struct XLorenzPlan {
  const double* __restrict__ p;
  const double* __restrict__ u0;
  const double* __restrict__ ts;
  int64_t n;
  float off[4], inv[4];
  struct Raw {
    double v[5];
  };
  __device__ Raw load(int64_t smp, int q) const {
    const int64_t s = smp < n ? smp : n - 1;
    const double* a = q ? u0 : p;
    Raw r;
    r.v[0] = a[s];
    r.v[1] = a[n + s];
    r.v[2] = a[2 * n + s];
    r.v[3] = ts[s];  // used by lane half 0 only; loaded by both, same lines
    r.v[4] = ts[n + s];
    return r;
  }
  __device__ half8 operator()(const Raw& r, int q) const {
    float v[8] = {};
    if (q == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = ((float)r.v[c] - off[c]) * inv[c];
      v[3] = (__logf((float)(r.v[4] - r.v[3])) - off[3]) * inv[3];
    } else {
      const float a = (float)r.v[0], b = (float)r.v[1], z = (float)r.v[2];
      const float rr = __builtin_sqrtf((a * a + b * b) + z * z);
      v[0] = rr * 1.1547005383792515f - 1.f;
      v[1] = z;
      v[2] = 1.f;
    }
    return pack8(v);
  }
};

template <class Sink>
__global__ void __launch_bounds__(256, 2) mlp_plan_f16(const float* __restrict__ params,
                                                       XLorenzPlan src, Sink sink, int64_t n) {
  __shared__ float sw[NPARAM];
  // layer 3's A operand, f16, per K step m and lane half q: rows 0 and 1 (W3) and a
  // zero row 2 that lanes 2..31 read
  __shared__ half8 sw3[4][2][3];
  load_params(sw, params);
  if (threadIdx.x < 24) {
    const int m = threadIdx.x / 6, hq = (threadIdx.x / 3) & 1, row = threadIdx.x % 3;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = row < NOUT ? sw[OFF_W3 + row * NH + unit_of_k(m, hq, e)] : 0.f;
    sw3[m][hq][row] = pack8(v);
  }
  __syncthreads();
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 31, q = l >> 5;
  // A operands of layers 1 and 2 (weights, fixed per kernel): lane l holds row li of
  // the tile and the eight K elements of lane half q; layer 2's biases (FP32) in the
  // output layout, the accumulator's initial value.
  half8 w1[2], w2[2][4];
  floatx16 b2[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int u = 32 * t + li;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int f = q == 0 ? (e < 4 ? e : -1) : (e < 2 ? 4 + e : (e == 2 ? NIN : -1));
      v[e] = f < 0 ? 0.f : (f < NIN ? sw[OFF_W1 + u * NIN + f] : sw[OFF_B1 + u]);
    }
    w1[t] = pack8(v);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = sw[OFF_W2 + u * NH + unit_of_k(m, q, e)];
      w2[t][m] = pack8(v);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) b2[t][r] = sw[OFF_B2 + 32 * t + 8 * (r >> 2) + 4 * q + (r & 3)];
  }
  const int row3 = li < NOUT ? li : NOUT;
  const float b3a = sw[OFF_B3], b3b = sw[OFF_B3 + 1];
  const floatx16 zero = {};
  const int64_t tiles = (n + 31) / 32, stride = (int64_t)gridDim.x * 4;
  auto tile_pass = [&](const XLorenzPlan::Raw& in, int64_t tile) {
    const int64_t smp = tile * 32 + li;
    const half8 x = src(in, q);
    floatx16 d1[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) d1[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(w1[t], x, zero, 0, 0, 0);
    half8 h1[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) h1[m] = relu_pack8(d1[m >> 1], 8 * (m & 1));
    floatx16 d2[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      floatx16 acc = b2[t];
#pragma unroll
      for (int m = 0; m < 4; ++m) acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(w2[t][m], h1[m], acc, 0, 0, 0);
      d2[t] = acc;
    }
    half8 h2[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) h2[m] = relu_pack8(d2[m >> 1], 8 * (m & 1));
    // layer 3: output rows 0 and 1 land in registers 0 and 1 of lanes 0..31
    floatx16 d3 = zero;
#pragma unroll
    for (int m = 0; m < 4; ++m)
      d3 = __builtin_amdgcn_mfma_f32_32x32x16_f16(sw3[m][q][row3], h2[m], d3, 0, 0, 0);
    if (q == 0 && smp < n) sink(smp, d3[0] + b3a, d3[1] + b3b);
  };
  // Three input buffers: while the wave computes a tile, the inputs of its next two
  // tiles are in flight.  The loop is unrolled by three so that each load writes the
  // buffer it refills (a register copy would wait for the load).
  int64_t tile = (int64_t)blockIdx.x * 4 + wv;
  XLorenzPlan::Raw r0 = src.load(tile * 32 + li, q);
  XLorenzPlan::Raw r1 = src.load((tile + stride) * 32 + li, q);
  XLorenzPlan::Raw r2 = src.load((tile + 2 * stride) * 32 + li, q);
  for (;;) {
    if (tile >= tiles) break;
    tile_pass(r0, tile);
    r0 = src.load((tile + 3 * stride) * 32 + li, q);
    tile += stride;
    if (tile >= tiles) break;
    tile_pass(r1, tile);
    r1 = src.load((tile + 3 * stride) * 32 + li, q);
    tile += stride;
    if (tile >= tiles) break;
    tile_pass(r2, tile);
    r2 = src.load((tile + 3 * stride) * 32 + li, q);
    tile += stride;
  }
}

template <class Src, class Sink>
__global__ void __launch_bounds__(256) mlp_forward_mfma(const float* __restrict__ params,
                                                        Src src, Sink sink, int64_t n) {
  __shared__ float sw[NPARAM];
  __shared__ float s_h2[4][32 * H2P];
  load_params(sw, params);
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 31, q = l >> 5;
  float wb[2][32];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int m = 0; m < 32; ++m) wb[t][m] = sw[OFF_W2 + (32 * t + li) * NH + 2 * m + q];
  const float b2a = sw[OFF_B2 + li], b2b = sw[OFF_B2 + 32 + li];
  float* hs = s_h2[wv];
  const int64_t tiles = (n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wv; tile < tiles;
       tile += (int64_t)gridDim.x * 4) {
    const int64_t smp = tile * 32 + li;
    const bool valid = smp < n;
    float xi[NIN];
    src(smp, valid, xi);
    float h1[32];
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      const int k = 2 * m + q;
      float z = sw[OFF_B1 + k];
#pragma unroll
      for (int j = 0; j < NIN; ++j) z = fmaf(sw[OFF_W1 + k * NIN + j], xi[j], z);
      h1[m] = relu(z);
    }
    floatx16 acc0 = {}, acc1 = {};
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(h1[m], wb[0][m], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(h1[m], wb[1][m], acc1, 0, 0, 0);
    }
    wave_sync_lds();  // the previous tile's reads of hs are done
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * q + (r & 3);
      hs[row * H2P + li] = relu(acc0[r] + b2a);
      hs[row * H2P + 32 + li] = relu(acc1[r] + b2b);
    }
    wave_sync_lds();
    float o0 = 0.f, o1 = 0.f;
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      const int k = 2 * m + q;
      const float h = hs[li * H2P + k];
      o0 = fmaf(sw[OFF_W3 + k], h, o0);
      o1 = fmaf(sw[OFF_W3 + NH + k], h, o1);
    }
    o0 += __shfl_xor(o0, 32);
    o1 += __shfl_xor(o1, 32);
    if (q == 0 && valid) sink(smp, o0 + sw[OFF_B3], o1 + sw[OFF_B3 + 1]);
  }
}

// ---- training pass on the matrix cores ---------------------------------------
// One wave per block; each wave takes 32-sample tiles and accumulates the gradient
// of its tiles.  The three 64 x 64 products run on v_mfma_f32_32x32x2f32 (layouts as
// in mlp_forward_mfma; A[i][k]: lane i + 32 (k % 2), B[k][j]: lane j + 32 (k % 2),
// D register r of lane l: row 8(r/4) + 4(l/32) + r%4, column l % 32):
//   forward  Z2[s][j] = sum_k H1[s][k] W2[j][k]      (M = samples, K = units)
//   backward D1[s][k] = sum_j D2[s][j] W2[j][k]      (M = samples, K = units)
//   gradient gW2[j][k] += sum_s D2[s][j] H1[s][k]    (M = units, K = samples)
// H1, H2, D2 and the output gradients of a tile pass through the wave's LDS tile
// to change layout.  Layers 1 and 3, the biases, the loss and the ReLU derivative
// (1 where the unit's output is positive, else 0) are vector work.  The wave's gradient (and squared-error sum) goes to
// partial[blockIdx.x]; mlp_grad_reduce adds the waves in a fixed order.
struct GradTile {
  float x[32 * NIN];
  float h1[32 * H2P];
  float h2[32 * H2P];
  float d2[32 * H2P];
  float g[32 * NOUT];
};

__global__ void __launch_bounds__(64) mlp_grad_mfma(const float* __restrict__ params,
                                                    const float* __restrict__ x,
                                                    const float* __restrict__ y, int64_t n,
                                                    float* partial) {
  __shared__ float sw[NPARAM];
  __shared__ GradTile L;
  load_params(sw, params);
  const int l = threadIdx.x, li = l & 31, q = l >> 5;
  const float inv = 1.0f / (float)(n * NOUT);
  float wb[2][32];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int m = 0; m < 32; ++m) wb[t][m] = sw[OFF_W2 + (32 * t + li) * NH + 2 * m + q];
  const float b2v[2] = {sw[OFF_B2 + li], sw[OFF_B2 + 32 + li]};
  const float w3c[2][2] = {{sw[OFF_W3 + li], sw[OFF_W3 + 32 + li]},
                           {sw[OFF_W3 + NH + li], sw[OFF_W3 + NH + 32 + li]}};
  floatx16 gw2[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) gw2[a][b] = floatx16{};
  float gw1[2][NIN], gb1[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gw3[NOUT][2], gb3[NOUT] = {0.f, 0.f};
  float gl = 0.f;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int j = 0; j < NIN; ++j) gw1[t][j] = 0.f;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) gw3[o][t] = 0.f;
  }
  const int64_t tiles = (n + 31) / 32;
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int64_t smp = tile * 32 + li;
    const bool valid = smp < n;
    float xi[NIN];
#pragma unroll
    for (int j = 0; j < NIN; ++j) xi[j] = valid ? x[smp * NIN + j] : 0.f;
    wave_sync_lds();  // the previous tile's LDS reads are done
    if (q == 0)
#pragma unroll
      for (int j = 0; j < NIN; ++j) L.x[li * NIN + j] = xi[j];
    // layer 1 in the A layout: lane (li, q) holds units k = 2m + q of sample li
    float h1a[32];
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      const int k = 2 * m + q;
      float z = sw[OFF_B1 + k];
#pragma unroll
      for (int j = 0; j < NIN; ++j) z = fmaf(sw[OFF_W1 + k * NIN + j], xi[j], z);
      h1a[m] = relu(z);
      L.h1[li * H2P + k] = h1a[m];
    }
    floatx16 z0 = {}, z1 = {};
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      z0 = __builtin_amdgcn_mfma_f32_32x32x2f32(h1a[m], wb[0][m], z0, 0, 0, 0);
      z1 = __builtin_amdgcn_mfma_f32_32x32x2f32(h1a[m], wb[1][m], z1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * q + (r & 3);
      L.h2[row * H2P + li] = relu(z0[r] + b2v[0]);
      L.h2[row * H2P + 32 + li] = relu(z1[r] + b2v[1]);
    }
    wave_sync_lds();
    // layer 3, error and output gradient of sample li (the two halves add up)
    float o0 = 0.f, o1 = 0.f;
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      const int k = 2 * m + q;
      const float h = L.h2[li * H2P + k];
      o0 = fmaf(sw[OFF_W3 + k], h, o0);
      o1 = fmaf(sw[OFF_W3 + NH + k], h, o1);
    }
    o0 += __shfl_xor(o0, 32);
    o1 += __shfl_xor(o1, 32);
    const float e0 = valid ? o0 + sw[OFF_B3] - y[smp * NOUT] : 0.f;
    const float e1 = valid ? o1 + sw[OFF_B3 + 1] - y[smp * NOUT + 1] : 0.f;
    if (q == 0) {
      const float g0 = 2.f * e0 * inv, g1 = 2.f * e1 * inv;
      L.g[li * NOUT] = g0;
      L.g[li * NOUT + 1] = g1;
      gb3[0] += g0;
      gb3[1] += g1;
      gl = fmaf(e0, e0, gl);
      gl = fmaf(e1, e1, gl);
    }
    wave_sync_lds();
    // D2 = (W3^T g) [H2 > 0] in the D layout (rows: samples, column: unit)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * q + (r & 3);
        const float g0 = L.g[row * NOUT], g1 = L.g[row * NOUT + 1];
        const float h = L.h2[row * H2P + 32 * t + li];
        const float d = h > 0.f ? fmaf(w3c[0][t], g0, w3c[1][t] * g1) : 0.f;
        L.d2[row * H2P + 32 * t + li] = d;
        gb2[t] += d;
        gw3[0][t] = fmaf(g0, h, gw3[0][t]);
        gw3[1][t] = fmaf(g1, h, gw3[1][t]);
      }
    wave_sync_lds();
    // D1 = D2 W2 on the matrix cores, then times [H1 > 0]
    floatx16 d0 = {}, d1 = {};
#pragma unroll
    for (int m = 0; m < 32; ++m) {
      const float a = L.d2[li * H2P + 2 * m + q];
      d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, sw[OFF_W2 + (2 * m + q) * NH + li], d0, 0, 0,
                                                0);
      d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, sw[OFF_W2 + (2 * m + q) * NH + 32 + li], d1,
                                                0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * q + (r & 3);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const float h = L.h1[row * H2P + 32 * t + li];
        const float dd = h > 0.f ? (t == 0 ? d0[r] : d1[r]) : 0.f;
        gb1[t] += dd;
#pragma unroll
        for (int j = 0; j < NIN; ++j) gw1[t][j] = fmaf(dd, L.x[row * NIN + j], gw1[t][j]);
      }
    }
    // gW2 += D2^T H1 on the matrix cores (K = the tile's samples)
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int srow = 2 * m + q;
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        const float a = L.d2[srow * H2P + 32 * tj + li];
#pragma unroll
        for (int tk = 0; tk < 2; ++tk)
          gw2[tj][tk] = __builtin_amdgcn_mfma_f32_32x32x2f32(
              a, L.h1[srow * H2P + 32 * tk + li], gw2[tj][tk], 0, 0, 0);
      }
    }
  }
  // the two halves of the wave hold different sample rows of the same units
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    gb1[t] += __shfl_xor(gb1[t], 32);
    gb2[t] += __shfl_xor(gb2[t], 32);
#pragma unroll
    for (int j = 0; j < NIN; ++j) gw1[t][j] += __shfl_xor(gw1[t][j], 32);
#pragma unroll
    for (int o = 0; o < NOUT; ++o) gw3[o][t] += __shfl_xor(gw3[o][t], 32);
  }
  // per-sample sums over the 32 sample lanes (the q == 1 lanes hold zeros)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    gb3[0] += __shfl_xor(gb3[0], off);
    gb3[1] += __shfl_xor(gb3[1], off);
    gl += __shfl_xor(gl, off);
  }
  float* pb = partial + (int64_t)blockIdx.x * (NPARAM + 1);
#pragma unroll
  for (int tj = 0; tj < 2; ++tj)
#pragma unroll
    for (int tk = 0; tk < 2; ++tk)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int jrow = 32 * tj + 8 * (r >> 2) + 4 * q + (r & 3);
        pb[OFF_W2 + jrow * NH + 32 * tk + li] = gw2[tj][tk][r];
      }
  if (q == 0) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int k = 32 * t + li;
#pragma unroll
      for (int j = 0; j < NIN; ++j) pb[OFF_W1 + k * NIN + j] = gw1[t][j];
      pb[OFF_B1 + k] = gb1[t];
      pb[OFF_B2 + k] = gb2[t];
#pragma unroll
      for (int o = 0; o < NOUT; ++o) pb[OFF_W3 + o * NH + k] = gw3[o][t];
    }
  }
  if (l == 0) {
    pb[OFF_B3] = gb3[0];
    pb[OFF_B3 + 1] = gb3[1];
    pb[NPARAM] = gl;
  }
}

// Sum the block partials in a fixed order (deterministic): accumulator r takes
// blocks r, r + 16, r + 32, ... in turn, then the 16 accumulators are added in
// order.  Sixteen independent loads per step keep the memory pipeline busy.
__global__ void mlp_grad_reduce(const float* __restrict__ partial, int nblocks, float* grad,
                                float* loss, float inv_count) {
  constexpr int R = 16;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > NPARAM) return;
  float acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.f;
  int b = 0;
  for (; b + R <= nblocks; b += R) {
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] += partial[(int64_t)(b + r) * (NPARAM + 1) + j];
  }
  for (int r = 0; b + r < nblocks; ++r) acc[r] += partial[(int64_t)(b + r) * (NPARAM + 1) + j];
  float s = acc[0];
#pragma unroll
  for (int r = 1; r < R; ++r) s += acc[r];
  if (j < NPARAM) grad[j] = s;
  else *loss = s * inv_count;
}

__global__ void mlp_adam(float* p, const float* __restrict__ g, float* m, float* v, float lr,
                         float b1, float b2, float eps, float c1, float c2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NPARAM) return;
  const float gj = g[j];
  const float mj = b1 * m[j] + (1.f - b1) * gj;
  const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
  m[j] = mj;
  v[j] = vj;
  const float mh = mj / c1;
  const float vh = vj / c2;
  p[j] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

// Resident blocks of a kernel on the current device, the grid of a forward launch:
// the blocks then loop over tiles with no partial last round of blocks.
static int64_t resident_grid(const void* kernel, int64_t n) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return -1;
  const int64_t resident = (int64_t)cus * per_cu;
  const int64_t want = ((n + 31) / 32 + 3) / 4;  // four waves of 32-sample tiles per block
  return want < resident ? want : resident;
}

template <class Src, class Sink>
static int64_t forward_grid(int64_t n) {
  // per process (initialised once, thread-safe); every MI355X has the same count
  static const int64_t full = resident_grid(
      reinterpret_cast<const void*>(&mlp_forward_mfma<Src, Sink>), int64_t(1) << 40);
  if (full <= 0) return -1;
  const int64_t want = ((n + 31) / 32 + 3) / 4;
  return want < full ? want : full;
}

extern "C" SDEH_HIDDEN int sdeh_mlp_forward_impl(const float* params, const float* x, float* y,
                                                 int64_t n, hipStream_t s) {
  if (n == 0) return 0;
  const int64_t g = forward_grid<XFromBuffer, YToBuffer>(n);
  if (g <= 0) return -2;
  hipLaunchKernelGGL((mlp_forward_mfma<XFromBuffer, YToBuffer>), dim3((unsigned)g), dim3(256), 0,
                     s, params, XFromBuffer{x}, YToBuffer{y}, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// The Lorenz cost model's binning step (features, forward, de-normalisation, sort
// key, dt0) in one launch; c[8] are the feature constants of XLorenz.
extern "C" SDEH_HIDDEN int sdeh_mlp_plan_lorenz_impl(const float* params, const double* p,
                                                     const double* u0, const double* tspan,
                                                     int64_t n, const double* c, float m0,
                                                     float m1, float s0, float s1, float* key,
                                                     double* dt0, float* pred, hipStream_t s) {
  if (n == 0) return 0;
  static const int64_t full =
      resident_grid(reinterpret_cast<const void*>(&mlp_plan_f16<YPlan>), int64_t(1) << 40);
  if (full <= 0) return -2;
  const int64_t want = ((n + 31) / 32 + 3) / 4;
  const int64_t g = want < full ? want : full;
  XLorenzPlan src{p, u0, tspan, n, {}, {}};
  for (int k = 0; k < 4; ++k) {
    src.off[k] = (float)c[2 * k];
    src.inv[k] = (float)(1.0 / c[2 * k + 1]);
  }
  hipLaunchKernelGGL((mlp_plan_f16<YPlan>), dim3((unsigned)g), dim3(256), 0, s, params, src,
                     YPlan{key, dt0, pred, m0, m1, s0, s1}, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Blocks of the gradient pass: one 64-lane block per 64 samples, at most 1024
// (the blocks then stride over the rest).  The partial sums need
// sdeh_mlp_grad_scratch_bytes(n) bytes of scratch.
static int64_t grad_blocks(int64_t n) {
  int64_t nb = (n + 31) / 32;  // one wave of 32-sample tiles per block
  return nb > 1024 ? 1024 : (nb < 1 ? 1 : nb);
}

extern "C" SDEH_HIDDEN size_t sdeh_mlp_grad_scratch_bytes(int64_t n) {
  return (size_t)grad_blocks(n) * (NPARAM + 1) * sizeof(float);
}

// `scratch` holds sdeh_mlp_grad_scratch_bytes(n) bytes.
extern "C" SDEH_HIDDEN int sdeh_mlp_loss_grad_impl(const float* params, const float* x,
                                                   const float* y, int64_t n, float* loss,
                                                   float* grad, float* scratch, hipStream_t s) {
  const int64_t nb = grad_blocks(n);
  hipLaunchKernelGGL(mlp_grad_mfma, dim3((unsigned)nb), dim3(64), 0, s, params, x, y, n,
                     scratch);
  if (hipGetLastError() != hipSuccess) return -2;
  hipLaunchKernelGGL(mlp_grad_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, scratch,
                     (int)nb, grad, loss, 1.0f / (float)(n * NOUT));
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" SDEH_HIDDEN int sdeh_mlp_adam_impl(float* p, const float* g, float* m, float* v,
                                              float lr, float b1, float b2, float eps, int32_t t,
                                              hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)t);
  const float c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, p, g, m, v, lr, b1,
                     b2, eps, c1, c2);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
