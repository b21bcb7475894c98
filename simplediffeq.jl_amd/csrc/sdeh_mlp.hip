// Learned cost model kernels: a 6 -> 64 -> 64 -> 2 MLP (tanh hidden layers,
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
// 8(r/4) + 4(l/32) + r%4 and unit l % 32 (+32 t); after tanh the tile goes through
// LDS (rows padded to 65 floats) so that layer 3 (64 -> 2) reads each sample's row.
// Inference uses a fast tanh (1 - 2/(e^{2x} + 1) with the hardware exp2 and
// reciprocal; absolute error about 1e-7); training keeps tanhf.
typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int H2P = NH + 1;

__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // e^{2x}
  return fmaf(-2.0f, __builtin_amdgcn_rcpf(e + 1.0f), 1.0f);
}

// tanh(x) from a = 2 x log2(e), the scaling already applied by the caller.
__device__ __forceinline__ float tanh_prescaled(float a) {
  return fmaf(-2.0f, __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(a) + 1.0f), 1.0f);
}

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
// activations rounded to f16 and every product accumulated in FP32 on
// v_mfma_f32_32x32x8_f16 (K = 8 per instruction).  Each layer is computed
// transposed, D^T = W X^T (M = units, N = 32 samples), because the MFMA output
// layout then hands the next layer its B operand without moving data: lane l holds
// sample l % 32 and, in register r of output tile t, unit 32 t + 8 (r / 4) +
// 4 (l / 32) + r % 4, which is exactly the B operand (rows k = 8 m + 4 (l / 32) + i,
// column l % 32) of the next layer's K step m = 4 t + r / 4, i = r % 4.  Biases are
// one more K step against a constant 1 input (layer 1: feature column 6), and layer
// 3 (two outputs) is rows 0 and 1 of a padded 32-row tile.  The six features of a
// sample are split between its two lanes: lane l < 32 computes the parameter and
// horizon features (K rows 0-3), lane l + 32 the initial-state ones (rows 4-5), the
// bias input 1 (row 6) and a zero (row 7).  Activations use the hardware exp and
// reciprocal (tanh = 1 - 2 / (e^2x + 1), absolute error about 1e-7 before the f16
// rounding), with the factor 2 log2(e) of the exponent folded into the layer 1 and
// layer 2 weights and biases before they are rounded to f16, so that each activation
// is exp2, add, reciprocal and fma.  The predictions only order and seed the solves;
// sdeh_mlp_forward keeps the FP32 forward pass.
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));

// Four floats as f16 (rounded toward zero, two per v_cvt_pkrtz_f16_f32).
__device__ __forceinline__ half4 pack4(float a, float b, float c, float d) {
  const fp16x2 lo = __builtin_amdgcn_cvt_pkrtz(a, b), hi = __builtin_amdgcn_cvt_pkrtz(c, d);
  half4 r;
  r[0] = (_Float16)lo[0];
  r[1] = (_Float16)lo[1];
  r[2] = (_Float16)hi[0];
  r[3] = (_Float16)hi[1];
  return r;
}

// Features of sample smp in FP32 for lane half q (see above): q = 0: (p_c - f[2c]) /
// f[2c+1] for c = 0, 1, 2 and (log(tf - t0) - f[6]) / f[7]; q = 1: |u0| / sqrt(3) *
// 2 - 1, u0_2, 1, 0.
struct XLorenzPlan {
  const double* __restrict__ p;
  const double* __restrict__ u0;
  const double* __restrict__ ts;
  int64_t n;
  float off[4], inv[4];
  __device__ half4 operator()(int64_t smp, bool valid, int q) const {
    if (!valid) return pack4(0.f, 0.f, q ? 1.f : 0.f, 0.f);
    if (q == 0) {
      const float x0 = ((float)p[smp] - off[0]) * inv[0];
      const float x1 = ((float)p[n + smp] - off[1]) * inv[1];
      const float x2 = ((float)p[2 * n + smp] - off[2]) * inv[2];
      const float x3 = (__logf((float)(ts[n + smp] - ts[smp])) - off[3]) * inv[3];
      return pack4(x0, x1, x2, x3);
    }
    const float a = (float)u0[smp], b = (float)u0[n + smp], z = (float)u0[2 * n + smp];
    const float r = __builtin_sqrtf((a * a + b * b) + z * z);
    return pack4(r * 1.1547005383792515f - 1.f, z, 1.f, 0.f);
  }
};

template <class Sink>
__global__ void __launch_bounds__(256) mlp_plan_f16(const float* __restrict__ params,
                                                    XLorenzPlan src, Sink sink, int64_t n) {
  __shared__ float sw[NPARAM];
  load_params(sw, params);
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = l & 31, q = l >> 5;
  // A operands (weights, fixed per kernel): lane l holds row li of the tile, K columns
  // 4q .. 4q + 3 of each K step; step 8 of layers 2 and 3 carries the bias (column 0).
  // Layers 1 and 2 are scaled by c = 2 log2(e): their outputs are the exponents of
  // tanh_prescaled.
  constexpr float c = 2.8853900817779268f;
  half4 w1[2], w2[2][9], w3[9];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int u = 32 * t + li;
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = 4 * q + i;
      v[i] = c * (k < NIN ? sw[OFF_W1 + u * NIN + k] : (k == NIN ? sw[OFF_B1 + u] : 0.f));
    }
    w1[t] = pack4(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const float* w = sw + OFF_W2 + u * NH + 8 * m + 4 * q;
      w2[t][m] = pack4(c * w[0], c * w[1], c * w[2], c * w[3]);
    }
    w2[t][8] = pack4(q == 0 ? c * sw[OFF_B2 + u] : 0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int k = 8 * m + 4 * q;
    const float* w = sw + OFF_W3 + (li < NOUT ? li : 0) * NH + k;
    w3[m] = li < NOUT ? pack4(w[0], w[1], w[2], w[3]) : pack4(0.f, 0.f, 0.f, 0.f);
  }
  w3[8] = pack4(li < NOUT && q == 0 ? sw[OFF_B3 + (li < NOUT ? li : 0)] : 0.f, 0.f, 0.f, 0.f);
  const half4 one = pack4(q == 0 ? 1.f : 0.f, 0.f, 0.f, 0.f);  // the bias step's input row
  const floatx16 zero = {};
  const int64_t tiles = (n + 31) / 32;
  for (int64_t tile = (int64_t)blockIdx.x * 4 + wv; tile < tiles;
       tile += (int64_t)gridDim.x * 4) {
    const int64_t smp = tile * 32 + li;
    const bool valid = smp < n;
    const half4 x = src(smp, valid, q);
    // layer 1: two 32-unit tiles
    floatx16 d1[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) d1[t] = __builtin_amdgcn_mfma_f32_32x32x8f16(w1[t], x, zero, 0, 0, 0);
    half4 h1[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const floatx16& d = d1[m >> 2];
      const int r = 4 * (m & 3);
      h1[m] = pack4(tanh_prescaled(d[r]), tanh_prescaled(d[r + 1]), tanh_prescaled(d[r + 2]),
                    tanh_prescaled(d[r + 3]));
    }
    // layer 2: two output tiles, eight K steps and the bias step each
    floatx16 d2[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      floatx16 acc = __builtin_amdgcn_mfma_f32_32x32x8f16(w2[t][8], one, zero, 0, 0, 0);
#pragma unroll
      for (int m = 0; m < 8; ++m) acc = __builtin_amdgcn_mfma_f32_32x32x8f16(w2[t][m], h1[m], acc, 0, 0, 0);
      d2[t] = acc;
    }
    half4 h2[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const floatx16& d = d2[m >> 2];
      const int r = 4 * (m & 3);
      h2[m] = pack4(tanh_prescaled(d[r]), tanh_prescaled(d[r + 1]), tanh_prescaled(d[r + 2]),
                    tanh_prescaled(d[r + 3]));
    }
    // layer 3: output rows 0 and 1 land in registers 0 and 1 of lanes 0..31
    floatx16 d3 = __builtin_amdgcn_mfma_f32_32x32x8f16(w3[8], one, zero, 0, 0, 0);
#pragma unroll
    for (int m = 0; m < 8; ++m) d3 = __builtin_amdgcn_mfma_f32_32x32x8f16(w3[m], h2[m], d3, 0, 0, 0);
    if (q == 0 && valid) sink(smp, d3[0], d3[1]);
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
      h1[m] = tanh_fast(z);
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
      hs[row * H2P + li] = tanh_fast(acc0[r] + b2a);
      hs[row * H2P + 32 + li] = tanh_fast(acc1[r] + b2b);
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
// to change layout.  Layers 1 and 3, the biases, the loss and the tanh derivative
// are vector work.  The wave's gradient (and squared-error sum) goes to
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
      h1a[m] = tanhf(z);
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
      L.h2[row * H2P + li] = tanhf(z0[r] + b2v[0]);
      L.h2[row * H2P + 32 + li] = tanhf(z1[r] + b2v[1]);
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
    // D2 = (W3^T g) (1 - H2^2) in the D layout (rows: samples, column: unit)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * q + (r & 3);
        const float g0 = L.g[row * NOUT], g1 = L.g[row * NOUT + 1];
        const float h = L.h2[row * H2P + 32 * t + li];
        const float d = fmaf(w3c[0][t], g0, w3c[1][t] * g1) * (1.f - h * h);
        L.d2[row * H2P + 32 * t + li] = d;
        gb2[t] += d;
        gw3[0][t] = fmaf(g0, h, gw3[0][t]);
        gw3[1][t] = fmaf(g1, h, gw3[1][t]);
      }
    wave_sync_lds();
    // D1 = D2 W2 on the matrix cores, then times (1 - H1^2)
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
        const float dd = (t == 0 ? d0[r] : d1[r]) * (1.f - h * h);
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
