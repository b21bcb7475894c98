// Learned cost model kernels: a 6 -> 64 -> 64 -> 2 MLP (tanh hidden layers,
// linear output, FP32) trained on device.  It predicts, from a trajectory's
// parameters, initial state and horizon, the log of its accepted-step count
// and the log of a good initial dt; the runtime's caller uses the predictions
// to bin trajectories into wavefronts of similar cost and to seed dt0.  The
// network has 4738 parameters, so every pass is tiny next to one ensemble
// solve; the kernels favour determinism (fixed reduction order) over speed.
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

// Forward pass of one sample from LDS weights.
__device__ inline void forward_one(const float* sw, const float (&x)[NIN], float (&h1)[NH],
                                   float (&h2)[NH], float (&out)[NOUT]) {
#pragma unroll 4
  for (int i = 0; i < NH; ++i) {
    float z = sw[OFF_B1 + i];
#pragma unroll
    for (int j = 0; j < NIN; ++j) z = fmaf(sw[OFF_W1 + i * NIN + j], x[j], z);
    h1[i] = tanhf(z);
  }
#pragma unroll 2
  for (int i = 0; i < NH; ++i) {
    float z = sw[OFF_B2 + i];
#pragma unroll
    for (int k = 0; k < NH; ++k) z = fmaf(sw[OFF_W2 + i * NH + k], h1[k], z);
    h2[i] = tanhf(z);
  }
#pragma unroll
  for (int o = 0; o < NOUT; ++o) {
    float z = sw[OFF_B3 + o];
#pragma unroll
    for (int k = 0; k < NH; ++k) z = fmaf(sw[OFF_W3 + o * NH + k], h2[k], z);
    out[o] = z;
  }
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
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(256) mlp_forward_mfma(const float* __restrict__ params,
                                                        const float* __restrict__ x, float* y,
                                                        int64_t n) {
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
#pragma unroll
    for (int j = 0; j < NIN; ++j) xi[j] = valid ? x[smp * NIN + j] : 0.f;
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
    if (q == 0 && valid) {
      y[smp * NOUT] = o0 + sw[OFF_B3];
      y[smp * NOUT + 1] = o1 + sw[OFF_B3 + 1];
    }
  }
}

// One wave per block.  Each pass takes 64 samples (one per lane), runs the
// forward and backward pass, stages the per-sample vectors in LDS, and lane i
// accumulates the gradient entries of hidden unit i over the 64 samples.
// Partial sums per block go to `partial` [gridDim.x][NPARAM + 1] (last = loss).
constexpr int LDP = NH + 1;  // padded row stride
__global__ void __launch_bounds__(64) mlp_grad_partial(const float* __restrict__ params,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ y, int64_t n,
                                                       float* partial) {
  __shared__ float sw[NPARAM];
  __shared__ float s_h1[64 * LDP], s_h2[64 * LDP], s_d1[64 * LDP], s_d2[64 * LDP];
  __shared__ float s_x[64 * NIN], s_g[64 * NOUT], s_l[64];
  load_params(sw, params);
  const int lane = threadIdx.x;
  const float inv = 1.0f / (float)(n * NOUT);  // loss = mean over n * NOUT outputs
  float gW2[NH], gW1[NIN], gb1 = 0.f, gb2 = 0.f, gW3[NOUT], gb3 = 0.f, gl = 0.f;
#pragma unroll
  for (int k = 0; k < NH; ++k) gW2[k] = 0.f;
#pragma unroll
  for (int j = 0; j < NIN; ++j) gW1[j] = 0.f;
#pragma unroll
  for (int o = 0; o < NOUT; ++o) gW3[o] = 0.f;

  for (int64_t base = (int64_t)blockIdx.x * 64; base < n; base += (int64_t)gridDim.x * 64) {
    const int64_t s = base + lane;
    const bool valid = s < n;
    float xi[NIN], h1[NH], h2[NH], out[NOUT], g[NOUT];
#pragma unroll
    for (int j = 0; j < NIN; ++j) xi[j] = valid ? x[s * NIN + j] : 0.f;
    forward_one(sw, xi, h1, h2, out);
    float l = 0.f;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
      const float e = valid ? out[o] - y[s * NOUT + o] : 0.f;
      l = fmaf(e, e, l);
      g[o] = 2.f * e * inv;
    }
    // dL/dz2 = (W3^T g) * (1 - h2^2)
    float d2[NH];
#pragma unroll
    for (int k = 0; k < NH; ++k) {
      float a = 0.f;
#pragma unroll
      for (int o = 0; o < NOUT; ++o) a = fmaf(sw[OFF_W3 + o * NH + k], g[o], a);
      d2[k] = a * (1.f - h2[k] * h2[k]);
    }
    // dL/dz1 = (W2^T d2) * (1 - h1^2)
#pragma unroll 2
    for (int k = 0; k < NH; ++k) {
      float a = 0.f;
#pragma unroll
      for (int i = 0; i < NH; ++i) a = fmaf(sw[OFF_W2 + i * NH + k], d2[i], a);
      s_d1[lane * LDP + k] = a * (1.f - h1[k] * h1[k]);
    }
#pragma unroll
    for (int k = 0; k < NH; ++k) {
      s_h1[lane * LDP + k] = h1[k];
      s_h2[lane * LDP + k] = h2[k];
      s_d2[lane * LDP + k] = d2[k];
    }
#pragma unroll
    for (int j = 0; j < NIN; ++j) s_x[lane * NIN + j] = xi[j];
#pragma unroll
    for (int o = 0; o < NOUT; ++o) s_g[lane * NOUT + o] = g[o];
    s_l[lane] = l;
    __syncthreads();
    // lane i owns hidden unit i: rows W2[i][:], W1[i][:], b1[i], b2[i], column W3[:][i]
    const int i = lane;
    for (int t = 0; t < 64; ++t) {
      const float d2ti = s_d2[t * LDP + i];
      const float d1ti = s_d1[t * LDP + i];
#pragma unroll
      for (int k = 0; k < NH; ++k) gW2[k] = fmaf(d2ti, s_h1[t * LDP + k], gW2[k]);
#pragma unroll
      for (int j = 0; j < NIN; ++j) gW1[j] = fmaf(d1ti, s_x[t * NIN + j], gW1[j]);
      gb1 += d1ti;
      gb2 += d2ti;
      const float h2ti = s_h2[t * LDP + i];
#pragma unroll
      for (int o = 0; o < NOUT; ++o) gW3[o] = fmaf(s_g[t * NOUT + o], h2ti, gW3[o]);
      if (i < NOUT) gb3 += s_g[t * NOUT + i];
      if (i == 0) gl += s_l[t];
    }
    __syncthreads();
  }
  float* pb = partial + (int64_t)blockIdx.x * (NPARAM + 1);
  const int i = lane;
#pragma unroll
  for (int k = 0; k < NH; ++k) pb[OFF_W2 + i * NH + k] = gW2[k];
#pragma unroll
  for (int j = 0; j < NIN; ++j) pb[OFF_W1 + i * NIN + j] = gW1[j];
  pb[OFF_B1 + i] = gb1;
  pb[OFF_B2 + i] = gb2;
#pragma unroll
  for (int o = 0; o < NOUT; ++o) pb[OFF_W3 + o * NH + i] = gW3[o];
  if (i < NOUT) pb[OFF_B3 + i] = gb3;
  if (i == 0) pb[NPARAM] = gl;
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

extern "C" SDEH_HIDDEN int sdeh_mlp_forward_impl(const float* params, const float* x, float* y,
                                                 int64_t n, hipStream_t s) {
  if (n == 0) return 0;
  int64_t g = ((n + 31) / 32 + 3) / 4;  // four waves of 32-sample tiles per block
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(mlp_forward_mfma, dim3((unsigned)g), dim3(256), 0, s, params, x, y, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Blocks of the gradient pass: one 64-lane block per 64 samples, at most 1024
// (the blocks then stride over the rest).  The partial sums need
// sdeh_mlp_grad_scratch_bytes(n) bytes of scratch.
static int64_t grad_blocks(int64_t n) {
  int64_t nb = (n + 63) / 64;
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
  hipLaunchKernelGGL(mlp_grad_partial, dim3((unsigned)nb), dim3(64), 0, s, params, x, y, n,
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
