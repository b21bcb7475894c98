// Small device utilities of the runtime: fills (so no hipMemsetAsync lands in
// a captured graph), the queue reset, and the per-solve counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__global__ void fill_f64(double* p, double v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void fill_f32(float* p, float v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// out[0] += sum(naccept + nreject), out[1] += sum(naccept),
// out[2] += count(retcode == 1), out[3] += count(retcode == 4).
// Grid-stride loop, wave reduction, block reduction through LDS, then one
// atomic per block per counter (few blocks, so little contention).
__global__ void __launch_bounds__(256) count_kernel(const int32_t* nacc, const int32_t* nrej,
                                                    const int32_t* ret, int64_t n,
                                                    unsigned long long* out) {
  __shared__ unsigned long long part[4][4];
  unsigned long long v[4] = {0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned long long a = (unsigned long long)nacc[i];
    v[1] += a;
    v[0] += a + (unsigned long long)nrej[i];
    v[2] += ret[i] == 1;
    v[3] += ret[i] == 4;
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 4; ++k) v[k] += __shfl_down(v[k], off);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 4; ++k) part[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long s = 0;
    for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) s += part[ww][threadIdx.x];
    atomicAdd(out + threadIdx.x, s);
  }
}

// Permutation check of a lane hand-out order: every entry in [0, n) exactly once.
__global__ void check_order_kernel(const int64_t* order, int64_t n, int32_t* seen,
                                   unsigned long long* bad) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = order[i];
    if (v < 0 || v >= n) atomicAdd(bad, 1ull);
    else if (atomicAdd(seen + v, 1) != 0) atomicAdd(bad, 1ull);
  }
}

inline int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_fill_f64(double* p, double v,
                                                                              int64_t n,
                                                                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_f64, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_fill_f32(float* p, float v,
                                                                              int64_t n,
                                                                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_f32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_zero_u64(
    unsigned long long* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_count(
    const int32_t* nacc, const int32_t* nrej, const int32_t* ret, int64_t n,
    unsigned long long* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  int g = grid_for(n);
  if (g > 512) g = 512;
  hipLaunchKernelGGL(count_kernel, dim3(g), dim3(256), 0, s, nacc, nrej, ret, n, out);
  return hipGetLastError();
}

// `seen` must hold n int32; `bad` must be zero on entry.
extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_check_order(
    const int64_t* order, int64_t n, int32_t* seen, unsigned long long* bad, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_f32, dim3(grid_for(n)), dim3(256), 0, s, (float*)seen, 0.0f, n);
  hipLaunchKernelGGL(check_order_kernel, dim3(grid_for(n)), dim3(256), 0, s, order, n, seen, bad);
  return hipGetLastError();
}
