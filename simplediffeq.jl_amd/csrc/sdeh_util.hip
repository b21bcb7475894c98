// Small device utilities of the runtime: fills (so no hipMemsetAsync lands in
// a captured graph), the queue reset, the per-solve counters, and the finishing
// passes of the packed save_everystep layout (offsets scan and compaction).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

// Device scratch of the allocation-free solve (sdeh_solve_static): words [0, 8) are
// the work-queue counter and the per-solve counts, the rest holds up to
// SDEH_STATIC_SAVEAT saveat times.  One copy per GPU, alive until the process ends;
// every sdeh_solve_static on a GPU shares it (calls that overlap on one GPU race).
// Host code never passes `scratch` itself to a kernel (in host code the name is a
// stand-in, not a device address): sdeh_util_static_scratch() looks the address up
// with hipGetSymbolAddress once per device, and callers use that pointer.
#define SDEH_STATIC_WORDS 8
#define SDEH_STATIC_SAVEAT 1024
static __device__ __align__(256) unsigned long long scratch[SDEH_STATIC_WORDS + SDEH_STATIC_SAVEAT];

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

// ---- packed save_everystep ------------------------------------------------
// Row counts cnt_i = (base ? base[i] : 0) + add are scanned in tiles of PK_TILE
// entries (256 threads x 4): pk_tile_sums -> pk_scan_tiles (one block) ->
// pk_offsets.
constexpr int PK_TILE = 1024;

// Exclusive scan over the 256 threads of a block; *total gets the block sum.
__device__ int64_t block_exclusive_scan(int64_t v, int64_t* total) {
  __shared__ int64_t sh[256];
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int64_t x = tid >= off ? sh[tid - off] : 0;
    __syncthreads();
    sh[tid] += x;
    __syncthreads();
  }
  const int64_t incl = sh[tid];
  *total = sh[255];
  __syncthreads();  // sh is reused by the next call
  return incl - v;
}

__device__ __forceinline__ int64_t pk_count(const int32_t* base, int32_t add, int64_t i) {
  return (base ? (int64_t)base[i] : 0) + add;
}

__global__ void __launch_bounds__(256) pk_tile_sums(const int32_t* base, int32_t add, int64_t n,
                                                    int64_t* tile_sum) {
  const int64_t i0 = (int64_t)blockIdx.x * PK_TILE + 4 * threadIdx.x;
  int64_t v = 0;
  for (int j = 0; j < 4; ++j)
    if (i0 + j < n) v += pk_count(base, add, i0 + j);
  int64_t total = 0;
  (void)block_exclusive_scan(v, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// One block: tile_sum[0..n_tiles) -> exclusive prefix sums, in place; *grand gets
// the sum of all counts.
__global__ void __launch_bounds__(256) pk_scan_tiles(int64_t* tile_sum, int64_t n_tiles,
                                                     unsigned long long* grand) {
  int64_t carry = 0;
  for (int64_t base = 0; base < n_tiles; base += 256) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < n_tiles ? tile_sum[i] : 0;
    int64_t total = 0;
    const int64_t ex = block_exclusive_scan(v, &total);
    if (i < n_tiles) tile_sum[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *grand = (unsigned long long)carry;
}

// off[i] = first row of trajectory i, clamped to the capacity; lim[i] = the rows
// it may write (its count, or what is left below the capacity); off[n] = rows used.
__global__ void __launch_bounds__(256) pk_offsets(const int32_t* base, int32_t add, int64_t n,
                                                  const int64_t* tile_off, int64_t cap,
                                                  int64_t* off, int32_t* lim) {
  const int64_t i0 = (int64_t)blockIdx.x * PK_TILE + 4 * threadIdx.x;
  int64_t c[4];
  int64_t v = 0;
  for (int j = 0; j < 4; ++j) {
    c[j] = i0 + j < n ? pk_count(base, add, i0 + j) : 0;
    v += c[j];
  }
  int64_t total = 0;
  int64_t o = tile_off[blockIdx.x] + block_exclusive_scan(v, &total);
  for (int j = 0; j < 4; ++j) {
    const int64_t i = i0 + j;
    if (i >= n) break;
    const int64_t room = cap > o ? cap - o : 0;
    lim[i] = (int32_t)(c[j] < room ? c[j] : room);
    off[i] = o < cap ? o : cap;
    o += c[j];
    if (i == n - 1) off[n] = o < cap ? o : cap;
  }
}

__global__ void write_i64(int64_t* p, int64_t v) {
  if (threadIdx.x == 0) *p = v;
}

// ---- longest span first ----------------------------------------------------
// A hand-out order by decreasing tf - t0, counting-sorted on a bucket of the span:
// 16 buckets per octave over 64 octaves (2^-30 .. 2^34; spans beyond clamp to the
// end buckets; spans <= 0 and NaN in bucket 0, last).  Each of SPAN_BLOCKS blocks
// takes a contiguous range of trajectories: span_hist counts its buckets in LDS,
// three scan passes (span_tile_sums, span_scan_tiles, span_tile_offsets) turn the
// counts into each (bucket, block)'s first slot (buckets in decreasing order, blocks
// in index order), and span_scatter places the block's
// trajectories, ranking them with LDS atomics.  No global atomics: the few buckets
// a narrow span distribution fills would serialise them.  Within a (bucket, block)
// the order is whatever the LDS atomics produce; no result depends on it.  The same
// sort runs on other weights (BucketSrc): a staged solve's expected remaining
// attempts (binary64 keys), and the cost model's predicted step counts, given as
// the plan's keys -log(count) in FP32.
constexpr int SPAN_BUCKETS = 1024;
constexpr int SPAN_BLOCKS = 256;

// Bucket of a positive weight w: 16 per octave, octave 0 = [2^-30, 2^-29).
__device__ __forceinline__ int weight_bucket(double w) {
  if (!(w > 0.0)) return 0;  // also NaN
  const unsigned long long bits = (unsigned long long)__double_as_longlong(w);
  int oct = (int)((bits >> 52) & 0x7ff) - (1023 - 30);
  oct = oct < 0 ? 0 : (oct > 63 ? 63 : oct);
  return oct * 16 + (int)((bits >> 48) & 15);
}

struct BucketSrc {
  const double* tspan;  // [2][n]: weight tf - t0
  const double* key;    // [n]: weight key[i]
  const float* lkey;    // [n]: weight exp(-lkey[i]), bucketed from lkey directly
  int64_t n;
  __device__ __forceinline__ int operator()(int64_t i) const {
    if (lkey) {
      // 16 * log2(exp(-k)) = -k * 16 / ln 2: 16 buckets per octave, evenly spaced
      // in log2 (weight_bucket spaces them evenly in the mantissa), octave 0 at
      // 2^-30 as there (bucket 480 starts at a count of 1)
      const float v = -lkey[i] * 23.083120654223414f;
      if (!(v == v)) return 0;  // NaN: last
      const float c = v < -480.0f ? -480.0f : (v > 543.0f ? 543.0f : v);
      const int b = (int)__builtin_floorf(c) + 480;
      return b < 1 ? 1 : b;
    }
    return weight_bucket(key ? key[i] : tspan[n + i] - tspan[i]);
  }
};

__device__ __forceinline__ void span_range(int64_t n, int64_t* lo, int64_t* hi) {
  const int64_t per = (n + SPAN_BLOCKS - 1) / SPAN_BLOCKS;
  *lo = (int64_t)blockIdx.x * per;
  *hi = *lo + per < n ? *lo + per : n;
}

// span_hist and span_scatter run SPAN_THREADS threads per block (16 waves per CU, so
// that key loads overlap the LDS atomics of other waves), each thread loading
// SPAN_BATCH keys before it ranks them.
constexpr int SPAN_THREADS = 1024;
constexpr int SPAN_BATCH = 4;

// bh[block][bucket] = trajectories of the block's range in the bucket.
__global__ void __launch_bounds__(SPAN_THREADS) span_hist(BucketSrc src, unsigned* bh) {
  __shared__ unsigned h[SPAN_BUCKETS];
  for (int k = threadIdx.x; k < SPAN_BUCKETS; k += blockDim.x) h[k] = 0u;
  __syncthreads();
  int64_t lo, hi;
  span_range(src.n, &lo, &hi);
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)SPAN_BATCH * blockDim.x) {
    int b[SPAN_BATCH];
#pragma unroll
    for (int j = 0; j < SPAN_BATCH; ++j) {
      const int64_t i = i0 + (int64_t)j * blockDim.x;
      b[j] = i < hi ? src(i) : -1;
    }
#pragma unroll
    for (int j = 0; j < SPAN_BATCH; ++j)
      if (b[j] >= 0) atomicAdd(h + b[j], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < SPAN_BUCKETS; k += blockDim.x)
    bh[(size_t)blockIdx.x * SPAN_BUCKETS + k] = h[k];
}

// The first slot of each (block, bucket): an exclusive scan of the counts taken in
// the order (bucket SPAN_BUCKETS-1 .. 0, then block 0 .. SPAN_BLOCKS-1), in three
// passes over tiles of SPAN_TILE entries of that order: the tile sums, their scan in
// one block, then each tile scanned and offset in place.  Entry e of the order is
// bucket SPAN_BUCKETS-1 - e / SPAN_BLOCKS of block e % SPAN_BLOCKS.
constexpr int SPAN_TILE = 1024;  // 256 threads x 4 entries
constexpr int SPAN_TILES = SPAN_BUCKETS * SPAN_BLOCKS / SPAN_TILE;
static_assert(SPAN_BUCKETS * SPAN_BLOCKS % SPAN_TILE == 0, "whole tiles");
static_assert(SPAN_TILES == 256, "one 256-thread block scans the tile sums");

__device__ __forceinline__ size_t span_entry(int64_t e) {
  const int64_t k = SPAN_BUCKETS - 1 - e / SPAN_BLOCKS, b = e % SPAN_BLOCKS;
  return (size_t)b * SPAN_BUCKETS + (size_t)k;
}

__global__ void __launch_bounds__(256) span_tile_sums(const unsigned* bh, unsigned* tsum) {
  const int64_t e0 = (int64_t)blockIdx.x * SPAN_TILE + 4 * threadIdx.x;
  int64_t v = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) v += bh[span_entry(e0 + j)];
  int64_t total = 0;
  (void)block_exclusive_scan(v, &total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = (unsigned)total;
}

__global__ void __launch_bounds__(256) span_scan_tiles(unsigned* tsum) {
  const int64_t v = tsum[threadIdx.x];
  int64_t total = 0;
  const int64_t ex = block_exclusive_scan(v, &total);
  tsum[threadIdx.x] = (unsigned)ex;
}

__global__ void __launch_bounds__(256) span_tile_offsets(unsigned* bh, const unsigned* tsum) {
  const int64_t e0 = (int64_t)blockIdx.x * SPAN_TILE + 4 * threadIdx.x;
  unsigned c[4];
  int64_t v = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    c[j] = bh[span_entry(e0 + j)];
    v += c[j];
  }
  int64_t total = 0;
  unsigned o = tsum[blockIdx.x] + (unsigned)block_exclusive_scan(v, &total);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    bh[span_entry(e0 + j)] = o;
    o += c[j];
  }
}

__global__ void __launch_bounds__(SPAN_THREADS) span_scatter(BucketSrc src, const unsigned* bh,
                                                             int64_t* order) {
  __shared__ unsigned cur[SPAN_BUCKETS];
  for (int k = threadIdx.x; k < SPAN_BUCKETS; k += blockDim.x)
    cur[k] = bh[(size_t)blockIdx.x * SPAN_BUCKETS + k];
  __syncthreads();
  int64_t lo, hi;
  span_range(src.n, &lo, &hi);
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += (int64_t)SPAN_BATCH * blockDim.x) {
    int b[SPAN_BATCH];
#pragma unroll
    for (int j = 0; j < SPAN_BATCH; ++j) {
      const int64_t i = i0 + (int64_t)j * blockDim.x;
      b[j] = i < hi ? src(i) : -1;
    }
#pragma unroll
    for (int j = 0; j < SPAN_BATCH; ++j)
      if (b[j] >= 0) order[atomicAdd(cur + b[j], 1u)] = i0 + (int64_t)j * blockDim.x;
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

// Offsets of the packed layout from the row counts (base[i] + add, or add when
// base is null): off [n + 1], lim [n] (rows each trajectory may write), *grand =
// total rows.  tiles: scratch of ceil(n / 1024) int64.
extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_pack_offsets(
    const int32_t* base, int32_t add, int64_t n, int64_t* tiles, int64_t cap, int64_t* off,
    int32_t* lim, unsigned long long* grand, hipStream_t s) {
  if (n <= 0) {
    hipLaunchKernelGGL(write_i64, dim3(1), dim3(64), 0, s, off, (int64_t)0);
    return hipGetLastError();
  }
  const int64_t n_tiles = (n + PK_TILE - 1) / PK_TILE;
  hipLaunchKernelGGL(pk_tile_sums, dim3((unsigned)n_tiles), dim3(256), 0, s, base, add, n, tiles);
  hipLaunchKernelGGL(pk_scan_tiles, dim3(1), dim3(256), 0, s, tiles, n_tiles, grand);
  hipLaunchKernelGGL(pk_offsets, dim3((unsigned)n_tiles), dim3(256), 0, s, base, add, n, tiles,
                     cap, off, lim);
  return hipGetLastError();
}

// order[0..n) = trajectories by decreasing weight (see span_hist).  bh: scratch of
// sdeh_util_span_scratch_bytes().  n must fit in 32 bits (the slot counters).
static hipError_t bucket_order(BucketSrc src, unsigned* bh, int64_t* order, hipStream_t s) {
  if (src.n <= 0) return hipSuccess;
  unsigned* tsum = bh + (size_t)SPAN_BLOCKS * SPAN_BUCKETS;
  hipLaunchKernelGGL(span_hist, dim3(SPAN_BLOCKS), dim3(SPAN_THREADS), 0, s, src, bh);
  hipLaunchKernelGGL(span_tile_sums, dim3(SPAN_TILES), dim3(256), 0, s, bh, tsum);
  hipLaunchKernelGGL(span_scan_tiles, dim3(1), dim3(SPAN_TILES), 0, s, tsum);
  hipLaunchKernelGGL(span_tile_offsets, dim3(SPAN_TILES), dim3(256), 0, s, bh, tsum);
  hipLaunchKernelGGL(span_scatter, dim3(SPAN_BLOCKS), dim3(SPAN_THREADS), 0, s, src, bh, order);
  return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_span_order(
    const double* tspan, int64_t n, unsigned* bh, int64_t* order, hipStream_t s) {
  return bucket_order(BucketSrc{tspan, nullptr, nullptr, n}, bh, order, s);
}

// order[0..n) = trajectories by decreasing key[i] (keys <= 0 and NaN last), on the
// same buckets and scratch as sdeh_util_span_order.
extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_key_order(
    const double* key, int64_t n, unsigned* bh, int64_t* order, hipStream_t s) {
  return bucket_order(BucketSrc{nullptr, key, nullptr, n}, bh, order, s);
}

// order[0..n) = trajectories by increasing lkey[i] (the cost model's -log predicted
// count: largest predicted cost first), on the 1/16-octave grid of exp(-lkey); NaN
// keys last.  Same scratch as sdeh_util_span_order.
extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_logkey_order(
    const float* lkey, int64_t n, unsigned* bh, int64_t* order, hipStream_t s) {
  return bucket_order(BucketSrc{nullptr, nullptr, lkey, n}, bh, order, s);
}

// Scratch of the bucket sorts: the per-block counts and the tile sums.
extern "C" __attribute__((visibility("hidden"))) size_t sdeh_util_span_scratch_bytes() {
  return ((size_t)SPAN_BLOCKS * SPAN_BUCKETS + SPAN_TILES) * sizeof(unsigned);
}

// The device address of `scratch` on the current device, looked up once per device
// (with that device current) and kept; *bytes = its size.  The first call on a
// device must come from the library's set-up (sdeh_create), before any solve uses it.
extern "C" __attribute__((visibility("hidden"))) hipError_t sdeh_util_static_scratch(
    void** ptr, size_t* bytes) {
  constexpr int kMaxDev = 64;
  static std::mutex mu;
  static void* addr[kMaxDev] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDev) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* q = nullptr;
    e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = q;
  }
  *ptr = addr[dev];
  *bytes = sizeof(scratch);
  return hipSuccess;
}

extern "C" __attribute__((visibility("hidden"))) int sdeh_util_static_saveat_max() {
  return SDEH_STATIC_SAVEAT;
}
