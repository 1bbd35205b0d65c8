// radix.hip -- stable LSD radix sort of (key, uint32 value) pairs for gfx950.
//
// One pass per 8-bit digit, three kernels per pass:
//   hist    : per 4096-element tile, an LDS histogram of the digit;
//   scan    : exclusive scan of the digit-major count matrix (3-phase);
//   scatter : per tile, stable ranks from wavefront ballots (a 64-lane match on the
//             8 digit bits), per-wave digit counters in LDS, then a prefix over the
//             tile's waves gives each element its global position.
// The element order inside a wave is (item slot, lane), so ranks follow the input
// order and the sort is stable.  Used by the k-d tree build (segmented value sort
// per level) and by order_by_depth.
#include <cstring>

#include "common.hpp"
#include "radix.hpp"

namespace ggp {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // 4096
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;

template <typename K>
__global__ __launch_bounds__(kThreads) void hist_kernel(const K* __restrict__ keys, int64_t n, int shift,
                                                        uint32_t* __restrict__ counts, int64_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile;
  for (int j = 0; j < kItems; ++j) {
    int64_t e = base + (int64_t)j * kThreads + threadIdx.x;
    if (e < n) atomicAdd(&h[(uint32_t)(keys[e] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  counts[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread (1024 threads).
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  if (wid == 0) {
    uint32_t w = lane < kScanThreads / 64 ? wsum[lane] : 0;
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t y = __shfl_up(w, o, 64);
      if (lane >= o) w += y;
    }
    if (lane < kScanThreads / 64) wsum[lane] = w;
  }
  __syncthreads();
  uint32_t before = wid > 0 ? wsum[wid - 1] : 0;
  if (total) *total = wsum[kScanThreads / 64 - 1];
  __syncthreads();
  return before + x - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ c, int64_t m,
                                                                   uint32_t* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  uint32_t s = 0;
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + (int64_t)threadIdx.x * kScanItems + j;
    if (e < m) s += c[e];
  }
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_top_kernel(uint32_t* __restrict__ sums, int64_t nchunks) {
  uint32_t carry = 0;
  for (int64_t b = 0; b < nchunks; b += kScanThreads) {
    int64_t e = b + threadIdx.x;
    uint32_t v = e < nchunks ? sums[e] : 0;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(v, &tot);
    if (e < nchunks) sums[e] = carry + ex;
    carry += tot;
  }
}

__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(uint32_t* __restrict__ c, int64_t m,
                                                                 const uint32_t* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanChunk;
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + (int64_t)threadIdx.x * kScanItems + j;
    v[j] = e < m ? c[e] : 0;
    s += v[j];
  }
  uint32_t ex = block_exclusive_scan(s, nullptr) + sums[blockIdx.x];
  for (int j = 0; j < kScanItems; ++j) {
    int64_t e = base + (int64_t)threadIdx.x * kScanItems + j;
    if (e < m) c[e] = ex;
    ex += v[j];
  }
}

template <typename K>
__global__ __launch_bounds__(kThreads) void scatter_kernel(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                           K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           int64_t n, int shift,
                                                           const uint32_t* __restrict__ offsets, int64_t nblocks) {
  __shared__ uint32_t cnt[kWaves][256];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)wid * 64 * kItems;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  K key[kItems];
  uint32_t val[kItems];
  uint32_t rank[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t e = base + (int64_t)j * 64 + lane;
    const bool valid = e < n;
    key[j] = valid ? kin[e] : K(0);
    val[j] = valid ? vin[e] : 0u;
    const uint32_t dig = (uint32_t)(key[j] >> shift) & 0xffu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (dig >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t before = cnt[wid][dig];
    rank[j] = before + (uint32_t)__popcll(peers & lt);
    const int leader = 63 - __clzll((long long)peers);
    if (valid && lane == leader) cnt[wid][dig] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // exclusive prefix over waves for every digit (thread t owns digit t)
  {
    const int dgt = threadIdx.x;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      uint32_t c = cnt[w][dgt];
      cnt[w][dgt] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t e = base + (int64_t)j * 64 + lane;
    if (e < n) {
      const uint32_t dig = (uint32_t)(key[j] >> shift) & 0xffu;
      const int64_t pos = (int64_t)offsets[(int64_t)dig * nblocks + blockIdx.x] + cnt[wid][dig] + rank[j];
      kout[pos] = key[j];
      vout[pos] = val[j];
    }
  }
}

}  // namespace

size_t radix_workspace(int64_t n) {
  const int64_t nblocks = (n + kTile - 1) / kTile;
  const int64_t m = nblocks * 256;
  const int64_t nchunks = (m + kScanChunk - 1) / kScanChunk;
  Workspace w(nullptr, 0);
  w.take<uint32_t>(m);
  w.take<uint32_t>(nchunks);
  return w.used + 256;
}

template <typename K>
int radix_sort_pairs(K** keys, uint32_t** vals, K* keys_alt, uint32_t* vals_alt, int64_t n, int begin_bit,
                     int end_bit, void* ws, size_t ws_bytes, hipStream_t s) {
  if (n <= 1 || end_bit <= begin_bit) return kOk;
  if (n > (int64_t)0xffffffffll) {
    set_error("radix sort: more than 2^32 elements");
    return kUnsupported;
  }
  const int64_t nblocks = (n + kTile - 1) / kTile;
  const int64_t m = nblocks * 256;
  const int64_t nchunks = (m + kScanChunk - 1) / kScanChunk;
  Workspace w(ws, ws_bytes);
  uint32_t* counts = w.take<uint32_t>(m);
  uint32_t* sums = w.take<uint32_t>(nchunks);
  if (!w.ok()) {
    set_error("radix sort: workspace too small");
    return kWorkspaceTooSmall;
  }
  K* kin = *keys;
  uint32_t* vin = *vals;
  K* kout = keys_alt;
  uint32_t* vout = vals_alt;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(hist_kernel<K>, dim3((unsigned)nblocks), dim3(kThreads), 0, s, kin, n, shift, counts,
                       nblocks);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nchunks), dim3(kScanThreads), 0, s, counts, m, sums);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kScanThreads), 0, s, sums, nchunks);
    hipLaunchKernelGGL(scan_down_kernel, dim3((unsigned)nchunks), dim3(kScanThreads), 0, s, counts, m, sums);
    hipLaunchKernelGGL(scatter_kernel<K>, dim3((unsigned)nblocks), dim3(kThreads), 0, s, kin, vin, kout, vout, n,
                       shift, counts, nblocks);
    GGP_LAUNCH_CHECK();
    K* tk = kin;
    kin = kout;
    kout = tk;
    uint32_t* tv = vin;
    vin = vout;
    vout = tv;
  }
  *keys = kin;
  *vals = vin;
  return kOk;
}

template int radix_sort_pairs<uint32_t>(uint32_t**, uint32_t**, uint32_t*, uint32_t*, int64_t, int, int, void*,
                                        size_t, hipStream_t);
template int radix_sort_pairs<uint64_t>(uint64_t**, uint32_t**, uint64_t*, uint32_t*, int64_t, int, int, void*,
                                        size_t, hipStream_t);

// ------------------------------------------------------------------ fills
__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u64_kernel(uint64_t* p, uint64_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int fill_u32(void* p, uint32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return kOk;
  unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(g), dim3(256), 0, s, static_cast<uint32_t*>(p), v, n);
  GGP_LAUNCH_CHECK();
  return kOk;
}
int read_to_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
  thread_local void* pinned = nullptr;  // one small buffer per host thread, kept until exit
  if (bytes > 4096) {
    set_error("read_to_host: at most 4096 bytes");
    return kInvalidArgument;
  }
  if (pinned == nullptr) GGP_HIP(hipHostMalloc(&pinned, 4096, hipHostMallocDefault));
  GGP_HIP(hipMemcpyAsync(pinned, src, bytes, hipMemcpyDeviceToHost, s));
  GGP_HIP(hipStreamSynchronize(s));
  std::memcpy(dst, pinned, bytes);
  return kOk;
}

int fill_u64(void* p, uint64_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return kOk;
  unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(g), dim3(256), 0, s, static_cast<uint64_t*>(p), v, n);
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace ggp
