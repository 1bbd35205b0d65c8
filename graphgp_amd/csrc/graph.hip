// graph.hip -- depth, depth ordering and batch offsets of the Vecchia graph (gfx950).
//
// References: graphgp/graph.py:101-120 (compute_depths), :123-137 (order_by_depth),
// :95-96 (offsets), :44-70 (check_graph).
//
// compute_depths: depth[i] = 1 + max depth[neighbours(i)] (depth 0 for the n0 seed
// points).  The reference iterates that map to its fixed point (one sweep per depth
// level).  Here it is a single dataflow pass: every neighbour precedes its point, so
// workgroups take tiles in increasing order from an atomic ticket and each lane
// publishes its depth as soon as all its neighbours' depths are published (-1 marks
// "not yet").  A lane only waits on lower tickets, which belong to workgroups that are
// already running, so the pass always completes; a bounded spin flags a malformed
// graph instead of hanging.  The result is the reference's unique fixed point.
//
// order_by_depth: stable sort by depth (the radix sort of radix.hip), gathers, and the
// neighbour remap through the inverse permutation.
// offsets: offsets[b] = #(depth <= b) by binary search over the sorted depths.
#include "common.hpp"
#include "radix.hpp"
#include "graphgp_amd_internal.hpp"

namespace ggp {

namespace {

constexpr int kDepThreads = 256;
constexpr uint32_t kSpinLimit = 1u << 24;

__global__ __launch_bounds__(kDepThreads) void depths_kernel(const int32_t* __restrict__ nbrs, int64_t n_ref, int k,
                                                            int64_t n0, int32_t* depths, uint32_t* ticket,
                                                            uint32_t* err) {
  __shared__ uint32_t s_tile;
  const int64_t ntiles = (n_ref + kDepThreads - 1) / kDepThreads;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile >= ntiles) return;
    const int64_t r = tile * kDepThreads + threadIdx.x;
    const int64_t i = n0 + r;
    bool done = r >= n_ref;
    if (!done) {
      for (int j = 0; j < k; ++j) {
        int32_t nb = nbrs[r * k + j];
        if (nb < 0 || (int64_t)nb >= i) {
          atomicOr(err, 1u);
          __hip_atomic_store(&depths[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          done = true;
          break;
        }
      }
    }
    for (uint32_t it = 0;; ++it) {
      if (!done) {
        int32_t m = -1;
        bool ready = true;
        for (int j = 0; j < k; ++j) {
          int32_t nb = nbrs[r * k + j];
          int32_t v = __hip_atomic_load(&depths[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ready = ready && v >= 0;
          m = v > m ? v : m;
        }
        if (ready) {
          __hip_atomic_store(&depths[i], m + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          done = true;
        }
      }
      if (__all(done)) break;
      if (it >= kSpinLimit) {
        if (!done) {
          atomicOr(err, 2u);
          __hip_atomic_store(&depths[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();
  }
}

__global__ void depth_init_kernel(int32_t* depths, int64_t n, int64_t n0) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) depths[i] = i < n0 ? 0 : -1;
}

__global__ void depth_keys_kernel(const int32_t* __restrict__ depths, int64_t n, uint32_t* __restrict__ keys,
                                  uint32_t* __restrict__ pay) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    keys[i] = (uint32_t)depths[i];
    pay[i] = (uint32_t)i;
  }
}

template <typename T>
__global__ void order_gather_kernel(const T* __restrict__ pts, int d, const int32_t* __restrict__ idx,
                                    const int32_t* __restrict__ depths, const uint32_t* __restrict__ order,
                                    int64_t n, T* __restrict__ pts_o, int32_t* __restrict__ idx_o,
                                    int32_t* __restrict__ dep_o, uint32_t* __restrict__ inv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t src = order[i];
  for (int j = 0; j < d; ++j) pts_o[i * d + j] = pts[(int64_t)src * d + j];
  idx_o[i] = idx ? idx[src] : (int32_t)src;
  dep_o[i] = depths[src];
  inv[src] = (uint32_t)i;
}

__global__ void order_neighbors_kernel(const int32_t* __restrict__ nbrs, int k, const uint32_t* __restrict__ order,
                                       const uint32_t* __restrict__ inv, int64_t n0, int64_t n_ref,
                                       int32_t* __restrict__ nbrs_o) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ref * k) return;
  const int64_t r = e / k;
  const int j = (int)(e - r * k);
  const int64_t src = (int64_t)order[n0 + r] - n0;
  nbrs_o[e] = (int32_t)inv[nbrs[src * k + j]];
}

__global__ void offsets_kernel(const int32_t* __restrict__ depths_sorted, int64_t n, int64_t n_off,
                               int64_t* __restrict__ offsets) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_off) return;
  // lower_bound(depths, b + 1)
  int64_t lo = 0, hi = n;
  const int32_t key = (int32_t)(b + 1);
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (depths_sorted[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  offsets[b] = lo;
}

// check_graph (graph.py:44-70) for every refined row: topological order and no
// neighbour inside the row's own batch.  err bit 1: offsets malformed; 2: topology;
// 4: same-batch neighbour.
__global__ void check_graph_kernel(const int32_t* __restrict__ nbrs, int64_t n_ref, int k, int64_t n,
                                   const int64_t* __restrict__ offsets, int64_t n_off, uint32_t* err) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) {
    if (offsets[0] != n - n_ref || offsets[n_off - 1] > n) atomicOr(err, 1u);
    for (int64_t b = 1; b < n_off; ++b)
      if (offsets[b] < offsets[b - 1]) atomicOr(err, 1u);
  }
  if (r >= n_ref) return;
  const int64_t i = offsets[0] + r;
  int32_t mx = -1;
  for (int j = 0; j < k; ++j) mx = max(mx, nbrs[r * k + j]);
  if ((int64_t)mx >= i) atomicOr(err, 2u);
  // batch of i: last b with offsets[b] <= i (searchsorted side='right' - 1)
  int64_t lo = 0, hi = n_off;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= i) lo = mid + 1;
    else hi = mid;
  }
  const int64_t b = lo - 1;
  if (b >= 0 && (int64_t)mx >= offsets[b]) atomicOr(err, 4u);
}

}  // namespace

size_t compute_depths_workspace() { return 512; }

int compute_depths(const int32_t* nbrs, int64_t n_ref, int k, int64_t n0, int32_t* depths, void* ws,
                   size_t ws_bytes, hipStream_t s) {
  GGP_CHECK_ARG(n0 >= 0 && n_ref >= 0 && k >= 1, "compute_depths: bad sizes");
  Workspace w(ws, ws_bytes);
  uint32_t* ctr = w.take<uint32_t>(2);
  if (!w.ok()) {
    set_error("compute_depths: workspace too small");
    return kWorkspaceTooSmall;
  }
  const int64_t n = n0 + n_ref;
  GGP_TRY(fill_u32(ctr, 0u, 2, s));
  hipLaunchKernelGGL(depth_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, depths, n, n0);
  GGP_LAUNCH_CHECK();
  if (n_ref == 0) return kOk;
  int dev = 0, cus = 256;
  GGP_HIP(hipGetDevice(&dev));
  GGP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t ntiles = (n_ref + kDepThreads - 1) / kDepThreads;
  const unsigned grid = (unsigned)std::min<int64_t>(ntiles, (int64_t)cus * 8);
  hipLaunchKernelGGL(depths_kernel, dim3(grid), dim3(kDepThreads), 0, s, nbrs, n_ref, k, n0, depths, ctr, ctr + 1);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int compute_depths_status(void* ws, uint32_t* host_err, hipStream_t s) {
  uint32_t* ctr = static_cast<uint32_t*>(ws);
  GGP_HIP(hipMemcpyAsync(host_err, ctr + 1, 4, hipMemcpyDeviceToHost, s));
  GGP_HIP(hipStreamSynchronize(s));
  return kOk;
}

size_t order_by_depth_workspace(int64_t n) {
  Workspace w(nullptr, 0);
  w.take<uint32_t>(n);
  w.take<uint32_t>(n);
  w.take<uint32_t>(n);
  w.take<uint32_t>(n);
  w.take<uint32_t>(n);
  w.take<char>(radix_workspace(n));
  return w.used + 256;
}

int order_by_depth(const void* pts, int64_t n, int d, int dtype, const int32_t* idx, const int32_t* nbrs,
                   int64_t n_ref, int k, const int32_t* depths, int32_t max_depth, void* pts_o, int32_t* idx_o,
                   int32_t* nbrs_o, int32_t* dep_o, void* ws, size_t ws_bytes, hipStream_t s) {
  GGP_CHECK_ARG(n >= 1 && d >= 1 && k >= 1 && n_ref >= 0 && n_ref <= n, "order_by_depth: bad sizes");
  GGP_CHECK_ARG(n < (int64_t(1) << 32), "order_by_depth: N must be below 2^32");
  const int64_t n0 = n - n_ref;
  Workspace w(ws, ws_bytes);
  uint32_t* k0 = w.take<uint32_t>(n);
  uint32_t* k1 = w.take<uint32_t>(n);
  uint32_t* p0 = w.take<uint32_t>(n);
  uint32_t* p1 = w.take<uint32_t>(n);
  uint32_t* inv = w.take<uint32_t>(n);
  const size_t rws = radix_workspace(n);
  void* rw = w.take<char>(rws);
  if (!w.ok()) {
    set_error("order_by_depth: workspace too small");
    return kWorkspaceTooSmall;
  }
  int bits = 0;
  while (bits < 31 && (int64_t(1) << bits) <= (int64_t)max_depth) ++bits;
  hipLaunchKernelGGL(depth_keys_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, depths, n, k0, p0);
  GGP_LAUNCH_CHECK();
  uint32_t* keys = k0;
  uint32_t* pay = p0;
  GGP_TRY(radix_sort_pairs<uint32_t>(&keys, &pay, k1, p1, n, 0, bits, rw, rws, s));
  if (dtype == 0)
    hipLaunchKernelGGL(order_gather_kernel<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const float*)pts, d,
                       idx, depths, pay, n, (float*)pts_o, idx_o, dep_o, inv);
  else
    hipLaunchKernelGGL(order_gather_kernel<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const double*)pts, d,
                       idx, depths, pay, n, (double*)pts_o, idx_o, dep_o, inv);
  GGP_LAUNCH_CHECK();
  if (n_ref > 0) {
    hipLaunchKernelGGL(order_neighbors_kernel, dim3(grid_for(n_ref * k, 256)), dim3(256), 0, s, nbrs, k, pay, inv,
                       n0, n_ref, nbrs_o);
    GGP_LAUNCH_CHECK();
  }
  return kOk;
}

int depth_offsets(const int32_t* depths_sorted, int64_t n, int64_t n_off, int64_t* offsets, hipStream_t s) {
  if (n_off <= 0) return kOk;
  hipLaunchKernelGGL(offsets_kernel, dim3(grid_for(n_off, 256)), dim3(256), 0, s, depths_sorted, n, n_off, offsets);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int check_graph(const int32_t* nbrs, int64_t n_ref, int k, int64_t n, const int64_t* offsets, int64_t n_off,
                uint32_t* err_dev, hipStream_t s) {
  GGP_CHECK_ARG(n_off >= 1, "check_graph: offsets must not be empty");
  GGP_TRY(fill_u32(err_dev, 0u, 1, s));
  hipLaunchKernelGGL(check_graph_kernel, dim3(grid_for(std::max<int64_t>(n_ref, 1), 256)), dim3(256), 0, s, nbrs,
                     n_ref, k, n, offsets, n_off, err_dev);
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace ggp
