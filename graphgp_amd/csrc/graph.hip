// graph.hip -- depth, depth ordering and batch offsets of the Vecchia graph (gfx950).
//
// References: graphgp/graph.py:101-120 (compute_depths), :123-137 (order_by_depth),
// :95-96 (offsets), :44-70 (check_graph).
//
// compute_depths: depth[i] = 1 + max depth[neighbours(i)] (depth 0 for the n0 seed
// points).  The reference iterates that map to its fixed point with Jacobi sweeps (one
// sweep per depth level).  Here each sweep is Gauss-Seidel in index order (in place,
// agent-coherent loads and stores, no waiting), so a dependency is usually resolved in
// the sweep that reaches it and a handful of sweeps suffice; the host stops at the first
// sweep that changes nothing, which is the reference's convergence test.  The fixed
// point of a graph in topological order is unique, so the result is the reference's.
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

// One Gauss-Seidel sweep of depth[i] = 1 + max depth[nbrs(i)] in index order.  Values
// written earlier in the same sweep may or may not be seen (no waiting, no ordering
// assumption); a sweep that changes nothing has reached the reference's fixed point.
// status bit 0: a neighbour that does not precede its point; bit 1: something changed.
__global__ __launch_bounds__(kDepThreads) void depths_sweep_kernel(const int32_t* __restrict__ nbrs, int64_t n_ref,
                                                                   int k, int64_t n0, int32_t* depths,
                                                                   uint32_t* status) {
  const int64_t r = (int64_t)blockIdx.x * kDepThreads + threadIdx.x;
  bool changed = false, bad = false;
  if (r < n_ref) {
    const int64_t i = n0 + r;
    int32_t m = -1;
    for (int j = 0; j < k; ++j) {
      const int32_t nb = nbrs[r * k + j];
      if (nb < 0 || (int64_t)nb >= i) {
        bad = true;
        break;
      }
      const int32_t v = __hip_atomic_load(&depths[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      m = v > m ? v : m;
    }
    if (!bad && m + 1 != depths[i]) {
      __hip_atomic_store(&depths[i], m + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      changed = true;
    }
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(status, 1u);
  if (__any(changed) && (threadIdx.x & 63) == 0) atomicOr(status, 2u);
}

__global__ void depth_init_kernel(int32_t* depths, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) depths[i] = 0;
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

// Sweeps until nothing changes (synchronises once per sweep).  The number of sweeps is
// at most max depth + 1 and in practice a handful: most neighbours lie far earlier in
// the order and are final before the sweep reaches their dependents.
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
  hipLaunchKernelGGL(depth_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, depths, n);
  GGP_LAUNCH_CHECK();
  if (n_ref == 0) return kOk;
  uint32_t status = 0;
  for (int64_t sweep = 0; sweep <= n_ref + 1; ++sweep) {
    GGP_TRY(fill_u32(ctr, 0u, 1, s));
    hipLaunchKernelGGL(depths_sweep_kernel, dim3(grid_for(n_ref, kDepThreads)), dim3(kDepThreads), 0, s, nbrs, n_ref,
                       k, n0, depths, ctr);
    GGP_LAUNCH_CHECK();
    GGP_HIP(hipMemcpyAsync(&status, ctr, 4, hipMemcpyDeviceToHost, s));
    GGP_HIP(hipStreamSynchronize(s));
    if (status & 1u) break;      // malformed graph: report through the status word
    if (!(status & 2u)) break;   // fixed point
  }
  uint32_t final_status = (status & 1u) ? 1u : 0u;
  GGP_HIP(hipMemcpyAsync(ctr + 1, &final_status, 4, hipMemcpyHostToDevice, s));
  GGP_HIP(hipStreamSynchronize(s));
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
