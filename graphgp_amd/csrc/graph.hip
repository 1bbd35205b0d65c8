// graph.hip -- depth, depth ordering and batch offsets of the Vecchia graph (gfx950).
//
// References: graphgp/graph.py:101-120 (compute_depths), :123-137 (order_by_depth),
// :95-96 (offsets), :44-70 (check_graph).
//
// compute_depths: depth[i] = 1 + max depth[neighbours(i)] (depth 0 for the n0 seed
// points).  The reference iterates that map to its fixed point with Jacobi sweeps (one
// sweep over all N points per depth level).  Here the longest-path depths come from
// Kahn's algorithm by rounds: the transposed graph (each point's children) is built once
// (count, scan, fill); round d then takes the frontier of points whose last unfinished
// neighbour finished in round d-1, so every edge is touched a constant number of times
// and a round costs one launch over its frontier.  A point leaves the frontier in the
// round equal to its longest path from the seeds, which is the reference's fixed point
// (unique for a graph in topological order).  When the caller's workspace cannot hold the
// transpose, Gauss-Seidel sweeps in index order are used instead (in place,
// agent-coherent loads and stores, no waiting) until a sweep changes nothing.
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

// ---- Kahn rounds.  ctr: [0] malformed flag, [1] final status, [2] head, [3] end, [4] tail,
// [5] depth of the current frontier, [6] finished workgroups of the current round.
constexpr int kCtrWords = 8;

// Validate row r, count its non-seed neighbours, count children, seed the first frontier.
__global__ void kahn_init_kernel(const int32_t* __restrict__ nbrs, int64_t n_ref, int k, int64_t n0,
                                 uint32_t* __restrict__ remaining, uint32_t* __restrict__ cnt,
                                 uint32_t* __restrict__ queue, int32_t* __restrict__ depths, uint32_t* ctr) {
  bool bad = false;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_ref; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = n0 + r;
    uint32_t rem = 0;
    bool row_bad = false;
    for (int j = 0; j < k; ++j) {
      const int32_t nb = nbrs[r * k + j];
      if (nb < 0 || (int64_t)nb >= i) {
        row_bad = true;
        break;
      }
      if (nb >= n0) ++rem;
    }
    bad = bad || row_bad;
    if (!row_bad) {
      for (int j = 0; j < k; ++j) {
        const int32_t nb = nbrs[r * k + j];
        if (nb >= n0) atomicAdd(&cnt[nb - n0], 1u);
      }
      remaining[r] = rem;
      if (rem == 0) {
        depths[i] = 1;
        queue[atomicAdd(&ctr[4], 1u)] = (uint32_t)r;
      }
    }
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&ctr[0], 1u);
}

// children of refined point j at child[rowptr[j] .. rowptr[j+1])
__global__ void kahn_fill_kernel(const int32_t* __restrict__ nbrs, int64_t n_ref, int k, int64_t n0,
                                 const int64_t* __restrict__ rowptr, uint32_t* __restrict__ cur,
                                 uint32_t* __restrict__ child) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_ref * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t nb = nbrs[e];
    if (nb < n0) continue;
    const int64_t j = nb - n0;
    child[rowptr[j] + atomicAdd(&cur[j], 1u)] = (uint32_t)(e / k);
  }
}

__global__ void kahn_start_kernel(uint32_t* ctr) {
  ctr[2] = 0;
  ctr[3] = ctr[4];
  ctr[5] = 1;
  ctr[6] = 0;
}

// One round: every child of a frontier point loses one unfinished neighbour; a child that
// reaches zero gets depth d + 1 and joins the next frontier.  Groups of G lanes take one
// frontier point each (64 / G points per wavefront and pass) and split its children, so a
// round costs a few atomic round trips rather than one per child (G >= k fits the average
// point in one step).  The next frontier is gathered without a hot counter: a wavefront
// keeps the points it frees in LDS and reserves queue space for them in batches of more
// than 64, and what is left at the end goes out with one reservation per workgroup.  The
// last workgroup to finish moves [head, end) to the points queued in this round.
template <int G>
__global__ __launch_bounds__(256) void kahn_round_kernel(const int64_t* __restrict__ rowptr,
                                                         const uint32_t* __restrict__ child, int64_t n0,
                                                         uint32_t* __restrict__ remaining, uint32_t* queue,
                                                         int32_t* __restrict__ depths, uint32_t* ctr) {
  constexpr int PPW = 64 / G;
  __shared__ uint32_t buf[4][128];
  __shared__ uint32_t wcnt[4], wbase[4];
  // The frontier window and depth were set by the previous launch and change only after
  // every workgroup of this one has read them (the last workgroup moves them), so plain
  // loads see them.
  const uint32_t head = ctr[2];
  const uint32_t end = ctr[3];
  const int32_t d = (int32_t)ctr[5];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int grp = lane / G, sub = lane % G;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + w);
  const uint32_t nwaves = gridDim.x * 4;
  uint32_t cnt = 0;  // points in buf[w], the same in every lane
  for (uint32_t qb = head + wave * PPW; qb < end; qb += nwaves * PPW) {
    const uint32_t q = qb + grp;
    int64_t p = 0, p1 = 0;
    if (q < end) {
      const uint32_t j = queue[q];
      p = rowptr[j] + sub;
      p1 = rowptr[j + 1];
    }
    while (__ballot(p < p1) != 0) {
      uint32_t i = 0;
      bool freed = false;
      if (p < p1) {
        i = child[p];
        freed = atomicSub(&remaining[i], 1u) == 1u;
      }
      const uint64_t m = __ballot(freed);
      if (m != 0) {
        if (freed) {
          depths[n0 + i] = d + 1;
          buf[w][cnt + (uint32_t)__popcll(m & below)] = i;
        }
        cnt += (uint32_t)__popcll(m);
        if (cnt > 64) {  // cnt <= 128: at most 64 before this step, plus at most 64
          __builtin_amdgcn_wave_barrier();
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(&ctr[4], cnt);
          base = __shfl(base, 0, 64);
          for (uint32_t e = lane; e < cnt; e += 64) queue[base + e] = buf[w][e];
          __builtin_amdgcn_wave_barrier();
          cnt = 0;
        }
      }
      p += G;
    }
  }
  if (lane == 0) wcnt[w] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    const uint32_t base = total ? atomicAdd(&ctr[4], total) : 0u;
    wbase[0] = base;
    wbase[1] = base + wcnt[0];
    wbase[2] = base + wcnt[0] + wcnt[1];
    wbase[3] = base + wcnt[0] + wcnt[1] + wcnt[2];
  }
  __syncthreads();
  for (uint32_t e = lane; e < cnt; e += 64) queue[wbase[w] + e] = buf[w][e];
  // The tail is only touched by returning atomics, all of which have returned before the
  // barrier, and the depths and queue entries are read by the next launch only; so the
  // finish count needs no fence (a device-scope fence per workgroup would write back the
  // L2 each time).
  __syncthreads();
  if (threadIdx.x == 0) {
    if (atomicAdd(&ctr[6], 1u) == gridDim.x - 1) {
      const uint32_t tail = __hip_atomic_load(&ctr[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctr[2], end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctr[3], tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctr[5], (uint32_t)(d + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ctr[6], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

int launch_round(const void* kernel, unsigned grid, hipStream_t s, const int64_t* rowptr, const uint32_t* child,
                 int64_t n0, uint32_t* remaining, uint32_t* queue, int32_t* depths, uint32_t* ctr) {
  void* args[] = {(void*)&rowptr, (void*)&child, (void*)&n0, (void*)&remaining,
                  (void*)&queue,  (void*)&depths, (void*)&ctr};
  GGP_HIP(hipLaunchKernel(kernel, dim3(grid), dim3(256), args, 0, s));
  return kOk;
}

// ---- exclusive scan of n uint32 counts into n + 1 int64 offsets (three phases)
constexpr int kScanTile = 2048;  // 256 threads x 8

__global__ __launch_bounds__(256) void scan_tile_sum_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                            int64_t* __restrict__ sums) {
  __shared__ int64_t red[4];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int64_t v = 0;
  for (int t = threadIdx.x; t < kScanTile; t += 256)
    if (base + t < n) v += in[base + t];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// in-place exclusive scan of the tile sums by one workgroup
__global__ __launch_bounds__(1024) void scan_sums_kernel(int64_t* sums, int64_t m) {
  __shared__ int64_t buf[1024];
  int64_t carry = 0;
  for (int64_t base = 0; base < m; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < m ? sums[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int64_t t = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < m) sums[i] = carry + buf[threadIdx.x] - v;
    const int64_t tot = buf[1023];
    __syncthreads();
    carry += tot;
  }
}

__global__ __launch_bounds__(256) void scan_tile_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                        const int64_t* __restrict__ sums, int64_t* __restrict__ out) {
  __shared__ int64_t buf[256];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * 8;
  uint32_t v[8];
  int64_t loc = 0;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    v[t] = base + t < n ? in[base + t] : 0u;
    loc += v[t];
  }
  buf[threadIdx.x] = loc;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int64_t t = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
    __syncthreads();
    buf[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = sums[blockIdx.x] + buf[threadIdx.x] - loc;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    if (base + t < n) out[base + t] = run;
    run += v[t];
  }
  if (base < n && base + 8 >= n) out[n] = run;
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
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_ref * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / k;
    const int j = (int)(e - r * k);
    const int64_t src = (int64_t)order[n0 + r] - n0;
    nbrs_o[e] = (int32_t)inv[nbrs[src * k + j]];
  }
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

namespace {

// ---- depths as a dataflow.  In a graph whose neighbours all precede their point (the
// order the preceding-neighbour query produces), depth_i = 1 + max over i's neighbours'
// depths, seeds 0 -- the reference's fixed point (graph.py:101-120), computed in one pass
// in index order instead of by rounds.  Persistent workgroups take tiles of 64 points from
// a ticket in index order; four lanes per point gather its neighbours' depths, waiting
// while one is still the sentinel -1 (a cached read first, the coherent re-read only while
// it is the sentinel), then publish 1 + their maximum.  Every dependency is on a lower
// index, so on a lower ticket held by a running workgroup: the lowest unfinished point can
// always finish.  A row with a neighbour that does not precede its point is malformed: it
// waits for nothing, publishes 0 and sets the status words (ctr[0], ctr[1]).
__global__ void depth_flow_init_kernel(int32_t* depths, int64_t n, int64_t n0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) depths[i] = i < n0 ? 0 : -1;
}

template <int RM>
__global__ __launch_bounds__(256) void depth_flow_kernel(const int32_t* __restrict__ nbrs, int64_t n_ref, int k,
                                                         int64_t n0, int32_t* depths, uint32_t* ctr, int64_t ntiles,
                                                         const uint32_t* yield, uint32_t keep) {
  constexpr int P = 4, GPT = 256 / P;
  __shared__ uint32_t s_tile;
  const int lane = threadIdx.x & (P - 1);
  const int gbase = (threadIdx.x & 63) & ~(P - 1);
  for (;;) {
    if (threadIdx.x == 0) s_tile = yield_now(yield, keep) ? 0xffffffffu : atomicAdd(&ctr[5], 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile >= ntiles) break;
    const int64_t r = tile * GPT + threadIdx.x / P;
    const bool valid = r < n_ref;
    const int64_t i = n0 + r;
    int32_t nb[RM];
    uint32_t pend = 0;
    bool row_bad = false;
#pragma unroll
    for (int q = 0; q < RM; ++q) {
      const int j = P * q + lane;
      nb[q] = 0;
      if (valid && j < k) {
        nb[q] = nbrs[r * k + j];
        if (nb[q] < 0 || (int64_t)nb[q] >= i) row_bad = true;
        else if (nb[q] >= n0) pend |= 1u << q;
      }
    }
    const bool bad = ((__ballot(row_bad) >> gbase) & 0xFull) != 0;  // the point's four lanes
    if (bad) pend = 0;
    int32_t mx = 0;
    bool done = !valid;
    for (uint32_t it = 0;; ++it) {
      if (!done) {
#pragma unroll
        for (int q = 0; q < RM; ++q) {
          if (pend & (1u << q)) {
            int32_t v = depths[nb[q]];
            if (v < 0) v = __hip_atomic_load(&depths[nb[q]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (v >= 0) {
              mx = v > mx ? v : mx;
              pend &= ~(1u << q);
            }
          }
        }
        // a point's lanes finish together, so its lanes are active for the shuffles
        if (((__ballot(pend == 0) >> gbase) & 0xFull) == 0xFull) {
          int32_t m = mx;
          const int32_t m1 = __shfl_xor(m, 1, P);
          m = m1 > m ? m1 : m;
          const int32_t m2 = __shfl_xor(m, 2, P);
          m = m2 > m ? m2 : m;
          if (lane == 0) {
            __hip_atomic_store(&depths[i], bad ? 0 : m + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (bad) {
              atomicOr(&ctr[0], 1u);
              atomicOr(&ctr[1], 1u);
            }
          }
          done = true;
        }
      }
      if (__all(done)) break;
      if (it > (1u << 22)) {  // not reached for a graph whose rows all passed the check above
        if (!done && lane == 0) {
          atomicOr(&ctr[1], 1u);
          __hip_atomic_store(&depths[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
}

// The dataflow launch (k <= 64); the workspace needs only the counters.
int depths_by_flow(const int32_t* nbrs, int64_t n_ref, int k, int64_t n0, int32_t* depths, uint32_t* ctr,
                   hipStream_t s) {
  const void* kern = k <= 8    ? (const void*)depth_flow_kernel<2>
                     : k <= 16 ? (const void*)depth_flow_kernel<4>
                     : k <= 32 ? (const void*)depth_flow_kernel<8>
                               : (const void*)depth_flow_kernel<16>;
  const int64_t n = n0 + n_ref;
  hipLaunchKernelGGL(depth_flow_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, depths, n, n0);
  GGP_LAUNCH_CHECK();
  const int64_t ntiles = (n_ref + 63) / 64;
  const int resident = device_cus() * std::max(1, blocks_per_cu(kern, 0));
  const int grid = (int)std::min<int64_t>(ntiles, persistent_grid(kern, resident, 0));
  const YieldCfg& y = yield_cfg();
  const uint32_t keep = yield_keep(grid);
  void* args[] = {(void*)&nbrs, (void*)&n_ref, (void*)&k, (void*)&n0, (void*)&depths, (void*)&ctr, (void*)&ntiles,
                  (void*)&y.dev, (void*)&keep};
  GGP_HIP(hipLaunchKernel(kern, dim3(grid), dim3(256), args, 0, s));
  return kOk;
}

struct KahnWs {
  uint32_t* ctr;
  uint32_t *remaining, *cnt, *queue, *child;
  int64_t *rowptr, *sums;
};

bool carve_kahn(Workspace& w, int64_t n_ref, int k, KahnWs& b) {
  b.ctr = w.take<uint32_t>(kCtrWords);
  b.remaining = w.take<uint32_t>(n_ref + 1);
  b.cnt = w.take<uint32_t>(n_ref + 1);
  b.queue = w.take<uint32_t>(n_ref + 1);
  b.rowptr = w.take<int64_t>(n_ref + 1);
  b.sums = w.take<int64_t>((n_ref + kScanTile - 1) / kScanTile + 1);
  b.child = w.take<uint32_t>(n_ref * k + 1);
  return w.ok();
}

// Gauss-Seidel sweeps until nothing changes (one host synchronisation per sweep).
int depths_by_sweeps(const int32_t* nbrs, int64_t n_ref, int k, int64_t n0, int32_t* depths, uint32_t* ctr,
                     hipStream_t s) {
  uint32_t status = 0;
  for (int64_t sweep = 0; sweep <= n_ref + 1; ++sweep) {
    GGP_TRY(fill_u32(ctr, 0u, 1, s));
    hipLaunchKernelGGL(depths_sweep_kernel, dim3(grid_for(n_ref, kDepThreads)), dim3(kDepThreads), 0, s, nbrs, n_ref,
                       k, n0, depths, ctr);
    GGP_LAUNCH_CHECK();
    GGP_TRY(read_to_host(&status, ctr, 4, s));
    if (status & 1u) break;     // malformed graph: report through the status word
    if (!(status & 2u)) break;  // fixed point
  }
  uint32_t final_status = (status & 1u) ? 1u : 0u;
  return fill_u32(ctr + 1, final_status, 1, s);
}

}  // namespace

size_t compute_depths_workspace(int64_t n_ref, int k) {
  Workspace w(nullptr, 0);
  KahnWs b;
  carve_kahn(w, n_ref, k, b);
  return w.used + 256;
}

// Depths by one of three methods: the dataflow (one launch, no host synchronisation; k <=
// 64, larger k falls back to one of the others), Kahn rounds (needs the transpose in the
// workspace, compute_depths_workspace; synchronises once after validation and once per 32
// rounds) or sweeps (the workspace needs only its first 32 bytes; one synchronisation per
// sweep).  All give the same depths; a malformed graph sets the status word.
int compute_depths(const int32_t* nbrs, int64_t n_ref, int k, int64_t n0, int32_t* depths, int method, void* ws,
                   size_t ws_bytes, hipStream_t s) {
  GGP_CHECK_ARG(method == kDepthsFlow || method == kDepthsRounds || method == kDepthsSweeps,
                "compute_depths: unknown method");
  GGP_CHECK_ARG(n0 >= 0 && n_ref >= 0 && k >= 1, "compute_depths: bad sizes");
  GGP_CHECK_ARG(n_ref < (int64_t(1) << 31), "compute_depths: at most 2^31 refined points");
  Workspace w(ws, ws_bytes);
  KahnWs b;
  const bool kahn = carve_kahn(w, n_ref, k, b);
  if (b.ctr == nullptr) {
    set_error("compute_depths: workspace too small");
    return kWorkspaceTooSmall;
  }
  const int64_t n = n0 + n_ref;
  GGP_TRY(fill_u32(b.ctr, 0u, kCtrWords, s));
  if (method == kDepthsFlow && k > 64) method = kahn ? kDepthsRounds : kDepthsSweeps;
  if (method == kDepthsRounds && !kahn) {
    set_error("compute_depths: workspace too small for Kahn rounds");
    return kWorkspaceTooSmall;
  }
  if (method == kDepthsFlow) return n_ref == 0 ? kOk : depths_by_flow(nbrs, n_ref, k, n0, depths, b.ctr, s);
  hipLaunchKernelGGL(depth_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, depths, n);
  GGP_LAUNCH_CHECK();
  if (n_ref == 0) return kOk;
  if (method == kDepthsSweeps) return depths_by_sweeps(nbrs, n_ref, k, n0, depths, b.ctr, s);

  GGP_TRY(fill_u32(b.cnt, 0u, n_ref, s));
  hipLaunchKernelGGL(kahn_init_kernel, dim3(budget_grid((const void*)kahn_init_kernel, grid_capped(n_ref, 256))),
                     dim3(256), 0, s, nbrs, n_ref, k, n0, b.remaining,
                     b.cnt, b.queue, depths, b.ctr);
  GGP_LAUNCH_CHECK();
  uint32_t bad = 0;
  GGP_TRY(read_to_host(&bad, b.ctr, 4, s));
  if (bad) {
    return fill_u32(b.ctr + 1, 1u, 1, s);
  }
  const int64_t tiles = (n_ref + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(scan_tile_sum_kernel, dim3((unsigned)tiles), dim3(256), 0, s, b.cnt, n_ref, b.sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, s, b.sums, tiles);
  hipLaunchKernelGGL(scan_tile_kernel, dim3((unsigned)tiles), dim3(256), 0, s, b.cnt, n_ref, b.sums, b.rowptr);
  GGP_LAUNCH_CHECK();
  GGP_TRY(fill_u32(b.cnt, 0u, n_ref, s));
  hipLaunchKernelGGL(kahn_fill_kernel, dim3(budget_grid((const void*)kahn_fill_kernel, grid_capped(n_ref * k, 256))),
                     dim3(256), 0, s, nbrs, n_ref, k, n0,
                     b.rowptr, b.cnt, b.child);
  hipLaunchKernelGGL(kahn_start_kernel, dim3(1), dim3(1), 0, s, b.ctr);
  GGP_LAUNCH_CHECK();
  // G lanes per frontier point with G >= k where possible (a point has k children on
  // average).  A frontier holds about n_ref / levels points and a few hundred levels is
  // typical; up to 2048 workgroups (the device's resident wavefronts) share it.
  const unsigned grid = (unsigned)std::min<int64_t>(std::max<int64_t>(n_ref / 1024, 256), 2048);
  const void* round_kernel = k <= 8    ? (const void*)kahn_round_kernel<8>
                             : k <= 16 ? (const void*)kahn_round_kernel<16>
                             : k <= 32 ? (const void*)kahn_round_kernel<32>
                                       : (const void*)kahn_round_kernel<64>;
  uint32_t hdr[3] = {0, 0, 0};
  for (int64_t round = 0; round <= n_ref + 1; round += 32) {
    for (int q = 0; q < 32; ++q)
      GGP_TRY(launch_round(round_kernel, grid, s, b.rowptr, b.child, n0, b.remaining, b.queue, depths, b.ctr));
    GGP_LAUNCH_CHECK();
    GGP_TRY(read_to_host(hdr, b.ctr + 2, 12, s));
    if (hdr[0] == hdr[1]) break;  // empty frontier
  }
  // every refined point must have left a frontier (always true for a graph in topological order)
  const uint32_t final_status = hdr[2] == (uint32_t)n_ref ? 0u : 1u;
  return fill_u32(b.ctr + 1, final_status, 1, s);
}

int compute_depths_status(void* ws, uint32_t* host_err, hipStream_t s) {
  uint32_t* ctr = static_cast<uint32_t*>(ws);
  GGP_TRY(read_to_host(host_err, ctr + 1, 4, s));
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
    hipLaunchKernelGGL(order_neighbors_kernel,
                       dim3(budget_grid((const void*)order_neighbors_kernel, grid_capped(n_ref * k, 256))), dim3(256),
                       0, s, nbrs, k, pay, inv,
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
