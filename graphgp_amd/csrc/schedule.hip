// schedule.hip -- a locality order for the dataflow over stored weights (gfx950).
//
// The refined points of one batch (depth level) do not depend on each other, so the
// dataflow may take a batch's points in any order; only the batch order is fixed
// (graph.py:95-96 batches; refine.py:123-145 processes them in turn).  Graph order lists a
// batch's points in k-d tree order, which interleaves subtrees: consecutive points are far
// apart in space, and a 64-point tile's neighbour values lie on about as many cache lines
// as the tile has neighbours.  The schedule orders every batch by the Morton code of its
// points (over their bounding box), and the dataflow stores its values in schedule order.
// A tile then holds points that are close in space, and their neighbours in an earlier
// batch are close in that batch's storage too, so a cache line serves several gathers.
//
//   sched[q]   graph index (0-based over the refined points) of schedule slot q
//   nbv[q, j]  the neighbours of sched[q] as value slots (seeds keep their graph position,
//              refined neighbour r moves to n0 + (slot of r)): position 0 the nearest
//              neighbour, positions 1 .. k-1 in increasing slot order
//   pts_sched  the points in slot order (optional): with nbv as neighbours, the refinement
//              kernel then works in slot space, its tiles' coordinate gathers local too
#include "common.hpp"
#include "radix.hpp"
#include "graphgp_amd_internal.hpp"

namespace ggp {

namespace {

constexpr int kMortonBits = 48;

template <typename T>
__global__ void bbox_kernel(const T* __restrict__ pts, int64_t n, int d, typename KeyOf<T>::type* __restrict__ lo,
                            typename KeyOf<T>::type* __restrict__ hi) {
  using K = typename KeyOf<T>::type;
  K mn[8], mx[8];
  for (int j = 0; j < d; ++j) { mn[j] = ~K(0); mx[j] = K(0); }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    for (int j = 0; j < d; ++j) {
      const K kk = KeyOf<T>::key(pts[i * d + j]);
      mn[j] = kk < mn[j] ? kk : mn[j];
      mx[j] = kk > mx[j] ? kk : mx[j];
    }
  for (int j = 0; j < d; ++j) {
    for (int o = 32; o > 0; o >>= 1) {
      const K a = (K)__shfl_xor(mn[j], o, 64), b = (K)__shfl_xor(mx[j], o, 64);
      mn[j] = a < mn[j] ? a : mn[j];
      mx[j] = b > mx[j] ? b : mx[j];
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&lo[j], mn[j]);
      atomicMax(&hi[j], mx[j]);
    }
  }
}

// key = (batch << 48) | Morton code of the point (48 / d bits per coordinate, the
// coordinates quantised over the bounding box); payload = refined index.
template <typename T>
__global__ void sched_keys_kernel(const T* __restrict__ pts, int d, int64_t n0, int64_t n_ref,
                                  const int64_t* __restrict__ offsets, int64_t n_off,
                                  const typename KeyOf<T>::type* __restrict__ lo,
                                  const typename KeyOf<T>::type* __restrict__ hi, uint64_t* __restrict__ keys,
                                  uint32_t* __restrict__ pay) {
  const int bits = kMortonBits / d;
  const double cells = (double)(uint64_t(1) << bits);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n_ref; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = n0 + p;
    // batch b holds graph positions [offsets[b], offsets[b + 1]); offsets[0] = n0
    int64_t lo_b = 0, hi_b = n_off - 1;  // largest b with offsets[b] <= g
    while (lo_b < hi_b) {
      const int64_t mid = (lo_b + hi_b + 1) / 2;
      if (offsets[mid] <= g) lo_b = mid;
      else hi_b = mid - 1;
    }
    uint64_t code = 0;
    for (int j = 0; j < d; ++j) {
      const double a = (double)KeyOf<T>::value(lo[j]), b = (double)KeyOf<T>::value(hi[j]);
      const double x = (double)pts[g * d + j];
      double c = b > a ? (x - a) / (b - a) * cells : 0.0;
      c = c < 0.0 ? 0.0 : (c > cells - 1.0 ? cells - 1.0 : c);
      const uint64_t qv = (uint64_t)c;
      for (int bit = 0; bit < bits; ++bit) code |= ((qv >> bit) & 1u) << (bit * d + j);
    }
    keys[p] = ((uint64_t)lo_b << kMortonBits) | code;
    pay[p] = (uint32_t)p;
  }
}

__global__ void sched_out_kernel(const uint32_t* __restrict__ sorted, int64_t n_ref, int32_t* __restrict__ sched,
                                 uint32_t* __restrict__ slot) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_ref; q += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = sorted[q];
    sched[q] = (int32_t)p;
    slot[p] = (uint32_t)q;
  }
}

__global__ void sched_nbrs_kernel(const int32_t* __restrict__ nbrs, int k, int64_t n0, int64_t n_ref,
                                  const int32_t* __restrict__ sched, const uint32_t* __restrict__ slot,
                                  int32_t* __restrict__ nbv) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_ref * k; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = e / k;
    const int j = (int)(e - q * k);
    const int32_t nb = nbrs[(int64_t)sched[q] * k + j];
    nbv[e] = nb < n0 ? nb : (int32_t)(n0 + slot[nb - n0]);
  }
}

// Every row of nbv with its positions 1 .. k-1 in increasing slot order (position 0, the
// nearest neighbour, stays first: the increment formulation is relative to it).  The
// refinement kernels read row q's neighbours across the lanes of one point, and the rows of
// a tile's points (close in space, in one batch) share many neighbours: in slot order the
// same lane of neighbouring points reads the same or a nearby slot, so a gather instruction
// touches fewer cache lines.  Rows of up to 64 entries, a bitonic network in registers.
template <int W>
__global__ void sort_rows_kernel(int32_t* __restrict__ nbv, int k, int64_t n_ref) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_ref; q += (int64_t)gridDim.x * blockDim.x) {
    int32_t* row = nbv + q * k;
    int32_t v[W];
#pragma unroll
    for (int i = 0; i < W; ++i) v[i] = (i + 1 < k) ? row[i + 1] : 0x7fffffff;
#pragma unroll
    for (int size = 2; size <= W; size <<= 1)
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
        for (int i = 0; i < W; ++i) {
          const int j = i ^ stride;
          if (j > i) {
            const bool up = (i & size) == 0;
            const int32_t a = v[i], b = v[j];
            if ((a > b) == up) {
              v[i] = b;
              v[j] = a;
            }
          }
        }
#pragma unroll
    for (int i = 0; i < W; ++i)
      if (i + 1 < k) row[i + 1] = v[i];
  }
}

// points in value-slot order: seeds keep their place, refined slot q holds point sched[q]
template <typename T>
__global__ void sched_points_kernel(const T* __restrict__ pts, int d, int64_t n0, int64_t n,
                                    const int32_t* __restrict__ sched, T* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * d; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e / d;
    const int j = (int)(e - s * d);
    const int64_t g = s < n0 ? s : n0 + sched[s - n0];
    out[e] = pts[g * d + j];
  }
}

struct SchedWs {
  uint64_t *key0, *key1;
  uint32_t *pay0, *pay1, *slot;
  void* bbox;
  void* sort_ws;
  size_t sort_ws_bytes;
};

bool carve_sched(Workspace& w, int64_t n_ref, SchedWs& b) {
  b.key0 = w.take<uint64_t>(n_ref);
  b.key1 = w.take<uint64_t>(n_ref);
  b.pay0 = w.take<uint32_t>(n_ref);
  b.pay1 = w.take<uint32_t>(n_ref);
  b.slot = w.take<uint32_t>(n_ref);
  b.bbox = w.take<uint64_t>(16);
  b.sort_ws_bytes = radix_workspace(n_ref);
  b.sort_ws = w.take<char>(b.sort_ws_bytes);
  return w.ok();
}

template <typename T>
int schedule_impl(const T* pts, int64_t n, int d, const int32_t* nbrs, int64_t n0, int k, const int64_t* offsets,
                  int64_t n_off, int32_t* sched, int32_t* nbv, T* pts_sched, void* ws, size_t ws_bytes,
                  hipStream_t s) {
  using K = typename KeyOf<T>::type;
  const int64_t n_ref = n - n0;
  GGP_CHECK_ARG(d >= 1 && d <= 8 && k >= 1 && n0 >= 1 && n_ref >= 0, "graph_schedule: bad sizes");
  GGP_CHECK_ARG(n < (int64_t(1) << 31), "graph_schedule: N must be below 2^31");
  GGP_CHECK_ARG(n_off >= 2 && n_off <= (int64_t(1) << (64 - kMortonBits)), "graph_schedule: 1 .. 65535 batches");
  if (n_ref == 0) return kOk;
  Workspace w(ws, ws_bytes);
  SchedWs b;
  if (!carve_sched(w, n_ref, b)) {
    set_error("graph_schedule: workspace too small");
    return kWorkspaceTooSmall;
  }
  K* lo = static_cast<K*>(b.bbox);
  K* hi = lo + 8;
  if (sizeof(K) == 4) {
    GGP_TRY(fill_u32(lo, 0xffffffffu, 8, s));
    GGP_TRY(fill_u32(hi, 0u, 8, s));
  } else {
    GGP_TRY(fill_u64(lo, ~0ull, 8, s));
    GGP_TRY(fill_u64(hi, 0ull, 8, s));
  }
  hipLaunchKernelGGL(bbox_kernel<T>, dim3(std::min<int64_t>(grid_for(n, 256), 2048)), dim3(256), 0, s, pts, n, d, lo,
                     hi);
  hipLaunchKernelGGL(sched_keys_kernel<T>, dim3(grid_capped(n_ref, 256)), dim3(256), 0, s, pts, d, n0, n_ref, offsets,
                     n_off, (const K*)lo, (const K*)hi, b.key0, b.pay0);
  GGP_LAUNCH_CHECK();
  uint64_t* keys = b.key0;
  uint32_t* pay = b.pay0;
  int end_bit = kMortonBits;
  while (end_bit < 64 && (int64_t(1) << (end_bit - kMortonBits)) < n_off) ++end_bit;
  GGP_TRY(radix_sort_pairs<uint64_t>(&keys, &pay, b.key1, b.pay1, n_ref, 0, end_bit, b.sort_ws, b.sort_ws_bytes, s));
  hipLaunchKernelGGL(sched_out_kernel, dim3(grid_capped(n_ref, 256)), dim3(256), 0, s, pay, n_ref, sched, b.slot);
  hipLaunchKernelGGL(sched_nbrs_kernel, dim3(grid_capped(n_ref * k, 256)), dim3(256), 0, s, nbrs, k, n0, n_ref,
                     (const int32_t*)sched, (const uint32_t*)b.slot, nbv);
  if (k >= 3 && k <= 65) {
    if (k <= 17)
      hipLaunchKernelGGL(sort_rows_kernel<16>, dim3(grid_capped(n_ref, 256)), dim3(256), 0, s, nbv, k, n_ref);
    else if (k <= 33)
      hipLaunchKernelGGL(sort_rows_kernel<32>, dim3(grid_capped(n_ref, 256)), dim3(256), 0, s, nbv, k, n_ref);
    else
      hipLaunchKernelGGL(sort_rows_kernel<64>, dim3(grid_capped(n_ref, 256)), dim3(256), 0, s, nbv, k, n_ref);
  }
  if (pts_sched)
    hipLaunchKernelGGL(sched_points_kernel<T>, dim3(grid_capped(n * d, 256)), dim3(256), 0, s, pts, d, n0, n,
                       (const int32_t*)sched, pts_sched);
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace

size_t graph_schedule_workspace(int64_t n_ref) {
  Workspace w(nullptr, 0);
  SchedWs b;
  carve_sched(w, n_ref, b);
  return w.used + 256;
}

int graph_schedule(const void* pts, int dtype, int64_t n, int d, const int32_t* nbrs, int64_t n0, int k,
                   const int64_t* offsets, int64_t n_off, int32_t* sched, int32_t* nbv, void* pts_sched, void* ws,
                   size_t ws_bytes, hipStream_t s) {
  if (dtype == 0)
    return schedule_impl<float>((const float*)pts, n, d, nbrs, n0, k, offsets, n_off, sched, nbv, (float*)pts_sched,
                                ws, ws_bytes, s);
  if (dtype == 1)
    return schedule_impl<double>((const double*)pts, n, d, nbrs, n0, k, offsets, n_off, sched, nbv,
                                 (double*)pts_sched, ws, ws_bytes, s);
  set_error("graph_schedule: dtype must be 0 (float32) or 1 (float64)");
  return kInvalidArgument;
}

}  // namespace ggp
