// knn.hip -- preceding k-nearest-neighbour search over the level-interleaved k-d tree.
//
// Reference: graphgp/tree.py:41-150.  One thread per query runs the reference's
// stack-free parent-pointer traversal (tree.py:108-150) exactly: the node is offered
// to the candidate list on first arrival from its parent, the near child is the left
// one iff q[dim] - p[dim] < 0, a far child is entered iff it exists (< max_index) and
// (q[dim] - p[dim])^2 <= r^2.  The candidate list keeps k slots (distance^2, id),
// initialised to (+inf, -1); a candidate replaces the FIRST slot holding the maximum
// iff it is strictly closer (tree.py:81-95).  The result is sorted by
// (sqrt(distance^2), id) (tree.py:102-104).  Every comparison uses the same separately
// rounded arithmetic as the reference (this file is built with -ffp-contract=off), so
// indices match bit for bit, ties included.
//
// Bounding-box pruning: each subtree's axis-aligned bounding box is built bottom-up, and
// a child is skipped when its box distance^2 is strictly greater than r^2.  Every point
// in such a subtree has distance^2 >= box distance^2 > r^2 (the box distance is summed in
// the same order, with the same rounding, and each term is a monotone lower bound of the
// point's term), so the reference would visit those points without inserting any of them,
// and r^2 never grows: the candidate list evolves exactly as in the reference, ties
// included.  The skip only removes work: near-empty visits around dense clusters, where
// split-plane distances alone cannot prune, dominate the reference's worst queries.
//
// MI355X mapping: the k slots live in VGPRs (k <= 64, compile-time slot count KMAX with
// slots >= k pinned at -inf so they never hold the maximum); the final order is a
// register bitonic network.  Larger k uses slots in a caller workspace.  The kernel is held
// to five wavefronts per SIMD up to k = 16 and three up to k = 32 (96 and 168 VGPRs; the
// compiler chose 103 and 169 VGPRs, one wavefront fewer each), and a visit's point and split
// loads carry no branch: at 16.8M clustered points, k = 30, 263 -> 225 ms.  Reading both
// children's boxes with the node (one round trip per visit instead of two) was measured and
// dropped: no gain at 16.8M (229 ms), and 0.45 -> 0.65 s at 67M uniform points, k = 16.
#include "common.hpp"
#include "graphgp_amd_internal.hpp"

namespace ggp {

namespace {

template <typename T, int D>
struct Query {
  T c[D];
  __device__ inline T get(int dim) const {
    T v = c[0];
#pragma unroll
    for (int j = 1; j < D; ++j)
      if (dim == j) v = c[j];
    return v;
  }
};

template <typename T, int D>
__device__ inline T dist2(const Query<T, D>& q, const T* __restrict__ p, int d) {
  T t = q.c[0] - p[0];
  T s = t * t;
#pragma unroll
  for (int j = 1; j < D; ++j) {
    if (j < d) {
      T u = q.c[j] - p[j];
      s = s + u * u;
    }
  }
  return s;
}

// box of node c: lo at box[c*2d + j], hi at box[c*2d + d + j]; distance^2 from q to the
// box, accumulated in the order of dist2 (the register kernel computes the same sum from
// boxes it has loaded with the node).
template <typename T>
__device__ inline T box_dist2_rt(const T* __restrict__ q, const T* __restrict__ b, int d) {
  T s = T(0);
  for (int j = 0; j < d; ++j) {
    const T lo = b[j] - q[j], hi = q[j] - b[d + j];
    const T g = lo > T(0) ? lo : (hi > T(0) ? hi : T(0));
    s = j == 0 ? g * g : s + g * g;
  }
  return s;
}

// boxes of the nodes of one tree level, from their own point and their children's boxes
template <typename T>
__global__ void box_level_kernel(const T* __restrict__ pts, int64_t n, int d, int64_t first, int64_t count,
                                 T* __restrict__ box) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t c = first + t;
  if (c >= n) return;
  const int64_t l = tree_left(c), r = tree_right(c);
  T* b = box + c * 2 * d;
  for (int j = 0; j < d; ++j) {
    T lo = pts[c * d + j], hi = lo;
    if (l < n) {
      lo = fmin(lo, box[l * 2 * d + j]);
      hi = fmax(hi, box[l * 2 * d + d + j]);
    }
    if (r < n) {
      lo = fmin(lo, box[r * 2 * d + j]);
      hi = fmax(hi, box[r * 2 * d + d + j]);
    }
    b[j] = lo;
    b[d + j] = hi;
  }
}

template <typename T>
__device__ inline bool pair_greater(T da, int32_t ia, T db, int32_t ib) {
  return da > db || (da == db && ia > ib);
}

// Query t: the k nearest preceding points of point qi, by the reference's traversal.
template <typename T, int D, int KMAX>
__device__ __forceinline__ void knn_query(const T* __restrict__ pts, const int32_t* __restrict__ split, int d,
                                          const int32_t* __restrict__ qidx, const int32_t* __restrict__ midx,
                                          int64_t q0, int64_t t, int k, int32_t* __restrict__ out, int64_t n,
                                          const T* __restrict__ box) {
  const int64_t qi = qidx ? (int64_t)qidx[t] : q0 + t;
  const int64_t max_index = min(midx ? (int64_t)midx[t] : qi, n);
  Query<T, D> q;
#pragma unroll
  for (int j = 0; j < D; ++j) q.c[j] = j < d ? pts[qi * d + j] : T(0);

  const T inf = __builtin_inf();
  T sd[KMAX];
  int32_t si[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    sd[j] = j < k ? inf : -inf;
    si[j] = -1;
  }
  int amax = 0;
  T r2 = inf;
  int64_t cur = 0, prev = -1;
  while (cur != -1) {
    const int64_t parent = tree_parent(cur);
    const int64_t l = tree_left(cur), r = tree_right(cur);
    // The node's point and split dimension, read together; coordinates j >= d, which the
    // instantiation carries beyond the data's dimension, read coordinate d - 1 and are
    // unused (a load under a branch would be waited for at its join).  A child's box is read
    // when it is tested.
    Query<T, D> p;
#pragma unroll
    for (int j = 0; j < D; ++j) p.c[j] = pts[cur * d + (j < d ? j : d - 1)];
    const int dim = split[cur];
    if (prev == parent) {
      const T s = dist2<T, D>(q, p.c, d);
      if (s < r2) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
          if (j == amax) { sd[j] = s; si[j] = (int32_t)cur; }
        amax = 0;
        r2 = sd[0];
#pragma unroll
        for (int j = 1; j < KMAX; ++j)
          if (sd[j] > r2) { r2 = sd[j]; amax = j; }
      }
    }
    const T sdist = q.get(dim) - p.get(dim);
    const bool left_near = sdist < T(0);
    const int64_t nearc = left_near ? l : r;
    const int64_t farc = left_near ? r : l;
    const bool far_in = sdist * sdist <= r2;
    auto child_box_dist2 = [&](bool near) -> T {
      return box_dist2_rt<T>(q.c, box + ((near == left_near) ? l : r) * 2 * d, d);
    };
    int64_t next;
    if (prev == nearc || (prev == parent && (nearc >= max_index || child_box_dist2(true) > r2)))
      next = (farc < max_index && far_in && !(child_box_dist2(false) > r2)) ? farc : parent;
    else
      next = prev == parent ? nearc : parent;
    prev = cur;
    cur = next;
  }
  // final order: (sqrt distance, index); padding slots sort last
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    if (j < k) sd[j] = sqrt(sd[j]);
    else { sd[j] = inf; si[j] = 0x7fffffff; }
  }
#pragma unroll
  for (int size = 2; size <= KMAX; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
      for (int j = 0; j < KMAX; ++j) {
        const int l = j ^ stride;
        if (l > j) {
          const bool asc = (j & size) == 0;
          const bool gt = pair_greater(sd[j], si[j], sd[l], si[l]);
          if (gt == asc) {
            T td = sd[j]; sd[j] = sd[l]; sd[l] = td;
            int32_t ti = si[j]; si[j] = si[l]; si[l] = ti;
          }
        }
      }
    }
  }
  int32_t* o = out + t * k;
#pragma unroll
  for (int j = 0; j < KMAX; ++j)
    if (j < k) o[j] = si[j];
}

// One lane per query.  Plain: workgroup b takes queries [256 b, 256 b + 256).  TICKET (a
// launch capped below the query count, so that it leaves room for other work on the
// device): workgroups take blocks of 256 queries from the ticket until none are left.
template <typename T, int D, int KMAX, bool TICKET>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KMAX <= 16 ? 5 : KMAX <= 32 ? 3 : 1)))
void knn_kernel(const T* __restrict__ pts, const int32_t* __restrict__ split,
                                                  int d, const int32_t* __restrict__ qidx,
                                                  const int32_t* __restrict__ midx, int64_t q0, int64_t nq, int k,
                                                  int32_t* __restrict__ out, int64_t n, const T* __restrict__ box,
                                                  uint32_t* ticket, const uint32_t* yield, uint32_t keep) {
  if constexpr (!TICKET) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nq) knn_query<T, D, KMAX>(pts, split, d, qidx, midx, q0, t, k, out, n, box);
  } else {
    __shared__ uint32_t s_blk;
    for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) s_blk = yield_now(yield, keep) ? 0xffffffffu : atomicAdd(ticket, 1u);
      __syncthreads();
      const uint32_t blk = s_blk;
      if ((int64_t)blk * 256 >= nq) break;
      const int64_t t = (int64_t)blk * 256 + threadIdx.x;
      if (t < nq) knn_query<T, D, KMAX>(pts, split, d, qidx, midx, q0, t, k, out, n, box);
    }
  }
}

// Generic k: slots in a per-query workspace region (k*(sizeof(T)+4) bytes per query).
template <typename T>
__global__ __launch_bounds__(256) void knn_generic_kernel(const T* __restrict__ pts,
                                                          const int32_t* __restrict__ split, int d,
                                                          const int32_t* __restrict__ qidx,
                                                          const int32_t* __restrict__ midx, int64_t q0, int64_t nq,
                                                          int k, int32_t* __restrict__ out, T* __restrict__ wsd,
                                                          int32_t* __restrict__ wsi, int64_t n,
                                                          const T* __restrict__ box) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const int64_t qi = qidx ? (int64_t)qidx[t] : q0 + t;
  const int64_t max_index = min(midx ? (int64_t)midx[t] : qi, n);
  const T* q = pts + qi * d;
  T* sd = wsd + t * k;
  int32_t* si = wsi + t * k;
  const T inf = __builtin_inf();
  for (int j = 0; j < k; ++j) { sd[j] = inf; si[j] = -1; }
  int amax = 0;
  T r2 = inf;
  int64_t cur = 0, prev = -1;
  while (cur != -1) {
    const int64_t parent = tree_parent(cur);
    const T* p = pts + cur * d;
    if (prev == parent) {
      T u = q[0] - p[0];
      T s = u * u;
      for (int j = 1; j < d; ++j) { u = q[j] - p[j]; s = s + u * u; }
      if (s < r2) {
        sd[amax] = s;
        si[amax] = (int32_t)cur;
        amax = 0;
        r2 = sd[0];
        for (int j = 1; j < k; ++j)
          if (sd[j] > r2) { r2 = sd[j]; amax = j; }
      }
    }
    const int dim = split[cur];
    const T sdist = q[dim] - p[dim];
    const int64_t l = tree_left(cur), r = tree_right(cur);
    const int64_t nearc = sdist < T(0) ? l : r;
    const int64_t farc = sdist < T(0) ? r : l;
    const bool far_in = sdist * sdist <= r2;
    int64_t next;
    if (prev == nearc || (prev == parent && (nearc >= max_index || box_dist2_rt<T>(q, box + nearc * 2 * d, d) > r2)))
      next = (farc < max_index && far_in && !(box_dist2_rt<T>(q, box + farc * 2 * d, d) > r2)) ? farc : parent;
    else
      next = prev == parent ? nearc : parent;
    prev = cur;
    cur = next;
  }
  for (int j = 0; j < k; ++j) sd[j] = sqrt(sd[j]);
  // insertion sort by (distance, index)
  for (int a = 1; a < k; ++a) {
    T dv = sd[a];
    int32_t iv = si[a];
    int b = a - 1;
    while (b >= 0 && pair_greater(sd[b], si[b], dv, iv)) {
      sd[b + 1] = sd[b];
      si[b + 1] = si[b];
      --b;
    }
    sd[b + 1] = dv;
    si[b + 1] = iv;
  }
  for (int j = 0; j < k; ++j) out[t * k + j] = si[j];
}

template <typename T, int D, int KMAX>
int launch_knn(const T* pts, const int32_t* split, int d, const int32_t* qidx, const int32_t* midx, int64_t q0,
               int64_t nq, int k, int32_t* out, int64_t n, const T* box, uint32_t* ticket, hipStream_t s) {
  const unsigned blocks = grid_for(nq, 256);
  const YieldCfg& y = yield_cfg();
  if (cu_budget() > 0 || y.dev != nullptr) {
    // a long kernel: within the calling thread's CU budget, or yielding while asked to
    const void* kern = (const void*)knn_kernel<T, D, KMAX, true>;
    const int cus = cu_budget() > 0 ? cu_budget() : device_cus();
    const int g = (int)std::min<int64_t>(blocks, (int64_t)cus * blocks_per_cu(kern, 0));
    if ((unsigned)g < blocks || y.dev != nullptr) {
      GGP_TRY(fill_u32(ticket, 0u, 1, s));
      hipLaunchKernelGGL((knn_kernel<T, D, KMAX, true>), dim3((unsigned)g), dim3(256), 0, s, pts, split, d, qidx,
                         midx, q0, nq, k, out, n, box, ticket, y.dev, yield_keep(g));
      return kOk;
    }
  }
  hipLaunchKernelGGL((knn_kernel<T, D, KMAX, false>), dim3(blocks), dim3(256), 0, s, pts, split, d, qidx, midx, q0,
                     nq, k, out, n, box, (uint32_t*)nullptr, (const uint32_t*)nullptr, 0u);
  return kOk;
}

template <typename T, int D>
int dispatch_k(const T* pts, const int32_t* split, int d, const int32_t* qidx, const int32_t* midx, int64_t q0,
               int64_t nq, int k, int32_t* out, int64_t n, const T* box, uint32_t* ticket, hipStream_t s) {
  if (k <= 8) return launch_knn<T, D, 8>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s);
  if (k <= 16) return launch_knn<T, D, 16>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s);
  if (k <= 32) return launch_knn<T, D, 32>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s);
  return launch_knn<T, D, 64>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s);
}

template <typename T>
int knn_impl(const T* pts, const int32_t* split, int64_t n, int d, const int32_t* qidx, const int32_t* midx,
             int64_t q0, int64_t nq, int k, int32_t* out, void* ws, size_t ws_bytes, hipStream_t s) {
  GGP_CHECK_ARG(d >= 1 && d <= 8, "query_neighbors: 1 <= d <= 8");
  GGP_CHECK_ARG(k >= 1, "query_neighbors: k >= 1");
  GGP_CHECK_ARG(n < (int64_t(1) << 31), "query_neighbors: N must be below 2^31");
  if (nq <= 0) return kOk;
  Workspace w(ws, ws_bytes);
  T* box = w.take<T>((size_t)n * 2 * d);
  uint32_t* ticket = w.take<uint32_t>(64);
  T* wsd = nullptr;
  int32_t* wsi = nullptr;
  if (k > 64) {
    wsd = w.take<T>((size_t)nq * k);
    wsi = w.take<int32_t>((size_t)nq * k);
  }
  if (!w.ok()) {
    set_error("query_neighbors: workspace too small");
    return kWorkspaceTooSmall;
  }
  // subtree boxes, deepest level first
  const int top = tree_level(n - 1);
  for (int lv = top; lv >= 0; --lv) {
    const int64_t first = (int64_t(1) << lv) - 1;
    const int64_t count = std::min<int64_t>(int64_t(1) << lv, n - first);
    hipLaunchKernelGGL(box_level_kernel<T>, dim3(grid_for(count, 256)), dim3(256), 0, s, pts, n, d, first, count, box);
  }
  GGP_LAUNCH_CHECK();
  if (k > 64) {
    hipLaunchKernelGGL(knn_generic_kernel<T>, dim3(grid_for(nq, 256)), dim3(256), 0, s, pts, split, d, qidx, midx,
                       q0, nq, k, out, wsd, wsi, n, (const T*)box);
  } else if (d == 1) {
    GGP_TRY((dispatch_k<T, 1>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s)));
  } else if (d == 2) {
    GGP_TRY((dispatch_k<T, 2>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s)));
  } else if (d == 3) {
    GGP_TRY((dispatch_k<T, 3>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s)));
  } else {
    GGP_TRY((dispatch_k<T, 8>(pts, split, d, qidx, midx, q0, nq, k, out, n, box, ticket, s)));
  }
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace

size_t query_neighbors_workspace(int64_t n, int d, int64_t nq, int k, int dtype) {
  Workspace w(nullptr, 0);
  w.take<char>((size_t)n * 2 * d * dtype_size(dtype));
  w.take<uint32_t>(64);
  if (k > 64) {
    w.take<char>((size_t)nq * k * dtype_size(dtype));
    w.take<int32_t>((size_t)nq * k);
  }
  return w.used + 256;
}

int query_neighbors(const void* pts, const int32_t* split, int64_t n, int d, int dtype, const int32_t* qidx,
                    const int32_t* midx, int64_t q0, int64_t nq, int k, int32_t* out, void* ws, size_t ws_bytes,
                    hipStream_t s) {
  if (dtype == 0)
    return knn_impl<float>((const float*)pts, split, n, d, qidx, midx, q0, nq, k, out, ws, ws_bytes, s);
  if (dtype == 1)
    return knn_impl<double>((const double*)pts, split, n, d, qidx, midx, q0, nq, k, out, ws, ws_bytes, s);
  set_error("query_neighbors: dtype must be 0 (float32) or 1 (float64)");
  return kInvalidArgument;
}

}  // namespace ggp
