// tree.hip -- k-d tree build in the reference's level-interleaved order (gfx950).
//
// Reference: graphgp/tree.py:153-234 (_build_tree / _update_nodes).  At level L the
// not-yet-placed points form 2^L contiguous segments whose sizes follow from N and L
// alone (q = rem / 2^L, r = rem % 2^L; the first r segments hold q+1 points).  For
// each segment the split dimension is the first argmax of the coordinate extent, the
// segment is stably sorted by that coordinate, its median (rank size/2) becomes node
// n_above + i, and the lower/upper halves become segments i and i + 2^L of level L+1.
// Stability chains across levels because every level sorts the previous order.
//
// MI355X mapping, per level:
//   extent  : contiguous chunks per thread, per-block LDS reduction when a block lies in
//             one segment, otherwise per-thread flushes with ordered-key atomics;
//   split   : one thread per segment;
//   keys    : (segment << 32) | order-preserving coordinate key, payload = position;
//   sort    : the stable LSD radix sort (radix.hip) over 32 + L key bits;
//   regroup : every sorted element goes straight to its final node id (medians) or to
//             its slot in the next level's segment, computed in closed form.
#include "common.hpp"
#include "radix.hpp"
#include "graphgp_amd_internal.hpp"

namespace ggp {

namespace {

struct Level {
  uint32_t n_above;  // first active position (= first node id of the level)
  uint32_t n_level;  // 2^L
  uint32_t q, r;     // segment sizes: first r segments hold q+1
  uint32_t n_act;    // N - n_above
  __host__ __device__ Level(uint32_t n, int L) {
    n_above = (1u << L) - 1u;
    n_level = 1u << L;
    n_act = n - n_above;
    q = n_act / n_level;
    r = n_act % n_level;
  }
  // segment index of relative position p (0 <= p < n_act)
  __device__ inline uint32_t seg_of(uint32_t p) const {
    const uint32_t big = r * (q + 1u);
    return p < big ? p / (q + 1u) : r + (p - big) / q;
  }
  __host__ __device__ inline uint32_t start(uint32_t i) const {
    return i < r ? i * (q + 1u) : r * (q + 1u) + (i - r) * q;
  }
  __host__ __device__ inline uint32_t size(uint32_t i) const { return q + (i < r ? 1u : 0u); }
};

template <typename K>
__device__ inline K kmin_(K a, K b) { return a < b ? a : b; }
template <typename K>
__device__ inline K kmax_(K a, K b) { return a < b ? b : a; }

constexpr int kExtThreads = 256;
constexpr int kExtItems = 8;

template <typename T>
__global__ __launch_bounds__(kExtThreads) void extent_kernel(const T* __restrict__ pts, int d,
                                                             const uint32_t* __restrict__ act, Level lv,
                                                             typename KeyOf<T>::type* __restrict__ kmin,
                                                             typename KeyOf<T>::type* __restrict__ kmax) {
  using K = typename KeyOf<T>::type;
  const uint32_t blk0 = blockIdx.x * (kExtThreads * kExtItems);
  if (blk0 >= lv.n_act) return;
  const uint32_t blk1 = min(lv.n_act, blk0 + kExtThreads * kExtItems) - 1u;
  const uint32_t s0 = lv.seg_of(blk0), s1 = lv.seg_of(blk1);
  const uint32_t p0 = blk0 + threadIdx.x * kExtItems;
  constexpr int kMaxD = 8;
  if (s0 == s1) {
    // whole block inside one segment: block reduction, one atomic per dimension
    __shared__ K smin[kExtThreads / 64][kMaxD], smax[kExtThreads / 64][kMaxD];
    K mn[kMaxD], mx[kMaxD];
    for (int j = 0; j < d; ++j) { mn[j] = ~K(0); mx[j] = K(0); }
    for (int t = 0; t < kExtItems; ++t) {
      uint32_t p = p0 + t;
      if (p < lv.n_act) {
        const T* x = pts + (size_t)act[p] * d;
        for (int j = 0; j < d; ++j) {
          K kk = KeyOf<T>::key(x[j]);
          mn[j] = kmin_(mn[j], kk);
          mx[j] = kmax_(mx[j], kk);
        }
      }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = 0; j < d; ++j) {
      K a = mn[j], b = mx[j];
      for (int o = 32; o > 0; o >>= 1) {
        a = kmin_(a, (K)__shfl_xor(a, o, 64));
        b = kmax_(b, (K)__shfl_xor(b, o, 64));
      }
      if (lane == 0) { smin[wid][j] = a; smax[wid][j] = b; }
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)d) {
      const int j = threadIdx.x;
      K a = smin[0][j], b = smax[0][j];
      for (int w = 1; w < kExtThreads / 64; ++w) { a = kmin_(a, smin[w][j]); b = kmax_(b, smax[w][j]); }
      atomicMin(&kmin[(size_t)s0 * d + j], a);
      atomicMax(&kmax[(size_t)s0 * d + j], b);
    }
    return;
  }
  // several segments in the block: per-thread runs, flushed on every segment change
  K mn[kMaxD], mx[kMaxD];
  uint32_t cur = 0xffffffffu;
  for (int t = 0; t < kExtItems; ++t) {
    uint32_t p = p0 + t;
    if (p >= lv.n_act) break;
    uint32_t s = lv.seg_of(p);
    if (s != cur) {
      if (cur != 0xffffffffu)
        for (int j = 0; j < d; ++j) {
          atomicMin(&kmin[(size_t)cur * d + j], mn[j]);
          atomicMax(&kmax[(size_t)cur * d + j], mx[j]);
        }
      cur = s;
      for (int j = 0; j < d; ++j) { mn[j] = ~K(0); mx[j] = K(0); }
    }
    const T* x = pts + (size_t)act[p] * d;
    for (int j = 0; j < d; ++j) {
      K kk = KeyOf<T>::key(x[j]);
      mn[j] = kmin_(mn[j], kk);
      mx[j] = kmax_(mx[j], kk);
    }
  }
  if (cur != 0xffffffffu)
    for (int j = 0; j < d; ++j) {
      atomicMin(&kmin[(size_t)cur * d + j], mn[j]);
      atomicMax(&kmax[(size_t)cur * d + j], mx[j]);
    }
}

// split dim = first argmax over dims of (max - min), evaluated in T like the reference
// (tree.py:162-165).  A segment of one point has zero extent everywhere -> 0.
template <typename T>
__global__ void split_kernel(int d, Level lv, const typename KeyOf<T>::type* __restrict__ kmin,
                             const typename KeyOf<T>::type* __restrict__ kmax, int32_t* __restrict__ split_dims,
                             uint32_t n_seg) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_seg) return;
  int best = 0;
  if (lv.size(i) > 1) {
    T bestv = KeyOf<T>::value(kmax[(size_t)i * d]) - KeyOf<T>::value(kmin[(size_t)i * d]);
    for (int j = 1; j < d; ++j) {
      T e = KeyOf<T>::value(kmax[(size_t)i * d + j]) - KeyOf<T>::value(kmin[(size_t)i * d + j]);
      if (e > bestv) { bestv = e; best = j; }
    }
  }
  split_dims[lv.n_above + i] = best;
}

// key = (segment << 32) | coordinate key (float points)
__global__ void keys_f32_kernel(const float* __restrict__ pts, int d, const uint32_t* __restrict__ act, Level lv,
                                const int32_t* __restrict__ split_dims, uint64_t* __restrict__ keys,
                                uint32_t* __restrict__ pay) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lv.n_act) return;
  uint32_t s = lv.seg_of(p);
  int dim = split_dims[lv.n_above + s];
  float v = pts[(size_t)act[p] * d + dim];
  keys[p] = ((uint64_t)s << 32) | float_key(v);
  pay[p] = p;
}

// double points: first pass sorts by the 64-bit coordinate key alone ...
__global__ void keys_f64_kernel(const double* __restrict__ pts, int d, const uint32_t* __restrict__ act, Level lv,
                                const int32_t* __restrict__ split_dims, uint64_t* __restrict__ keys,
                                uint32_t* __restrict__ pay) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lv.n_act) return;
  uint32_t s = lv.seg_of(p);
  int dim = split_dims[lv.n_above + s];
  keys[p] = double_key(pts[(size_t)act[p] * d + dim]);
  pay[p] = p;
}
// ... then a stable pass by the segment of each element's original position.
__global__ void keys_seg_kernel(Level lv, const uint32_t* __restrict__ pay, uint64_t* __restrict__ keys) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= lv.n_act) return;
  keys[p] = lv.seg_of(pay[p]);
}

// Place every sorted element: the median of segment i is node n_above + i; the lower
// half goes to next-level segment i, the upper half to segment i + 2^L.
__global__ void regroup_kernel(Level lv, Level nx, bool last, const uint32_t* __restrict__ pay,
                               const uint32_t* __restrict__ act, uint32_t* __restrict__ next_act,
                               int32_t* __restrict__ final_idx) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= lv.n_act) return;
  const uint32_t id = act[pay[s]];
  const uint32_t i = lv.seg_of(s);
  const uint32_t rank = s - lv.start(i);
  const uint32_t mid = lv.size(i) / 2u;
  if (rank == mid) {
    final_idx[lv.n_above + i] = (int32_t)id;
  } else if (!last) {
    uint32_t dst = rank < mid ? nx.start(i) + rank : nx.start(i + lv.n_level) + (rank - mid - 1u);
    next_act[dst] = id;
  }
}

// ---- the last levels of every segment in one workgroup each.  Once a level's segments
// hold at most kSubCap points, the whole subtree of each segment is built in LDS by one
// workgroup: its descendant segments at level L are segments i + m 2^L0 (m < 2^(L-L0)),
// whose sizes follow from N and L as above.  Sub-segment m lives in slot [m P, m P + size)
// with P the next power of two of the level's largest segment (M P <= kSubCap for every
// N: the subtree of a segment of at most 2^12 points is at most 12 levels deep).  Per
// level: extents and split dimension per sub-segment, keys (coordinate key, position in
// the previous order), a bitonic sort inside every slot -- the previous position makes the
// keys distinct, so this is the reference's stable sort -- and the median / halves
// regroup into the next level's slots.
constexpr int kSubCap = 4096;
constexpr int kSubThreads = 256;  // small enough to fit beside a CU-budgeted large launch

__host__ __device__ inline uint32_t pow2_at_least(uint32_t x) {
  uint32_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// first level whose segments all fit one workgroup (levels if the slots would not fit)
inline int subtree_level(uint32_t n, int levels) {
  int L0 = levels;
  for (int L = 0; L < levels; ++L) {
    const Level lv(n, L);
    if (lv.q + (lv.r ? 1u : 0u) <= (uint32_t)kSubCap) {
      L0 = L;
      break;
    }
  }
  for (int L = L0; L < levels; ++L) {
    const Level lv(n, L);
    if ((uint64_t(1) << (L - L0)) * pow2_at_least(lv.q + (lv.r ? 1u : 0u)) > (uint64_t)kSubCap) return levels;
  }
  return L0;
}

template <typename T>
constexpr size_t subtree_lds() {
  // ids (two buffers), split dims, sort keys (+ positions for 64-bit coordinate keys)
  return 2 * kSubCap * 4 + kSubCap + kSubCap * 8 + (sizeof(T) == 8 ? kSubCap * 4 : 0);
}

template <typename T>
__global__ __launch_bounds__(kSubThreads) void subtree_kernel(const T* __restrict__ pts, int d, uint32_t n, int L0,
                                                              int levels, const uint32_t* __restrict__ act,
                                                              int32_t* __restrict__ split_dims,
                                                              int32_t* __restrict__ final_idx) {
  using K = typename KeyOf<T>::type;
  constexpr bool kWide = sizeof(T) == 8;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  uint64_t* skey = reinterpret_cast<uint64_t*>(smem_raw);
  uint32_t* spos = reinterpret_cast<uint32_t*>(skey + kSubCap);  // 64-bit keys only
  uint32_t* cur = kWide ? spos + kSubCap : spos;
  uint32_t* nxt = cur + kSubCap;
  uint8_t* spl = reinterpret_cast<uint8_t*>(nxt + kSubCap);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t stride0 = 1u << L0;
  const Level l0(n, L0);
  const uint32_t n_seg0 = min(l0.n_level, l0.n_act);
  for (uint32_t i = blockIdx.x; i < n_seg0; i += gridDim.x) {
    {
      const uint32_t st = l0.start(i), sz = l0.size(i);
      for (uint32_t t = tid; t < sz; t += kSubThreads) cur[t] = act[st + t];
    }
    __syncthreads();
    for (int L = L0; L < levels; ++L) {
      const Level lv(n, L);
      const bool last = L == levels - 1;
      const uint32_t M = 1u << (L - L0);
      const uint32_t P = pow2_at_least(lv.q + (lv.r ? 1u : 0u));
      const int lgP = __builtin_ctz(P);
      // extents and split dimension (first argmax of max - min, in T), as split_kernel
      auto split_of = [&](const K* mn, const K* mx) {
        int best = 0;
        T bestv = KeyOf<T>::value(mx[0]) - KeyOf<T>::value(mn[0]);
        for (int j = 1; j < d; ++j) {
          const T e = KeyOf<T>::value(mx[j]) - KeyOf<T>::value(mn[j]);
          if (e > bestv) { bestv = e; best = j; }
        }
        return best;
      };
      if (P >= 64) {
        for (uint32_t m = wave; m < M; m += kSubThreads / 64) {
          const uint32_t sg = i + m * stride0, sz = lv.size(sg);
          K mn[8], mx[8];
          for (int j = 0; j < d; ++j) { mn[j] = ~K(0); mx[j] = K(0); }
          for (uint32_t t = lane; t < sz; t += 64) {
            const T* x = pts + (size_t)cur[(m << lgP) + t] * d;
            for (int j = 0; j < d; ++j) {
              const K kk = KeyOf<T>::key(x[j]);
              mn[j] = kmin_(mn[j], kk);
              mx[j] = kmax_(mx[j], kk);
            }
          }
          for (int j = 0; j < d; ++j)
            for (int o = 32; o > 0; o >>= 1) {
              mn[j] = kmin_(mn[j], (K)__shfl_xor(mn[j], o, 64));
              mx[j] = kmax_(mx[j], (K)__shfl_xor(mx[j], o, 64));
            }
          if (lane == 0) {
            const int best = sz > 1 ? split_of(mn, mx) : 0;
            spl[m] = (uint8_t)best;
            split_dims[lv.n_above + sg] = best;
          }
        }
      } else {
        for (uint32_t m = tid; m < M; m += kSubThreads) {
          const uint32_t sg = i + m * stride0, sz = lv.size(sg);
          if (sz == 0) continue;
          int best = 0;
          if (sz > 1) {
            K mn[8], mx[8];
            for (int j = 0; j < d; ++j) { mn[j] = ~K(0); mx[j] = K(0); }
            for (uint32_t t = 0; t < sz; ++t) {
              const T* x = pts + (size_t)cur[(m << lgP) + t] * d;
              for (int j = 0; j < d; ++j) {
                const K kk = KeyOf<T>::key(x[j]);
                mn[j] = kmin_(mn[j], kk);
                mx[j] = kmax_(mx[j], kk);
              }
            }
            best = split_of(mn, mx);
          }
          spl[m] = (uint8_t)best;
          split_dims[lv.n_above + sg] = best;
        }
      }
      __syncthreads();
      const uint32_t total = M << lgP;
      if (P > 1) {
        // keys: (coordinate key, position in the previous order); empty slot entries last
        for (uint32_t e = tid; e < total; e += kSubThreads) {
          const uint32_t m = e >> lgP, t = e & (P - 1);
          const uint32_t sz = lv.size(i + m * stride0);
          if (t < sz) {
            const K kk = KeyOf<T>::key(pts[(size_t)cur[e] * d + spl[m]]);
            if constexpr (kWide) {
              skey[e] = (uint64_t)kk;
              spos[e] = t;
            } else {
              skey[e] = ((uint64_t)kk << 32) | t;
            }
          } else {
            skey[e] = ~uint64_t(0);
            if constexpr (kWide) spos[e] = ~0u;
          }
        }
        __syncthreads();
        // bitonic sort of every slot of P entries, ascending
        for (uint32_t k = 2; k <= P; k <<= 1) {
          for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t e = tid; e < total; e += kSubThreads) {
              const uint32_t x = e ^ j;
              if (x > e) {
                const bool asc = ((e & (P - 1)) & k) == 0;
                const uint64_t a = skey[e], b = skey[x];
                bool gt;
                if constexpr (kWide) gt = a > b || (a == b && spos[e] > spos[x]);
                else gt = a > b;
                if (gt == asc) {
                  skey[e] = b;
                  skey[x] = a;
                  if constexpr (kWide) {
                    const uint32_t pa = spos[e];
                    spos[e] = spos[x];
                    spos[x] = pa;
                  }
                }
              }
            }
            __syncthreads();
          }
        }
      }
      // median -> node id; lower half -> slot m, upper half -> slot m + M of the next level
      const uint32_t Pn = last ? 1u : pow2_at_least(Level(n, L + 1).q + (Level(n, L + 1).r ? 1u : 0u));
      for (uint32_t e = tid; e < total; e += kSubThreads) {
        const uint32_t m = e >> lgP, t = e & (P - 1);
        const uint32_t sg = i + m * stride0, sz = lv.size(sg);
        if (t >= sz) continue;
        uint32_t pos = t;
        if (P > 1) pos = kWide ? spos[e] : (uint32_t)(skey[e] & 0xffffffffu);
        const uint32_t id = cur[(m << lgP) + pos];
        const uint32_t mid = sz / 2u;
        if (t == mid) final_idx[lv.n_above + sg] = (int32_t)id;
        else if (!last) nxt[t < mid ? m * Pn + t : (m + M) * Pn + (t - mid - 1u)] = id;
      }
      __syncthreads();
      uint32_t* sw = cur;
      cur = nxt;
      nxt = sw;
    }
  }
}

__global__ void iota_kernel(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

template <typename T>
__global__ void gather_points_kernel(const T* __restrict__ pts, int d, const int32_t* __restrict__ idx,
                                     T* __restrict__ out, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * d; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / d;
    const int j = (int)(e - i * d);
    out[e] = pts[(int64_t)idx[i] * d + j];
  }
}

struct TreeBufs {
  uint32_t *act0, *act1, *pay0, *pay1;
  uint64_t *key0, *key1;
  void *kmin, *kmax;
  void* sort_ws;
  size_t sort_ws_bytes;
};

bool carve_tree(Workspace& w, int64_t n, int d, int dtype, TreeBufs& b) {
  const size_t ks = dtype_size(dtype);
  b.act0 = w.take<uint32_t>(n);
  b.act1 = w.take<uint32_t>(n);
  b.pay0 = w.take<uint32_t>(n);
  b.pay1 = w.take<uint32_t>(n);
  b.key0 = w.take<uint64_t>(n);
  b.key1 = w.take<uint64_t>(n);
  b.kmin = w.take<char>((size_t)n * d * ks);
  b.kmax = w.take<char>((size_t)n * d * ks);
  b.sort_ws_bytes = radix_workspace(n);
  b.sort_ws = w.take<char>(b.sort_ws_bytes);
  return w.ok();
}

}  // namespace

size_t build_tree_workspace(int64_t n, int d, int dtype) {
  Workspace w(nullptr, 0);
  TreeBufs b;
  carve_tree(w, n, d, dtype, b);
  return w.used + 256;
}

template <typename T>
static int build_tree_impl(const T* pts, int64_t n64, int d, T* pts_out, int32_t* split_dims, int32_t* indices,
                           void* ws, size_t ws_bytes, hipStream_t s) {
  using K = typename KeyOf<T>::type;
  GGP_CHECK_ARG(n64 >= 1, "build_tree: need at least one point");
  GGP_CHECK_ARG(n64 < (int64_t(1) << 31), "build_tree: N must be below 2^31");
  GGP_CHECK_ARG(d >= 1 && d <= 8, "build_tree: 1 <= d <= 8");
  const uint32_t n = (uint32_t)n64;
  Workspace w(ws, ws_bytes);
  TreeBufs b;
  if (!carve_tree(w, n64, d, sizeof(T) == 8 ? 1 : 0, b)) {
    set_error("build_tree: workspace too small");
    return kWorkspaceTooSmall;
  }
  int levels = 0;
  while ((n >> levels) != 0) ++levels;
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, b.act0, n);
  GGP_LAUNCH_CHECK();
  uint32_t* act = b.act0;
  uint32_t* nxt = b.act1;
  // levels from L0 on: one workgroup per segment (LDS); the levels above: device-wide
  const int L0 = subtree_level(n, levels);
  for (int L = 0; L < L0; ++L) {
    Level lv(n, L);
    const bool last = (L == levels - 1);
    Level nx(n, last ? L : L + 1);
    const uint32_t n_seg = min(lv.n_level, lv.n_act);
    // split dimensions
    if (lv.q >= 2 || (lv.q == 1 && lv.r > 0)) {
      if (sizeof(K) == 4) {
        GGP_TRY(fill_u32(b.kmin, 0xffffffffu, (int64_t)n_seg * d, s));
        GGP_TRY(fill_u32(b.kmax, 0u, (int64_t)n_seg * d, s));
      } else {
        GGP_TRY(fill_u64(b.kmin, ~0ull, (int64_t)n_seg * d, s));
        GGP_TRY(fill_u64(b.kmax, 0ull, (int64_t)n_seg * d, s));
      }
      hipLaunchKernelGGL(extent_kernel<T>, dim3(grid_for(lv.n_act, kExtThreads * kExtItems)), dim3(kExtThreads), 0,
                         s, pts, d, act, lv, (K*)b.kmin, (K*)b.kmax);
    }
    hipLaunchKernelGGL(split_kernel<T>, dim3(grid_for(n_seg, 256)), dim3(256), 0, s, d, lv, (const K*)b.kmin,
                       (const K*)b.kmax, split_dims, n_seg);
    GGP_LAUNCH_CHECK();
    // stable sort of every segment by its split coordinate
    uint64_t* keys = b.key0;
    uint32_t* pay = b.pay0;
    if (lv.q >= 2 || (lv.q == 1 && lv.r > 0)) {
      if (sizeof(T) == 4) {
        hipLaunchKernelGGL(keys_f32_kernel, dim3(grid_for(lv.n_act, 256)), dim3(256), 0, s, (const float*)pts, d,
                           act, lv, split_dims, keys, pay);
        GGP_LAUNCH_CHECK();
        GGP_TRY(radix_sort_pairs<uint64_t>(&keys, &pay, keys == b.key0 ? b.key1 : b.key0,
                                           pay == b.pay0 ? b.pay1 : b.pay0, lv.n_act, 0, 32 + L, b.sort_ws,
                                           b.sort_ws_bytes, s));
      } else {
        hipLaunchKernelGGL(keys_f64_kernel, dim3(grid_for(lv.n_act, 256)), dim3(256), 0, s, (const double*)pts, d,
                           act, lv, split_dims, keys, pay);
        GGP_LAUNCH_CHECK();
        GGP_TRY(radix_sort_pairs<uint64_t>(&keys, &pay, keys == b.key0 ? b.key1 : b.key0,
                                           pay == b.pay0 ? b.pay1 : b.pay0, lv.n_act, 0, 64, b.sort_ws,
                                           b.sort_ws_bytes, s));
        if (L > 0) {
          hipLaunchKernelGGL(keys_seg_kernel, dim3(grid_for(lv.n_act, 256)), dim3(256), 0, s, lv, pay, keys);
          GGP_LAUNCH_CHECK();
          GGP_TRY(radix_sort_pairs<uint64_t>(&keys, &pay, keys == b.key0 ? b.key1 : b.key0,
                                             pay == b.pay0 ? b.pay1 : b.pay0, lv.n_act, 0, L, b.sort_ws,
                                             b.sort_ws_bytes, s));
        }
      }
    } else {
      // every segment holds at most one point: the order is already final
      hipLaunchKernelGGL(iota_kernel, dim3(grid_for(lv.n_act, 256)), dim3(256), 0, s, pay, lv.n_act);
      GGP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(regroup_kernel, dim3(grid_for(lv.n_act, 256)), dim3(256), 0, s, lv, nx, last, pay, act, nxt,
                       indices);
    GGP_LAUNCH_CHECK();
    uint32_t* t = act;
    act = nxt;
    nxt = t;
  }
  if (L0 < levels) {
    const Level l0(n, L0);
    constexpr size_t lds = subtree_lds<T>();
    static bool lds_set = false;  // the attribute is per kernel and process
    if (!lds_set) {
      GGP_HIP(hipFuncSetAttribute((const void*)subtree_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
      lds_set = true;
    }
    // one workgroup per segment, fewer under a CU budget or while asked to yield
    unsigned grid = std::min(l0.n_level, l0.n_act);
    if (cu_budget() > 0 || yield_cfg().host != nullptr) {
      const int per_cu = blocks_per_cu((const void*)subtree_kernel<T>, lds);
      int cus = cu_budget() > 0 ? cu_budget() : device_cus();
      if (yield_cfg().host != nullptr && *yield_cfg().host != 0u) cus = std::min(cus, yield_cfg().keep_cus);
      grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t)cus * per_cu));
    }
    hipLaunchKernelGGL(subtree_kernel<T>, dim3(grid), dim3(kSubThreads), lds, s, pts, d, n, L0, levels, act,
                       split_dims, indices);
    GGP_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(gather_points_kernel<T>, dim3(grid_capped(n64 * d, 256)), dim3(256), 0, s, pts, d, indices,
                     pts_out, n64);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int build_tree(const void* pts, int64_t n, int d, int dtype, void* pts_out, int32_t* split_dims, int32_t* indices,
               void* ws, size_t ws_bytes, hipStream_t s) {
  if (dtype == 0)
    return build_tree_impl<float>((const float*)pts, n, d, (float*)pts_out, split_dims, indices, ws, ws_bytes, s);
  if (dtype == 1)
    return build_tree_impl<double>((const double*)pts, n, d, (double*)pts_out, split_dims, indices, ws, ws_bytes, s);
  set_error("build_tree: dtype must be 0 (float32) or 1 (float64)");
  return kInvalidArgument;
}

}  // namespace ggp
