// refine.hip -- fused Vecchia refinement for gfx950: generate, inverse and log-det.
//
// Reference: graphgp/refine.py:79-159 (refine), :198-231 (refine_inv), :251-275
// (refine_logdet), :288-305 (compute_cov_matrix / cov_lookup).  For refined point i
// with neighbours n_1..n_k the joint covariance K over (n_1..n_k, i) (neighbours first,
// fine point last) is built from the tabulated covariance by linear interpolation of
// r = ||a - b||; with L = chol(K), mean = L[k,:k] . L_A^{-1} v[nbrs] and
// std = L[k,k]; generate writes mean + std * xi, the inverse (v - mean) / std and the
// log-det sums log std.
//
// MI355X mapping (one kernel per call, no per-batch launches):
//  * P lanes cooperate on one point; row i of the (k+1)^2 system lives in lane i % P,
//    slot i / P, holding columns 0..P*slot+P-1 (VGPRs, compile-time indexed).
//  * Matrix entries are produced on the fly: the neighbour coordinates are gathered
//    once (one row per lane) and broadcast across the P lanes with DPP moves (quad_perm,
//    plus a row shift for P = 8, row_newbcast for P = 16); the covariance table sits in LDS as
//    {x_j, f_j, slope_j, x_j+1} segments and r is located by a log2 guess plus an
//    exact fix-up (binary search when the bins are not log spaced).
//  * Right-looking Cholesky on the row-cyclic layout; the forward substitution of the
//    neighbour values runs on the same factor; the fine point's row gives mean/std.
//  * generate: workgroups take tiles from an atomic ticket in index order; a refined
//    value is published with an agent-scope store and a lane waits (polling, never
//    blocking the rest of its wavefront) only for its own neighbours' values, which
//    belong to lower tickets.  Any graph in topological order completes; one batch
//    ordering (offsets) is not needed, and the result equals the reference's
//    batch-serial sweep because every value is a deterministic function of earlier
//    values.  A bounded spin turns a malformed graph into an error, never a hang.
//  * Precision mode "increment" (FP32 only): the system is rewritten in increments
//    relative to the nearest neighbour, B_ij = g(r_i0) + g(r_0j) - g(r_ij) with the
//    variogram table g = f_0 - f, so that the entries keep their relative precision
//    when k(r) ~ k(0) (clustered points); the value is v_0 + mean(increment) + std*xi.
#include <optional>

#include "common.hpp"
#include "graphgp_amd_internal.hpp"
#include "refine_impl.hpp"
#include "dense.hpp"
#include "refine_ops.hpp"
#include "generic_slab.hpp"

namespace ggp {

namespace {

// ---------------------------------------------------------------- table prep
// Table layout: nb-1 segments {c, s} (value c + s r) followed by the nb edges (bins).
template <typename T>
__global__ void prep_table_kernel(const double* __restrict__ bins, const double* __restrict__ vals, int nb,
                                  int increment, TabE<T>* __restrict__ tab) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  T* edges = reinterpret_cast<T*>(tab + (nb - 1));
  if (j < nb) edges[j] = (T)bins[j];
  if (j >= nb - 1) return;
  const double f0 = vals[0];
  double x0 = bins[j], x1 = bins[j + 1];
  double fa = vals[j], fb = vals[j + 1];
  if (increment) {
    fa = f0 - fa;
    fb = f0 - fb;
  }
  // jnp.interp: fp[i-1] + (x - xp[i-1]) / dx * df, constant fp[i-1] when dx is ~0
  double dx = x1 - x0;
  double s = (fabs(dx) <= 4.930380657631324e-32) ? 0.0 : (fb - fa) / dx;
  TabE<T> e;
  e.s = (T)s;
  e.c = sizeof(T) == 4 ? (T)fa : (T)(fa - s * x0);
  tab[j] = e;
}

// The packed FP32 table (refine_impl.hpp, TabP): nseg segments and the constant tail.
__global__ void prep_packed_kernel(const double* __restrict__ bins, const double* __restrict__ vals, int nb,
                                   int increment, TabP* __restrict__ tab) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const double f0 = vals[0];
  TabP e;
  if (j == nb - 1) {
    e.c = (float)(increment ? f0 - vals[nb - 1] : vals[nb - 1]);
    e.i = e.c;
    e.s = 0.f;
    e.e = (float)bins[nb - 1];
  } else {
    double fa = vals[j], fb = vals[j + 1];
    if (increment) {
      fa = f0 - fa;
      fb = f0 - fb;
    }
    const double dx = bins[j + 1] - bins[j];
    const double sl = (fabs(dx) <= 4.930380657631324e-32) ? 0.0 : (fb - fa) / dx;
    e.c = (float)fa;
    e.s = (float)sl;
    e.e = (float)bins[j];
    e.i = (float)(fa - sl * (double)e.e);
  }
  tab[j] = e;
}

__global__ void prep_meta_kernel(const double* __restrict__ bins, const double* __restrict__ vals, int nb,
                                 int increment, TabMeta* __restrict__ meta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  TabMeta m;
  m.nseg = nb - 1;
  m.x_first = bins[0];
  m.x_last = bins[nb - 1];
  m.f_first = increment ? 0.0 : vals[0];
  m.f_last = increment ? vals[0] - vals[nb - 1] : vals[nb - 1];
  m.f_zero = vals[0];
  // log-spacing guess over bins[1..nb-1]
  if (nb >= 3 && bins[1] > 0.0 && bins[nb - 1] > bins[1]) {
    double l1 = log2(bins[1]), l2 = log2(bins[nb - 1]);
    m.lb1 = (float)l1;
    m.inv_step = (float)((double)(nb - 2) / (l2 - l1));
    m.x1 = bins[1];
    const double is = (double)(nb - 2) / (l2 - l1);
    m.ka = (float)(0.5 * is);
    m.kb = (float)(1.0 - l1 * is);
  } else {
    m.lb1 = 0.f;
    m.inv_step = 0.f;
    m.ka = 0.f;
    m.kb = 0.f;
    m.x1 = bins[nb > 1 ? 1 : 0];
  }
  m.log_exact = m.inv_step > 0.f ? 1 : 0;
  *meta = m;
}

// Certifies the fast lookup: for every segment j >= 1 the guess at the geometric middle
// of the segment must be j (as computed by the kernel, from r^2 in FP32).
__global__ void prep_check_kernel(const double* __restrict__ bins, int nb, TabMeta* __restrict__ meta) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (j >= nb - 1) return;
  const float lb1 = meta->lb1, inv_step = meta->inv_step;
  const double mid = sqrt(bins[j] * bins[j + 1]);
  int g = log_guess((float)(mid * mid), lb1, inv_step);
  g = g < 1 ? 1 : (g > nb - 2 ? nb - 2 : g);
  // the packed FP32 guess, at the same point
  const int gp = packed_index((float)(mid * mid), meta->ka, meta->kb, nb - 1);
  if (g != j || gp != j || !(bins[j] > 0.0)) atomicAnd(&meta->log_exact, 0);
}

}  // namespace

size_t table_workspace(int nb, int t64) { return (size_t)nb * (t64 ? 24 : 12) + 512; }

int prep_packed(const double* bins, const double* vals, int nb, int increment, TabP* tab, hipStream_t s) {
  hipLaunchKernelGGL(prep_packed_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, s, bins, vals, nb, increment, tab);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int prep_table(const double* bins, const double* vals, int nb, int t64, int increment, void* tab, TabMeta* meta,
               hipStream_t s) {
  GGP_CHECK_ARG(nb >= 2, "covariance table needs at least two bins");
  if (t64)
    hipLaunchKernelGGL(prep_table_kernel<double>, dim3(grid_for(nb, 256)), dim3(256), 0, s, bins, vals, nb,
                       increment, (TabE<double>*)tab);
  else
    hipLaunchKernelGGL(prep_table_kernel<float>, dim3(grid_for(nb, 256)), dim3(256), 0, s, bins, vals, nb,
                       increment, (TabE<float>*)tab);
  hipLaunchKernelGGL(prep_meta_kernel, dim3(1), dim3(64), 0, s, bins, vals, nb, increment, meta);
  if (nb > 2) hipLaunchKernelGGL(prep_check_kernel, dim3(grid_for(nb, 256)), dim3(256), 0, s, bins, nb, meta);
  GGP_LAUNCH_CHECK();
  return kOk;
}


// ================================================================== generic k
// Any k (and d up to 8): one lane per point, the (k+1)^2 system in a per-lane slab of
// the workspace.  Used for k + 1 > 40 (e.g. k = n0 = 200 in the dense-limit check).
namespace {

template <typename T>
__global__ __launch_bounds__(64) void refine_generic_kernel(RefineArgs<T> a, T* __restrict__ scratch) {
  const int k = a.k, m = k + 1;
  // the block's systems interleaved by lane (generic_slab.hpp): K (m x m) then u (m)
  const LaneSlab<T> S{scratch + (size_t)blockIdx.x * 64 * ((size_t)m * m + m) + threadIdx.x};
  const size_t uo = (size_t)m * m;
  auto K = [&](size_t e) -> T& { return S[e]; };
  auto u = [&](int i) -> T& { return S[uo + i]; };
  const TabMeta meta = *a.meta;
  const T x1 = (T)meta.x1, x_last = (T)meta.x_last, f_last = (T)meta.f_last;
  __shared__ uint32_t s_tile;
  const bool incr = a.increment != 0;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile >= a.ntiles) break;
    const int64_t pt = tile * 64 + threadIdx.x;
    const bool valid = pt < a.n_ref;
    const int64_t fine = a.n0 + pt;
    auto pid = [&](int i) -> int64_t { return i < k ? (int64_t)a.nbrs[pt * k + i] : fine; };
    auto coord = [&](int64_t id, int t) { return load_as<T>(a.pts, a.pts_f64, id * a.d + t); };
    auto dist = [&](int i, int j) {
      int64_t p = pid(i), q = pid(j);
      T s = T(0);
      for (int t = 0; t < a.d; ++t) {
        T dd = coord(p, t) - coord(q, t);
        s += dd * dd;
      }
      return sqrt(s);
    };
    const T* gedges = reinterpret_cast<const T*>(a.tab + meta.nseg);
    auto look = [&](T r) { return table_lookup<T>(a.tab, gedges, meta.nseg, x_last, f_last, r); };
    T sigma = T(1);
    if (valid) {
      for (int i = 0; i < m; ++i) {
        T gi = incr ? look(dist(i, 0)) : T(0);
        for (int j = 0; j <= i; ++j) {
          T v = look(dist(i, j));
          if (incr) {
            T gj = look(dist(j, 0));
            if (j == 0) v = (i == 0) ? (T)meta.f_zero : -gi;
            else v = (gi + gj) - v;
          }
          K((size_t)i * m + j) = v;
        }
      }
      // Cholesky of the whole joint system: its last pivot is sigma
      slab_cholesky(S, m, m);
      sigma = K((size_t)k * m + k);
    }
    // u[k] - L[k, :k] . L_A^{-1} u[:k]
    auto forward = [&]() -> T {
      slab_solve_lower(S, k, m, uo);
      return u(k) - slab_dot(S, (size_t)k * m, uo, k);
    };
    if (a.mode == kLogdet) {
      double ld = valid ? log((double)sigma) : 0.0;
      if (valid && ld != ld) atomicOr(&a.flags[0], 1u);
      for (int o = 32; o > 0; o >>= 1) ld += __shfl_xor(ld, o, 64);
      if (threadIdx.x == 0) a.partials[tile] = ld;
      continue;
    }
    const int64_t dst = valid ? (a.perm ? (int64_t)a.perm[fine] : fine) : 0;
    if (a.mode == kInverse) {
      if (valid) {
        for (int i = 0; i < k; ++i) u(i) = a.vals[pid(i)];
        u(k) = a.vals[fine];
        if (incr)
          for (int i = 1; i <= k; ++i) u(i) -= u(0);
        T xi = forward() / sigma;
        if (xi != xi) atomicOr(&a.flags[0], 1u);
        if (a.out_f64) ((double*)a.out)[dst - a.io_base] = (double)xi;
        else ((float*)a.out)[dst - a.io_base] = (float)xi;
      }
      continue;
    }
    const T xi = valid ? load_as<T>(a.xi, a.xi_f64, dst - a.io_base) : T(0);
    bool done = !valid;
    for (uint32_t it = 0;; ++it) {
      if (!done) {
        bool ok = true;
        for (int i = 0; i < k; ++i) {
          T v = __hip_atomic_load(&a.vals[pid(i)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = ok && !Sentinel<T>::is(v);
          u(i) = v;
        }
        if (ok) {
          u(k) = T(0);
          T v0 = T(0);
          if (incr) {
            v0 = u(0);
            for (int i = 1; i < k; ++i) u(i) -= v0;
          }
          T v = v0 - forward() + sigma * xi;
          if (v != v) {
            atomicOr(&a.flags[0], 1u);
            v = __builtin_nan("");
          }
          __hip_atomic_store(&a.vals[fine], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (a.out) {
            if (a.out_f64) ((double*)a.out)[dst] = (double)v;
            else ((float*)a.out)[dst] = (float)v;
          }
          done = true;
        }
      }
      if (__all(done)) break;
      if (it > (1u << 24)) {
        if (!done) {
          atomicOr(&a.flags[1], 1u);
          __hip_atomic_store(&a.vals[fine], (T)__builtin_nan(""), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
}

// Blocks of the generic kernel: at most 256, and at most 512 MB of slabs.
template <typename T>
int generic_blocks(int k) {
  const size_t m = (size_t)k + 1;
  const size_t per_block = (size_t)64 * (m * m + m) * sizeof(T);
  const size_t cap = (size_t)512 << 20;
  size_t b = cap / per_block;
  return (int)std::max<size_t>(1, std::min<size_t>(256, b));
}

template <typename T>
size_t generic_scratch_bytes(int k) {
  const size_t m = (size_t)k + 1;
  return (size_t)generic_blocks<T>(k) * 64 * (m * m + m) * sizeof(T);
}

// ------------------------------------------------------------------ small kernels
template <typename T>
__global__ void seed_fill_kernel(T* __restrict__ vals, const void* __restrict__ init, int init_f64, int64_t n0,
                                 int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vals[i] = i < n0 ? load_as<T>(init, init_f64, i) : Sentinel<T>::value();
}

template <typename T>
__global__ void gather_vals_kernel(const void* __restrict__ src, int src_f64, const int32_t* __restrict__ perm,
                                   int64_t n, T* __restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] = load_as<T>(src, src_f64, perm ? (int64_t)perm[i] : i);
}

__global__ void gather_seed_xi_kernel(const void* __restrict__ xi, int xi_f64, const int32_t* __restrict__ perm,
                                      int64_t n0, double* __restrict__ xs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n0) return;
  xs[i] = load_as<double>(xi, xi_f64, perm ? (int64_t)perm[i] : i);
}

template <typename T>
__global__ void seed_out_kernel(const double* __restrict__ sv, const int32_t* __restrict__ perm, int64_t n0,
                                T* __restrict__ vals, void* __restrict__ out, int out_f64, uint32_t* flags) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n0) return;
  const double v = sv[i];
  if (v != v) atomicOr(&flags[0], 2u);  // bit 2: the seed block failed (bit 1: a refined value)
  if (vals) vals[i] = (T)v;
  if (out) {
    const int64_t dst = perm ? (int64_t)perm[i] : i;
    if (out_f64) ((double*)out)[dst] = v;
    else ((float*)out)[dst] = (float)v;
  }
}

template <typename T>
__global__ void seed_vals_to_f64_kernel(const T* __restrict__ vals, int64_t n0, double* __restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n0) v[i] = (double)vals[i];
}

// FP64 fallback of the two-phase generate.  When more points need the FP64 recomputation
// than its list holds (error bit 2, set by the weights phase), the FP32-increment values
// are not within their bound, so the whole refinement is recomputed in FP64 (the
// precision="fp64" kernel) into a value buffer of its own, and its values, rounded, are the
// output.  Each kernel of the fallback returns at once unless the bit is set.
__global__ void fallback_fill_kernel(double* __restrict__ vals64, const double* __restrict__ seeds64,
                                     const void* __restrict__ init, int init_f64, int64_t n0, int64_t n,
                                     uint32_t* __restrict__ flags) {
  if ((flags[1] & 2u) == 0) return;
  const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (first == 0) atomicAnd(&flags[0], ~1u);  // refined values are recomputed; seed failure stays
  for (int64_t i = first; i < n; i += (int64_t)gridDim.x * blockDim.x)
    vals64[i] = i < n0 ? (seeds64 ? seeds64[i] : load_as<double>(init, init_f64, i)) : Sentinel<double>::value();
}

template <typename T>
__global__ void fallback_copy_kernel(const double* __restrict__ vals64, T* __restrict__ vals, int64_t n0, int64_t n,
                                     const uint32_t* __restrict__ flags) {
  if ((flags[1] & 2u) == 0) return;
  for (int64_t i = n0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    vals[i] = (T)vals64[i];
}

__global__ void nan_fill_kernel(void* __restrict__ out, int out_f64, int64_t n, const uint32_t* __restrict__ flags) {
  if (flags[0] == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (out_f64) ((double*)out)[i] = __builtin_nan("");
    else ((float*)out)[i] = __builtin_nanf("");
  }
}

__global__ void logdet_sum_kernel(const double* __restrict__ partials, int64_t np, const double* __restrict__ seed,
                                  double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += 1024) s += partials[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += red[w];
    *out = seed ? t + *seed : t;
  }
}

constexpr double kDefaultFixTau = 1e-3;

// Pivot threshold of the FP64 recomputation: a point is recomputed when a Cholesky pivot
// (or the fine point's Schur complement) falls below tau times its original diagonal,
// i.e. when FP32 would lose more than ~log10(1/tau) digits.  GGP_FIX_TAU overrides it.
double fix_tau() {
  const char* e = getenv("GGP_FIX_TAU");
  return e ? atof(e) : kDefaultFixTau;
}

size_t max_lds() {
  static size_t lds = 0;
  if (lds == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      lds = (size_t)v;
    else
      lds = 65536;
  }
  return lds;
}

// workspace layout shared by all refine flows
template <typename T>
struct RefineWs {
  TabE<T>* tab = nullptr;
  TabP* tabp = nullptr;  // FP32: the packed table of the register kernel
  TabMeta* meta = nullptr;
  TabE<double>* tab64 = nullptr;  // FP64 direct table for the FP64 recomputation (FP32 increment mode)
  TabMeta* meta64 = nullptr;
  uint32_t* ctr = nullptr;  // [0] ticket, [1] NaN flag, [2] error flag, [3] FP64-recomputed points
  uint32_t* gtk = nullptr;  // the dataflow kernel's group tickets, one cache line each
  double* partials = nullptr;
  T* vals = nullptr;  // graph-order values (N)
  double* dense = nullptr;
  double* dvec = nullptr;  // n0
  double* dout = nullptr;  // n0 (+2 scalars)
  T* scratch = nullptr;
  // two-phase generate (FP32 increment mode)
  T* wts = nullptr;
  T* sig = nullptr;
  double* vals64 = nullptr;  // the FP64 fallback's values
  int32_t* fixslot = nullptr;
  int32_t* dpos = nullptr;
  uint32_t* fix_list = nullptr;
  uint32_t fix_cap = 0;
  double* w64 = nullptr;
  double* sig64 = nullptr;
};

inline bool generic_config(const RefineProblem& p) { return p.k + 1 > 40 || p.d > 8 || (p.d > 3 && p.k + 1 > 16); }

// generate in the FP32 increment mode runs in three phases (weights, FP64 recomputation of
// the flagged points, dataflow over the stored weights)
template <typename T>
bool two_phase(const RefineProblem& p, int mode) {
  return sizeof(T) == 4 && p.increment && mode == kGenerate && !generic_config(p) && p.n_ref > 0;
}

// list slots for the FP64 recomputation: a quarter of the points (the workloads measured here
// flag about 1.4%); beyond that error bit 2 is set and the FP64 fallback recomputes everything
inline uint32_t fix_capacity(int64_t n_ref) { return (uint32_t)std::max<int64_t>(4096, n_ref / 4); }

template <typename T>
bool carve(Workspace& w, const RefineProblem& p, int mode, RefineWs<T>& b) {
  // counters, first, so ggp_refine_counters finds them at offset 0: [0] ticket, [1] and [2]
  // flags, [3] FP64-recomputed count, then the tickets of the two-phase generate's FP64 list
  // [4] and FP64 fallback [5]; the dataflow's group tickets follow, so that one fill at the
  // start clears them all
  b.ctr = w.take<uint32_t>(8);
  if (two_phase<T>(p, mode)) b.gtk = w.take<uint32_t>(kMaxGroups * 32);
  b.tab = w.take<TabE<T>>(p.nb);  // nb-1 segments + nb edges (of T) fit in nb entries + nb/2
  w.take<T>(p.nb + 2);
  if (sizeof(T) == 4) b.tabp = w.take<TabP>(p.nb);
  b.meta = w.take<TabMeta>(1);
  if (sizeof(T) == 4 && p.increment) {
    b.tab64 = w.take<TabE<double>>(p.nb);
    w.take<double>(p.nb + 2);
    b.meta64 = w.take<TabMeta>(1);
  }
  const int64_t n0 = p.n - p.n_ref;
  b.partials = w.take<double>((p.n_ref + 31) / 32 + 1);
  if (mode != kLogdet) b.vals = w.take<T>(p.n);  // the log-det reads no values
  if (p.with_seed) {
    b.dense = w.take<double>((size_t)n0 * n0 + 4096);
    b.dvec = w.take<double>(n0 + 1);
    b.dout = w.take<double>(n0 + 4);
  }
  if (generic_config(p)) b.scratch = w.take<T>(generic_scratch_bytes<T>(p.k) / sizeof(T));
  if (two_phase<T>(p, mode)) {
    b.fix_cap = (uint32_t)std::min<int64_t>(fix_capacity(p.n_ref), p.n_ref);
    b.wts = w.take<T>((size_t)p.n_ref * p.k);
    b.sig = w.take<T>(p.n_ref);
    b.fixslot = w.take<int32_t>(p.n_ref);
    b.dpos = w.take<int32_t>(p.n_ref);
    b.fix_list = w.take<uint32_t>(b.fix_cap);
    b.w64 = w.take<double>((size_t)b.fix_cap * p.k);
    b.sig64 = w.take<double>(b.fix_cap);
    b.vals64 = w.take<double>(p.n);
  }
  return w.ok();
}

}  // namespace

// Optional events recorded around the refinement kernel (profiling hook, see
// ggp_profile_events); null when unused.
thread_local hipEvent_t g_prof_start = nullptr, g_prof_stop = nullptr;

namespace {

template <typename T>
int run_kernel_inner(RefineArgs<T> a, const RefineProblem& p, RefineWs<T>& b, hipStream_t s);

template <typename T>
int run_kernel(RefineArgs<T> a, const RefineProblem& p, RefineWs<T>& b, hipStream_t s) {
  if (g_prof_start) GGP_HIP(hipEventRecord(g_prof_start, s));
  int rc = run_kernel_inner<T>(a, p, b, s);
  if (g_prof_stop) GGP_HIP(hipEventRecord(g_prof_stop, s));
  return rc;
}

// Problem and workspace fields of the kernel arguments (the caller sets mode, values, I/O).
template <typename T>
void setup_args(RefineArgs<T>& a, const RefineProblem& p, RefineWs<T>& b) {
  a.pts = p.pts;
  a.pts_f64 = p.pts_f64;
  a.d = p.d;
  a.nbrs = p.nbrs;
  a.k = p.k;
  a.n0 = p.n - p.n_ref;
  a.n_ref = p.n_ref;
  a.tab = b.tab;
  a.tabp = b.tabp;
  a.meta = b.meta;
  a.ticket = b.ctr;
  a.flags = b.ctr + 1;
  a.partials = b.partials;
  a.increment = p.increment;
  a.tab64 = b.tab64;
  a.meta64 = b.meta64;
  a.fix_count = b.ctr + 3;
  a.fix_tau = fix_tau();
  a.wts = b.wts;
  a.sig = b.sig;
  a.fixslot = b.fixslot;
  a.dpos = b.dpos;
  a.fix_list = b.fix_list;
  a.fix_cap = b.fix_cap;
  a.w64 = b.w64;
  a.sig64 = b.sig64;
}

template <typename T>
int run_kernel_inner(RefineArgs<T> a, const RefineProblem& p, RefineWs<T>& b, hipStream_t s) {
  setup_args<T>(a, p, b);
  if (p.n_ref == 0) return kOk;
  const int kk = p.k + 1;
  const bool generic = kk > 40 || p.d > 8 || (p.d > 3 && kk > 16);
  if (generic) {
    a.ntiles = (p.n_ref + 63) / 64;
    const int grid = (int)std::min<int64_t>(a.ntiles, generic_blocks<T>(p.k));
    hipLaunchKernelGGL(refine_generic_kernel<T>, dim3(grid), dim3(64), 0, s, a, b.scratch);
    GGP_LAUNCH_CHECK();
    return kOk;
  }
  // LDS: FP32 the packed table (nb entries), FP64 the segments and the edges
  const size_t lds = sizeof(T) == 4 ? (size_t)p.nb * sizeof(TabP)
                                    : (size_t)(p.nb - 1) * sizeof(TabE<T>) + (size_t)p.nb * sizeof(T);
  a.fix_lds_offset = (int)((lds + 15) & ~size_t(15));
  const int grid = device_cus() * 4;
  int rc = launch_refine<T>(a, kk, s, lds, grid);
  if (rc == kUnsupported && ggp::last_error()[0] == 0) set_error("refine: unsupported configuration");
  return rc;
}

// A second stream per host thread and device, for the seed block that runs beside the
// weights phase, and the events that fork it from and join it to the caller's stream.
// Default priority, non-blocking; created on first use and kept for the thread's lifetime.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

int side_stream(SideStream** out) {
  thread_local SideStream side[16];
  int dev = 0;
  GGP_HIP(hipGetDevice(&dev));
  SideStream& x = side[dev & 15];
  if (x.s == nullptr) {
    GGP_HIP(hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking));
    GGP_HIP(hipEventCreateWithFlags(&x.fork, hipEventDisableTiming));
    GGP_HIP(hipEventCreateWithFlags(&x.join, hipEventDisableTiming));
  }
  *out = &x;
  return kOk;
}

// Refined points from which the seed block of the two-phase generate runs beside the
// weights phase (below 600k points the weights phase is shorter than what it would hide;
// at 800k points, k = 30, it saves 0.23 ms, but the stream is shared by the thread's calls,
// and a served request of that size then queued behind a large request's work on it: p99
// latency of requests up to 1e6 points 32 -> 110 ms), and
// the workgroup slots the weights phase leaves free for it meanwhile: one per eight CUs (32
// on MI355X; at 16.8M points the seed block then takes 1.4 ms beside the weights phase,
// with 24 free slots 12 ms).  GGP_SEED_RESERVE overrides it; 0 runs the seed block before
// the weights phase.
constexpr int64_t kSeedBesideMinRefined = int64_t(1) << 20;
int seed_reserve_wgs() {
  const char* e = getenv("GGP_SEED_RESERVE");
  return e ? atoi(e) : std::max(1, device_cus() / 8);
}

// The three phases of generate in the FP32 increment mode: the weights kernel, the FP64
// recomputation of the listed points, and the dataflow over the stored weights.  a.vals
// already holds the seeds and sentinels, or the seed block is running on `side`: then the
// weights kernel leaves seed_reserve_wgs() workgroup slots free for it, a second launch of
// the weights kernel with that many workgroups follows the seed block on the side stream
// and takes the tiles left by then from the same ticket, and the FP64 list waits for both.
template <typename T>
int run_two_phase(RefineArgs<T> a, const RefineProblem& p, RefineWs<T>& b, hipStream_t s,
                  SideStream* side = nullptr) {
  if (g_prof_start) GGP_HIP(hipEventRecord(g_prof_start, s));
  setup_args<T>(a, p, b);  // the later phases launch with this copy of the arguments
  std::optional<StageScope> stage;
  stage.emplace("weights", s);
  // with a schedule, the weights and FP64 phases work in slot space (points in slot order,
  // neighbours as value slots), so their weights, sigma and list are slot-indexed
  const void* w_pts = p.sched ? p.sched_pts : p.pts;
  const int32_t* w_nbrs = p.sched ? p.sched_nbrs : p.nbrs;
  RefineProblem pw = p;  // run_kernel_inner sets the kernel's inputs from its problem
  pw.pts = w_pts;
  pw.nbrs = w_nbrs;
  // the weights phases store sigma xi per slot (xi through the schedule and the permutation)
  a.sched = p.sched;
  a.mode = kWeights;
  if (side != nullptr) {
    a.reserve_wgs = seed_reserve_wgs();
    GGP_TRY(run_kernel_inner<T>(a, pw, b, s));
    a.reserve_wgs = 0;
    a.grid_limit = seed_reserve_wgs();
    GGP_TRY(run_kernel_inner<T>(a, pw, b, side->s));
    a.grid_limit = 0;
    GGP_HIP(hipEventRecord(side->join, side->s));
    GGP_HIP(hipStreamWaitEvent(s, side->join, 0));  // all weights tiles, and the seeds
  } else {
    GGP_TRY(run_kernel_inner<T>(a, pw, b, s));
  }
  // the profiling events bracket the dominant kernel: the weights phase
  if (g_prof_stop) GGP_HIP(hipEventRecord(g_prof_stop, s));
  stage.emplace("fp64_list", s);
  // FP64 recomputation of the listed points with the FP64 register kernel (weights mode over
  // the list, FP64 direct table)
  {
    RefineArgs<double> d{};
    d.pts = w_pts;
    d.pts_f64 = p.pts_f64;
    d.d = p.d;
    d.nbrs = w_nbrs;
    d.k = p.k;
    d.n0 = p.n - p.n_ref;
    d.n_ref = p.n_ref;
    d.tab = b.tab64;
    d.meta = b.meta64;
    d.ticket = b.ctr + 4;  // cleared at the start
    d.flags = b.ctr + 1;
    d.mode = kWeights;
    d.increment = 0;
    d.fix_count = b.ctr + 3;
    d.fix_cap = b.fix_cap;
    d.w64 = b.w64;
    d.sig64 = b.sig64;
    d.list_in = b.fix_list;
    d.sched = p.sched;
    d.perm = p.perm;
    d.xi = a.xi;
    d.xi_f64 = a.xi_f64;
    d.io_base = a.io_base;
    const size_t lds64 = (size_t)(p.nb - 1) * sizeof(TabE<double>) + (size_t)p.nb * sizeof(double);
    d.fix_lds_offset = (int)((lds64 + 15) & ~size_t(15));
    GGP_TRY(launch_refine<double>(d, p.k + 1, s, lds64, device_cus() * 4));
  }
  const int cus = device_cus();
  stage.emplace("dataflow", s);
  a.mode = kGenerate;
  a.pts = p.pts;
  a.nbrs = p.nbrs;
  a.nbv = p.sched_nbrs;
  // Lanes per point and rows per lane: four lanes up to k = 16 (2 or 4 rows), eight lanes
  // beyond (4 or 5 rows, k <= 40 here), with four and six workgroups per CU.  At 16.8M
  // points, k = 30 (ms): four lanes with 3/4/6/8 workgroups per CU 3.60/3.37/3.45/3.45,
  // eight lanes 3.58/-/3.22/3.22, sixteen lanes 4.49/-/3.73/3.58; the poll interval made no
  // difference.
  const bool wide = p.k > 16;
  const int lanes = wide ? 8 : 4;
  a.ntiles = (p.n_ref + 256 / lanes - 1) / (256 / lanes);
  auto kern = !wide ? (p.k <= 8 ? apply_kernel<T, 4, 2> : apply_kernel<T, 4, 4>)
                    : (p.k <= 32 ? apply_kernel<T, 8, 4> : apply_kernel<T, 8, 5>);
  const int grid = (int)std::min<int64_t>(a.ntiles, persistent_grid((const void*)kern, cus * (wide ? 6 : 4), 0));
  a.yield = yield_cfg().dev;
  a.groups = std::min(grid, kMaxGroups);
  // every group keeps a workgroup that does not yield (the first `keep` ones do not)
  a.keep = std::max<uint32_t>(yield_keep(grid), (uint32_t)a.groups);
  a.group_ticket = b.gtk;  // cleared at the start
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, a);
  GGP_LAUNCH_CHECK();
  return kOk;
}

// The FP64 fallback of the two-phase generate (see fallback_fill_kernel): three launches
// that return at once unless the FP64 list overflowed.
template <typename T>
int run_fallback(const RefineArgs<T>& a, const RefineProblem& p, const RefineIO& io, RefineWs<T>& b, hipStream_t s) {
  StageScope stage("fp64_fallback", s);
  const int64_t n0 = p.n - p.n_ref;
  hipLaunchKernelGGL(fallback_fill_kernel, dim3(grid_for(std::min<int64_t>(p.n, 1 << 20), 256)), dim3(256), 0, s,
                     b.vals64, p.with_seed ? (const double*)b.dout : nullptr, io.init, io.init_f64, n0, p.n,
                     b.ctr + 1);
  GGP_LAUNCH_CHECK();
  RefineArgs<double> f{};
  f.pts = p.sched ? p.sched_pts : p.pts;
  f.pts_f64 = p.pts_f64;
  f.d = p.d;
  f.nbrs = p.sched ? p.sched_nbrs : p.nbrs;
  f.k = p.k;
  f.n0 = n0;
  f.n_ref = p.n_ref;
  f.perm = p.perm;
  f.xi = io.xi;
  f.xi_f64 = io.xi_f64;
  f.vals = b.vals64;
  f.out = a.out;  // null for the raw refine, whose graph-order values are copied below
  f.out_f64 = io.out_f64;
  f.io_base = a.io_base;
  f.tab = b.tab64;
  f.meta = b.meta64;
  f.ticket = b.ctr + 5;  // cleared at the start
  f.flags = b.ctr + 1;
  f.mode = kGenerate;
  f.increment = 0;
  f.sched = p.sched;
  f.only_if = b.ctr + 2;
  const size_t lds64 = (size_t)(p.nb - 1) * sizeof(TabE<double>) + (size_t)p.nb * sizeof(double);
  f.fix_lds_offset = (int)((lds64 + 15) & ~size_t(15));
  GGP_TRY(launch_refine<double>(f, p.k + 1, s, lds64, device_cus() * 4));
  if (!p.with_seed) {
    hipLaunchKernelGGL(fallback_copy_kernel<T>, dim3(grid_for(std::min<int64_t>(p.n_ref, 1 << 20), 256)), dim3(256),
                       0, s, b.vals64, static_cast<T*>(io.out), n0, p.n, b.ctr + 1);
    GGP_LAUNCH_CHECK();
  }
  return kOk;
}

template <typename T>
int refine_flow(const RefineProblem& p, int mode, const RefineIO& io, void* ws, size_t ws_bytes, hipStream_t s) {
  GGP_CHECK_ARG(p.n >= 1 && p.n_ref >= 0 && p.n_ref <= p.n && p.k >= 1 && p.d >= 1, "refine: bad sizes");
  GGP_CHECK_ARG(p.nb >= 2, "refine: covariance table needs at least two bins");
  GGP_CHECK_ARG(p.n < (int64_t(1) << 31), "refine: N must be below 2^31");
  GGP_CHECK_ARG(!(p.increment && sizeof(T) == 8), "refine: the increment formulation is an FP32 mode");
  Workspace w(ws, ws_bytes);
  RefineWs<T> b;
  if (!carve<T>(w, p, mode, b)) {
    set_error("refine: workspace too small");
    return kWorkspaceTooSmall;
  }
  const int64_t n0 = p.n - p.n_ref;
  // the seed block of a large two-phase generate runs on the side stream beside the weights
  // phase (which needs no seed), and so do the value buffer's fill and the FP64 table, which
  // only the later phases read; the dataflow waits for them
  SideStream* side = nullptr;
  if (mode == kGenerate && p.with_seed && two_phase<T>(p, mode) && p.n_ref >= kSeedBesideMinRefined &&
      cu_budget() == 0 && yield_cfg().dev == nullptr && seed_reserve_wgs() > 0)
    GGP_TRY(side_stream(&side));
  {
    StageScope stage("tables", s);
    // the counters and the dataflow's group tickets (see carve)
    const size_t nctr = b.gtk != nullptr ? (size_t)(b.gtk + kMaxGroups * 32 - b.ctr) : 8;
    GGP_TRY(fill_u32(b.ctr, 0u, nctr, s));
    GGP_TRY(prep_table(p.bins, p.vals, p.nb, sizeof(T) == 8, p.increment, b.tab, b.meta, s));
    if (b.tabp) GGP_TRY(prep_packed(p.bins, p.vals, p.nb, p.increment, b.tabp, s));
    if (b.tab64 && side == nullptr) GGP_TRY(prep_table(p.bins, p.vals, p.nb, 1, 0, b.tab64, b.meta64, s));
  }
  RefineArgs<T> a{};
  a.mode = mode;
  a.perm = p.perm;
  if (mode == kGenerate) {
    // graph-order buffer: seeds (given or from the dense block) then sentinels
    if (p.with_seed) {
      hipStream_t ss = s;
      if (side != nullptr) {
        GGP_HIP(hipEventRecord(side->fork, s));
        GGP_HIP(hipStreamWaitEvent(side->s, side->fork, 0));
        ss = side->s;
        GGP_TRY(prep_table(p.bins, p.vals, p.nb, 1, 0, b.tab64, b.meta64, ss));
      }
      hipLaunchKernelGGL(seed_fill_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, ss, b.vals, (const void*)nullptr,
                         0, (int64_t)0, p.n);
      GGP_LAUNCH_CHECK();
      StageScope stage("seed_block", ss);
      GGP_TRY(dense_factor(p.pts, p.pts_f64, p.d, n0, p.bins, p.vals, p.nb, b.dense, ss));
      hipLaunchKernelGGL(gather_seed_xi_kernel, dim3(grid_for(n0, 256)), dim3(256), 0, ss, io.xi, io.xi_f64, p.perm,
                         n0, b.dvec);
      GGP_LAUNCH_CHECK();
      GGP_TRY(dense_apply(b.dense, n0, b.dvec, b.dout, ss));
      hipLaunchKernelGGL(seed_out_kernel<T>, dim3(grid_for(n0, 256)), dim3(256), 0, ss, b.dout, p.perm, n0, b.vals,
                         io.out, io.out_f64, b.ctr + 1);
      GGP_LAUNCH_CHECK();
      a.vals = b.vals;
      a.out = io.out;
      a.io_base = 0;
    } else {
      T* vals = static_cast<T*>(io.out);  // raw refine: output is the graph-order values array
      hipLaunchKernelGGL(seed_fill_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, s, vals, io.init, io.init_f64,
                         n0, p.n);
      GGP_LAUNCH_CHECK();
      a.vals = vals;
      a.out = nullptr;
      a.io_base = n0;
    }
    a.xi = io.xi;
    a.xi_f64 = io.xi_f64;
    a.out_f64 = io.out_f64;
    if (two_phase<T>(p, mode)) {
      GGP_TRY(run_two_phase<T>(a, p, b, s, side));
      GGP_TRY(run_fallback<T>(a, p, io, b, s));
    } else if (p.sched != nullptr && !generic_config(p)) {
      // the fused dataflow in slot space: points and neighbours in schedule order, values
      // stored in slot order, xi and output through the schedule
      RefineProblem pw = p;
      pw.pts = p.sched_pts;
      pw.nbrs = p.sched_nbrs;
      a.sched = p.sched;
      StageScope stage("refine", s);
      GGP_TRY(run_kernel<T>(a, pw, b, s));
    } else {
      StageScope stage("refine", s);
      GGP_TRY(run_kernel<T>(a, p, b, s));
    }
    if (p.with_seed) {
      StageScope stage("nan_fill", s);
      hipLaunchKernelGGL(nan_fill_kernel, dim3(grid_for(std::min<int64_t>(p.n, 1 << 20), 256)), dim3(256), 0, s,
                         io.out, io.out_f64, p.n, b.ctr + 1);
      GGP_LAUNCH_CHECK();
    }
    return kOk;
  }
  if (mode == kInverse) {
    hipLaunchKernelGGL(gather_vals_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, s, io.values, io.values_f64,
                       p.perm, p.n, b.vals);
    GGP_LAUNCH_CHECK();
    a.vals = b.vals;
    a.out = io.out;
    a.out_f64 = io.out_f64;
    a.io_base = p.with_seed ? 0 : n0;
    {
      StageScope stage("refine", s);
      GGP_TRY(run_kernel<T>(a, p, b, s));
    }
    if (p.with_seed) {
      StageScope stage("seed_block", s);
      GGP_TRY(dense_factor(p.pts, p.pts_f64, p.d, n0, p.bins, p.vals, p.nb, b.dense, s));
      hipLaunchKernelGGL(seed_vals_to_f64_kernel<T>, dim3(grid_for(n0, 256)), dim3(256), 0, s, b.vals, n0, b.dvec);
      GGP_LAUNCH_CHECK();
      GGP_TRY(dense_solve(b.dense, n0, b.dvec, s));
      hipLaunchKernelGGL(seed_out_kernel<T>, dim3(grid_for(n0, 256)), dim3(256), 0, s, b.dvec, p.perm, n0,
                         (T*)nullptr, io.out, io.out_f64, b.ctr + 1);
      GGP_LAUNCH_CHECK();
      hipLaunchKernelGGL(nan_fill_kernel, dim3(grid_for(std::min<int64_t>(p.n, 1 << 20), 256)), dim3(256), 0, s,
                         io.out, io.out_f64, p.n, b.ctr + 1);
      GGP_LAUNCH_CHECK();
    } else if (io.init) {
      // raw refine_inv: the initial values are the first n0 values
      hipLaunchKernelGGL(seed_vals_to_f64_kernel<T>, dim3(grid_for(n0, 256)), dim3(256), 0, s, b.vals, n0,
                         (double*)io.init);
      GGP_LAUNCH_CHECK();
    }
    return kOk;
  }
  // log-det
  a.vals = nullptr;
  {
    StageScope stage("refine", s);
    GGP_TRY(run_kernel<T>(a, p, b, s));
  }
  const int64_t np = p.n_ref == 0 ? 0 : (p.k + 1 > 40 || p.d > 8 || (p.d > 3 && p.k + 1 > 16)) ? (p.n_ref + 63) / 64 : -1;
  int64_t ntiles = np;
  if (ntiles < 0) {
    // tiles of the register kernel: points per tile = 256 / P for the chosen configuration
    ntiles = (p.n_ref + refine_points_per_tile(sizeof(T) == 8, p.k + 1, p.d) - 1) /
             refine_points_per_tile(sizeof(T) == 8, p.k + 1, p.d);
  }
  const double* seed = nullptr;
  if (p.with_seed && n0 > 0) {
    StageScope stage("seed_block", s);
    GGP_TRY(dense_factor(p.pts, p.pts_f64, p.d, n0, p.bins, p.vals, p.nb, b.dense, s));
    GGP_TRY(dense_logdiag(b.dense, n0, b.dout, s));
    GGP_TRY(dense_logdet_fallback(p.pts, p.pts_f64, p.d, n0, p.bins, p.vals, p.nb, b.dense, b.dout, s));
    seed = b.dout;
  }
  StageScope stage("logdet_sum", s);
  hipLaunchKernelGGL(logdet_sum_kernel, dim3(1), dim3(1024), 0, s, b.partials, ntiles, seed, (double*)io.out);
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace

size_t refine_lds_limit() { return max_lds(); }

int refine_points_per_tile(int t64, int kk, int d) {
  if (kk > 40 || d > 8 || (d > 3 && kk > 16)) return 64;
  if (!t64) {
    if (d > 3) return kk <= 8 ? 256 : 128;
    if (kk <= 12) return 256;
    if (kk <= 20) return 128;
    if (kk <= 32) return 64;
    return 32;
  }
  if (d > 3) return kk <= 8 ? 128 : 64;
  if (kk <= 4) return 256;
  if (kk <= 12) return 128;
  if (kk <= 24) return 64;
  return 32;
}

size_t refine_workspace(const RefineProblem& p, int mode, int t64) {
  Workspace w(nullptr, 0);
  if (t64) {
    RefineWs<double> b;
    carve<double>(w, p, mode, b);
  } else {
    RefineWs<float> b;
    carve<float>(w, p, mode, b);
  }
  return w.used + 256;
}

int refine_run(const RefineProblem& p, int mode, int t64, const RefineIO& io, void* ws, size_t ws_bytes,
               hipStream_t s) {
  return t64 ? refine_flow<double>(p, mode, io, ws, ws_bytes, s) : refine_flow<float>(p, mode, io, ws, ws_bytes, s);
}

// Dense-only flows (generate_dense / generate_dense_inv / generate_dense_logdet).
size_t dense_flow_workspace(int64_t n0) {
  Workspace w(nullptr, 0);
  w.take<double>((size_t)n0 * n0 + 4096);
  w.take<double>(n0 + 1);
  w.take<double>(n0 + 4);
  w.take<uint32_t>(4);
  return w.used + 256;
}

int dense_flow(const void* pts, int pts_f64, int d, int64_t n0, const double* bins, const double* vals, int nb,
               int mode, const void* in, int in_f64, void* out, int out_f64, void* ws, size_t ws_bytes,
               hipStream_t s) {
  GGP_CHECK_ARG(n0 >= 1 && d >= 1 && nb >= 2, "dense: bad sizes");
  Workspace w(ws, ws_bytes);
  double* A = w.take<double>((size_t)n0 * n0 + 4096);
  double* v = w.take<double>(n0 + 1);
  double* o = w.take<double>(n0 + 4);
  uint32_t* flags = w.take<uint32_t>(4);
  if (!w.ok()) {
    set_error("dense: workspace too small");
    return kWorkspaceTooSmall;
  }
  GGP_TRY(fill_u32(flags, 0u, 4, s));
  GGP_TRY(dense_factor(pts, pts_f64, d, n0, bins, vals, nb, A, s));
  if (mode == kLogdet) {
    GGP_TRY(dense_logdiag(A, n0, (double*)out, s));
    return dense_logdet_fallback(pts, pts_f64, d, n0, bins, vals, nb, A, (double*)out, s);
  }
  hipLaunchKernelGGL(gather_seed_xi_kernel, dim3(grid_for(n0, 256)), dim3(256), 0, s, in, in_f64,
                     (const int32_t*)nullptr, n0, v);
  GGP_LAUNCH_CHECK();
  if (mode == kGenerate) {
    GGP_TRY(dense_apply(A, n0, v, o, s));
  } else {
    GGP_TRY(dense_solve(A, n0, v, s));
    o = v;
  }
  hipLaunchKernelGGL(seed_out_kernel<double>, dim3(grid_for(n0, 256)), dim3(256), 0, s, o, (const int32_t*)nullptr,
                     n0, (double*)nullptr, out, out_f64, flags);
  GGP_LAUNCH_CHECK();
  // a failed factorisation (any NaN) makes the whole output NaN, as jnp.linalg.cholesky's
  // all-NaN factor does in the reference (refine.py:74, :193)
  hipLaunchKernelGGL(nan_fill_kernel, dim3(grid_for(std::min<int64_t>(n0, 1 << 20), 256)), dim3(256), 0, s, out,
                     out_f64, n0, flags);
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace ggp
