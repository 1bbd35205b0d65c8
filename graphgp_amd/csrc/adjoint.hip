// adjoint.hip -- vector-Jacobian products of generate, generate_inv and generate_logdet
// with respect to their inputs and the covariance table (the reference differentiates
// them with JAX autodiff, refine.py:21-56, 162-183, 234-240; README.md:36 mentions custom
// primitives in its CUDA extension).
//
// Refined points: adjoint_kernel (adjoint_impl.hpp) for d <= 3 and k + 1 <= 32, else
// adjoint_generic_kernel (one lane per point, its system in a slab of the workspace).
// generate is one reverse dataflow launch; the inverse and the log-det have no chain.
// Seed block (values = L xi, L = chol(K) over the n0 seeds), in FP64:
//   generate: xibar = L^T vbar.  With Lbar = tril(vbar xi^T), Phi(L^T Lbar) is the lower
//   triangle of u xi^T with its diagonal halved (u = L^T vbar).  The Cholesky adjoint
//   (Murray 2016, "Differentiation of the Cholesky decomposition") is
//   S = L^{-T} Phi(L^T Lbar) L^{-1}: the derivative with respect to the shared entry
//   K_ab = K_ba is S_ab + S_ba (S_aa on the diagonal).  With X = L^{-1} and
//   T[a][c] = u_a (sum_{b<a} xi_b X[b][c] + xi_a X[a][c] / 2), S = X^T T, and the lookup of
//   every ordered pair (a, b) feeds S[a][b] to the table.
//   inverse (xi_s = L^{-1} v_s, cotangent t): vbar_s += L^{-T} t = X^T t, and
//   Lbar = -tril((X^T t) xi_s^T), so u = -t with xi_s in place of xi above.
//   log-det (log|det K| / 2, cotangent l): S = (l / 2) K^{-1} = (l / 2) X^T X.
#include "adjoint_impl.hpp"
#include "dense.hpp"
#include "generic_slab.hpp"
#include "graphgp_amd_internal.hpp"
#include "refine_ops.hpp"

namespace ggp {

namespace {

template <typename T>
__global__ void gather_kernel(const void* __restrict__ src, int src_f64, const int32_t* __restrict__ perm, int64_t n,
                              T* __restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = load_as<T>(src, src_f64, perm ? (int64_t)perm[i] : i);
}

__global__ void count_children_kernel(const int32_t* __restrict__ nbrs, int64_t m, uint32_t* __restrict__ cnt) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[nbrs[e]], 1u);
}

template <typename T>
__global__ void seed_vbar_kernel(const T* __restrict__ vbar, const void* __restrict__ xi, int xi_f64,
                                 const int32_t* __restrict__ perm, int64_t n0, double* __restrict__ vb,
                                 double* __restrict__ xs) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n0) return;
  vb[i] = (double)vbar[i];
  xs[i] = load_as<double>(xi, xi_f64, perm ? (int64_t)perm[i] : i);
}

// u = L^T vb; one wavefront per column
__global__ __launch_bounds__(256) void trmv_t_kernel(const double* __restrict__ L, int64_t n,
                                                     const double* __restrict__ vb, double* __restrict__ u,
                                                     const int32_t* __restrict__ perm, void* __restrict__ gxi,
                                                     int gxi_f64) {
  const int64_t a = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (a >= n) return;
  double s = 0.0;
  for (int64_t i = a + lane; i < n; i += 64) s += L[i * n + a] * vb[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) {
    u[a] = s;
    const int64_t dst = perm ? (int64_t)perm[a] : a;
    if (gxi_f64) ((double*)gxi)[dst] = s;
    else ((float*)gxi)[dst] = (float)s;
  }
}

// X = L^{-1}; one wavefront per column c, the column held in LDS:
// x_i = (delta_ic - sum_{m=c}^{i-1} L[i][m] x_m) / L[i][i], the sum split over the 64 lanes
__global__ __launch_bounds__(64) void tri_inv_kernel(const double* __restrict__ L, int64_t n, double* __restrict__ X) {
  extern __shared__ double xc[];
  const int64_t c = blockIdx.x;
  const int lane = threadIdx.x;
  for (int64_t i = lane; i < c; i += 64) X[i * n + c] = 0.0;
  for (int64_t i = c; i < n; ++i) {
    const double* row = L + i * n;
    double s = 0.0;
    for (int64_t m = c + lane; m < i; m += 64) s += row[m] * xc[m - c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    const double xi = ((i == c ? 1.0 : 0.0) - s) / row[i];
    if (lane == 0) {
      xc[i - c] = xi;
      X[i * n + c] = xi;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// T[a][c] = u_a (sum_{b<a} xi_b X[b][c] + xi_a X[a][c] / 2); one thread per column
__global__ void phi_kernel(const double* __restrict__ X, int64_t n, const double* __restrict__ u,
                           const double* __restrict__ xs, double* __restrict__ Tm) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double run = 0.0;
  for (int64_t a = 0; a < n; ++a) {
    const double x = X[a * n + c];
    Tm[a * n + c] = u[a] * (run + 0.5 * xs[a] * x);
    run += xs[a] * x;
  }
}

// S = X^T Tm (64 x 64 tiles)
__global__ __launch_bounds__(256) void gemm_tn_kernel(const double* __restrict__ X, const double* __restrict__ Tm,
                                                      int64_t n, double* __restrict__ S) {
  __shared__ double Xs[64][65];
  __shared__ double Ts[64][65];
  const int64_t d0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  const int t = threadIdx.x;
  const int tr = (t / 16) * 4, tc = (t % 16) * 4;
  double acc[4][4] = {};
  for (int64_t a0 = 0; a0 < n; a0 += 64) {
    for (int q = t; q < 64 * 64; q += 256) {
      const int r = q / 64, cc = q % 64;
      const int64_t a = a0 + r;
      Xs[r][cc] = (a < n && d0 + cc < n) ? X[a * n + d0 + cc] : 0.0;
      Ts[r][cc] = (a < n && c0 + cc < n) ? Tm[a * n + c0 + cc] : 0.0;
    }
    __syncthreads();
    for (int r = 0; r < 64; ++r) {
      double xv[4], tv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) { xv[u] = Xs[r][tr + u]; tv[u] = Ts[r][tc + u]; }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += xv[u] * tv[v];
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t d = d0 + tr + u, c = c0 + tc + v;
      if (d < n && c < n) S[d * n + c] = acc[u][v];
    }
}

// table gradient of the seed block: every pair (a, b) feeds S[a][b] through its lookup
__global__ __launch_bounds__(256) void seed_table_grad_kernel(const void* __restrict__ pts, int pts_f64, int d,
                                                              int64_t n, const double* __restrict__ S,
                                                              const double* __restrict__ bins, int nb,
                                                              double* __restrict__ gtab) {
  extern __shared__ double hist[];
  for (int j = threadIdx.x; j < nb; j += blockDim.x) hist[j] = 0.0;
  __syncthreads();
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const double g = S[e];
    if (g == 0.0) continue;
    const int64_t a = e / n, b = e - a * n;
    double s2 = 0.0;
    for (int t = 0; t < d; ++t) {
      const double u = load_as<double>(pts, pts_f64, a * d + t) - load_as<double>(pts, pts_f64, b * d + t);
      s2 += u * u;
    }
    const double r = sqrt(s2);
    if (r > bins[nb - 1]) {
      atomicAdd(&hist[nb - 1], g);
      continue;
    }
    int lo = 0, hi = nb;  // upper_bound
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (bins[mid] <= r) lo = mid + 1;
      else hi = mid;
    }
    const int i = lo < 1 ? 1 : (lo > nb - 1 ? nb - 1 : lo);
    const double dx = bins[i] - bins[i - 1];
    const double t = dx > 0.0 ? (r - bins[i - 1]) / dx : 0.0;
    atomicAdd(&hist[i - 1], g * (1.0 - t));
    if (t != 0.0) atomicAdd(&hist[i], g * t);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nb; j += blockDim.x)
    if (hist[j] != 0.0) atomicAdd(&gtab[j], hist[j]);
}

// u <- -t (the inverse's seed-block form of u) and a = X^T t (X = L^{-1}, lower) added to
// the seeds' vbar; one wavefront per column
template <typename T>
__global__ __launch_bounds__(256) void seed_inv_kernel(const double* __restrict__ X, int64_t n,
                                                       const double* __restrict__ t, double* __restrict__ u,
                                                       T* __restrict__ vbar) {
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n) return;
  double s = 0.0;
  for (int64_t i = c + lane; i < n; i += 64) s += X[i * n + c] * t[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) {
    u[c] = -t[c];
    vbar[c] = (T)((double)vbar[c] + s);
  }
}

__global__ void scale_kernel(double* __restrict__ S, int64_t m, const double* __restrict__ l, double f) {
  const double g = f * *l;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
    S[e] *= g;
}

template <typename T>
__global__ void scatter_out_kernel(const T* __restrict__ vbar, const int32_t* __restrict__ perm, int64_t n,
                                   void* __restrict__ out, int out_f64) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t dst = perm ? (int64_t)perm[i] : i;
  if (out_f64) ((double*)out)[dst] = (double)vbar[i];
  else ((float*)out)[dst] = (float)vbar[i];
}

// a malformed graph (a wait gave up): every gradient is undefined, so all of it is NaN
__global__ void vjp_nan_fill_kernel(void* __restrict__ out, int out_f64, int64_t n, double* __restrict__ gtab, int nb,
                                    const uint32_t* __restrict__ flags) {
  if ((flags[0] | flags[1]) == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (out_f64) ((double*)out)[i] = __builtin_nan("");
    else ((float*)out)[i] = __builtin_nanf("");
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x)
    gtab[i] = __builtin_nan("");
}

// ------------------------------------------------------------------ any k, any d
// One lane per refined point; its (k+1)^2 system, w, z and the k+1 points' coordinates in a
// slab of the workspace (interleaved by lane, so the lanes' accesses coalesce).  Same
// formulas as adjoint_kernel (see adjoint_impl.hpp), exact binary-search table lookups
// (as refine_generic_kernel).  The table and the gradient histogram sit in LDS when they
// fit (else the table is read from the workspace and the gradient goes to global FP64
// atomics).  Points are taken in decreasing order; in generate mode a lane waits (without
// blocking its wavefront) until its point's children are done.
template <typename T>
__host__ __device__ inline size_t generic_slab(int k, int d) {
  const size_t m = (size_t)k + 1;
  return m * m + 2 * m + m * (size_t)d;
}

template <typename T>
__global__ __launch_bounds__(64) void adjoint_generic_kernel(AdjointArgs<T> a, int mode, T* __restrict__ scratch,
                                                             int lds_mode) {
  extern __shared__ __attribute__((aligned(16))) double ghist[];
  const int k = a.k, m = k + 1, dd = a.d;
  const int hist_lds = lds_mode >= 1, tab_lds = lds_mode >= 2;
  const LaneSlab<T> S{scratch + (size_t)blockIdx.x * 64 * generic_slab<T>(k, dd) + threadIdx.x};
  const size_t wo = (size_t)m * m, zo = wo + m, co = zo + m;  // w, z, coordinates after K
  auto Kat = [&](size_t e) -> T& { return S[e]; };
  auto wat = [&](int i) -> T& { return S[wo + i]; };
  auto zat = [&](int i) -> T& { return S[zo + i]; };
  auto cat = [&](int i, int t) -> T& { return S[co + (size_t)i * dd + t]; };
  const TabMeta meta = *a.meta;
  const int nseg = meta.nseg;
  const T x_last = (T)meta.x_last, f_last = (T)meta.f_last;
  const TabE<T>* tab = a.tab;
  const T* edges = reinterpret_cast<const T*>(a.tab + nseg);
  if (hist_lds) {
    for (int j = threadIdx.x; j < a.nbins; j += blockDim.x) ghist[j] = 0.0;
  }
  if (tab_lds) {
    TabE<T>* ltab = reinterpret_cast<TabE<T>*>(ghist + ((a.nbins + 1) & ~1));
    T* ledges = reinterpret_cast<T*>(ltab + nseg);
    for (int j = threadIdx.x; j < nseg; j += blockDim.x) ltab[j] = tab[j];
    for (int j = threadIdx.x; j <= nseg; j += blockDim.x) ledges[j] = edges[j];
    tab = ltab;
    edges = ledges;
  }
  __syncthreads();
  __shared__ uint32_t s_tile;
  auto seg_of = [&](T r) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (edges[mid] <= r) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  };
  auto hist_add = [&](int j, double g) {
    if (hist_lds) atomicAdd(&ghist[j], g);
    else atomicAdd(&a.gtab[j], g);
  };
  auto grad_entry = [&](T r, double g) {
    if (r >= x_last) {
      hist_add(nseg, g);
      return;
    }
    const int j = seg_of(r);
    const double x0 = (double)edges[j], xn = (double)edges[j + 1];
    const double dx = xn - x0;
    const double t = dx > 0.0 ? ((double)r - x0) / dx : 0.0;
    hist_add(j, g * (1.0 - t));
    if (t != 0.0) hist_add(j + 1, g * t);
  };
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile >= a.ntiles) break;
    const int64_t pt = a.n_ref - 1 - (tile * 64 + threadIdx.x);
    const bool valid = pt >= 0;
    const int64_t fine = a.n0 + (valid ? pt : 0);
    auto pid = [&](int i) -> int64_t { return i < k ? (int64_t)a.nbrs[pt * k + i] : fine; };
    if (valid)
      for (int i = 0; i < m; ++i) {
        const int64_t q = pid(i);
        for (int t = 0; t < dd; ++t) cat(i, t) = load_as<T>(a.pts, a.pts_f64, q * dd + t);
      }
    auto dist = [&](int i, int j) {
      T s2 = T(0);
      for (int t = 0; t < dd; ++t) {
        const T u = cat(i, t) - cat(j, t);
        s2 += u * u;
      }
      return sqrt(s2);
    };
    auto look = [&](T r) { return table_lookup<T>(tab, edges, nseg, x_last, f_last, r); };
    double sigma = 1.0;
    if (valid) {
      for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) Kat((size_t)i * m + j) = look(i == j ? T(0) : dist(i, j));
      // Cholesky of A (k x k, lower, in place); b stays in row k
      slab_cholesky(S, k, m);
    }
    // x <- A^{-1} x with the factor: L y = x, then L^T x = y
    auto solve = [&](size_t xo) {
      slab_solve_lower(S, k, m, xo);
      slab_solve_upper(S, k, m, xo);
    };
    if (valid) {
      for (int i = 0; i < k; ++i) wat(i) = Kat((size_t)k * m + i);
      solve(wo);
      double bw = 0.0;
      for (int i = 0; i < k; ++i) bw += (double)Kat((size_t)k * m + i) * (double)wat(i);
      sigma = sqrt((double)Kat((size_t)k * m + k) - bw);
      for (int i = 0; i < k; ++i) zat(i) = mode != kLogdet ? a.vals[pid(i)] : T(0);
      if (mode != kLogdet) solve(zo);
    }
    // cotangent coefficients of this point (see adjoint_impl.hpp): ybar = al z, and be the
    // cotangent of sigma^2; the scatter runs after the point's dependency work, for all
    // lanes together, so that it is off the chain of the generate sweep
    double al = 0.0, be = 0.0;
    if (mode == kGenerate) {
      const double xi = valid ? load_as<double>(a.xi, a.xi_f64, a.perm ? (int64_t)a.perm[fine] : fine) : 0.0;
      bool done = !valid;
      for (uint32_t it = 0;; ++it) {
        if (!done && __hip_atomic_load(&a.children[fine], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
          const double s = (double)__hip_atomic_load(&a.vbar[fine], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const int64_t dst = a.perm ? (int64_t)a.perm[fine] : fine;
          if (a.gxi_f64) ((double*)a.gxi)[dst] = sigma * s;
          else ((float*)a.gxi)[dst] = (float)(sigma * s);
          for (int i = 0; i < k; ++i)
            __hip_atomic_fetch_add(&a.vbar[pid(i)], (T)(s * (double)wat(i)), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          __builtin_amdgcn_s_waitcnt(0);
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          for (int i = 0; i < k; ++i)
            __hip_atomic_fetch_sub(&a.children[pid(i)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          al = s;
          be = s * xi / (2.0 * sigma);
          done = true;
        }
        if (__all(done)) break;
        if (it > (1u << 24)) {
          if (!done) atomicOr(&a.flags[1], 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    } else if (mode == kInverse) {
      if (valid) {
        double bz = 0.0;
        for (int i = 0; i < k; ++i) bz += (double)Kat((size_t)k * m + i) * (double)zat(i);
        const double xo = ((double)a.vals[fine] - bz) / sigma;
        const double t = load_as<double>(a.cot, a.cot_f64, a.perm ? (int64_t)a.perm[fine] : fine);
        al = -t / sigma;
        be = -t * xo / (2.0 * sigma * sigma);
        __hip_atomic_fetch_add(&a.vbar[fine], (T)(t / sigma), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int i = 0; i < k; ++i)
          __hip_atomic_fetch_add(&a.vbar[pid(i)], (T)(al * (double)wat(i)), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
      }
    } else if (valid) {
      be = *a.ld_bar / (2.0 * sigma * sigma);
    }
    if (valid && (al != 0.0 || be != 0.0)) {
      for (int i = 0; i < k; ++i) {
        const double wi = (double)wat(i), zi = (double)zat(i);
        for (int j = 0; j <= i; ++j) {
          const double wj = (double)wat(j), zj = (double)zat(j);
          const double g = (i == j) ? (-al * wi * zi + be * wi * wi) : (-al * (wi * zj + wj * zi) + 2.0 * be * wi * wj);
          grad_entry(i == j ? T(0) : dist(i, j), g);
        }
        grad_entry(dist(i, k), al * zi - 2.0 * be * wi);
      }
      grad_entry(T(0), be);
    }
  }
  if (hist_lds) {
    __syncthreads();
    for (int j = threadIdx.x; j < a.nbins; j += blockDim.x)
      if (ghist[j] != 0.0) atomicAdd(&a.gtab[j], ghist[j]);
  }
}

// Blocks of the generic adjoint: at most 256, and at most 512 MB of slabs.
template <typename T>
int generic_adjoint_blocks(int k, int d) {
  const size_t per_block = (size_t)64 * generic_slab<T>(k, d) * sizeof(T);
  const size_t b = ((size_t)512 << 20) / per_block;
  return (int)std::max<size_t>(1, std::min<size_t>(256, b));
}

inline bool register_adjoint(const RefineProblem& p) { return p.d <= 3 && p.k + 1 <= 32; }

template <typename T>
struct VjpWs {
  uint32_t* ctr;
  TabE<T>* tab;
  TabMeta* meta;
  T* vals;
  T* vbar;
  uint32_t* children;
  double *L, *X, *Tm, *S, *vb, *xs, *u;
  T* slab;
};

template <typename T>
bool carve_vjp(Workspace& w, const RefineProblem& p, VjpWs<T>& b) {
  const int64_t n0 = p.n - p.n_ref;
  b.ctr = w.take<uint32_t>(4);
  b.tab = w.take<TabE<T>>(p.nb);
  w.take<T>(p.nb + 2);
  b.meta = w.take<TabMeta>(1);
  b.vals = w.take<T>(p.n);
  b.vbar = w.take<T>(p.n);
  b.children = w.take<uint32_t>(p.n);
  b.L = w.take<double>((size_t)n0 * n0 + 4096);
  b.X = w.take<double>((size_t)n0 * n0);
  b.Tm = w.take<double>((size_t)n0 * n0);
  b.S = w.take<double>((size_t)n0 * n0);
  b.vb = w.take<double>(n0 + 1);
  b.xs = w.take<double>(n0 + 1);
  b.u = w.take<double>(n0 + 1);
  b.slab = register_adjoint(p) ? nullptr
                               : w.take<T>((size_t)generic_adjoint_blocks<T>(p.k, p.d) * 64 *
                                           generic_slab<T>(p.k, p.d));
  return w.ok();
}

int cus() {
  int dev = 0, c = 256;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    c = 256;
  return c;
}

template <typename T>
int launch_register(AdjointArgs<T> a, int mode, int kk, hipStream_t s, size_t lds, int grid) {
  if (mode == kGenerate) return launch_adjoint_mode<T, kGenerate>(a, kk, s, lds, grid);
  if (mode == kInverse) return launch_adjoint_mode<T, kInverse>(a, kk, s, lds, grid);
  return launch_adjoint_mode<T, kLogdet>(a, kk, s, lds, grid);
}

template <typename T>
int vjp_flow(const RefineProblem& p, int mode, const VjpIO& io, void* ws, size_t ws_bytes, hipStream_t s) {
  GGP_CHECK_ARG(p.n >= 1 && p.n_ref >= 0 && p.n_ref < p.n && p.k >= 1 && p.d >= 1, "vjp: bad sizes");
  GGP_CHECK_ARG(p.nb >= 2, "vjp: covariance table needs at least two bins");
  GGP_CHECK_ARG(mode == kGenerate || mode == kInverse || mode == kLogdet, "vjp: unknown mode");
  Workspace w(ws, ws_bytes);
  VjpWs<T> b;
  if (!carve_vjp<T>(w, p, b)) {
    set_error("vjp: workspace too small");
    return kWorkspaceTooSmall;
  }
  const int64_t n0 = p.n - p.n_ref;
  const size_t xl = (size_t)n0 * sizeof(double);
  GGP_CHECK_ARG(xl <= 160 * 1024, "vjp: seed block too large for the LDS column inverse (n0 <= 20480)");
  GGP_TRY(fill_u32(b.ctr, 0u, 4, s));
  GGP_TRY(prep_table(p.bins, p.vals, p.nb, sizeof(T) == 8, 0, b.tab, b.meta, s));
  if (mode != kLogdet)
    hipLaunchKernelGGL(gather_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, s, io.values, io.values_f64, p.perm,
                       p.n, b.vals);
  if (mode == kGenerate) {
    hipLaunchKernelGGL(gather_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, s, io.gvalues, io.g_f64, p.perm,
                       p.n, b.vbar);
  } else if (mode == kInverse) {
    if (sizeof(T) == 8) GGP_TRY(fill_u64(b.vbar, 0ull, p.n, s));
    else GGP_TRY(fill_u32(b.vbar, 0u, p.n, s));
  }
  GGP_LAUNCH_CHECK();
  GGP_TRY(fill_u64(io.gtab, 0ull, p.nb, s));
  if (p.n_ref > 0) {
    if (mode == kGenerate) {
      GGP_TRY(fill_u32(b.children, 0u, p.n, s));
      hipLaunchKernelGGL(count_children_kernel, dim3(grid_capped(p.n_ref * p.k, 256)), dim3(256), 0, s, p.nbrs,
                         p.n_ref * (int64_t)p.k, b.children);
      GGP_LAUNCH_CHECK();
    }
    AdjointArgs<T> a{};
    a.pts = p.pts;
    a.pts_f64 = p.pts_f64;
    a.d = p.d;
    a.nbrs = p.nbrs;
    a.k = p.k;
    a.n0 = n0;
    a.n_ref = p.n_ref;
    a.perm = p.perm;
    a.xi = io.xi;
    a.xi_f64 = io.xi_f64;
    a.vals = b.vals;
    a.vbar = b.vbar;
    a.children = b.children;
    a.gxi = io.gxi;
    a.gxi_f64 = io.gxi_f64;
    a.gtab = io.gtab;
    a.cot = io.gvalues;
    a.cot_f64 = io.g_f64;
    a.ld_bar = io.gscalar;
    a.tab = b.tab;
    a.meta = b.meta;
    a.ticket = b.ctr;
    a.flags = b.ctr + 1;
    a.nbins = p.nb;
    const size_t tab_bytes = (size_t)(p.nb - 1) * sizeof(TabE<T>) + (size_t)p.nb * sizeof(T);
    const size_t lds = ((tab_bytes + 15) & ~size_t(15)) + (size_t)p.nb * sizeof(double);
    int rc = kUnsupported;
    if (register_adjoint(p) && lds <= 160 * 1024) rc = launch_register<T>(a, mode, p.k + 1, s, lds, cus() * 4);
    if (rc == kUnsupported) {
      if (b.slab == nullptr) {
        set_error("vjp: the register adjoint does not fit (covariance table too large for LDS)");
        return kUnsupported;
      }
      a.ntiles = (p.n_ref + 63) / 64;
      const int grid = (int)std::min<int64_t>(a.ntiles, generic_adjoint_blocks<T>(p.k, p.d));
      // LDS: the histogram (n_bins doubles, even count), then the table (tab_bytes)
      const size_t hl = (size_t)((p.nb + 1) & ~1) * sizeof(double);
      const int lds_mode = hl + tab_bytes <= 64 * 1024 ? 2 : (hl <= 64 * 1024 ? 1 : 0);
      const size_t glds = lds_mode == 2 ? hl + tab_bytes : (lds_mode == 1 ? hl : 0);
      hipLaunchKernelGGL(adjoint_generic_kernel<T>, dim3(grid), dim3(64), glds, s, a, mode, b.slab, lds_mode);
      GGP_LAUNCH_CHECK();
    } else if (rc != kOk) {
      return rc;
    }
  }
  // seed block
  GGP_TRY(dense_factor(p.pts, p.pts_f64, p.d, n0, p.bins, p.vals, p.nb, b.L, s));
  if (xl > 65536)
    GGP_HIP(hipFuncSetAttribute((const void*)tri_inv_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)xl));
  if (mode == kGenerate) {
    hipLaunchKernelGGL(seed_vbar_kernel<T>, dim3(grid_for(n0, 256)), dim3(256), 0, s, b.vbar, io.xi, io.xi_f64,
                       p.perm, n0, b.vb, b.xs);
    hipLaunchKernelGGL(trmv_t_kernel, dim3(grid_for(n0, 4)), dim3(256), 0, s, b.L, n0, b.vb, b.u, p.perm, io.gxi,
                       io.gxi_f64);
    hipLaunchKernelGGL(tri_inv_kernel, dim3((unsigned)n0), dim3(64), xl, s, b.L, n0, b.X);
    hipLaunchKernelGGL(phi_kernel, dim3(grid_for(n0, 64)), dim3(64), 0, s, b.X, n0, b.u, b.xs, b.Tm);
    hipLaunchKernelGGL(gemm_tn_kernel, dim3(grid_for(n0, 64), grid_for(n0, 64)), dim3(256), 0, s, b.X, b.Tm, n0, b.S);
  } else if (mode == kInverse) {
    // xs = the seeds' xi = L^{-1} v_s; t = the seeds' cotangents (vb); u = -t, vbar_s += X^T t
    hipLaunchKernelGGL(seed_vbar_kernel<T>, dim3(grid_for(n0, 256)), dim3(256), 0, s, b.vals, io.gvalues, io.g_f64,
                       p.perm, n0, b.xs, b.vb);
    GGP_LAUNCH_CHECK();
    GGP_TRY(dense_solve(b.L, n0, b.xs, s));
    hipLaunchKernelGGL(tri_inv_kernel, dim3((unsigned)n0), dim3(64), xl, s, b.L, n0, b.X);
    hipLaunchKernelGGL(seed_inv_kernel<T>, dim3(grid_for(n0, 4)), dim3(256), 0, s, b.X, n0, b.vb, b.u, b.vbar);
    hipLaunchKernelGGL(phi_kernel, dim3(grid_for(n0, 64)), dim3(64), 0, s, b.X, n0, b.u, b.xs, b.Tm);
    hipLaunchKernelGGL(gemm_tn_kernel, dim3(grid_for(n0, 64), grid_for(n0, 64)), dim3(256), 0, s, b.X, b.Tm, n0, b.S);
    hipLaunchKernelGGL(scatter_out_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, s, b.vbar, p.perm, p.n, io.gxi,
                       io.gxi_f64);
  } else {
    hipLaunchKernelGGL(tri_inv_kernel, dim3((unsigned)n0), dim3(64), xl, s, b.L, n0, b.X);
    hipLaunchKernelGGL(gemm_tn_kernel, dim3(grid_for(n0, 64), grid_for(n0, 64)), dim3(256), 0, s, b.X, b.X, n0, b.S);
    hipLaunchKernelGGL(scale_kernel, dim3(grid_capped(n0 * n0, 256)), dim3(256), 0, s, b.S, n0 * n0, io.gscalar, 0.5);
  }
  GGP_LAUNCH_CHECK();
  const size_t hl = (size_t)p.nb * sizeof(double);
  if (hl > 65536) GGP_HIP(hipFuncSetAttribute((const void*)seed_table_grad_kernel,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)hl));
  hipLaunchKernelGGL(seed_table_grad_kernel, dim3((unsigned)std::min<int64_t>(grid_for(n0 * n0, 256), 1024)),
                     dim3(256), hl, s, p.pts, p.pts_f64, p.d, n0, b.S, p.bins, p.nb, io.gtab);
  GGP_LAUNCH_CHECK();
  hipLaunchKernelGGL(vjp_nan_fill_kernel, dim3(grid_for(std::min<int64_t>(p.n, 1 << 20), 256)), dim3(256), 0, s,
                     mode == kLogdet ? nullptr : io.gxi, io.gxi_f64, mode == kLogdet ? 0 : p.n, io.gtab, p.nb,
                     b.ctr + 1);
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace

size_t vjp_workspace(const RefineProblem& p, int t64) {
  Workspace w(nullptr, 0);
  if (t64) {
    VjpWs<double> b;
    carve_vjp<double>(w, p, b);
  } else {
    VjpWs<float> b;
    carve_vjp<float>(w, p, b);
  }
  return w.used + 256;
}

int vjp_run(const RefineProblem& p, int mode, int t64, const VjpIO& io, void* ws, size_t ws_bytes, hipStream_t s) {
  return t64 ? vjp_flow<double>(p, mode, io, ws, ws_bytes, s) : vjp_flow<float>(p, mode, io, ws, ws_bytes, s);
}

}  // namespace ggp
