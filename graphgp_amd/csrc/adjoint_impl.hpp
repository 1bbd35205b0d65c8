// adjoint_impl.hpp -- reverse-mode (vector-Jacobian) kernel of the refinement.
//
// For L = sum_i g_i values_i through generate (refine.py:21-56, 79-159), with value
// v_i = b^T A^{-1} v[nbrs] + sigma xi_i, sigma^2 = c - b^T A^{-1} b, w = A^{-1} b,
// z = A^{-1} v[nbrs] and s = vbar_i (the adjoint of v_i, complete once every point that
// uses i as a neighbour has been processed):
//   xibar_i = sigma s;   vbar[nbr_j] += s w_j;
//   dL/dA_aa = -s w_a z_a + s xi w_a^2 / (2 sigma)
//   dL/dA_ab = -s (w_a z_b + w_b z_a) + s xi w_a w_b / sigma       (a < b < k)
//   dL/db_a  =  s z_a - s xi w_a / sigma
//   dL/dc    =  s xi / (2 sigma)
// and each covariance entry interp(r) feeds (1 - t) to vals[j] and t to vals[j+1] of its
// segment j (t = (r - x_j) / (x_{j+1} - x_j)).
//
// MI355X mapping: the same P-lane row-cyclic register factorisation as the forward
// kernel, recomputed (nothing stored between passes); two forward and two backward
// substitutions on it give w and z; the table gradient goes to an FP64 histogram in
// LDS (ds_add_f64) flushed once per workgroup.  Points are taken in decreasing order
// from an atomic ticket; a point starts its scatter once its count of unfinished
// children reaches zero, so the sweep is one launch and dependencies always lie on lower
// tickets.  vbar is only ever changed by device-scope atomics and read by a device-scope
// atomic load, so no cache maintenance is needed: a child waits for its additions to
// complete (s_waitcnt) before it decrements the counts, and a parent reads vbar only after
// it has observed its count at zero.
#pragma once
#include "refine_impl.hpp"

namespace ggp {

template <typename T>
struct AdjointArgs {
  const void* pts;
  int pts_f64;
  int d;
  const int32_t* nbrs;
  int k;
  int64_t n0, n_ref;
  const int32_t* perm;
  const void* xi;  // caller order
  int xi_f64;
  const T* vals;       // generate: forward values; inverse: its input values (graph order)
  T* vbar;             // adjoint of the values, graph order (accumulated)
  uint32_t* children;  // unfinished children per point
  void* gxi;           // caller order
  int gxi_f64;
  double* gtab;  // (n_bins) table gradient, FP64
  const void* cot;       // inverse: cotangent of xi (caller order)
  int cot_f64;
  const double* ld_bar;  // log-det: cotangent of the log-det (device scalar)
  const TabE<T>* tab;
  const TabMeta* meta;
  uint32_t* ticket;
  uint32_t* flags;
  int64_t ntiles;
  int nbins;
};

template <typename T, int KK, int P, int D, bool FAST, int MODE>
__global__ __launch_bounds__(256, 2) void adjoint_kernel(AdjointArgs<T> a) {
  static_assert(KK % P == 0, "KK must be a multiple of P");
  constexpr int R = KK / P;
  constexpr int S = P * R * (R + 1) / 2;
  constexpr int GPT = 256 / P;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  const TabMeta meta = *a.meta;
  if ((meta.log_exact != 0) != FAST) return;
  TabE<T>* tab = reinterpret_cast<TabE<T>*>(smem_raw);
  T* edges = reinterpret_cast<T*>(tab + meta.nseg);
  double* hist = reinterpret_cast<double*>(smem_raw + ((((size_t)meta.nseg * sizeof(TabE<T>) +
                                                         (size_t)(meta.nseg + 1) * sizeof(T)) + 15) & ~size_t(15)));
  __shared__ uint32_t s_tile;
  const T* gedges = reinterpret_cast<const T*>(a.tab + meta.nseg);
  for (int j = threadIdx.x; j < meta.nseg; j += blockDim.x) tab[j] = a.tab[j];
  for (int j = threadIdx.x; j <= meta.nseg; j += blockDim.x) edges[j] = gedges[j];
  for (int j = threadIdx.x; j < a.nbins; j += blockDim.x) hist[j] = 0.0;
  __syncthreads();
  const T x1 = (T)meta.x1, x_last = (T)meta.x_last, f_last = (T)meta.f_last;
  const int nseg = meta.nseg;
  const float lb1 = meta.lb1, inv_step = meta.inv_step;
  const int lane = threadIdx.x & (P - 1);
  const int grp = threadIdx.x / P;
  const int k_arg = a.k;

  // segment of r (same rule as the forward lookup) and its value
  auto seg_of = [&](T r, T r2) -> int {
    int j;
    if constexpr (FAST) {
      j = log_guess((float)r2 * (sizeof(T) == 8 ? 0.99999f : 1.0f), lb1, inv_step);
      j = j < 1 ? 1 : (j > nseg - 1 ? nseg - 1 : j);
      j = (r < x1) ? 0 : j;
      if constexpr (sizeof(T) == 8) {
        if (j + 1 <= nseg - 1 && r >= edges[j + 1]) j += 1;
      }
    } else {
      int lo = 0, hi = nseg - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (edges[mid] <= r) lo = mid;
        else hi = mid - 1;
      }
      j = lo;
    }
    return j;
  };
  auto lookup2 = [&](T r2) -> T {
    const T r = fast_sqrt<T>(r2);
    if (r >= x_last) return f_last;
    const int j = seg_of(r, r2);
    return seg_value<T>(tab[j], edges[j], r);
  };
  // scatter dL/d(entry) = g into the table histogram
  auto grad_entry = [&](T r2, double g) {
    const T r = fast_sqrt<T>(r2);
    if (r >= x_last) {
      atomicAdd(&hist[nseg], g);
      return;
    }
    const int j = seg_of(r, r2);
    const double x0 = (double)edges[j], xn = (double)edges[j + 1];
    const double dx = xn - x0;
    const double t = dx > 0.0 ? ((double)r - x0) / dx : 0.0;
    atomicAdd(&hist[j], g * (1.0 - t));
    if (t != 0.0) atomicAdd(&hist[j + 1], g * t);
  };

  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile >= a.ntiles) break;
    // k anew in every tile, opaque to the compiler: the unrolled column tests stay scalar
    // compares where they are used instead of hoisted lane masks, which spill
    int k = k_arg;
    asm volatile("" : "+s"(k));
    // tiles in decreasing index order
    const int64_t pt = a.n_ref - 1 - (tile * GPT + grp);
    const bool valid = pt >= 0;
    const int64_t fine = a.n0 + pt;

    int32_t nid[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = P * r + lane;
      nid[r] = (valid && i < k) ? a.nbrs[pt * k + i] : -1;
    }
    // coordinates of the rows and of the fine point; loaded again (L2-resident) for the
    // gradient phase instead of being kept live through the factorisation
    auto load_rows = [&](T (&cc)[R][D], T (&ccf)[D]) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int j = 0; j < D; ++j)
          cc[r][j] = (nid[r] >= 0 && j < a.d) ? load_as<T>(a.pts, a.pts_f64, (int64_t)nid[r] * a.d + j) : T(0);
#pragma unroll
      for (int j = 0; j < D; ++j) ccf[j] = (valid && j < a.d) ? load_as<T>(a.pts, a.pts_f64, fine * a.d + j) : T(0);
    };
    T c[R][D], cf[D];
    load_rows(c, cf);

    // joint covariance of the neighbours (rows < k) and the cross-covariance b
    T m[S];
    T bv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) bv[r] = (P * r + lane < k) ? lookup2(pair_dist2<T, D>(c[r], cf)) : T(0);
    static_for<0, KK>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      constexpr int js = j % P, jr = j / P;
      if (j < k) {
        T cj[D];
#pragma unroll
        for (int t = 0; t < D; ++t) cj[t] = bcast<P, js>(c[jr][t]);
#pragma unroll
        for (int r = jr; r < R; ++r) m[P * r * (r + 1) / 2 + j] = lookup2(pair_dist2<T, D>(c[r], cj));
      }
    });
    // Cholesky of A (k x k)
    T rdl[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rdl[r] = T(0);
    static_for<0, KK>([&](auto cc) {
      constexpr int col = decltype(cc)::value;
      constexpr int cs = col / P, cl = col % P;
      if (col < k) {
        const T dg = bcast<P, cl>(m[P * cs * (cs + 1) / 2 + col]);
        const T rs = rsqrt_t<T>(dg);
        if (lane == cl) rdl[cs] = rs;
#pragma unroll
        for (int r = cs; r < R; ++r) m[P * r * (r + 1) / 2 + col] *= rs;
        static_for<col + 1, KK>([&](auto mc) {
          constexpr int mm = decltype(mc)::value;
          constexpr int ms = mm / P, ml = mm % P;
          if (mm < k) {
            const T bb = bcast<P, ml>(m[P * ms * (ms + 1) / 2 + col]);
#pragma unroll
            for (int r = ms; r < R; ++r) m[P * r * (r + 1) / 2 + mm] -= m[P * r * (r + 1) / 2 + col] * bb;
          }
        });
      }
    });
    // x <- A^{-1} x for a row-distributed vector: L y = x, then L^T x = y
    auto solve = [&](T (&x)[R]) {
      static_for<0, KK>([&](auto cc) {
        constexpr int col = decltype(cc)::value;
        constexpr int cs = col / P, cl = col % P;
        if (col < k) {
          const T y = bcast<P, cl>(x[cs] * rdl[cs]);
          if (lane == cl) x[cs] = y;
#pragma unroll
          for (int r = cs; r < R; ++r)
            if (P * r + lane > col) x[r] -= m[P * r * (r + 1) / 2 + col] * y;
        }
      });
      static_for<0, KK>([&](auto cc) {
        constexpr int col = KK - 1 - decltype(cc)::value;
        constexpr int cs = col / P, cl = col % P;
        if (col < k) {
          const T xi_ = bcast<P, cl>(x[cs] * rdl[cs]);
          if (lane == cl) x[cs] = xi_;
          // x_m -= L[col][m] * x_col for m < col: L[col][m] sits in row col's storage
          static_for<0, col>([&](auto mc) {
            constexpr int mm = decltype(mc)::value;
            constexpr int ms = mm / P, ml = mm % P;
            const T lcm = bcast<P, cl>(m[P * cs * (cs + 1) / 2 + mm]);
            if (lane == ml) x[ms] -= lcm * xi_;
          });
        }
      });
    };
    T w[R], z[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      w[r] = bv[r];
      z[r] = (MODE != kLogdet && valid && P * r + lane < k) ? a.vals[nid[r]] : T(0);
    }
    solve(w);
    if constexpr (MODE != kLogdet) solve(z);
    // sigma^2 = c - b . w  (group sum over the P lanes)
    double bw = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (P * r + lane < k) bw += (double)bv[r] * (double)w[r];
    bw = group_sum<P>(bw);
    const double c0 = (double)lookup2(T(0));
    const double sigma = sqrt(c0 - bw);
    const int gbase = (threadIdx.x & 63) & ~(P - 1);
    const uint64_t gmask = (P == 64) ? ~0ull : ((1ull << P) - 1ull);

    // table gradient of one point from the cotangents of its weights and variance: with
    // ybar = A^{-1} wbar = al * z and q = beta the cotangent of sigma^2,
    //   dA_aa = -al w_a z_a + be w_a^2,  dA_ab = -al (w_a z_b + w_b z_a) + 2 be w_a w_b,
    //   db_a = al z_a - 2 be w_a,  dc = be.
    // Called by all lanes of the point's group together (cross-lane reads stay inside it).
    auto scatter = [&](double al, double be) {
      T c[R][D], cf[D];
      load_rows(c, cf);
      // entries of A (diagonal and below), then b, then c.  Unrolled over the row slot jr
      // only; the column's lane js is a run-time loop (ds_bpermute broadcast), which keeps
      // the register footprint of this phase small.
      static_for<0, R>([&](auto jrc) {
        constexpr int jr = decltype(jrc)::value;
#pragma unroll 1
        for (int js = 0; js < P; ++js) {
          const int j = P * jr + js;
          if (j < k) {
            const int src = gbase + js;
            T cj[D];
#pragma unroll
            for (int t = 0; t < D; ++t) cj[t] = __shfl(c[jr][t], src, 64);
            const double wj = (double)__shfl(w[jr], src, 64);
            const double zj = (double)__shfl(z[jr], src, 64);
#pragma unroll
            for (int r = jr; r < R; ++r) {
              const int i = P * r + lane;
              if (i >= j && i < k) {
                const double wi = (double)w[r], zi = (double)z[r];
                const double g = (i == j) ? (-al * wi * zi + be * wi * wi)
                                          : (-al * (wi * zj + wj * zi) + 2.0 * be * wi * wj);
                grad_entry(pair_dist2<T, D>(c[r], cj), g);
              }
            }
          }
        }
      });
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (P * r + lane < k) grad_entry(pair_dist2<T, D>(c[r], cf), al * (double)z[r] - 2.0 * be * (double)w[r]);
      if (lane == 0) atomicAdd(&hist[0], be);
    };

    if constexpr (MODE == kGenerate) {
      const double xi = valid ? load_as<double>(a.xi, a.xi_f64, a.perm ? (int64_t)a.perm[fine] : fine) : 0.0;
      // wait until every child has added its share to vbar[fine]; the polling loop is
      // wavefront-uniform so a point whose child sits in the same wavefront never blocks it
      bool done = !valid;
      for (uint32_t it = 0;; ++it) {
        if (!done) {
          const uint32_t left = __hip_atomic_load(&a.children[fine], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const uint64_t bal = __ballot(left == 0);
          if (((bal >> gbase) & gmask) == gmask) {
            const double s = (double)__hip_atomic_load(&a.vbar[fine], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (lane == 0) {
              const double gx = sigma * s;
              const int64_t dst = a.perm ? (int64_t)a.perm[fine] : fine;
              if (a.gxi_f64) ((double*)a.gxi)[dst] = gx;
              else ((float*)a.gxi)[dst] = (float)gx;
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (P * r + lane < k)
                __hip_atomic_fetch_add(&a.vbar[nid[r]], (T)(s * (double)w[r]), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            // every addition has completed at its device-scope destination before any count drops
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (P * r + lane < k)
                __hip_atomic_fetch_sub(&a.children[nid[r]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // generate: v = w . u + sigma xi, so wbar = s u (ybar = s z) and the cotangent of
            // sigma^2 is s xi / (2 sigma)
            if (s != 0.0) scatter(s, s * xi / (2.0 * sigma));
            done = true;
          }
        }
        if (__all(done)) break;
        if (it > (1u << 22)) {
          // a malformed graph: the gradient is undefined; the host turns it all into NaN
          if (!done) atomicOr(&a.flags[1], 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    } else if constexpr (MODE == kInverse) {
      // xi_i = (v_i - w . u) / sigma with w . u = b . z; cotangent t of xi_i:
      //   vbar_i += t / sigma, vbar[nbrs] += -t w / sigma, ybar = -(t / sigma) z,
      //   cotangent of sigma^2 = -t xi_i / (2 sigma^2)
      double bz = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (P * r + lane < k) bz += (double)bv[r] * (double)z[r];
      bz = group_sum<P>(bz);
      if (valid) {
        const double vi = (double)a.vals[fine];
        const double xo = (vi - bz) / sigma;
        const double t = load_as<double>(a.cot, a.cot_f64, a.perm ? (int64_t)a.perm[fine] : fine);
        const double al = -t / sigma;
        if (lane == 0)
          __hip_atomic_fetch_add(&a.vbar[fine], (T)(t / sigma), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (P * r + lane < k)
            __hip_atomic_fetch_add(&a.vbar[nid[r]], (T)(al * (double)w[r]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        if (t != 0.0) scatter(al, -t * xo / (2.0 * sigma * sigma));
      }
    } else {
      // log-det: sum log sigma_i with cotangent l: sigma^2 gets l / (2 sigma^2), w none
      if (valid) {
        const double l = *a.ld_bar;
        if (l != 0.0) scatter(0.0, l / (2.0 * sigma * sigma));
      }
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < a.nbins; j += blockDim.x)
    if (hist[j] != 0.0) atomicAdd(&a.gtab[j], hist[j]);
}

// Per (type, mode): the register configurations (d <= 3, k + 1 <= 32); kUnsupported
// otherwise.  Instantiated in adjoint_k_*.hip, one file per type and mode.
template <typename T, int MODE>
int launch_adjoint_mode(AdjointArgs<T> a, int kk, hipStream_t s, size_t lds, int grid);

// ------------------------------------------------------------------ launch
template <typename T, int KK, int P, int D, int MODE>
int launch_adjoint_config(AdjointArgs<T> a, hipStream_t s, size_t lds, int grid) {
  constexpr int64_t gpt = 256 / P;
  a.ntiles = (a.n_ref + gpt - 1) / gpt;
  if (a.ntiles == 0) return kOk;
  if ((int64_t)grid > a.ntiles) grid = (int)a.ntiles;
  auto k1 = adjoint_kernel<T, KK, P, D, true, MODE>;
  auto k0 = adjoint_kernel<T, KK, P, D, false, MODE>;
  if (lds > 65536) {
    GGP_HIP(hipFuncSetAttribute((const void*)k1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    GGP_HIP(hipFuncSetAttribute((const void*)k0, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  hipLaunchKernelGGL(k1, dim3(persistent_grid((const void*)k1, grid, lds)), dim3(256), lds, s, a);
  hipLaunchKernelGGL(k0, dim3(persistent_grid((const void*)k0, grid, lds)), dim3(256), lds, s, a);
  GGP_LAUNCH_CHECK();
  return kOk;
}

// at most four rows per lane: the adjoint keeps w, z and the factor live together
#define GGP_ADJOINT_INSTANTIATE(T, MODE)                                                          \
  template <>                                                                                     \
  int launch_adjoint_mode<T, MODE>(AdjointArgs<T> a, int kk, hipStream_t s, size_t lds, int grid) { \
    if (a.d > 3) return kUnsupported;                                                             \
    if (kk <= 8) return launch_adjoint_config<T, 8, 2, 3, MODE>(a, s, lds, grid);                 \
    if (kk <= 12) return launch_adjoint_config<T, 12, 4, 3, MODE>(a, s, lds, grid);               \
    if (kk <= 16) return launch_adjoint_config<T, 16, 4, 3, MODE>(a, s, lds, grid);               \
    if (kk <= 24) return launch_adjoint_config<T, 24, 8, 3, MODE>(a, s, lds, grid);               \
    if (kk <= 32) return launch_adjoint_config<T, 32, 16, 3, MODE>(a, s, lds, grid);              \
    return kUnsupported;                                                                          \
  }

}  // namespace ggp
