// refine_impl.hpp -- the fused refinement kernel template (see refine.hip for the design).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace ggp {

template <typename T>
struct alignas(4 * sizeof(T)) TabE {
  T x0, f0, s, x1;  // segment [x0, x1): value f0 + (r - x0) * s
};

struct TabMeta {
  int nseg;
  double x_first, x_last, f_first, f_last, f_zero, x1;
  float lb1, inv_step;
  int log_exact;  // 1: the log2 guess alone finds the segment (log-spaced bins)
};

// Segment guess used by the fast lookup; the table check in the prep kernel uses the
// same expression, so "log_exact" certifies exactly what the kernel computes.
__device__ __forceinline__ int log_guess(float r2, float lb1, float inv_step) {
  const float lr = 0.5f * __builtin_amdgcn_logf(r2);
  return 1 + (int)((lr - lb1) * inv_step);
}

enum RefineMode : int { kGenerate = 0, kInverse = 1, kLogdet = 2 };

template <typename T>
struct RefineArgs {
  const void* pts;
  int pts_f64;
  int d;
  const int32_t* nbrs;
  int k;
  int64_t n0, n_ref;
  const int32_t* perm;  // graph.indices (caller order <-> graph order) or null
  const void* xi;       // generate: xi in caller order
  int xi_f64;
  T* vals;    // graph-order values: generate output (pre-filled seeds + sentinels); inverse input
  void* out;  // generate: values in caller order; inverse: xi in caller order (refined rows); may be null
  int out_f64;
  double* partials;  // logdet: one FP64 sum per tile
  const TabE<T>* tab;
  const TabMeta* meta;
  uint32_t* ticket;
  uint32_t* flags;  // [0] NaN produced, [1] malformed graph / wait limit
  int mode;
  int increment;
  int64_t ntiles;
  int64_t io_base;  // subtracted from the caller-order index of xi (generate) / out (inverse)
};

// ------------------------------------------------------------------ helpers
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Broadcast lane S of every P-lane group to the whole group.
template <int P, int S>
__device__ __forceinline__ int bcast_i32(int x) {
  if constexpr (P == 1) {
    return x;
  } else if constexpr (P == 2) {
    constexpr int q = S | (S << 2) | ((2 + S) << 4) | ((2 + S) << 6);
    return __builtin_amdgcn_update_dpp(0, x, q, 0xF, 0xF, false);
  } else if constexpr (P == 4) {
    return __builtin_amdgcn_update_dpp(0, x, S * 0x55, 0xF, 0xF, false);
  } else {
    static_assert(P == 8, "P must be 1, 2, 4 or 8");
    return __builtin_amdgcn_ds_swizzle(x, 0x18 | (S << 5));
  }
}
template <int P, int S>
__device__ __forceinline__ float bcast(float x) {
  return __int_as_float(bcast_i32<P, S>(__float_as_int(x)));
}
template <int P, int S>
__device__ __forceinline__ double bcast(double x) {
  long long v = __double_as_longlong(x);
  int lo = bcast_i32<P, S>((int)(v & 0xffffffffll));
  int hi = bcast_i32<P, S>((int)(v >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename T>
struct Sentinel;
template <>
struct Sentinel<float> {
  __device__ static float value() { return __int_as_float(-1); }
  __device__ static bool is(float x) { return __float_as_int(x) == -1; }
};
template <>
struct Sentinel<double> {
  __device__ static double value() { return __longlong_as_double(-1ll); }
  __device__ static bool is(double x) { return __double_as_longlong(x) == -1ll; }
};

template <typename T>
__device__ __forceinline__ T rsqrt_t(T x);
template <>
__device__ __forceinline__ float rsqrt_t<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <>
__device__ __forceinline__ double rsqrt_t<double>(double x) { return 1.0 / sqrt(x); }

template <typename T>
__device__ __forceinline__ T load_as(const void* p, int f64, int64_t i) {
  return f64 ? (T)((const double*)p)[i] : (T)((const float*)p)[i];
}

// Table lookup: interp(r, bins, vals) (refine.py:305) on the LDS segment table.
template <typename T>
__device__ __forceinline__ T table_lookup(const TabE<T>* __restrict__ tab, int nseg, T x1, T x_last, T f_last,
                                          float lb1, float inv_step, T r) {
  if (r >= x_last) return f_last;
  int j;
  if (r < x1) {
    j = 0;
  } else {
    float lr = __log2f((float)r);
    j = 1 + (int)((lr - lb1) * inv_step);
    j = j < 1 ? 1 : (j > nseg - 1 ? nseg - 1 : j);
  }
  TabE<T> e = tab[j];
#pragma unroll 1
  for (int it = 0; it < 3 && (r < e.x0 || r >= e.x1); ++it) {
    j += (r < e.x0) ? -1 : 1;
    j = j < 0 ? 0 : (j > nseg - 1 ? nseg - 1 : j);
    e = tab[j];
  }
  if (r < e.x0 || r >= e.x1) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (tab[mid].x0 <= r) lo = mid;
      else hi = mid - 1;
    }
    e = tab[lo];
  }
  return e.f0 + (r - e.x0) * e.s;
}

// Lookup for log-spaced bins: one LDS read.  The FP32 guess is off by one segment only
// for r within ~1e-7 (relative) of a bin edge, where the neighbouring segment's line
// differs by a second-order term far below FP32 rounding, so FP32 uses it as is; FP64
// corrects the segment with one compare-and-reload.
template <typename T>
__device__ __forceinline__ T table_lookup_fast(const TabE<T>* __restrict__ tab, int nseg, T x1, T x_last, T f_last,
                                               float lb1, float inv_step, T r, T r2) {
  int j = log_guess((float)r2, lb1, inv_step);
  j = j < 1 ? 1 : (j > nseg - 1 ? nseg - 1 : j);
  j = (r < x1) ? 0 : j;
  TabE<T> e = tab[j];
  if constexpr (sizeof(T) == 8) {
    // FP64 keeps the exact segment: the guess is at most one segment off
    if (r < e.x0 || r >= e.x1) {
      j += (r < e.x0) ? -1 : 1;
      j = j < 0 ? 0 : (j > nseg - 1 ? nseg - 1 : j);
      e = tab[j];
    }
  }
  const T v = e.f0 + (r - e.x0) * e.s;
  return (r >= x_last) ? f_last : v;
}

template <typename T>
__device__ __forceinline__ T fast_sqrt(T x) {
  return sqrt(x);
}
template <>
__device__ __forceinline__ float fast_sqrt<float>(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

template <typename T, int D>
__device__ __forceinline__ T pair_dist2(const T (&a)[D], const T (&b)[D]) {
  T t = a[0] - b[0];
  T s = t * t;
#pragma unroll
  for (int j = 1; j < D; ++j) {
    T u = a[j] - b[j];
    s += u * u;
  }
  return s;
}

template <typename T, int D>
__device__ __forceinline__ T pair_dist(const T (&a)[D], const T (&b)[D]) {
  T t = a[0] - b[0];
  T s = t * t;
#pragma unroll
  for (int j = 1; j < D; ++j) {
    T u = a[j] - b[j];
    s += u * u;
  }
  return sqrt(s);
}

// ------------------------------------------------------------------ the kernel
template <typename T, int KK, int P, int D, bool FAST>
__global__ __launch_bounds__(256) void refine_kernel(RefineArgs<T> a) {
  static_assert(KK % P == 0, "KK must be a multiple of P");
  constexpr int R = KK / P;
  constexpr int S = P * R * (R + 1) / 2;
  constexpr int GPT = 256 / P;
  extern __shared__ char smem_raw[];
  TabE<T>* tab = reinterpret_cast<TabE<T>*>(smem_raw);
  __shared__ uint32_t s_tile;
  __shared__ double s_red[4];

  const TabMeta meta = *a.meta;
  // two variants are launched; the one that does not match the table exits at once
  if ((meta.log_exact != 0) != FAST) return;
  for (int j = threadIdx.x; j < meta.nseg; j += blockDim.x) tab[j] = a.tab[j];
  __syncthreads();
  const T x1 = (T)meta.x1, x_last = (T)meta.x_last, f_last = (T)meta.f_last, f_zero = (T)meta.f_zero;
  const int nseg = meta.nseg;
  const float lb1 = meta.lb1, inv_step = meta.inv_step;
  auto lookup2 = [&](T r2) {
    if constexpr (FAST) {
      const T r = fast_sqrt<T>(r2);
      return table_lookup_fast<T>(tab, nseg, x1, x_last, f_last, lb1, inv_step, r, r2);
    } else {
      return table_lookup<T>(tab, nseg, x1, x_last, f_last, lb1, inv_step, sqrt(r2));
    }
  };

  const int lane = threadIdx.x & (P - 1);
  const int grp = threadIdx.x / P;
  const int k = a.k;
  const bool incr = a.increment != 0;

  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile >= a.ntiles) break;
    const int64_t pt = tile * GPT + grp;
    const bool valid = pt < a.n_ref;
    const int64_t fine = a.n0 + pt;

    // ---- gather this lane's rows: neighbours first, fine point at row k
    T c[R][D];
    T g0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = P * r + lane;
      int64_t id = -1;
      if (valid) id = i < k ? (int64_t)a.nbrs[pt * k + i] : (i == k ? fine : -1);
#pragma unroll
      for (int j = 0; j < D; ++j) c[r][j] = (id >= 0 && j < a.d) ? load_as<T>(a.pts, a.pts_f64, id * a.d + j) : T(0);
    }
    if (incr) {
      T c0[D];
#pragma unroll
      for (int j = 0; j < D; ++j) c0[j] = bcast<P, 0>(c[0][j]);
#pragma unroll
      for (int r = 0; r < R; ++r) g0[r] = lookup2(pair_dist2<T, D>(c[r], c0));
    }

    // ---- covariance entries (rows of this lane, all columns up to the slot width)
    T m[S];
    static_for<0, KK>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      constexpr int js = j % P, jr = j / P;
      if (j <= k) {
        T cj[D];
#pragma unroll
        for (int t = 0; t < D; ++t) cj[t] = bcast<P, js>(c[jr][t]);
        T gj = T(0);
        if (incr) gj = bcast<P, js>(g0[jr]);
#pragma unroll
        for (int r = jr; r < R; ++r) {
          const int i = P * r + lane;
          T v = lookup2(pair_dist2<T, D>(c[r], cj));
          if (incr) {
            if (j == 0) v = (i == 0) ? f_zero : -g0[r];
            else v = (i == 0) ? -gj : (g0[r] + gj) - v;
          }
          m[P * r * (r + 1) / 2 + j] = v;
        }
      }
    });

    // ---- right-looking Cholesky on the row-cyclic layout
    T rdl[R];  // 1 / L_ii for this lane's rows
#pragma unroll
    for (int r = 0; r < R; ++r) rdl[r] = T(0);
    static_for<0, KK - 1>([&](auto cc) {
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
          if (mm <= k) {
            const T b = bcast<P, ml>(m[P * ms * (ms + 1) / 2 + col]);
#pragma unroll
            for (int r = ms; r < R; ++r) m[P * r * (r + 1) / 2 + mm] -= m[P * r * (r + 1) / 2 + col] * b;
          }
        });
      }
    });

    // Schur complement of the fine row: owned by lane k % P at slot k / P
    T skk = T(0);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (lane == q && P * r + q == k) skk = m[P * r * (r + 1) / 2 + P * r + q];
    const bool owner = (lane == (k % P));

    // ---- forward substitution of the (transformed) values; returns row k's residual
    auto forward = [&](T (&u)[R]) -> T {
      static_for<0, KK - 1>([&](auto cc) {
        constexpr int col = decltype(cc)::value;
        constexpr int cs = col / P, cl = col % P;
        if (col < k) {
          const T y = bcast<P, cl>(u[cs] * rdl[cs]);
#pragma unroll
          for (int r = cs; r < R; ++r) u[r] -= m[P * r * (r + 1) / 2 + col] * y;
        }
      });
      T uk = T(0);
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (P * r + lane == k) uk = u[r];
      return uk;
    };

    if (a.mode == kLogdet) {
      double ld = 0.0;
      if (valid && owner) {
        ld = 0.5 * log((double)skk);
        if (ld != ld) atomicOr(&a.flags[0], 1u);
      }
      // block reduction in a fixed order (deterministic)
      for (int o = 32; o > 0; o >>= 1) ld += __shfl_xor(ld, o, 64);
      if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = ld;
      __syncthreads();
      if (threadIdx.x == 0) a.partials[tile] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
      __syncthreads();
      continue;
    }

    if (a.mode == kInverse) {
      T u[R];
      T v0 = T(0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = P * r + lane;
        u[r] = T(0);
        if (valid && i < k) u[r] = a.vals[a.nbrs[pt * k + i]];
        if (valid && i == k) u[r] = a.vals[fine];
      }
      if (incr) {
        v0 = bcast<P, 0>(u[0]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = P * r + lane;
          if (i != 0 && i <= k) u[r] -= v0;
        }
      }
      const T uk = forward(u);
      if (valid && owner) {
        const T xi = uk / sqrt(skk);
        if (xi != xi) atomicOr(&a.flags[0], 1u);
        const int64_t dst = (a.perm ? (int64_t)a.perm[fine] : fine) - a.io_base;
        if (a.out_f64) ((double*)a.out)[dst] = (double)xi;
        else ((float*)a.out)[dst] = (float)xi;
      }
      continue;
    }

    // ---- generate: wait for the neighbours' values, then publish this value
    T xi = T(0);
    int64_t dst = fine;
    if (valid && owner) {
      dst = a.perm ? (int64_t)a.perm[fine] : fine;
      xi = load_as<T>(a.xi, a.xi_f64, dst - a.io_base);
    }
    bool done = !valid;
    // neighbour values are fetched incrementally: a row is reloaded only while it still
    // holds the "not yet published" sentinel, so a waiting lane issues few loads
    T u[R];
    uint32_t pend = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      u[r] = T(0);
      if (valid && P * r + lane < k) pend |= 1u << r;
    }
    for (uint32_t it = 0;; ++it) {
      if (!done) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (pend & (1u << r)) {
            const T v = __hip_atomic_load(&a.vals[a.nbrs[pt * k + P * r + lane]], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
            if (!Sentinel<T>::is(v)) {
              u[r] = v;
              pend &= ~(1u << r);
            }
          }
        }
        const uint64_t bal = __ballot(pend == 0);
        const int gbase = (threadIdx.x & 63) & ~(P - 1);
        const uint64_t gmask = (P == 64) ? ~0ull : ((1ull << P) - 1ull);
        if (((bal >> gbase) & gmask) == gmask) {
          T v0 = T(0);
          if (incr) {
            v0 = bcast<P, 0>(u[0]);
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const int i = P * r + lane;
              if (i != 0 && i < k) u[r] -= v0;
            }
          }
          const T uk = forward(u);
          if (owner) {
            T v = v0 - uk + sqrt(skk) * xi;
            if (v != v) {
              atomicOr(&a.flags[0], 1u);
              v = __builtin_nan("");
            }
            __hip_atomic_store(&a.vals[fine], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (a.out) {
              if (a.out_f64) ((double*)a.out)[dst] = (double)v;
              else ((float*)a.out)[dst] = (float)v;
            }
          }
          done = true;
        }
      }
      if (__all(done)) break;
      if (it > (1u << 22)) {
        if (!done) {
          atomicOr(&a.flags[1], 1u);
          atomicOr(&a.flags[0], 1u);  // the whole output becomes NaN
          if (owner) __hip_atomic_store(&a.vals[fine], (T)__builtin_nan(""), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
  }
}

// Launch one configuration; defined in the refine_*.hip instantiation units.
template <typename T>
int launch_refine(const RefineArgs<T>& a, int kk_needed, hipStream_t s, size_t lds_bytes, int grid);

}  // namespace ggp
