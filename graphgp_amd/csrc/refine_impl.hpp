// refine_impl.hpp -- the fused refinement kernel template (see refine.hip for the design).
#pragma once

#include <type_traits>

#include "common.hpp"
#include "refine_dpp.hpp"

namespace ggp {

// Segment j of the interpolation table on [edge_j, edge_j+1).  FP32: c is the value at
// edge_j and the segment is c + (r - edge_j) s (r - edge_j is exact, so entries keep their
// relative precision); FP64: c is the intercept and the segment is c + s r.
template <typename T>
struct alignas(2 * sizeof(T)) TabE {
  T c, s;
};

template <typename T>
__device__ __forceinline__ T seg_value(const TabE<T>& e, T edge, T r) {
  if constexpr (sizeof(T) == 4) return e.c + (r - edge) * e.s;
  else return e.c + e.s * r;
}

struct TabMeta {
  int nseg;
  double x_first, x_last, f_first, f_last, f_zero, x1;
  float lb1, inv_step;
  float ka, kb;   // packed FP32 lookup: segment = clamp(trunc(log2(r^2) * ka + kb), 0, nseg)
  int log_exact;  // 1: the log2 guesses alone find the segment (log-spaced bins)
};

// Segment guess used by the FP64 fast lookup; the table check in the prep kernel uses the
// same expression, so "log_exact" certifies exactly what the kernel computes.
__device__ __forceinline__ int log_guess(float r2, float lb1, float inv_step) {
  const float lr = 0.5f * __builtin_amdgcn_logf(r2);
  return 1 + (int)((lr - lb1) * inv_step);
}

// The FP32 table, packed: entry j < nseg is segment j on [e_j, e_{j+1}) with value
// c + (r - e) s (c the value at e, so r - e is exact) = i + r s (i = c - e s, the
// intercept, rounded from FP64); entry nseg is the constant tail {f_last, 0, x_last,
// f_last} for r >= x_last.  The log-spaced lookup keeps only {i, s} in LDS, 8 bytes a
// segment (one 8-byte LDS read and one FMA per lookup, and twice as many distinct LDS
// banks among a wavefront's random segments as with 16-byte entries); the exact lookup
// keeps the whole entry.
struct alignas(16) TabP {
  float i, s, e, c;
};

// Segment of r (given r^2) in a log-spaced packed table: t = 1 + (log2(r) - log2(x_1)) /
// step folded into one FMA of log2(r^2); truncation equals floor wherever the clamp to
// [0, nseg] does not decide (t < 1 means r < x_1: segment 0; r >= x_last: the tail).
// The prep kernel certifies this very expression at every segment's middle.
__device__ __forceinline__ int packed_index(float r2, float ka, float kb, int nseg) {
  // clamped in FP32 (one v_med3), then truncated: -inf (r = 0) gives 0, NaN any valid index
  return (int)__builtin_amdgcn_fmed3f(__builtin_amdgcn_logf(r2) * ka + kb, 0.0f, (float)nseg);
}

// Packed FP32 lookups: the log-spaced guess, and an exact binary search over the edges
// (any monotone bins); both land on the tail entry for r >= x_last.
__device__ __forceinline__ float packed_lookup_fast(const float2* __restrict__ tab, float ka, float kb, int nseg,
                                                    float r2) {
  const float2 e = tab[packed_index(r2, ka, kb, nseg)];
  return __builtin_fmaf(__builtin_amdgcn_sqrtf(r2), e.y, e.x);
}
__device__ __forceinline__ float packed_lookup(const TabP* __restrict__ tab, int nseg, float r) {
  int lo = 0, hi = nseg;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].e <= r) lo = mid;
    else hi = mid - 1;
  }
  const TabP e = tab[lo];
  return e.c + (r - e.e) * e.s;
}

// kWeights: first phase of the two-phase generate (per-point weights and sigma, no waiting)
enum RefineMode : int { kGenerate = 0, kInverse = 1, kLogdet = 2, kWeights = 3 };

// Wavefronts of a workgroup that recompute flagged points in FP64, each with an LDS slab
// (flagged points are about 1% of a tile's 64 points; two slabs keep the fused kernels at
// two workgroups per CU with the 32 KB packed FP32 table)
constexpr int kFixWaves = 2;

// consecutive tiles one XCD group of the dataflow kernel takes at a time (at 16.8M points,
// k = 30, four lanes per point: 4 tiles 3.46 ms, 8 tiles 3.43, 16 tiles 3.39, 32 and 64
// tiles 3.36; eight lanes: 16 tiles 3.24, 32 tiles 3.22, 64 tiles 3.21)
constexpr uint32_t kGroupTiles = 32;
constexpr int kMaxGroups = 8;  // XCDs of an MI355X

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
  const TabP* tabp;  // FP32: the packed table (nseg + 1 entries), which the FP32 kernel uses
  const TabMeta* meta;
  uint32_t* ticket;
  uint32_t* flags;  // [0] NaN produced, [1] malformed graph / wait limit
  int mode;
  int increment;
  int64_t ntiles;
  int64_t io_base;  // subtracted from the caller-order index of xi (generate) / out (inverse)
  // FP64 recomputation of ill-conditioned points (FP32 modes): FP64 direct table, the
  // pivot threshold, a counter of recomputed points and the byte offset of the per-wave
  // LDS slabs after the table.  tab64 == nullptr switches it off.
  const TabE<double>* tab64;
  const TabMeta* meta64;
  double fix_tau;
  uint32_t* fix_count;
  int fix_lds_offset;
  // two-phase generate: weights w (n_ref x k) and sigma xi of every point (mean = w . u,
  // with u the neighbour values, in increments of neighbour 0 in the increment
  // formulation); the points needing the FP64 recomputation, their list slot (-1: none),
  // FP64 weights and sigma xi
  T* wts;
  T* sig;
  int32_t* fixslot;
  // output position of every slot (the caller-order index through the schedule and the
  // permutation), stored by the weights phase, which gathers it for xi anyway, so the
  // dataflow reads it in slot order instead of gathering the schedule and the permutation;
  // null: the dataflow gathers them
  int32_t* dpos;
  uint32_t* fix_list;
  uint32_t fix_cap;
  double* w64;
  double* sig64;
  // weights mode over a list: tile slot i is point list_in[i] for i < min(*fix_count,
  // fix_cap); outputs go to w64 / sig64 at slot i
  const uint32_t* list_in;
  // yield: while *yield is nonzero, workgroups from `keep` on stop taking tiles
  const uint32_t* yield;
  uint32_t keep;
  // XCD-aware tiles of the dataflow kernel: workgroup b belongs to group b % groups (the
  // device places consecutive workgroups on its XCDs in turn), and group g takes blocks of
  // kGroupTiles consecutive tiles, blocks g, g + groups, ..., from its own counter
  // (group_ticket[g * 32], one cache line each), so neighbouring tiles, which share
  // neighbour values, run behind one L2.  Every group takes its tiles in increasing order,
  // which keeps the dataflow's progress argument (the lowest unfinished tile is always
  // taken, and its dependencies are finished).
  uint32_t* group_ticket;
  int groups;
  // dataflow over stored weights in schedule order (or null): slot q takes refined point
  // sched[q], reads its weights and sigma at slot q (the weights phase then ran in slot
  // space), its neighbours' values at the slots nbv[q, .], and stores its value at slot
  // n0 + q
  const int32_t* sched;
  const int32_t* nbv;
  // a launch that runs only if bit 2 of *only_if is set (the FP64 fallback of the
  // two-phase generate); null: always
  const uint32_t* only_if;
  // workgroup slots a persistent launch leaves free (for the seed block running beside it),
  // and a cap on its grid (0: none)
  int reserve_wgs;
  int grid_limit;
};

// ------------------------------------------------------------------ helpers
// Sets bits of an error word from a per-point error path.  An all-NaN output reaches it from
// every point, and atomics on one address serialise at the memory (one per point made a
// 4M-point all-NaN generate take several milliseconds more); once the bits are set, the
// read finds them and no atomic is issued.
__device__ __forceinline__ void raise_flag(uint32_t* f, uint32_t bits) {
  if ((__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bits) != bits) atomicOr(f, bits);
}

// FP32, four lanes per point: the trailing update of column COL in column MM (slot MS,
// lane ML) for the lane's rows in slots MS..R-1, m[r][MM] -= m[r][COL] * (lane ML's
// m[MS][COL]), as one statement of DPP-fused FMAs.  Only the column's first statement
// (MM = COL + 1) starts with s_nop: every source of the column, m[MS][COL], is also a
// multiplicand of that first statement (its rows are all rows from slot (COL + 1) / P on),
// so the column scaling that wrote it precedes the first statement, and the first
// statement's s_nop and FMAs separate it from every later one.  The compiler sees the
// statements' writes, so its own DPP reads after them (the next pivot) get their wait
// states from the compiler.
template <int ML, int MS, int P, int COL, int MM, bool NOP, int S, int... Is>
__device__ __forceinline__ void update_rows_dpp(float (&m)[S], std::integer_sequence<int, Is...>) {
  FmacDpp<ML, sizeof...(Is), NOP>::run(m[P * MS * (MS + 1) / 2 + COL], m[P * (MS + Is) * (MS + Is + 1) / 2 + MM]...,
                                  m[P * (MS + Is) * (MS + Is + 1) / 2 + COL]...);
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Broadcast lane S of every P-lane group to the whole group.  The quad_perm sources are
// always inside the lane's own quad, so no lane reads an invalid one; with bound_ctrl set
// the DPP move needs no old value, and no copy into its destination first.  For 8-lane
// groups a second DPP move copies the source quad's value into the group's other quad:
// row_shr:4 into banks 1 and 3 (lanes 4-7 and 12-15 of each row read lanes 0-3 and 8-11),
// or row_shl:4 into banks 0 and 2; the lanes of the other banks keep the first move's value.
// A 16-lane group is a DPP row: row_newbcast copies its lane S to the row.
template <int P, int S>
__device__ __forceinline__ int bcast_i32(int x) {
  if constexpr (P == 1) {
    return x;
  } else if constexpr (P == 2) {
    constexpr int q = S | (S << 2) | ((2 + S) << 4) | ((2 + S) << 6);
    return __builtin_amdgcn_update_dpp(0, x, q, 0xF, 0xF, true);
  } else if constexpr (P == 4) {
    return __builtin_amdgcn_update_dpp(0, x, S * 0x55, 0xF, 0xF, true);
  } else if constexpr (P == 8) {
    if constexpr (S < 4) {
      const int t = __builtin_amdgcn_update_dpp(0, x, S * 0x55, 0xF, 0xF, true);
      return __builtin_amdgcn_update_dpp(t, t, 0x114, 0xF, 0xA, true);  // row_shr:4, banks 1, 3
    } else {
      const int t = __builtin_amdgcn_update_dpp(0, x, (S - 4) * 0x55, 0xF, 0xF, true);
      return __builtin_amdgcn_update_dpp(t, t, 0x104, 0xF, 0x5, true);  // row_shl:4, banks 0, 2
    }
  } else {
    static_assert(P == 16, "P must be 1, 2, 4, 8 or 16");
    return __builtin_amdgcn_update_dpp(0, x, 0x150 + S, 0xF, 0xF, true);  // row_newbcast:S
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

// Sum over a P-lane group, the butterfly x += x[lane ^ o] for o = 1, 2, 4, 8, all DPP moves
// inside a 16-lane row, with no LDS round trip: offsets 1 and 2 are quad permutations
// ([1,0,3,2], [2,3,0,1]); offset 4 is row_shl:4 into banks 0 and 2 (lanes read lane + 4)
// and row_shr:4 into banks 1 and 3 (lane - 4); offset 8 is row_ror:8.  Same additions in
// the same order as the plain shuffle butterfly.
template <int O>
__device__ __forceinline__ int xor_lane_i32(int x) {
  if constexpr (O == 1 || O == 2) {
    return __builtin_amdgcn_update_dpp(0, x, O == 1 ? 0xB1 : 0x4E, 0xF, 0xF, true);
  } else if constexpr (O == 4) {
    const int t = __builtin_amdgcn_update_dpp(0, x, 0x104, 0xF, 0x5, true);
    return __builtin_amdgcn_update_dpp(t, x, 0x114, 0xF, 0xA, true);
  } else {
    static_assert(O == 8, "offsets inside a 16-lane row only");
    return __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, true);
  }
}
template <int O>
__device__ __forceinline__ float xor_lane(float x) {
  return __int_as_float(xor_lane_i32<O>(__float_as_int(x)));
}
template <int O>
__device__ __forceinline__ double xor_lane(double x) {
  long long v = __double_as_longlong(x);
  int lo = xor_lane_i32<O>((int)(v & 0xffffffffll));
  int hi = xor_lane_i32<O>((int)(v >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int P, typename T>
__device__ __forceinline__ T group_sum(T x) {
  static_assert(P >= 1 && P <= 16, "groups of at most one DPP row");
  if constexpr (P >= 2) x += xor_lane<1>(x);
  if constexpr (P >= 4) x += xor_lane<2>(x);
  if constexpr (P >= 8) x += xor_lane<4>(x);
  if constexpr (P >= 16) x += xor_lane<8>(x);
  return x;
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

// Exact table lookup: interp(r, bins, vals) (refine.py:305) with the segment found by
// binary search over the edges (any monotone bins).
template <typename T>
__device__ __forceinline__ T table_lookup(const TabE<T>* __restrict__ tab, const T* __restrict__ edges, int nseg,
                                          T x_last, T f_last, T r) {
  if (r >= x_last) return f_last;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (edges[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return seg_value<T>(tab[lo], edges[lo], r);
}

// Lookup for log-spaced bins.  FP32: the log2 guess alone picks the segment (an 8-byte and
// a 4-byte LDS read); it is off by one only for r within ~1e-6 (relative) of a bin edge, where
// the neighbouring segment's line differs by a second-order term below FP32 rounding.
// FP64: the guess is biased low so the segment is j or j+1, chosen by comparing with the
// edge between them; the three reads are independent of each other (no branch).
template <typename T>
__device__ __forceinline__ T table_lookup_fast(const TabE<T>* __restrict__ tab, const T* __restrict__ edges,
                                               int nseg, T x1, T x_last, T f_last, float lb1, float inv_step, T r,
                                               T r2) {
  if constexpr (sizeof(T) == 4) {
    // rounding down: below the first positive bin (r < x1, and r = 0) the guess is <= 0,
    // so one clamp to [0, nseg - 1] also finds segment 0 (for j >= 1 this equals the
    // certified log_guess, whose argument is then positive)
    const float lr = 0.5f * __builtin_amdgcn_logf((float)r2);
    int j = 1 + (int)floorf((lr - lb1) * inv_step);
    j = max(0, min(j, nseg - 1));
    const T v = seg_value<T>(tab[j], edges[j], r);
    return (r >= x_last) ? f_last : v;
  } else {
    int j = log_guess((float)r2 * 0.99999f, lb1, inv_step);
    j = j < 1 ? 1 : (j > nseg - 1 ? nseg - 1 : j);
    j = (r < x1) ? 0 : j;
    const int j1 = j + 1 > nseg - 1 ? nseg - 1 : j + 1;
    const TabE<T> e0 = tab[j], e1 = tab[j1];
    const T b = edges[j + 1];
    const T v = (r >= b) ? (e1.c + e1.s * r) : (e0.c + e0.s * r);
    return (r >= x_last) ? f_last : v;
  }
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

// ------------------------------------------------------------------ FP64 recomputation
// One wavefront recomputes one point's conditional mean and standard deviation in FP64:
// lane i < k+1 owns row i of the joint covariance (neighbours first, fine point last) in
// an LDS slab, the table lookup is the exact FP64 one, the factorisation is a
// right-looking Cholesky with wavefront fences between its phases.  Used for the rare
// points whose FP32 factorisation loses accuracy (small pivots or a small Schur
// complement relative to the diagonal).  Runs before the point's neighbour values are
// needed (it depends on the geometry only): returns sigma and, if want_w, writes the
// FP64 weights w (mean = w . v[nbrs]) to wslot, so that only a dot product is left
// on the dataflow's critical path.
template <typename T, int D, bool FAST>
__device__ __forceinline__ double fix64(const RefineArgs<T>& a, int64_t pt, bool want_w, double* slab, int ld,
                                        double* wslot) {
  const int lane = threadIdx.x & 63;
  const int k = a.k, m = k + 1;
  const int64_t fine = a.n0 + pt;
  const TabMeta mt = *a.meta64;
  const double* edges = reinterpret_cast<const double*>(a.tab64 + mt.nseg);
  int64_t id = -1;
  if (lane < k) id = a.nbrs[pt * k + lane];
  else if (lane == k) id = fine;
  double c[D];
#pragma unroll
  for (int t = 0; t < D; ++t) c[t] = (id >= 0 && t < a.d) ? load_as<double>(a.pts, a.pts_f64, id * a.d + t) : 0.0;
  // pass 1: squared distances of this lane's row into the slab (no memory latency)
  for (int j = 0; j < m; ++j) {
    double cj[D];
#pragma unroll
    for (int t = 0; t < D; ++t) cj[t] = __shfl(c[t], j, 64);
    if (lane >= j && lane < m) slab[lane * ld + j] = pair_dist2<double, D>(c, cj);
  }
  // pass 2: table lookups in batches of 8 independent (global, L2-resident) reads
  constexpr int B = 8;
  for (int j0 = 0; j0 < m; j0 += B) {
    double v[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int j = j0 + q;
      v[q] = 0.0;
      if (j <= lane && lane < m) {
        const double r2 = slab[lane * ld + j];
        if constexpr (FAST)
          v[q] = table_lookup_fast<double>(a.tab64, edges, mt.nseg, mt.x1, mt.x_last, mt.f_last, mt.lb1,
                                           mt.inv_step, sqrt(r2), r2);
        else
          v[q] = table_lookup<double>(a.tab64, edges, mt.nseg, mt.x_last, mt.f_last, sqrt(r2));
      }
    }
#pragma unroll
    for (int q = 0; q < B; ++q)
      if (j0 + q <= lane && lane < m) slab[lane * ld + j0 + q] = v[q];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // right-looking Cholesky, lane i owns row i; after the scaled column is written, the
  // trailing update of a row runs in batches of 8 (all loads of a batch before its stores)
  for (int col = 0; col < k; ++col) {
    const double lc = sqrt(slab[col * ld + col]);
    const double lcol = (lane > col && lane < m) ? slab[lane * ld + col] / lc : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane == col) slab[col * ld + col] = lc;
    if (lane > col && lane < m) slab[lane * ld + col] = lcol;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int mm0 = col + 1; mm0 < m; mm0 += 8) {
      double lmc[8], row[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int mm = mm0 + q;
        const bool act = mm <= lane && lane < m;
        lmc[q] = act ? slab[mm * ld + col] : 0.0;
        row[q] = act ? slab[lane * ld + mm] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int mm = mm0 + q;
        if (mm <= lane && lane < m) slab[lane * ld + mm] = row[q] - lcol * lmc[q];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  const double sigma = sqrt(slab[k * ld + k]);
  if (want_w) {
    // weights w = L_A^{-T} z with z = L[k, :k]: mean = w . v[nbrs]
    double zj = (lane < k) ? slab[k * ld + lane] : 0.0;
    for (int i = k - 1; i >= 0; --i) {
      const double wi = __shfl(zj, i, 64) / slab[i * ld + i];
      if (lane == i) zj = wi;
      if (lane < i) zj -= slab[i * ld + lane] * wi;
    }
    if (lane < k) wslot[lane] = zj;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  return sigma;
}

// ------------------------------------------------------------------ the kernel
// MODE >= 0 fixes the mode at compile time (the code of the other modes is dropped);
// MODE = -1 takes it from a.mode.
template <typename T, int KK, int P, int D, bool FAST, int MODE = -1>
__global__ __launch_bounds__(256, 2) void refine_kernel(RefineArgs<T> a) {
  static_assert(KK % P == 0, "KK must be a multiple of P");
  const int mode = MODE >= 0 ? MODE : a.mode;
  constexpr int R = KK / P;
  constexpr int S = P * R * (R + 1) / 2;
  constexpr int GPT = 256 / P;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  constexpr bool PACKED = sizeof(T) == 4;  // FP32 reads the packed table
  TabE<T>* tab = reinterpret_cast<TabE<T>*>(smem_raw);
  TabP* ptab = reinterpret_cast<TabP*>(smem_raw);
  float2* qtab = reinterpret_cast<float2*>(smem_raw);  // FP32 log-spaced: {intercept, slope}
  __shared__ uint32_t s_tile;
  __shared__ double s_red[4];
  __shared__ int s_nfix;
  __shared__ int s_fixlist[GPT];
  __shared__ double s_fixsig[GPT];

  if (a.only_if != nullptr && (__hip_atomic_load(a.only_if, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2u) == 0)
    return;
  const TabMeta meta = *a.meta;
  // two variants are launched; the one that does not match the table exits at once
  if ((meta.log_exact != 0) != FAST) return;
  T* edges = reinterpret_cast<T*>(tab + meta.nseg);
  if constexpr (PACKED) {
    if constexpr (FAST) {
      for (int j = threadIdx.x; j <= meta.nseg; j += blockDim.x) qtab[j] = make_float2(a.tabp[j].i, a.tabp[j].s);
    } else {
      for (int j = threadIdx.x; j <= meta.nseg; j += blockDim.x) ptab[j] = a.tabp[j];
    }
  } else {
    const T* gedges = reinterpret_cast<const T*>(a.tab + meta.nseg);
    for (int j = threadIdx.x; j < meta.nseg; j += blockDim.x) tab[j] = a.tab[j];
    for (int j = threadIdx.x; j <= meta.nseg; j += blockDim.x) edges[j] = gedges[j];
  }
  __syncthreads();
  const T x1 = (T)meta.x1, x_last = (T)meta.x_last, f_last = (T)meta.f_last, f_zero = (T)meta.f_zero;
  const int nseg = meta.nseg;
  const float lb1 = meta.lb1, inv_step = meta.inv_step, ka = meta.ka, kb = meta.kb;
  auto lookup2 = [&](T r2) -> T {
    if constexpr (PACKED) {
      if constexpr (FAST) return packed_lookup_fast(qtab, ka, kb, nseg, r2);
      else return packed_lookup(ptab, nseg, sqrt(r2));
    } else if constexpr (FAST) {
      const T r = fast_sqrt<T>(r2);
      return table_lookup_fast<T>(tab, edges, nseg, x1, x_last, f_last, lb1, inv_step, r, r2);
    } else {
      return table_lookup<T>(tab, edges, nseg, x_last, f_last, sqrt(r2));
    }
  };

  const int lane = threadIdx.x & (P - 1);
  const int grp = threadIdx.x / P;
  const int k_arg = a.k;
  // The weights phase builds and factorises all KK rows and columns without run-time
  // column tests: rows i > k (padding, zero coordinates) and columns j > k hold garbage,
  // and in a right-looking Cholesky row i only ever updates rows and columns >= i, so
  // they never reach the k + 1 real rows; only the pivot scale of columns >= k (which
  // would rescale the fine point's Schur complement) and the recurrence's writes for
  // i >= k are masked.  Without the tests (one scalar compare and branch per column and
  // per update, a third of the instructions issued) the weights phase takes 11.5 ms
  // instead of 14.4 ms at 16.8M points, k = 30.  The fused modes, the exact-lookup
  // variant and FP64 keep the tests: without them the compiler's schedule of their single
  // large block spills hundreds of VGPRs.
  constexpr bool GUARD = !(MODE == kWeights && FAST && sizeof(T) == 4);
  constexpr bool LISTABLE = MODE == kWeights && sizeof(T) == 8;
  // the weights phases run in one formulation each (FP32: increments; FP64 list: plain),
  // fixed at compile time there
  const bool incr = MODE == kWeights ? sizeof(T) == 4 : a.increment != 0;
  const bool fixing = a.tab64 != nullptr;
  const T tau = (T)a.fix_tau;
  // FP64 recomputation: the first kFixWaves wavefronts have an LDS slab each
  double* slab = reinterpret_cast<double*>(smem_raw + a.fix_lds_offset) +
                 ((threadIdx.x >> 6) % kFixWaves) * (KK * (KK + 1));
  // FP64 weights of the recomputed points: one KK-slot per group of the workgroup
  double* wgrp =
      reinterpret_cast<double*>(smem_raw + a.fix_lds_offset) + kFixWaves * (KK * (KK + 1)) + (threadIdx.x / P) * KK;
  const int gbase = (threadIdx.x & 63) & ~(P - 1);
  const uint64_t gmask = (P == 64) ? ~0ull : ((1ull << P) - 1ull);

  // Tickets are taken one tile ahead: thread 0 takes the next tile's ticket at the start of
  // the current tile, so its round trip overlaps the tile's work (left uninitialised in the
  // other threads, so that no value meets the atomic's result at the branch's end, which
  // would make the wavefront wait for it there).  In the dataflow modes a workgroup then
  // holds its current tile and a higher one; the lowest unfinished tile is still held by a
  // workgroup whose earlier tiles are finished, so every dependency still completes.
  uint32_t next_ticket;
  if (threadIdx.x == 0) next_ticket = yield_now(a.yield, a.keep) ? 0xffffffffu : atomicAdd(a.ticket, 1u);
  for (;;) {
    __syncthreads();  // every thread is done with the previous tile (and its ticket read)
    if (threadIdx.x == 0) s_tile = next_ticket;
    __syncthreads();
    const int64_t tile = s_tile;
    if (tile >= a.ntiles) break;
    if (threadIdx.x == 0) next_ticket = yield_now(a.yield, a.keep) ? 0xffffffffu : atomicAdd(a.ticket, 1u);
    // k anew in every tile, opaque to the compiler (an SGPR it may not look through): the
    // unrolled column tests below are then scalar compares where they are used, instead of
    // being hoisted out of the tile loop as one lane mask each, which spill.
    int k = k_arg;
    asm volatile("" : "+s"(k));
    int64_t pt = tile * GPT + grp;
    bool valid = pt < a.n_ref;
    const int64_t slot_in = pt;
    // the list form is the FP64 recomputation's (compiled out of the FP32 kernels, which then
    // keep fewer argument registers live)
    if (LISTABLE && a.list_in) {
      uint32_t cnt = __hip_atomic_load(a.fix_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cnt > a.fix_cap) cnt = a.fix_cap;
      if (tile * GPT >= (int64_t)cnt) break;  // tiles are taken in order: nothing left
      valid = slot_in < (int64_t)cnt;
      pt = valid ? (int64_t)a.list_in[slot_in] : 0;
    }
    const int64_t fine = a.n0 + pt;
    // weights phases: xi of the point, issued now and used at the end (sigma xi is stored);
    // pt is the point's slot, whose graph position comes from the schedule
    T xw = T(0);
    if (mode == kWeights && valid) {
      const int64_t g = a.n0 + (a.sched != nullptr ? (int64_t)a.sched[pt] : pt);
      const int64_t gi = a.perm != nullptr ? (int64_t)a.perm[g] : g;
      xw = load_as<T>(a.xi, a.xi_f64, gi - a.io_base);
      if (a.dpos != nullptr && lane == 0) a.dpos[pt] = (int32_t)gi;
    }

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
    // original diagonal of this lane's rows (for the conditioning test): the table at r = 0
    // (f_zero), and in the increment formulation g(r_i0) + g(r_0i) - g(0) = 2 g(r_i0) for
    // i > 0 (the increment table is exactly 0 at r = 0)
    T d0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) d0[r] = (incr && P * r + lane != 0) ? g0[r] + g0[r] : f_zero;
    static_for<0, KK>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      constexpr int js = j % P, jr = j / P;
      if (!GUARD || j <= k) {
        T cj[D];
#pragma unroll
        for (int t = 0; t < D; ++t) cj[t] = bcast<P, js>(c[jr][t]);
        T gj = T(0);
        if (incr) gj = bcast<P, js>(g0[jr]);
        T tv[R];  // the table values of this column's entries
        if constexpr (PACKED && FAST) {
          // in three passes, so that the column's LDS reads are in flight together: the
          // segments, the reads, the interpolations
          T r2v[R];
          int ix[R];
#pragma unroll
          for (int r = jr; r < R; ++r) {
            r2v[r] = pair_dist2<T, D>(c[r], cj);
            ix[r] = packed_index(r2v[r], ka, kb, nseg);
          }
          float2 e[R];
#pragma unroll
          for (int r = jr; r < R; ++r) e[r] = qtab[ix[r]];
#pragma unroll
          for (int r = jr; r < R; ++r) tv[r] = __builtin_fmaf(__builtin_amdgcn_sqrtf(r2v[r]), e[r].y, e[r].x);
        } else {
#pragma unroll
          for (int r = jr; r < R; ++r) tv[r] = lookup2(pair_dist2<T, D>(c[r], cj));
        }
#pragma unroll
        for (int r = jr; r < R; ++r) {
          const int i = P * r + lane;
          T v = tv[r];
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
    bool bad = false;  // a pivot lost most of its diagonal: FP32 is not accurate here
    static_for<0, KK - 1>([&](auto cc) {
      constexpr int col = decltype(cc)::value;
      constexpr int cs = col / P, cl = col % P;
      if (!GUARD || col < k) {
        const T dg = bcast<P, cl>(m[P * cs * (cs + 1) / 2 + col]);
        const T rs = (GUARD || col < k) ? rsqrt_t<T>(dg) : T(1);
        if (lane == cl) rdl[cs] = rs;
#pragma unroll
        for (int r = cs; r < R; ++r) m[P * r * (r + 1) / 2 + col] *= rs;
        static_for<col + 1, KK>([&](auto mc) {
          constexpr int mm = decltype(mc)::value;
          constexpr int ms = mm / P, ml = mm % P;
          if (!GUARD || mm <= k) {
            if constexpr (sizeof(T) == 4 && P == 4) {
              update_rows_dpp<ml, ms, P, col, mm, mm == col + 1>(m, std::make_integer_sequence<int, R - ms>());
            } else {
              const T b = bcast<P, ml>(m[P * ms * (ms + 1) / 2 + col]);
#pragma unroll
              for (int r = ms; r < R; ++r) m[P * r * (r + 1) / 2 + mm] -= m[P * r * (r + 1) / 2 + col] * b;
            }
          }
        });
      }
    });

    // a pivot that lost most of its diagonal (FP32 is not accurate there): pivot_i =
    // 1 / rdl_i^2 < tau d0_i, tested by the lane that owns row i (a NaN pivot fails too)
    if (fixing) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (P * r + lane < k) bad = bad || !(rdl[r] * rdl[r] * (tau * d0[r]) <= T(1));
    }
    // Schur complement of the fine row: owned by lane k % P at slot k / P
    T skk = T(0);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (lane == q && P * r + q == k) skk = m[P * r * (r + 1) / 2 + P * r + q];
    const bool owner = (lane == (k % P));
    bool flagged = false;
    if (fixing) {
      T dk = T(1);
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (P * r + lane == k) dk = d0[r];
      const uint64_t bb = __ballot((owner && !(skk >= tau * dk)) || bad);
      flagged = valid && ((bb >> gbase) & gmask) != 0;
    }
    // FP64 recomputation of the tile's flagged points before any waiting: the flagged
    // groups are listed in LDS and the workgroup's four wavefronts take them in turn
    double sig64 = 1.0;
    if (fixing && mode != kWeights) {
      if (threadIdx.x == 0) s_nfix = 0;
      __syncthreads();
      if (flagged && lane == 0) s_fixlist[atomicAdd(&s_nfix, 1)] = grp;
      __syncthreads();
      const int nfix = s_nfix;
      for (int it = threadIdx.x >> 6; it < nfix && (threadIdx.x >> 6) < kFixWaves; it += kFixWaves) {
        const int g = s_fixlist[it];
        double* wslot = reinterpret_cast<double*>(smem_raw + a.fix_lds_offset) + kFixWaves * (KK * (KK + 1)) + g * KK;
        const double sg = fix64<T, D, FAST>(a, tile * GPT + g, mode != kLogdet, slab, KK + 1, wslot);
        if ((threadIdx.x & 63) == 0) {
          s_fixsig[g] = sg;
          atomicAdd(a.fix_count, 1u);
        }
      }
      __syncthreads();
      if (flagged) sig64 = s_fixsig[grp];
    }
    if (mode == kWeights) {
      // w = L_A^{-T} z with z = L[k, :k] (so that mean = z . L_A^{-1} u = w . u): with
      // w_k = -1, w_i = -(sum_{i<m<=k} L[m][i] w_m) / L[i][i], rows m spread over the
      // group's lanes, one group sum per column
      T wv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) wv[r] = (P * r + lane == k) ? T(-1) : T(0);
      static_for<0, KK - 1>([&](auto ic) {
        constexpr int i = KK - 2 - decltype(ic)::value;
        constexpr int is = i / P, il = i % P;
        if (!GUARD || i < k) {
          T part = T(0);
#pragma unroll
          for (int r = is; r < R; ++r) {
            const int row = P * r + lane;
            if (row > i && row <= k) part += m[P * r * (r + 1) / 2 + i] * wv[r];
          }
          part = group_sum<P>(part);
          const T wi = -part * bcast<P, il>(rdl[is]);
          if (lane == il && i < k) wv[is] = wi;
        }
      });
      if (LISTABLE && valid && a.list_in) {
        // FP64 recomputation of a listed point: FP64 weights and sigma at its list slot
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (P * r + lane < k) a.w64[slot_in * k + P * r + lane] = (double)wv[r];
        if (owner) a.sig64[slot_in] = (double)sqrt(skk) * (double)xw;
      } else if (valid) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (P * r + lane < k) a.wts[pt * k + P * r + lane] = wv[r];
        if (owner) {
          // a NaN here is not yet an error: a flagged point's FP32 sigma is replaced by its
          // FP64 recomputation, and the published values are checked in the last phase
          a.sig[pt] = sqrt(skk) * xw;
          int32_t slot = -1;
          if (flagged) {
            const uint32_t q = atomicAdd(a.fix_count, 1u);
            if (q < a.fix_cap) {
              a.fix_list[q] = (uint32_t)pt;
              slot = (int32_t)q;
            } else {
              atomicOr(&a.flags[1], 2u);  // more points to recompute than list slots
            }
          }
          a.fixslot[pt] = slot;
        }
      }
      continue;
    }

    // FP64 mean of a recomputed point from this lane's rows' raw neighbour values
    auto fixed_mean = [&](const T (&uv)[R]) -> double {
      double part = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (P * r + lane < k) part += wgrp[P * r + lane] * (double)uv[r];
      return group_sum<P>(part);
    };

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

    if (mode == kLogdet) {
      double ld = 0.0;
      if (valid && owner) ld = flagged ? log(sig64) : 0.5 * log((double)skk);
      if (valid && owner && ld != ld) raise_flag(&a.flags[0], 1u);
      // block reduction in a fixed order (deterministic)
      for (int o = 32; o > 0; o >>= 1) ld += __shfl_xor(ld, o, 64);
      if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = ld;
      __syncthreads();
      if (threadIdx.x == 0) a.partials[tile] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
      __syncthreads();
      continue;
    }

    if (mode == kInverse) {
      T u[R];
      T v0 = T(0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = P * r + lane;
        u[r] = T(0);
        if (valid && i < k) u[r] = a.vals[a.nbrs[pt * k + i]];
        if (valid && i == k) u[r] = a.vals[fine];
      }
      const double mean64 = flagged ? fixed_mean(u) : 0.0;
      if (incr) {
        v0 = bcast<P, 0>(u[0]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = P * r + lane;
          if (i != 0 && i <= k) u[r] -= v0;
        }
      }
      const T uk = forward(u);
      double xi = (double)(uk / sqrt(skk));
      if (flagged && owner) xi = ((double)a.vals[fine] - mean64) / sig64;
      if (valid && owner) {
        if (xi != xi) raise_flag(&a.flags[0], 1u);
        const int64_t dst = (a.perm ? (int64_t)a.perm[fine] : fine) - a.io_base;
        if (a.out_f64) ((double*)a.out)[dst] = xi;
        else ((float*)a.out)[dst] = (float)xi;
      }
      continue;
    }

    // ---- generate: wait for the neighbours' values, then publish this value
    T xi = T(0);
    int64_t dst = fine;
    if (valid && owner) {
      // in slot space (a.sched) the point's graph position comes from the schedule
      const int64_t g = a.sched != nullptr ? a.n0 + (int64_t)a.sched[pt] : fine;
      dst = a.perm ? (int64_t)a.perm[g] : g;
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
    auto publish = [&](T v) {
      if (v != v) {
        raise_flag(&a.flags[0], 1u);
        v = __builtin_nan("");
      }
      __hip_atomic_store(&a.vals[fine], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (a.out) {
        if (a.out_f64) ((double*)a.out)[dst] = (double)v;
        else ((float*)a.out)[dst] = (float)v;
      }
    };
    for (uint32_t it = 0;; ++it) {
      if (!done) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (pend & (1u << r)) {
            // a published value never changes, so a cached read that is not the sentinel
            // is final; only the sentinel needs the coherent re-read
            const int32_t nb = a.nbrs[pt * k + P * r + lane];
            T v = a.vals[nb];
            if (Sentinel<T>::is(v)) v = __hip_atomic_load(&a.vals[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!Sentinel<T>::is(v)) {
              u[r] = v;
              pend &= ~(1u << r);
            }
          }
        }
        const uint64_t bal = __ballot(pend == 0);
        if (((bal >> gbase) & gmask) == gmask) {
          if (flagged) {
            const double mean64 = fixed_mean(u);
            if (owner) publish((T)(mean64 + sig64 * (double)xi));
            done = true;
          } else {
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
            if (owner) publish(v0 - uk + sqrt(skk) * xi);
            done = true;
          }
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

// Last phase of the two-phase generate: the dataflow over the stored weights, P lanes per
// point, lane j % P gathers neighbour j.  A point waits until all its neighbours' values are
// published (sentinel = unpublished; a cached read first, the coherent re-read only while it
// is still the sentinel), then publishes v = [v0] + w . u + sigma xi (FP64 weights and sum
// for the recomputed points); the weights phases stored sigma xi per slot, so no xi is
// gathered here.  Persistent workgroups take tiles from an atomic ticket in index order, so
// every dependency is on a lower ticket held by a running workgroup.
//
// The loads of a row stay behind that row's own test: a variant that issued every row's
// loads at once (clamped addresses, one wait per batch, coherent re-reads of all pending rows
// per poll round) took 4.2-4.7 ms against 3.6 ms for this form at 16.8M points, k = 30; the
// phase is bound by memory traffic, and the extra coherent reads only added to it.
template <typename T, int P, int RM>
__global__ __launch_bounds__(256) void apply_kernel(RefineArgs<T> a) {
  constexpr int GPT = 256 / P;
  __shared__ uint32_t s_tile;
  const int lane = threadIdx.x & (P - 1);
  const int grp = threadIdx.x / P;
  const int k = a.k;
  const bool incr = a.increment != 0;
  const int gbase = (threadIdx.x & 63) & ~(P - 1);
  const uint64_t gmask = (P == 64) ? ~0ull : ((1ull << P) - 1ull);
  const uint32_t group = blockIdx.x % (uint32_t)a.groups;
  for (;;) {
    if (threadIdx.x == 0) {
      if (yield_now(a.yield, a.keep)) {
        s_tile = 0xffffffffu;
      } else {
        const uint32_t t = atomicAdd(&a.group_ticket[group * 32], 1u);
        const uint64_t tl = ((uint64_t)(t / kGroupTiles) * a.groups + group) * kGroupTiles + t % kGroupTiles;
        s_tile = tl < 0xffffffffull ? (uint32_t)tl : 0xffffffffu;
      }
    }
    __syncthreads();
    const int64_t tile = s_tile;
    __syncthreads();
    if (tile >= a.ntiles) break;
    const int64_t q = tile * GPT + grp;  // slot: the refined point's index, or its schedule slot
    const bool valid = q < a.n_ref;
    // the refined point (graph order); with dpos only its output position is needed
    const int64_t pt = (a.sched != nullptr && a.dpos == nullptr && valid) ? (int64_t)a.sched[q] : q;
    const int64_t fine = a.n0 + pt;  // graph position: output through perm
    const int64_t store = a.n0 + q;  // value slot; weights, sigma xi and fixslot are slot-indexed
    const int32_t slot = valid ? a.fixslot[q] : -1;
    int32_t nb[RM];
    double wd[RM];
    uint32_t pend = 0;
#pragma unroll
    for (int r = 0; r < RM; ++r) {
      const int j = P * r + lane;
      nb[r] = 0;
      wd[r] = 0.0;
      if (valid && j < k) {
        nb[r] = a.nbv != nullptr ? a.nbv[q * k + j] : a.nbrs[pt * k + j];
        wd[r] = slot >= 0 ? a.w64[(int64_t)slot * k + j] : (double)a.wts[q * k + j];
        pend |= 1u << r;
      }
    }
    T sx = T(0);
    int64_t dst = fine;
    const bool owner = lane == 0;
    if (valid && owner) {
      dst = a.dpos != nullptr ? (int64_t)a.dpos[q] : a.perm ? (int64_t)a.perm[fine] : fine;
      sx = slot >= 0 ? (T)0 : a.sig[q];
    }
    T u[RM];
#pragma unroll
    for (int r = 0; r < RM; ++r) u[r] = T(0);
    bool done = !valid;
    for (uint32_t it = 0;; ++it) {
      if (!done) {
#pragma unroll
        for (int r = 0; r < RM; ++r) {
          if (pend & (1u << r)) {
            T v = a.vals[nb[r]];
            if (Sentinel<T>::is(v)) v = __hip_atomic_load(&a.vals[nb[r]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!Sentinel<T>::is(v)) {
              u[r] = v;
              pend &= ~(1u << r);
            }
          }
        }
        const uint64_t bal = __ballot(pend == 0);
        if (((bal >> gbase) & gmask) == gmask) {
          T v;
          if (slot >= 0) {
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < RM; ++r) part += wd[r] * (double)u[r];
            // a group's lanes finish together, so the group sum's DPP sources are active
            part = group_sum<P>(part);
            v = (T)(part + a.sig64[slot]);
          } else {
            // neighbour 0 is lane 0's first slot
            const T v0 = incr ? bcast<P, 0>(u[0]) : T(0);
            T part = T(0);
#pragma unroll
            for (int r = 0; r < RM; ++r) {
              const int j = P * r + lane;
              const T uj = (incr && j != 0) ? u[r] - v0 : u[r];
              part += (T)wd[r] * uj;
            }
            part = group_sum<P>(part);
            v = v0 + part + sx;
          }
          if (owner) {
            if (v != v) {
              raise_flag(&a.flags[0], 1u);
              v = __builtin_nan("");
            }
            __hip_atomic_store(&a.vals[store], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
          atomicOr(&a.flags[0], 1u);
          if (owner) __hip_atomic_store(&a.vals[store], (T)__builtin_nan(""), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
