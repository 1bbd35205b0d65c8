// graphgp_oracle.cpp -- CPU restatement of the reference GraphGP hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker (or as the timed
// CPU baseline).  The product path (graphgp_amd) never links, loads or calls it.
//
// Every function restates the reference algorithm, read as text, with the reference
// file:line it follows.  Arithmetic is in the dtype the reference would use: FP64 for
// the refinement (the reference's tests run with jax_enable_x64), and the points'
// own dtype (float or double) for the tree build and the kNN search, whose results
// are integer permutations that depend on the exact rounding of the coordinate
// arithmetic.  Compiled with -ffp-contract=off so that a*b+c is never fused.
//
// Assumptions that could only be confirmed by running JAX (absent here):
//  * squared distances are summed left to right: ((d0^2 + d1^2) + d2^2);
//  * lax.sort's float canonicalisation: -0.0 sorts equal to +0.0, NaN after all;
//  * jnp.interp = fp[i-1] + (x - xp[i-1]) / dx * df with i = searchsorted(xp, x,
//    'right') clipped to [1, n-1], fp[0] below xp[0] and fp[-1] above xp[-1].

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- tree helpers
// tree.py:187-208 (_compute_left/_compute_right/_compute_parent).  The reference
// takes floor(log2(c+1)) through frexp; the exact intent (the commented clz at
// tree.py:189) is used here, which is identical for every id below 2^53.
inline int64_t tree_level(int64_t c) { return 63 - __builtin_clzll((uint64_t)(c + 1)); }
inline int64_t tree_left(int64_t c) { return c + (int64_t(1) << tree_level(c)); }
inline int64_t tree_right(int64_t c) { return c + (int64_t(2) << tree_level(c)); }
inline int64_t tree_parent(int64_t c) {
  if (c == 0) return -1;
  int64_t lvl = tree_level(c);
  int64_t n_above = (int64_t(1) << lvl) - 1;
  int64_t npl = int64_t(1) << (lvl - 1);
  return c < n_above + npl ? c - npl : c - 2 * npl;
}

// lax.sort float key (see header): -0.0 == +0.0, NaN greatest.
template <typename T>
inline bool key_less(T a, T b) {
  bool an = std::isnan(a), bn = std::isnan(b);
  if (an || bn) return !an && bn;
  return a < b;
}
template <typename T>
inline bool key_eq(T a, T b) {
  bool an = std::isnan(a), bn = std::isnan(b);
  if (an || bn) return an && bn;
  return a == b;
}

// tree.py:211-234 (_update_nodes)
void update_nodes(std::vector<int64_t>& nodes, int64_t level) {
  const int64_t n = (int64_t)nodes.size();
  const int64_t n_above = (int64_t(1) << level) - 1;
  const int64_t n_level = int64_t(1) << level;
  const int64_t n_rem = n - n_above;
  const int64_t q = n_rem / n_level, r = n_rem % n_level;
  for (int64_t idx = 0; idx < n; ++idx) {
    int64_t nd = nodes[idx];
    int64_t i = nd - n_above;
    int64_t mid = (i < r ? i * (q + 1) + (q + 1) / 2 : r * (q + 1) + (i - r) * q + q / 2) + n_above;
    if (idx < n_above || idx == mid) continue;
    nodes[idx] = idx < mid ? nd + n_level : nd + 2 * n_level;
  }
}

// tree.py:153-184 (_build_tree): per level, split dim = first argmax of per-node
// extent (empty / single nodes give 0), then a sort by (node, value, position).
template <typename T>
void build_tree(const T* pts_in, int64_t n, int d, T* pts_out, int32_t* split_out, int64_t* idx_out) {
  std::vector<T> pts(pts_in, pts_in + n * d), tmp(n * d);
  std::vector<int64_t> nodes(n, 0), indices(n), order(n), tmpi(n), tmpn(n);
  std::vector<int32_t> split(n, -1), new_split(n);
  std::iota(indices.begin(), indices.end(), 0);
  std::vector<T> vmax(n * d), vmin(n * d), val(n);
  int levels = 0;
  while ((n >> levels) != 0) ++levels;  // int.bit_length()
  const T inf = std::numeric_limits<T>::infinity();
  for (int level = 0; level < levels; ++level) {
    std::fill(vmax.begin(), vmax.end(), -inf);
    std::fill(vmin.begin(), vmin.end(), inf);
    for (int64_t e = 0; e < n; ++e)
      for (int j = 0; j < d; ++j) {
        T x = pts[e * d + j];
        T& mx = vmax[nodes[e] * d + j];
        T& mn = vmin[nodes[e] * d + j];
        if (x > mx) mx = x;
        if (x < mn) mn = x;
      }
    const int64_t n_above = (int64_t(1) << level) - 1;
    for (int64_t nd = 0; nd < n; ++nd) {
      int best = 0;
      T bestv = vmax[nd * d] - vmin[nd * d];
      for (int j = 1; j < d; ++j) {
        T ext = vmax[nd * d + j] - vmin[nd * d + j];
        if (ext > bestv) { bestv = ext; best = j; }
      }
      if (nd >= n_above) split[nd] = best;
    }
    for (int64_t e = 0; e < n; ++e) val[e] = pts[e * d + split[nodes[e]]];
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      if (nodes[a] != nodes[b]) return nodes[a] < nodes[b];
      if (!key_eq(val[a], val[b])) return key_less(val[a], val[b]);
      return a < b;
    });
    for (int64_t e = 0; e < n; ++e) {
      int64_t s = order[e];
      std::memcpy(&tmp[e * d], &pts[s * d], sizeof(T) * d);
      tmpi[e] = indices[s];
      tmpn[e] = nodes[s];
    }
    pts.swap(tmp);
    indices.swap(tmpi);
    nodes.swap(tmpn);
    update_nodes(nodes, level);
  }
  std::memcpy(pts_out, pts.data(), sizeof(T) * n * d);
  for (int64_t i = 0; i < n; ++i) { split_out[i] = split[i]; idx_out[i] = indices[i]; }
}

// The same sort as build_tree above, decomposed.  Before each level's sort, every node's
// members form one contiguous run of the array: the array is sorted by the previous
// level's node ids, and _update_nodes (tree.py:211-234) splits a node's run into a left
// part, its median and a right part.  Sorting by (node, value, array position) therefore
// places each run, sorted by (value, position), at the offset given by the number of
// members of the smaller node ids, and the runs can be sorted independently (in
// parallel).  The split dimension per node is the same first argmax of the extents, over
// the same members.  Results equal build_tree's bit for bit (tests compare both).
template <typename T>
void build_tree_runs(const T* pts_in, int64_t n, int d, T* pts_out, int32_t* split_out, int64_t* idx_out) {
  std::vector<T> pts(pts_in, pts_in + n * d), tmp(n * d);
  std::vector<int64_t> nodes(n, 0), indices(n), tmpi(n), tmpn(n), count(n + 1);
  std::vector<int32_t> split(n, -1);
  std::iota(indices.begin(), indices.end(), 0);
  int levels = 0;
  while ((n >> levels) != 0) ++levels;
  const T inf = std::numeric_limits<T>::infinity();
  std::vector<int64_t> run_start;
  for (int level = 0; level < levels; ++level) {
    const int64_t n_above = (int64_t(1) << level) - 1;
    // runs of equal node id
    run_start.clear();
    for (int64_t e = 0; e < n; ++e)
      if (e == 0 || nodes[e] != nodes[e - 1]) run_start.push_back(e);
    const int64_t nruns = (int64_t)run_start.size();
    run_start.push_back(n);
    // nodes >= n_above without members get split 0 (empty extents); the others below
#pragma omp parallel for schedule(static)
    for (int64_t nd = n_above; nd < n; ++nd) split[nd] = 0;
    std::fill(count.begin(), count.end(), 0);
    for (int64_t r = 0; r < nruns; ++r) count[nodes[run_start[r]] + 1] += run_start[r + 1] - run_start[r];
    for (int64_t nd = 0; nd < n; ++nd) count[nd + 1] += count[nd];  // count[nd] = first slot of node nd
#pragma omp parallel
    {
      std::vector<std::pair<T, int64_t>> buf;
#pragma omp for schedule(dynamic, 1)
      for (int64_t r = 0; r < nruns; ++r) {
        const int64_t a = run_start[r], b = run_start[r + 1], nd = nodes[a];
        int best = 0;
        if (nd >= n_above) {
          T bestv = T(0);
          for (int j = 0; j < d; ++j) {
            T mx = -inf, mn = inf;
            for (int64_t e = a; e < b; ++e) {
              const T x = pts[e * d + j];
              if (x > mx) mx = x;
              if (x < mn) mn = x;
            }
            const T ext = mx - mn;
            if (j == 0) {
              bestv = ext;
            } else if (ext > bestv) {
              bestv = ext;
              best = j;
            }
          }
          split[nd] = best;
        }
        const int dim = split[nd] < 0 ? 0 : split[nd];
        buf.resize(b - a);
        for (int64_t e = a; e < b; ++e) buf[e - a] = {pts[e * d + dim], e};
        if (b - a > 1)
          std::sort(buf.begin(), buf.end(), [](const std::pair<T, int64_t>& x, const std::pair<T, int64_t>& y) {
            if (!key_eq(x.first, y.first)) return key_less(x.first, y.first);
            return x.second < y.second;
          });
        const int64_t dst = count[nd];
        for (int64_t q = 0; q < b - a; ++q) {
          const int64_t src = buf[q].second;
          std::memcpy(&tmp[(dst + q) * d], &pts[src * d], sizeof(T) * d);
          tmpi[dst + q] = indices[src];
          tmpn[dst + q] = nd;
        }
      }
    }
    pts.swap(tmp);
    indices.swap(tmpi);
    nodes.swap(tmpn);
    update_nodes(nodes, level);
  }
  std::memcpy(pts_out, pts.data(), sizeof(T) * n * d);
  for (int64_t i = 0; i < n; ++i) {
    split_out[i] = split[i];
    idx_out[i] = indices[i];
  }
}

template <typename T>
inline T sqdist(const T* a, const T* b, int d) {
  T s = (a[0] - b[0]) * (a[0] - b[0]);
  for (int j = 1; j < d; ++j) {
    T t = a[j] - b[j];
    s = s + t * t;
  }
  return s;
}

// tree.py:79-150 (_single_query_neighbors + _traverse_tree).
template <typename T>
void query_one(const T* pts, const int32_t* split, int d, int64_t qi, int64_t max_index, int k,
               int32_t* out, int64_t* visited) {
  std::vector<int64_t> nb(k, -1);
  std::vector<T> d2(k, std::numeric_limits<T>::infinity());
  T r2 = std::numeric_limits<T>::infinity();
  const T* q = pts + qi * d;
  int64_t cur = 0, prev = -1, nvis = 0;
  while (cur != -1) {
    int64_t parent = tree_parent(cur);
    if (prev == parent) {  // update_func, tree.py:81-95
      ++nvis;
      T s = sqdist(q, pts + cur * d, d);
      int m = 0;
      for (int j = 1; j < k; ++j)
        if (d2[j] > d2[m]) m = j;  // first argmax
      if (s < d2[m]) { nb[m] = cur; d2[m] = s; }
      r2 = d2[0];
      for (int j = 1; j < k; ++j)
        if (d2[j] > r2) r2 = d2[j];
    }
    int dim = split[cur];
    T sd = q[dim] - pts[cur * d + dim];
    int64_t l = tree_left(cur), r = tree_right(cur);
    int64_t nearc = sd < 0 ? l : r, farc = sd < 0 ? r : l;
    bool far_in = sd * sd <= r2;
    int64_t next;
    if (prev == nearc || (prev == parent && nearc >= max_index))
      next = (farc < max_index && far_in) ? farc : parent;
    else
      next = prev == parent ? nearc : parent;
    prev = cur;
    cur = next;
  }
  // tree.py:102-105: sort by (fp distance, index)
  std::vector<std::pair<T, int64_t>> fin(k);
  for (int j = 0; j < k; ++j) {
    int64_t idx = nb[j];
    int64_t src = idx < 0 ? idx + 0 : idx;  // -1 never occurs when n0 >= k
    T dist = src >= 0 ? std::sqrt(sqdist(pts + src * d, q, d)) : std::numeric_limits<T>::infinity();
    fin[j] = {dist, idx};
  }
  std::sort(fin.begin(), fin.end(), [](const std::pair<T, int64_t>& a, const std::pair<T, int64_t>& b) {
    if (!key_eq(a.first, b.first)) return key_less(a.first, b.first);
    return a.second < b.second;
  });
  for (int j = 0; j < k; ++j) out[j] = (int32_t)fin[j].second;
  if (visited) *visited = nvis;
}

// ---------------------------------------------------------------- covariance
// refine.py:298-305 (cov_lookup = jnp.interp) -- formula documented in the header.
inline double interp(double x, const double* xp, const double* fp, int n) {
  if (x < xp[0]) return fp[0];
  if (x > xp[n - 1]) return fp[n - 1];
  // searchsorted side='right' (the first i with xp[i] > x), as a binary search without
  // data-dependent branches: the same index as std::upper_bound
  const double* base = xp;
  int len = n;
  while (len > 1) {
    const int half = len >> 1;
    base = (base[half] <= x) ? base + half : base;
    len -= half;
  }
  int i = (int)(base - xp) + (*base <= x ? 1 : 0);
  if (i < 1) i = 1;
  if (i > n - 1) i = n - 1;
  double df = fp[i] - fp[i - 1], dx = xp[i] - xp[i - 1], delta = x - xp[i - 1];
  constexpr double eps = 4.930380657631324e-32;  // np.spacing(np.finfo(float).eps)
  if (std::fabs(dx) <= eps) return fp[i - 1];
  return fp[i - 1] + (delta / dx) * df;
}

// refine.py:288-295: r = ||a - b|| (sqrt of summed squares), then the table lookup.
inline double pdist(const double* a, const double* b, int d) {
  double s = 0.0;
  for (int j = 0; j < d; ++j) { double t = a[j] - b[j]; s = s + t * t; }
  return std::sqrt(s);
}

// In-place lower Cholesky of an n x n row-major matrix (jnp.linalg.cholesky).  On failure
// (a pivot that is not positive, or NaN input) the whole factor is NaN: JAX's CPU cholesky
// replaces the factor by NaN when LAPACK's potrf reports an error.
void cholesky(double* A, int n) {
  bool failed = false;
  for (int j = 0; j < n && !failed; ++j) {
    double s = A[j * n + j];
    for (int m = 0; m < j; ++m) s -= A[j * n + m] * A[j * n + m];
    if (!(s > 0.0)) {
      failed = true;
      break;
    }
    double djj = std::sqrt(s);
    A[j * n + j] = djj;
    for (int i = j + 1; i < n; ++i) {
      double t = A[i * n + j];
      for (int m = 0; m < j; ++m) t -= A[i * n + m] * A[j * n + m];
      A[i * n + j] = t / djj;
    }
    for (int i = 0; i < j; ++i) A[i * n + j] = 0.0;
  }
  if (failed)
    for (int64_t e = 0; e < (int64_t)n * n; ++e) A[e] = std::numeric_limits<double>::quiet_NaN();
}

// log|det A| of an n x n row-major matrix by Gaussian elimination with partial pivoting
// (LAPACK getrf's rule: the first row of largest |a_ij| in the column), which is what
// jnp.linalg.slogdet computes: the sum of log|u_jj|, -inf for a singular matrix, NaN for
// NaN input.  A is overwritten.
double lu_logabsdet(double* A, int64_t n) {
  double s = 0.0;
  for (int64_t j = 0; j < n; ++j) {
    int64_t p = j;
    double best = std::fabs(A[j * n + j]);
    for (int64_t i = j + 1; i < n; ++i) {
      const double a = std::fabs(A[i * n + j]);
      if (a > best) {
        best = a;
        p = i;
      }
    }
    if (best != best) return std::numeric_limits<double>::quiet_NaN();
    if (best == 0.0) return -std::numeric_limits<double>::infinity();
    if (p != j)
      for (int64_t c = 0; c < n; ++c) std::swap(A[j * n + c], A[p * n + c]);
    const double piv = A[j * n + j];
    s += std::log(std::fabs(piv));
    for (int64_t i = j + 1; i < n; ++i) {
      const double l = A[i * n + j] / piv;
      if (l == 0.0) continue;
      for (int64_t c = j + 1; c < n; ++c) A[i * n + c] -= l * A[j * n + c];
    }
  }
  return s;
}

struct Table { const double* bins; const double* vals; int n; };

// Joint covariance of (neighbours..., fine point), refine.py:129-132.
void joint_cov(const double* pts, int d, const int64_t* nbr, int k, int64_t fine, const Table& t, double* K) {
  const int n = k + 1;
  for (int i = 0; i < n; ++i) {
    const double* a = pts + (i < k ? nbr[i] : fine) * d;
    for (int j = 0; j <= i; ++j) {
      const double* b = pts + (j < k ? nbr[j] : fine) * d;
      double v = interp(pdist(a, b, d), t.bins, t.vals, t.n);
      K[i * n + j] = v;
      K[j * n + i] = v;
    }
  }
}

// fast_jit form, refine.py:133-135: L = chol(K); mean_vec = solve(L_A^T, L[k,:k]); std = L[k,k].
void conditional_weights(double* K, int k, double* mean_vec, double* std_out) {
  const int n = k + 1;
  cholesky(K, n);
  for (int i = 0; i < k; ++i) mean_vec[i] = K[k * n + i];
  for (int i = k - 1; i >= 0; --i) {  // back substitution with L_A^T
    double s = mean_vec[i];
    for (int j = i + 1; j < k; ++j) s -= K[j * n + i] * mean_vec[j];
    mean_vec[i] = s / K[i * n + i];
  }
  *std_out = K[k * n + k];
}

}  // namespace

extern "C" {

int ora_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void ora_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// ------------------------------------------------------------------ tree.py
void ora_build_tree_f32(const float* p, int64_t n, int d, float* po, int32_t* s, int64_t* idx) {
  build_tree_runs<float>(p, n, d, po, s, idx);
}
void ora_build_tree_f64(const double* p, int64_t n, int d, double* po, int32_t* s, int64_t* idx) {
  build_tree_runs<double>(p, n, d, po, s, idx);
}
// the literal whole-array sort per level (tree.py:153-184), for pinning build_tree_runs
void ora_build_tree_literal_f32(const float* p, int64_t n, int d, float* po, int32_t* s, int64_t* idx) {
  build_tree<float>(p, n, d, po, s, idx);
}
void ora_build_tree_literal_f64(const double* p, int64_t n, int d, double* po, int32_t* s, int64_t* idx) {
  build_tree<double>(p, n, d, po, s, idx);
}

// tree.py:66-76 query_neighbors (vmapped traversal); preceding = query_idx == max_idx.
void ora_query_neighbors_f32(const float* p, const int32_t* split, int64_t n, int d, const int64_t* qidx,
                             const int64_t* midx, int64_t nq, int k, int32_t* out, int64_t* visited) {
  (void)n;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < nq; ++i)
    query_one<float>(p, split, d, qidx[i], midx[i], k, out + i * k, visited ? visited + i : nullptr);
}
void ora_query_neighbors_f64(const double* p, const int32_t* split, int64_t n, int d, const int64_t* qidx,
                             const int64_t* midx, int64_t nq, int k, int32_t* out, int64_t* visited) {
  (void)n;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < nq; ++i)
    query_one<double>(p, split, d, qidx[i], midx[i], k, out + i * k, visited ? visited + i : nullptr);
}

// Brute force for tests/test_tree.py:44-59's property at any size: for each query q (a
// tree position), the k smallest squared distances among points [0, q) (the preceding
// points), with the kNN's own FP32 arithmetic (sqdist, separately rounded), ordered by
// (distance^2, index).  One pass over the points serves all queries: every thread keeps
// a k-list per query for its share of the points, and the lists are merged at the end.
void ora_brute_preceding_knn_f32(const float* p, int64_t n, int d, const int64_t* queries, int64_t nq, int k,
                                 float* out_d2, int64_t* out_idx) {
  using Item = std::pair<float, int64_t>;
  auto less = [](const Item& a, const Item& b) { return a.first < b.first || (a.first == b.first && a.second < b.second); };
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  std::vector<std::vector<Item>> lists((size_t)nthreads * nq);
  int64_t qmax = 0;
  for (int64_t j = 0; j < nq; ++j) qmax = std::max(qmax, queries[j]);
  const int64_t limit = std::min(n, qmax);
#pragma omp parallel
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    std::vector<Item>* mine = &lists[(size_t)t * nq];
    std::vector<float> worst(nq, std::numeric_limits<float>::infinity());
    std::vector<int64_t> worst_idx(nq, std::numeric_limits<int64_t>::max());
#pragma omp for schedule(static)
    for (int64_t i = 0; i < limit; ++i) {
      const float* x = p + i * d;
      for (int64_t j = 0; j < nq; ++j) {
        if (i >= queries[j]) continue;
        const float s = sqdist(x, p + queries[j] * d, d);
        if (s > worst[j] || (s == worst[j] && i > worst_idx[j])) continue;
        std::vector<Item>& L = mine[j];
        if ((int)L.size() < k) {
          L.push_back({s, i});
          std::push_heap(L.begin(), L.end(), less);
        } else {
          std::pop_heap(L.begin(), L.end(), less);
          L.back() = {s, i};
          std::push_heap(L.begin(), L.end(), less);
        }
        if ((int)L.size() == k) {
          worst[j] = L.front().first;
          worst_idx[j] = L.front().second;
        }
      }
    }
  }
  for (int64_t j = 0; j < nq; ++j) {
    std::vector<Item> all;
    for (int t = 0; t < nthreads; ++t) all.insert(all.end(), lists[(size_t)t * nq + j].begin(), lists[(size_t)t * nq + j].end());
    std::sort(all.begin(), all.end(), less);
    for (int c = 0; c < k; ++c) {
      out_d2[j * k + c] = c < (int)all.size() ? all[c].first : std::numeric_limits<float>::infinity();
      out_idx[j * k + c] = c < (int)all.size() ? all[c].second : -1;
    }
  }
}

// ------------------------------------------------------------------ graph.py
// graph.py:101-120 compute_depths.  The reference iterates depth[i] = 1 + max
// depth[nbrs] to a fixed point; for a topologically ordered graph that fixed point
// is the longest-path depth, computed here in one forward pass.  Returns 0, or -1
// if a neighbour does not precede its point (where the fixed point is not unique).
int ora_compute_depths(const int32_t* nbrs, int64_t n_ref, int k, int64_t n0, int32_t* depths) {
  for (int64_t i = 0; i < n0; ++i) depths[i] = 0;
  for (int64_t r = 0; r < n_ref; ++r) {
    int32_t m = -1;
    for (int j = 0; j < k; ++j) {
      int64_t nb = nbrs[r * k + j];
      if (nb < 0 || nb >= n0 + r) return -1;
      m = std::max(m, depths[nb]);
    }
    depths[n0 + r] = m + 1;
  }
  return 0;
}

// graph.py:123-137 order_by_depth: stable argsort by depth, gathers, neighbour remap.
void ora_order_by_depth(const double* pts, int64_t n, int d, const int64_t* indices, const int32_t* nbrs, int k,
                        const int32_t* depths, double* pts_o, int64_t* idx_o, int32_t* nbrs_o, int32_t* depths_o,
                        int64_t n0) {
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return depths[a] < depths[b]; });
  std::vector<int64_t> inv(n);
  for (int64_t i = 0; i < n; ++i) inv[order[i]] = i;
  for (int64_t i = 0; i < n; ++i) {
    std::memcpy(pts_o + i * d, pts + order[i] * d, sizeof(double) * d);
    idx_o[i] = indices[order[i]];
    depths_o[i] = depths[order[i]];
  }
  for (int64_t r = 0; r < n - n0; ++r) {
    int64_t src = order[n0 + r] - n0;
    for (int j = 0; j < k; ++j) nbrs_o[r * k + j] = (int32_t)inv[nbrs[src * k + j]];
  }
}

// graph.py:95-96: offsets = searchsorted(depths, 1..max+1); returns the count.
int64_t ora_offsets(const int32_t* depths_sorted, int64_t n, int64_t* offsets, int64_t cap) {
  int32_t mx = 0;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, depths_sorted[i]);
  int64_t cnt = (int64_t)mx + 1;
  for (int64_t b = 1; b <= mx + 1 && b - 1 < cap; ++b)
    offsets[b - 1] = std::lower_bound(depths_sorted, depths_sorted + n, (int32_t)b) - depths_sorted;
  return cnt;
}

// ------------------------------------------------------------------ refine.py
double ora_interp(double x, const double* xp, const double* fp, int n) { return interp(x, xp, fp, n); }

// refine.py:288-295 compute_cov_matrix for explicit point sets.
void ora_cov_matrix(const double* a, int64_t na, const double* b, int64_t nb, int d, const double* bins,
                    const double* vals, int nbins, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < na; ++i)
    for (int64_t j = 0; j < nb; ++j) out[i * nb + j] = interp(pdist(a + i * d, b + j * d, d), bins, vals, nbins);
}

// refine.py:59-76 generate_dense: values = chol(K) @ xi.  Also returns L (n x n) if asked.
void ora_generate_dense(const double* pts, int64_t n, int d, const double* bins, const double* vals, int nbins,
                        const double* xi, double* values, double* L_out) {
  std::vector<double> K(n * n);
  ora_cov_matrix(pts, n, pts, n, d, bins, vals, nbins, K.data());
  cholesky(K.data(), (int)n);
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t j = 0; j <= i; ++j) s += K[i * n + j] * xi[j];
    values[i] = s;
  }
  if (L_out) std::memcpy(L_out, K.data(), sizeof(double) * n * n);
}

// refine.py:186-195 generate_dense_inv: xi = solve(L, values).
void ora_generate_dense_inv(const double* pts, int64_t n, int d, const double* bins, const double* vals, int nbins,
                            const double* values, double* xi) {
  std::vector<double> K(n * n);
  ora_cov_matrix(pts, n, pts, n, d, bins, vals, nbins, K.data());
  cholesky(K.data(), (int)n);
  for (int64_t i = 0; i < n; ++i) {
    double s = values[i];
    for (int64_t j = 0; j < i; ++j) s -= K[i * n + j] * xi[j];
    xi[i] = s / K[i * n + i];
  }
}

// refine.py:243-248 generate_dense_logdet = slogdet(K)[1] / 2: log|det K| / 2 by the
// partially pivoted LU that slogdet uses, so an indefinite K still gives a finite value.
double ora_generate_dense_logdet(const double* pts, int64_t n, int d, const double* bins, const double* vals,
                                 int nbins) {
  std::vector<double> K(n * n);
  ora_cov_matrix(pts, n, pts, n, d, bins, vals, nbins, K.data());
  return 0.5 * lu_logabsdet(K.data(), n);
}

// refine.py:128-135 for every refined point: mean_vec (n_ref x k) and std (n_ref).
void ora_refine_weights(const double* pts, int64_t n, int d, const int32_t* nbrs, int k, int64_t n0,
                        const double* bins, const double* vals, int nbins, double* mean_vec, double* stdv) {
  Table t{bins, vals, nbins};
  const int64_t n_ref = n - n0;
#pragma omp parallel
  {
    std::vector<double> K((k + 1) * (k + 1));
    std::vector<int64_t> nb(k);
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n_ref; ++r) {
      for (int j = 0; j < k; ++j) nb[j] = nbrs[r * k + j];
      joint_cov(pts, d, nb.data(), k, n0 + r, t, K.data());
      conditional_weights(K.data(), k, mean_vec + r * k, stdv + r);
    }
  }
}

// refine.py:123-145 (fast_jit): batch-serial values = mean_vec . values[nbrs] + std * xi.
// For a graph that passes check_graph each value depends only on earlier batches, so
// the point-serial sweep below is the batch sweep.
void ora_refine(const double* pts, int64_t n, int d, const int32_t* nbrs, int k, int64_t n0, const double* bins,
                const double* vals, int nbins, const double* initial_values, const double* xi, double* values) {
  const int64_t n_ref = n - n0;
  std::vector<double> mv(n_ref * k), sd(n_ref);
  ora_refine_weights(pts, n, d, nbrs, k, n0, bins, vals, nbins, mv.data(), sd.data());
  for (int64_t i = 0; i < n0; ++i) values[i] = initial_values[i];
  for (int64_t r = 0; r < n_ref; ++r) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += mv[r * k + j] * values[nbrs[r * k + j]];
    values[n0 + r] = s + sd[r] * xi[r];
  }
}

// refine.py:147-157 + 278-285 (loop form): mean = L[k,:k] @ solve(L_A, v), batch by batch.
void ora_refine_loop(const double* pts, int64_t n, int d, const int32_t* nbrs, int k, int64_t n0,
                     const int64_t* offsets, int64_t n_off, const double* bins, const double* vals, int nbins,
                     const double* initial_values, const double* xi, double* values) {
  Table t{bins, vals, nbins};
  for (int64_t i = 0; i < n0; ++i) values[i] = initial_values[i];
  for (int64_t b = 1; b < n_off; ++b) {
    int64_t s0 = offsets[b - 1], s1 = offsets[b];
#pragma omp parallel
    {
      std::vector<double> K((k + 1) * (k + 1)), y(k);
      std::vector<int64_t> nb(k);
#pragma omp for schedule(dynamic, 256)
      for (int64_t i = s0; i < s1; ++i) {
        int64_t r = i - n0;
        for (int j = 0; j < k; ++j) nb[j] = nbrs[r * k + j];
        joint_cov(pts, d, nb.data(), k, i, t, K.data());
        cholesky(K.data(), k + 1);
        const int m = k + 1;
        for (int a = 0; a < k; ++a) {
          double s = values[nb[a]];
          for (int c = 0; c < a; ++c) s -= K[a * m + c] * y[c];
          y[a] = s / K[a * m + a];
        }
        double mean = 0.0;
        for (int a = 0; a < k; ++a) mean += K[k * m + a] * y[a];
        values[i] = mean + K[k * m + k] * xi[r];
      }
    }
  }
}

// refine.py:198-231 refine_inv: xi = (values - mean) / std for all refined points.
void ora_refine_inv(const double* pts, int64_t n, int d, const int32_t* nbrs, int k, int64_t n0, const double* bins,
                    const double* vals, int nbins, const double* values, double* xi) {
  const int64_t n_ref = n - n0;
  std::vector<double> mv(n_ref * k), sd(n_ref);
  ora_refine_weights(pts, n, d, nbrs, k, n0, bins, vals, nbins, mv.data(), sd.data());
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_ref; ++r) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += mv[r * k + j] * values[nbrs[r * k + j]];
    xi[r] = (values[n0 + r] - s) / sd[r];
  }
}

// refine.py:251-275 refine_logdet = sum log std.
double ora_refine_logdet(const double* pts, int64_t n, int d, const int32_t* nbrs, int k, int64_t n0,
                         const double* bins, const double* vals, int nbins) {
  const int64_t n_ref = n - n0;
  std::vector<double> mv(n_ref * k), sd(n_ref);
  ora_refine_weights(pts, n, d, nbrs, k, n0, bins, vals, nbins, mv.data(), sd.data());
  double s = 0.0;
  for (int64_t r = 0; r < n_ref; ++r) s += std::log(sd[r]);
  return s;
}

// Local conditional of one refined point given explicit neighbour values (used to check
// single points of a large GPU run without re-running the whole chain).
void ora_local_conditional(const double* pts, int d, const int32_t* nbr_row, int k, int64_t fine, const double* bins,
                           const double* vals, int nbins, const double* nbr_values, double* mean, double* stdv) {
  Table t{bins, vals, nbins};
  std::vector<double> K((k + 1) * (k + 1)), mv(k);
  std::vector<int64_t> nb(k);
  for (int j = 0; j < k; ++j) nb[j] = nbr_row[j];
  joint_cov(pts, d, nb.data(), k, fine, t, K.data());
  conditional_weights(K.data(), k, mv.data(), stdv);
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += mv[j] * nbr_values[j];
  *mean = s;
}

// ------------------------------------------------------------------ adjoint
// Reverse-mode derivative of L = sum_i w_i * values_i through refine (points in graph
// order) for fixed seed values: returns dL/dxi (refined part), dL/d(initial_values)
// and the accumulated dL/d(cov_vals).  This is the reference's JAX autodiff of
// refine.py:123-145 written out by hand (no reference adjoint code exists,
// README.md:36); it is pinned by central finite differences in the tests.
//   per point: value = b^T A^{-1} v + sigma * xi,  sigma = sqrt(c - b^T A^{-1} b)
//   w = A^{-1} b, z = A^{-1} v, s = dL/dvalue
//   dA_ii: -s w_i z_i + s xi w_i^2 / (2 sigma);  dA_ij (i<j): -s (w_i z_j + w_j z_i) + s xi w_i w_j / sigma
//   db_i: s z_i - s xi w_i / sigma;  dc: s xi / (2 sigma);  dv_nbr: s w
void ora_refine_vjp(const double* pts, int64_t n, int d, const int32_t* nbrs, int k, int64_t n0, const double* bins,
                    const double* vals, int nbins, const double* values, const double* xi, const double* gvalues,
                    double* gxi, double* ginit, double* gvals) {
  const int64_t n_ref = n - n0;
  Table t{bins, vals, nbins};
  std::vector<double> vbar(gvalues, gvalues + n);
  for (int b = 0; b < nbins; ++b) gvals[b] = 0.0;
  const int m = k + 1;
  std::vector<double> K(m * m), Kc(m * m), w(k), z(k), v(k), tmp(k);
  std::vector<int64_t> nb(k);
  auto lookup_grad = [&](double r, double g) {
    // d interp / d fp: (1 - frac) to fp[i-1], frac to fp[i]; clamps feed the end values.
    if (r < bins[0]) { gvals[0] += g; return; }
    if (r > bins[nbins - 1]) { gvals[nbins - 1] += g; return; }
    int i = (int)(std::upper_bound(bins, bins + nbins, r) - bins);
    if (i < 1) i = 1;
    if (i > nbins - 1) i = nbins - 1;
    double dx = bins[i] - bins[i - 1];
    const double eps = std::nextafter(std::numeric_limits<double>::epsilon(), 1.0) -
                       std::numeric_limits<double>::epsilon();
    if (std::fabs(dx) <= eps) { gvals[i - 1] += g; return; }
    double f = (r - bins[i - 1]) / dx;
    gvals[i - 1] += g * (1.0 - f);
    gvals[i] += g * f;
  };
  for (int64_t r = n_ref - 1; r >= 0; --r) {
    const int64_t i = n0 + r;
    const double s = vbar[i];
    for (int j = 0; j < k; ++j) nb[j] = nbrs[r * k + j];
    joint_cov(pts, d, nb.data(), k, i, t, K.data());
    Kc = K;
    cholesky(K.data(), m);
    // w = A^{-1} b : forward with L_A then back with L_A^T
    for (int a = 0; a < k; ++a) {
      double acc = Kc[k * m + a];
      for (int c = 0; c < a; ++c) acc -= K[a * m + c] * tmp[c];
      tmp[a] = acc / K[a * m + a];
    }
    for (int a = k - 1; a >= 0; --a) {
      double acc = tmp[a];
      for (int c = a + 1; c < k; ++c) acc -= K[c * m + a] * w[c];
      w[a] = acc / K[a * m + a];
    }
    for (int j = 0; j < k; ++j) v[j] = values[nb[j]];
    for (int a = 0; a < k; ++a) {
      double acc = v[a];
      for (int c = 0; c < a; ++c) acc -= K[a * m + c] * tmp[c];
      tmp[a] = acc / K[a * m + a];
    }
    for (int a = k - 1; a >= 0; --a) {
      double acc = tmp[a];
      for (int c = a + 1; c < k; ++c) acc -= K[c * m + a] * z[c];
      z[a] = acc / K[a * m + a];
    }
    const double sigma = K[k * m + k];
    const double x = xi[r];
    gxi[r] = sigma * s;
    for (int j = 0; j < k; ++j) vbar[nb[j]] += s * w[j];
    if (s == 0.0) continue;
    auto P = [&](int a) { return pts + (a < k ? nb[a] : i) * d; };
    for (int a = 0; a < k; ++a) {
      lookup_grad(pdist(P(a), P(a), d), -s * w[a] * z[a] + s * x * w[a] * w[a] / (2.0 * sigma));
      for (int c = a + 1; c < k; ++c)
        lookup_grad(pdist(P(a), P(c), d), -s * (w[a] * z[c] + w[c] * z[a]) + s * x * w[a] * w[c] / sigma);
      lookup_grad(pdist(P(a), P(k), d), s * z[a] - s * x * w[a] / sigma);
    }
    lookup_grad(pdist(P(k), P(k), d), s * x / (2.0 * sigma));
  }
  for (int64_t i = 0; i < n0; ++i) ginit[i] = vbar[i];
}

// Reverse mode of refine_inv (mode 1, refine.py:198-231) or refine_logdet (mode 2,
// refine.py:251-275) for fixed tables: every refined point on its own (no chain).  For point
// i with w = A^{-1} b, z = A^{-1} u (u = values[nbrs]), sigma = L[k][k]:
//   mode 1, cotangent t of xi_i = (v_i - b.z) / sigma:  gvalues[i] += t / sigma,
//     gvalues[nbrs] += -t w / sigma, and the table sees al = -t / sigma (wbar = al u) and
//     be = -t xi_i / (2 sigma^2) (the cotangent of sigma^2);
//   mode 2, cotangent l of sum log sigma_i:  al = 0, be = l / (2 sigma^2);
// dA_aa = -al w_a z_a + be w_a^2, dA_ab = -al (w_a z_b + w_b z_a) + 2 be w_a w_b,
// db_a = al z_a - 2 be w_a, dc = be, each through its entry's lookup.  Written out by hand
// from the reference's JAX autodiff; pinned by central differences in the tests.
void ora_refine_local_vjp(const double* pts, int64_t n, int d, const int32_t* nbrs, int k, int64_t n0,
                          const double* bins, const double* vals, int nbins, int mode, const double* values,
                          const double* cot, double* gvalues, double* gvals) {
  const int64_t n_ref = n - n0;
  Table t{bins, vals, nbins};
  for (int b = 0; b < nbins; ++b) gvals[b] = 0.0;
  if (mode == 1)
    for (int64_t i = 0; i < n; ++i) gvalues[i] = 0.0;
  const int m = k + 1;
  std::vector<double> K(m * m), Kc(m * m), w(k), z(k), tmp(k);
  std::vector<int64_t> nb(k);
  auto lookup_grad = [&](double r, double g) {
    if (r < bins[0]) { gvals[0] += g; return; }
    if (r > bins[nbins - 1]) { gvals[nbins - 1] += g; return; }
    int i = (int)(std::upper_bound(bins, bins + nbins, r) - bins);
    if (i < 1) i = 1;
    if (i > nbins - 1) i = nbins - 1;
    const double dx = bins[i] - bins[i - 1];
    if (std::fabs(dx) <= 4.930380657631324e-32) { gvals[i - 1] += g; return; }
    const double f = (r - bins[i - 1]) / dx;
    gvals[i - 1] += g * (1.0 - f);
    gvals[i] += g * f;
  };
  auto solve_a = [&](const double* rhs, double* out) {  // out = A^{-1} rhs with L_A
    for (int a = 0; a < k; ++a) {
      double acc = rhs[a];
      for (int c = 0; c < a; ++c) acc -= K[a * m + c] * tmp[c];
      tmp[a] = acc / K[a * m + a];
    }
    for (int a = k - 1; a >= 0; --a) {
      double acc = tmp[a];
      for (int c = a + 1; c < k; ++c) acc -= K[c * m + a] * out[c];
      out[a] = acc / K[a * m + a];
    }
  };
  std::vector<double> bvec(k), u(k);
  for (int64_t r = 0; r < n_ref; ++r) {
    const int64_t i = n0 + r;
    for (int j = 0; j < k; ++j) nb[j] = nbrs[r * k + j];
    joint_cov(pts, d, nb.data(), k, i, t, K.data());
    Kc = K;
    cholesky(K.data(), m);
    for (int a = 0; a < k; ++a) bvec[a] = Kc[k * m + a];
    solve_a(bvec.data(), w.data());
    const double sigma = K[k * m + k];
    double al = 0.0, be = 0.0;
    for (int a = 0; a < k; ++a) z[a] = 0.0;
    if (mode == 1) {
      for (int j = 0; j < k; ++j) u[j] = values[nb[j]];
      solve_a(u.data(), z.data());
      double bz = 0.0;
      for (int a = 0; a < k; ++a) bz += bvec[a] * z[a];
      const double xo = (values[i] - bz) / sigma;
      const double tt = cot[r];
      al = -tt / sigma;
      be = -tt * xo / (2.0 * sigma * sigma);
      gvalues[i] += tt / sigma;
      for (int j = 0; j < k; ++j) gvalues[nb[j]] += al * w[j];
    } else {
      be = cot[0] / (2.0 * sigma * sigma);
    }
    auto P = [&](int a) { return pts + (a < k ? nb[a] : i) * d; };
    for (int a = 0; a < k; ++a) {
      lookup_grad(pdist(P(a), P(a), d), -al * w[a] * z[a] + be * w[a] * w[a]);
      for (int c = a + 1; c < k; ++c)
        lookup_grad(pdist(P(a), P(c), d), -al * (w[a] * z[c] + w[c] * z[a]) + 2.0 * be * w[a] * w[c]);
      lookup_grad(pdist(P(a), P(k), d), al * z[a] - 2.0 * be * w[a]);
    }
    lookup_grad(pdist(P(k), P(k), d), be);
  }
}

}  // extern "C"
