// dense.hip -- the n0 x n0 seed block in FP64 (gfx950).
//
// Reference: graphgp/refine.py:59-76 (generate_dense: chol(K) @ xi), :186-195
// (generate_dense_inv: solve(L, values)), :243-248 (generate_dense_logdet:
// slogdet(K)/2 = sum log diag L).  The seed block is always factorised in FP64.
//
// Kernels: covariance assembly (one thread per entry, binary-search interp on the FP64
// table); a right-looking blocked Cholesky with 64-wide block columns, two launches per
// block column: the panel as a 64x64 GEMM against the inverse of the diagonal factor,
// and the trailing lower-triangle update in 64x64 LDS tiles whose diagonal-owner
// workgroup also factorises and inverts the next diagonal block in LDS; a
// wavefront-per-row triangular matvec, a blocked forward substitution in one
// workgroup, and a fixed-order log-diagonal sum.
#include "common.hpp"
#include "dense.hpp"

namespace ggp {

namespace {

constexpr int NB = 64;

__device__ inline double interp_dev(double x, const double* __restrict__ xp, const double* __restrict__ fp, int n) {
  if (x < xp[0]) return fp[0];
  if (x > xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n;  // upper_bound
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (xp[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  int i = lo < 1 ? 1 : (lo > n - 1 ? n - 1 : lo);
  double df = fp[i] - fp[i - 1], dx = xp[i] - xp[i - 1], delta = x - xp[i - 1];
  if (fabs(dx) <= 4.930380657631324e-32) return fp[i - 1];
  return fp[i - 1] + (delta / dx) * df;
}

// With `gate` non-null the kernel runs only when *gate is not finite (the log-det
// fallback below rebuilds K after a failed Cholesky).
__global__ void dense_cov_kernel(const void* __restrict__ pts, int pts_f64, int d, int64_t n,
                                 const double* __restrict__ bins, const double* __restrict__ vals, int nb,
                                 double* __restrict__ A, const double* __restrict__ gate) {
  if (gate != nullptr && isfinite(*gate)) return;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e - i * n;
    double s = 0.0;
    for (int t = 0; t < d; ++t) {
      double a = pts_f64 ? ((const double*)pts)[i * d + t] : (double)((const float*)pts)[i * d + t];
      double b = pts_f64 ? ((const double*)pts)[j * d + t] : (double)((const float*)pts)[j * d + t];
      double u = a - b;
      s += u * u;
    }
    A[e] = interp_dev(sqrt(s), bins, vals, nb);
  }
}

__global__ void cross_cov_kernel(const void* __restrict__ pa, int64_t na, const void* __restrict__ pb, int64_t nb_,
                                 int d, int f64, const double* __restrict__ bins, const double* __restrict__ vals,
                                 int nbins, double* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < na * nb_;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / nb_, j = e - i * nb_;
    double s = 0.0;
    for (int t = 0; t < d; ++t) {
      double a = f64 ? ((const double*)pa)[i * d + t] : (double)((const float*)pa)[i * d + t];
      double b = f64 ? ((const double*)pb)[j * d + t] : (double)((const float*)pb)[j * d + t];
      double u = a - b;
      s += u * u;
    }
    out[e] = interp_dev(sqrt(s), bins, vals, nbins);
  }
}

// T = the nr x nc block of A at (r0, c0), zero-padded to 64 x 64.  Every thread issues its
// sixteen loads (clamped addresses) before the first LDS store, so the tile costs one
// memory round trip.
__device__ __forceinline__ void load_tile(double (*T)[NB + 1], const double* __restrict__ A, int64_t n, int64_t r0,
                                          int64_t c0, int nr, int nc) {
  double v[NB * NB / 256];
#pragma unroll
  for (int u = 0; u < NB * NB / 256; ++u) {
    const int q = threadIdx.x + 256 * u, r = q / NB, c = q % NB;
    v[u] = A[(r0 + min(r, nr - 1)) * n + c0 + min(c, nc - 1)];
  }
#pragma unroll
  for (int u = 0; u < NB * NB / 256; ++u) {
    const int q = threadIdx.x + 256 * u, r = q / NB, c = q % NB;
    T[r][c] = (r < nr && c < nc) ? v[u] : 0.0;
  }
}

// Factorise the diagonal block D (nbj x nbj, lower, in LDS) and write its inverse to Di.
// Rows beyond nbj are padded with the identity.  Lane r of wavefront w holds row r's
// columns 16w .. 16w + 15 of the trailing matrix (a) and of the inverse being formed (b)
// in registers.  Column c is published by its owning wavefront, together with row c of
// b, in a double-buffered LDS exchange (one workgroup barrier per column): every lane
// then applies the rank-1 update a[r][m] -= a[r][c] a[m][c] / a[c][c] and the matching
// forward-substitution step b[r] -= a[r][c] b[c] / a[c][c], so L^{-1} comes out of the
// same sweep (row r of L^{-1} is b[r] / L[r][r]).  A column's unscaled values go to D as
// it is published and its registers are not read again, so the update runs over all
// sixteen registers unmasked; the square roots and the scaling follow in one pass.  The
// exchange buffers live in Di, which is written only after the sweep, and the square
// roots in D's padding column, so the routine needs no LDS beyond the two tiles.
__device__ __forceinline__ void factor_invert_lds(double (*D)[NB + 1], double (*Di)[NB + 1], int nbj) {
  static_assert(NB == 64, "one lane per row of the block");
  constexpr int Q = NB / 4;
  const int t = threadIdx.x;
  const int row = t & (NB - 1);
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  __syncthreads();
  double a[Q], b[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = Q * w + q;
    a[q] = row < nbj ? (m < nbj ? D[row][m] : 0.0) : (m == row ? 1.0 : 0.0);
    b[q] = m == row ? 1.0 : 0.0;
  }
  __syncthreads();  // D is rewritten column by column below
  // exchange: 2 x [column c (64) | row c of b (64)] at the start of Di
  double* xbuf = &Di[0][0];
  for (int c0 = 0; c0 < NB; c0 += Q) {
    const bool owner = w == c0 / Q;
#pragma unroll
    for (int qc = 0; qc < Q; ++qc) {
      const int c = c0 + qc;
      double* cb = xbuf + (qc & 1) * 2 * NB;
      double* rb = cb + NB;
      if (owner) cb[row] = a[qc];
      if (row == c) {
#pragma unroll
        for (int q = 0; q < Q; ++q) rb[Q * w + q] = b[q];
      }
      __syncthreads();
      double cm[Q], rq[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        cm[q] = cb[Q * w + q];
        rq[q] = rb[Q * w + q];
      }
      const double dcc = cb[c];
      const double x = cb[row];
      const double s = row > c ? x / dcc : 0.0;
      // a changes only right of the pivot (columns > c), b only at columns <= c (row c of
      // b is zero beyond c): a wavefront whose sixteen columns lie on one side skips the
      // other update (a uniform test)
      if (c0 <= Q * w) {
#pragma unroll
        for (int q = 0; q < Q; ++q) a[q] -= s * cm[q];
      }
      if (c0 >= Q * w) {
#pragma unroll
        for (int q = 0; q < Q; ++q) b[q] -= s * rq[q];
      }
      if (owner) D[row][c] = row >= c ? x : 0.0;
    }
  }
  __syncthreads();
  if (t < NB) D[t][NB] = sqrt(D[t][t]);  // padding column: never part of the block
  __syncthreads();
  const double dr = D[row][NB];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int m = Q * w + q;
    const double v = D[row][m];
    D[row][m] = row > m ? v / D[m][NB] : (row == m ? dr : 0.0);
    Di[row][m] = b[q] / dr;
  }
  __syncthreads();
}

// Block jb's diagonal factor (first step only; later steps factor inside update).
__global__ __launch_bounds__(256) void diag_factor_kernel(double* __restrict__ A, int64_t n, int jb,
                                                          double* __restrict__ Linv) {
  __shared__ double D[NB][NB + 1];
  __shared__ double Di[NB][NB + 1];
  const int64_t j0 = (int64_t)jb * NB;
  const int nbj = (int)min<int64_t>(NB, n - j0);
  load_tile(D, A, n, j0, j0, nbj, nbj);
  factor_invert_lds(D, Di, nbj);
  for (int q = threadIdx.x; q < NB * NB; q += 256) {
    const int r = q / NB, c = q % NB;
    if (r < nbj && c < nbj) A[(j0 + r) * n + j0 + c] = c <= r ? D[r][c] : 0.0;
    Linv[q] = Di[r][c];
  }
}

// Panel: L[ib][jb] = A[ib][jb] L_jj^{-T}, one 64 x 64 block per workgroup (blockIdx.x + 1
// blocks below the diagonal).
__global__ __launch_bounds__(256) void panel_gemm_kernel(double* __restrict__ A, int64_t n, int jb,
                                                         const double* __restrict__ Linv) {
  __shared__ double X[NB][NB + 1];
  __shared__ double Li[NB][NB + 1];
  const int ib = jb + 1 + blockIdx.x;
  const int64_t j0 = (int64_t)jb * NB, i0 = (int64_t)ib * NB;
  const int nbj = (int)min<int64_t>(NB, n - j0);
  const int t = threadIdx.x;
  load_tile(X, A, n, i0, j0, (int)min<int64_t>(NB, n - i0), nbj);
  load_tile(Li, Linv, NB, 0, 0, NB, NB);
  __syncthreads();
  const int tr = (t / 16) * 4, tc = (t % 16) * 4;
  double acc[4][4] = {};
  for (int m = 0; m < nbj; ++m) {
    double xr[4], lc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { xr[u] = X[tr + u][m]; lc[u] = Li[tc + u][m]; }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] += xr[u] * lc[v];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = i0 + tr + u;
      const int c = tc + v;
      if (r < n && c < nbj) A[r * n + j0 + c] = acc[u][v];
    }
}

// Trailing update A[ib][mb] -= L[ib][jb] L[mb][jb]^T for jb < mb <= ib.  The workgroup
// that owns the next diagonal block (jb+1, jb+1) also factorises and inverts it.
__global__ __launch_bounds__(256) void update_kernel(double* __restrict__ A, int64_t n, int jb, int nrem,
                                                     double* __restrict__ Linv) {
  __shared__ double Li[NB][NB + 1];
  __shared__ double Lm[NB][NB + 1];
  int e = blockIdx.x;
  int aa = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
  while ((aa + 1) * (aa + 2) / 2 <= e) ++aa;
  while (aa * (aa + 1) / 2 > e) --aa;
  const int bb = e - aa * (aa + 1) / 2;
  const int ib = jb + 1 + aa, mb = jb + 1 + bb;
  const int64_t j0 = (int64_t)jb * NB, i0 = (int64_t)ib * NB, m0 = (int64_t)mb * NB;
  const int nbj = (int)min<int64_t>(NB, n - j0);
  const int t = threadIdx.x;
  const int tr = (t / 16) * 4, tc = (t % 16) * 4;
  // the target block's current values, fetched with the tiles (one memory round trip)
  double old[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) old[u][v] = A[min(i0 + tr + u, n - 1) * n + min(m0 + tc + v, n - 1)];
  load_tile(Li, A, n, i0, j0, (int)min<int64_t>(NB, n - i0), nbj);
  load_tile(Lm, A, n, m0, j0, (int)min<int64_t>(NB, n - m0), nbj);
  __syncthreads();
  double acc[4][4] = {};
  for (int c = 0; c < nbj; ++c) {
    double li[4], lm[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { li[u] = Li[tr + u][c]; lm[u] = Lm[tc + u][c]; }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] += li[u] * lm[v];
  }
  if (e != 0) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = i0 + tr + u, c = m0 + tc + v;
        if (r < n && c < n && c <= r) A[r * n + c] = old[u][v] - acc[u][v];
      }
    return;
  }
  // next diagonal block: update in LDS, factorise, invert, write back
  const int nb2 = (int)min<int64_t>(NB, n - i0);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = tr + u, c = tc + v;
      Li[r][c] = (r < nb2 && c < nb2 && c <= r) ? old[u][v] - acc[u][v] : 0.0;
    }
  factor_invert_lds(Li, Lm, nb2);
  for (int q = t; q < NB * NB; q += 256) {
    const int r = q / NB, c = q % NB;
    if (r < nb2 && c < nb2) A[(i0 + r) * n + i0 + c] = c <= r ? Li[r][c] : 0.0;
    Linv[q] = Lm[r][c];
  }
}

// values[i] = sum_{j<=i} L[i][j] xs[j]; one wavefront per row.
__global__ __launch_bounds__(256) void trmv_kernel(const double* __restrict__ L, int64_t n,
                                                   const double* __restrict__ xs, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t j = lane; j <= i; j += 64) s += L[i * n + j] * xs[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) out[i] = s;
}

// Solve L x = v for x (in place on v); one workgroup of 1024 threads, 64-row blocks.
__global__ __launch_bounds__(1024) void trsv_kernel(const double* __restrict__ L, int64_t n, double* __restrict__ v) {
  __shared__ double xb[NB];
  const int t = threadIdx.x;
  for (int64_t b0 = 0; b0 < n; b0 += NB) {
    const int nbb = (int)min<int64_t>(NB, n - b0);
    if (t < 64) {
      // sequential solve of the diagonal block by one wavefront (lane c finalises x_c)
      double vi = t < nbb ? v[b0 + t] : 0.0;
      for (int c = 0; c < nbb; ++c) {
        double xc = __shfl(vi, c, 64) / L[(b0 + c) * n + b0 + c];
        if (t == c) vi = xc;
        if (t > c && t < nbb) vi -= L[(b0 + t) * n + b0 + c] * xc;
        if (t == c) xb[c] = xc;
      }
      if (t < nbb) v[b0 + t] = vi;
    }
    __syncthreads();
    for (int64_t i = b0 + nbb + t; i < n; i += 1024) {
      double s = v[i];
      const double* row = L + i * n + b0;
      for (int c = 0; c < nbb; ++c) s -= row[c] * xb[c];
      v[i] = s;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void logdiag_kernel(const double* __restrict__ L, int64_t n, double* out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += log(L[i * n + i]);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tsum = 0.0;
    for (int w = 0; w < 16; ++w) tsum += red[w];
    *out = tsum;
  }
}

// log|det K| / 2 by Gaussian elimination with partial pivoting (the first row of largest
// |a_ij| in the column, LAPACK getrf's rule), which is what jnp.linalg.slogdet computes
// (refine.py:248): finite for an indefinite K, -inf for a singular one, NaN for NaN input.
// Runs only when *logdet (the Cholesky's sum of log L_jj) is not finite, i.e. when the
// Cholesky failed; then K has been rebuilt in A by the gated dense_cov_kernel.  One
// workgroup: per column a pivot search, a row swap (columns >= j only; the multipliers are
// not kept) and the trailing update, each wavefront taking rows and its lanes columns.
// Column j and row j are only read during the update, so no staging is needed.
__global__ __launch_bounds__(1024) void lu_logdet_kernel(double* __restrict__ A, int64_t n, double* __restrict__ logdet) {
  if (isfinite(*logdet)) return;
  __shared__ double s_best[16];
  __shared__ int64_t s_arg[16];
  __shared__ int64_t s_piv;
  __shared__ int s_stop;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double acc = 0.0;  // thread 0 only
  for (int64_t j = 0; j < n; ++j) {
    double best = -1.0;
    int64_t arg = j;
    for (int64_t i = j + t; i < n; i += 1024) {
      const double a = fabs(A[i * n + j]);
      if (a != a) {
        best = a;  // NaN wins the search and stops the elimination
        arg = i;
        break;
      }
      if (a > best) {
        best = a;
        arg = i;
      }
    }
    // reduction: NaN first, then the largest, then the lowest row
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int64_t oa = __shfl_xor(arg, o, 64);
      const bool take = (ob != ob) ? !(best != best && arg < oa) : (best == best && (ob > best || (ob == best && oa < arg)));
      if (take) {
        best = ob;
        arg = oa;
      }
    }
    if (lane == 0) {
      s_best[w] = best;
      s_arg[w] = arg;
    }
    __syncthreads();
    if (t == 0) {
      double b = s_best[0];
      int64_t a = s_arg[0];
      for (int q = 1; q < 16; ++q) {
        const double ob = s_best[q];
        const int64_t oa = s_arg[q];
        const bool take = (ob != ob) ? !(b != b && a < oa) : (b == b && (ob > b || (ob == b && oa < a)));
        if (take) {
          b = ob;
          a = oa;
        }
      }
      int stop = 0;
      if (b != b) {
        acc = __builtin_nan("");
        stop = 1;
      } else if (b == 0.0) {
        acc = -__builtin_inf();
        stop = 1;
      }
      s_piv = a;
      s_stop = stop;
    }
    __syncthreads();
    if (s_stop) break;
    const int64_t p = s_piv;
    if (p != j)
      for (int64_t c = j + t; c < n; c += 1024) {
        const double x = A[j * n + c];
        A[j * n + c] = A[p * n + c];
        A[p * n + c] = x;
      }
    __syncthreads();
    const double piv = A[j * n + j];
    if (t == 0) acc += log(fabs(piv));
    for (int64_t i = j + 1 + w; i < n; i += 16) {
      const double l = A[i * n + j] / piv;
      if (l == 0.0) continue;
      for (int64_t c = j + 1 + lane; c < n; c += 64) A[i * n + c] -= l * A[j * n + c];
    }
    __syncthreads();
  }
  if (t == 0) *logdet = 0.5 * acc;
}

}  // namespace

size_t dense_workspace(int64_t n0) { return (size_t)(n0 * n0 + NB * NB) * sizeof(double) + 256; }

int dense_factor(const void* pts, int pts_f64, int d, int64_t n0, const double* bins, const double* vals, int nb,
                 double* A, hipStream_t s) {
  if (n0 <= 0) return kOk;
  hipLaunchKernelGGL(dense_cov_kernel, dim3(grid_capped(n0 * n0, 256)), dim3(256), 0, s, pts, pts_f64, d, n0, bins,
                     vals, nb, A, (const double*)nullptr);
  GGP_LAUNCH_CHECK();
  const int nblk = (int)((n0 + NB - 1) / NB);
  double* Linv = A + (size_t)n0 * n0;  // 64 x 64 scratch after the matrix
  hipLaunchKernelGGL(diag_factor_kernel, dim3(1), dim3(256), 0, s, A, n0, 0, Linv);
  GGP_LAUNCH_CHECK();
  for (int jb = 0; jb < nblk - 1; ++jb) {
    const int nrem = nblk - jb - 1;
    hipLaunchKernelGGL(panel_gemm_kernel, dim3(nrem), dim3(256), 0, s, A, n0, jb, Linv);
    hipLaunchKernelGGL(update_kernel, dim3(nrem * (nrem + 1) / 2), dim3(256), 0, s, A, n0, jb, nrem, Linv);
    GGP_LAUNCH_CHECK();
  }
  return kOk;
}

int dense_apply(const double* L, int64_t n0, const double* xs, double* out, hipStream_t s) {
  if (n0 <= 0) return kOk;
  hipLaunchKernelGGL(trmv_kernel, dim3(grid_for(n0, 4)), dim3(256), 0, s, L, n0, xs, out);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int dense_solve(const double* L, int64_t n0, double* v, hipStream_t s) {
  if (n0 <= 0) return kOk;
  hipLaunchKernelGGL(trsv_kernel, dim3(1), dim3(1024), 0, s, L, n0, v);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int cov_matrix(const void* a, int64_t na, const void* b, int64_t nb_, int d, int f64, const double* bins,
               const double* vals, int nbins, double* out, hipStream_t s) {
  if (na * nb_ == 0) return kOk;
  hipLaunchKernelGGL(cross_cov_kernel, dim3(grid_capped(na * nb_, 256)), dim3(256), 0, s, a, na, b, nb_, d, f64, bins,
                     vals, nbins, out);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int dense_logdiag(const double* L, int64_t n0, double* out, hipStream_t s) {
  hipLaunchKernelGGL(logdiag_kernel, dim3(1), dim3(1024), 0, s, L, n0, out);
  GGP_LAUNCH_CHECK();
  return kOk;
}

int dense_logdet_fallback(const void* pts, int pts_f64, int d, int64_t n0, const double* bins, const double* vals,
                          int nb, double* A, double* logdet, hipStream_t s) {
  if (n0 <= 0) return kOk;
  hipLaunchKernelGGL(dense_cov_kernel, dim3(grid_capped(n0 * n0, 256)), dim3(256), 0, s, pts, pts_f64, d, n0, bins,
                     vals, nb, A, (const double*)logdet);
  GGP_LAUNCH_CHECK();
  hipLaunchKernelGGL(lu_logdet_kernel, dim3(1), dim3(1024), 0, s, A, n0, logdet);
  GGP_LAUNCH_CHECK();
  return kOk;
}

}  // namespace ggp
