// adjoint.hip -- vector-Jacobian product of generate with respect to xi and the
// covariance table (the reference differentiates generate with JAX autodiff,
// refine.py:21-159; README.md:36 mentions custom primitives in its CUDA extension).
//
// Refined points: adjoint_kernel (adjoint_impl.hpp), one reverse dataflow launch.
// Seed block (values = L xi, L = chol(K) over the n0 seeds), in FP64:
//   xibar = L^T vbar.  With Lbar = tril(vbar xi^T), Phi(L^T Lbar) is the lower triangle of
//   u xi^T with its diagonal halved (u = L^T vbar).  The Cholesky adjoint (Murray 2016,
//   "Differentiation of the Cholesky decomposition") is S = L^{-T} Phi(L^T Lbar) L^{-1}:
//   the derivative with respect to the shared entry K_ab = K_ba is S_ab + S_ba (S_aa on the
//   diagonal).  With X = L^{-1} and T[a][c] = u_a (sum_{b<a} xi_b X[b][c] + xi_a X[a][c] / 2),
//   S = X^T T, and the lookup of every ordered pair (a, b) feeds S[a][b] to the table.
#include "adjoint_impl.hpp"
#include "dense.hpp"
#include "graphgp_amd_internal.hpp"
#include "refine_ops.hpp"

namespace ggp {

template <>
int launch_adjoint<float>(AdjointArgs<float> a, int kk, hipStream_t s, size_t lds, int grid) {
  if (a.d > 3) return kUnsupported;
  // at most four rows per lane: the adjoint keeps w, z and the factor live together
  if (kk <= 8) return launch_adjoint_config<float, 8, 2, 3>(a, s, lds, grid);
  if (kk <= 12) return launch_adjoint_config<float, 12, 4, 3>(a, s, lds, grid);
  if (kk <= 16) return launch_adjoint_config<float, 16, 4, 3>(a, s, lds, grid);
  if (kk <= 24) return launch_adjoint_config<float, 24, 8, 3>(a, s, lds, grid);
  if (kk <= 32) return launch_adjoint_config<float, 32, 16, 3>(a, s, lds, grid);
  return kUnsupported;
}

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

template <typename T>
struct VjpWs {
  uint32_t* ctr;
  TabE<T>* tab;
  TabMeta* meta;
  T* vals;
  T* vbar;
  uint32_t* children;
  double *L, *X, *Tm, *S, *vb, *xs, *u;
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
  return w.ok();
}

int cus() {
  int dev = 0, c = 256;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    c = 256;
  return c;
}

template <typename T>
int vjp_flow(const RefineProblem& p, const VjpIO& io, void* ws, size_t ws_bytes, hipStream_t s) {
  GGP_CHECK_ARG(p.n >= 1 && p.n_ref >= 0 && p.n_ref < p.n && p.k >= 1 && p.d >= 1 && p.d <= 3,
                "generate_vjp: bad sizes (1 <= d <= 3)");
  GGP_CHECK_ARG(p.k <= 32, "generate_vjp: k <= 32");
  GGP_CHECK_ARG(p.nb >= 2, "generate_vjp: covariance table needs at least two bins");
  Workspace w(ws, ws_bytes);
  VjpWs<T> b;
  if (!carve_vjp<T>(w, p, b)) {
    set_error("generate_vjp: workspace too small");
    return kWorkspaceTooSmall;
  }
  const int64_t n0 = p.n - p.n_ref;
  GGP_TRY(fill_u32(b.ctr, 0u, 4, s));
  GGP_TRY(prep_table(p.bins, p.vals, p.nb, sizeof(T) == 8, 0, b.tab, b.meta, s));
  hipLaunchKernelGGL(gather_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, s, io.values, io.values_f64, p.perm,
                     p.n, b.vals);
  hipLaunchKernelGGL(gather_kernel<T>, dim3(grid_for(p.n, 256)), dim3(256), 0, s, io.gvalues, io.g_f64, p.perm, p.n,
                     b.vbar);
  GGP_LAUNCH_CHECK();
  GGP_TRY(fill_u32(b.children, 0u, p.n, s));
  GGP_TRY(fill_u64(io.gtab, 0ull, p.nb, s));
  if (p.n_ref > 0) {
    hipLaunchKernelGGL(count_children_kernel, dim3(grid_capped(p.n_ref * p.k, 256)), dim3(256), 0, s, p.nbrs,
                       p.n_ref * (int64_t)p.k, b.children);
    GGP_LAUNCH_CHECK();
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
    a.tab = b.tab;
    a.meta = b.meta;
    a.ticket = b.ctr;
    a.flags = b.ctr + 1;
    a.nbins = p.nb;
    const size_t tab_bytes = (size_t)(p.nb - 1) * sizeof(TabE<T>) + (size_t)p.nb * sizeof(T);
    const size_t lds = ((tab_bytes + 15) & ~size_t(15)) + (size_t)p.nb * sizeof(double);
    GGP_CHECK_ARG(lds <= 160 * 1024, "generate_vjp: covariance table too large for the LDS-resident adjoint");
    int rc = launch_adjoint<T>(a, p.k, s, lds, cus() * 4);
    if (rc != kOk) {
      if (rc == kUnsupported) set_error("generate_vjp: unsupported configuration (k <= 32, d <= 3)");
      return rc;
    }
  }
  // seed block
  GGP_TRY(dense_factor(p.pts, p.pts_f64, p.d, n0, p.bins, p.vals, p.nb, b.L, s));
  hipLaunchKernelGGL(seed_vbar_kernel<T>, dim3(grid_for(n0, 256)), dim3(256), 0, s, b.vbar, io.xi, io.xi_f64, p.perm,
                     n0, b.vb, b.xs);
  hipLaunchKernelGGL(trmv_t_kernel, dim3(grid_for(n0, 4)), dim3(256), 0, s, b.L, n0, b.vb, b.u, p.perm, io.gxi,
                     io.gxi_f64);
  const size_t xl = (size_t)n0 * sizeof(double);
  GGP_CHECK_ARG(xl <= 160 * 1024, "generate_vjp: seed block too large for the LDS column inverse (n0 <= 20480)");
  if (xl > 65536)
    GGP_HIP(hipFuncSetAttribute((const void*)tri_inv_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)xl));
  hipLaunchKernelGGL(tri_inv_kernel, dim3((unsigned)n0), dim3(64), xl, s, b.L, n0, b.X);
  hipLaunchKernelGGL(phi_kernel, dim3(grid_for(n0, 64)), dim3(64), 0, s, b.X, n0, b.u, b.xs, b.Tm);
  hipLaunchKernelGGL(gemm_tn_kernel, dim3(grid_for(n0, 64), grid_for(n0, 64)), dim3(256), 0, s, b.X, b.Tm, n0, b.S);
  GGP_LAUNCH_CHECK();
  const size_t hl = (size_t)p.nb * sizeof(double);
  if (hl > 65536) GGP_HIP(hipFuncSetAttribute((const void*)seed_table_grad_kernel,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)hl));
  hipLaunchKernelGGL(seed_table_grad_kernel, dim3((unsigned)std::min<int64_t>(grid_for(n0 * n0, 256), 1024)),
                     dim3(256), hl, s, p.pts, p.pts_f64, p.d, n0, b.S, p.bins, p.nb, io.gtab);
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

int vjp_run(const RefineProblem& p, int t64, const VjpIO& io, void* ws, size_t ws_bytes, hipStream_t s) {
  return t64 ? vjp_flow<double>(p, io, ws, ws_bytes, s) : vjp_flow<float>(p, io, ws, ws_bytes, s);
}

}  // namespace ggp
