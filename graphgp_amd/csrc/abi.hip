// abi.hip -- the extern "C" entry points declared in include/graphgp_amd.h.
#include <optional>
#include <string>

#include "../../include/graphgp_amd.h"
#include "common.hpp"
#include "graphgp_amd_internal.hpp"
#include "refine_ops.hpp"
#include "dense.hpp"

namespace ggp {

extern thread_local hipEvent_t g_prof_start, g_prof_stop;

namespace {
thread_local std::string g_error;
thread_local int g_cu_budget = 0;
}

void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }
int& cu_budget() { return g_cu_budget; }
thread_local YieldCfg g_yield;
YieldCfg& yield_cfg() { return g_yield; }

int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, c = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        c <= 0)
      c = 256;
    cus = c;
  }
  return cus;
}

namespace {

inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

int check_dtype(int dt, const char* what) {
  if (dt != GGP_F32 && dt != GGP_F64) {
    set_error(std::string(what) + ": dtype must be GGP_F32 or GGP_F64");
    return kInvalidArgument;
  }
  return kOk;
}

__global__ void max_kernel(const int32_t* __restrict__ v, int64_t n, int32_t* __restrict__ out) {
  __shared__ int32_t red[16];
  int32_t m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, v[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = max(t, red[w]);
    atomicMax(out, t);
  }
}

struct GraphBufs {
  char* tpts;
  int32_t* split;
  int32_t* idx;
  int32_t* nbrs;
  int32_t* depths;
  int32_t* dmax;
  void* sub;
  size_t sub_bytes;
};

size_t graph_sub_bytes(int64_t n, int d, int dtype, int64_t n0, int k) {
  size_t a = build_tree_workspace(n, d, dtype);
  size_t b = query_neighbors_workspace(n, d, n - n0, k, dtype);
  size_t c = compute_depths_workspace(n - n0, k);
  size_t e = order_by_depth_workspace(n);
  return std::max(std::max(a, b), std::max(c, e));
}

bool carve_graph(Workspace& w, int64_t n, int d, int dtype, int64_t n0, int k, GraphBufs& g) {
  g.tpts = w.take<char>((size_t)n * d * dtype_size(dtype));
  g.split = w.take<int32_t>(n);
  g.idx = w.take<int32_t>(n);
  g.nbrs = w.take<int32_t>((size_t)(n - n0) * k);
  g.depths = w.take<int32_t>(n);
  g.dmax = w.take<int32_t>(64);
  g.sub_bytes = graph_sub_bytes(n, d, dtype, n0, k);
  g.sub = w.take<char>(g.sub_bytes);
  return w.ok();
}

RefineProblem problem(const void* pts, int pdtype, int64_t n, int d, const int32_t* nbrs, int64_t n0, int k,
                      const int32_t* perm, const double* bins, const double* vals, int nb, int precision,
                      int with_seed) {
  RefineProblem p;
  p.pts = pts;
  p.pts_f64 = pdtype == GGP_F64;
  p.n = n;
  p.d = d;
  p.nbrs = nbrs;
  p.n_ref = n - n0;
  p.k = k;
  p.perm = perm;
  p.bins = bins;
  p.vals = vals;
  p.nb = nb;
  p.increment = precision == GGP_PREC_FP32_INC;
  p.with_seed = with_seed;
  return p;
}

int check_refine_args(int64_t n, int64_t n0, int k, int precision, int pdtype) {
  GGP_CHECK_ARG(n >= 1 && n0 >= 1 && n0 <= n, "refine: need 1 <= n0 <= n");
  GGP_CHECK_ARG(k >= 1, "refine: k must be positive");
  GGP_CHECK_ARG(precision == GGP_PREC_FP64 || precision == GGP_PREC_FP32 || precision == GGP_PREC_FP32_INC,
                "refine: unknown precision");
  return check_dtype(pdtype, "refine points");
}

inline int t64_of(int precision) { return precision == GGP_PREC_FP64 ? 1 : 0; }

}  // namespace
}  // namespace ggp

using namespace ggp;

extern "C" {

const char* ggp_last_error(void) { return ggp::last_error(); }

int ggp_set_cu_budget(int cus) {
  const int prev = ggp::cu_budget();
  ggp::cu_budget() = cus > 0 ? cus : 0;
  return prev;
}
int ggp_set_yield(const uint32_t* flag, int keep_cus) {
  YieldCfg& y = ggp::yield_cfg();
  if (flag == nullptr) {
    y = YieldCfg{};
    return kOk;
  }
  GGP_CHECK_ARG(keep_cus >= 1, "set_yield: keep_cus must be at least 1");
  void* dev = nullptr;
  GGP_HIP(hipHostGetDevicePointer(&dev, const_cast<uint32_t*>(flag), 0));
  y.host = flag;
  y.dev = static_cast<const uint32_t*>(dev);
  y.keep_cus = keep_cus;
  return kOk;
}
int ggp_abi_version(void) { return 1; }

size_t ggp_build_tree_workspace(int64_t n, int d, int dtype) { return build_tree_workspace(n, d, dtype); }

int ggp_build_tree(const void* points, int64_t n, int d, int dtype, void* points_out, int32_t* split_dims,
                   int32_t* indices, void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(dtype, "build_tree"));
  return build_tree(points, n, d, dtype, points_out, split_dims, indices, workspace, workspace_bytes, S(stream));
}

size_t ggp_query_neighbors_workspace(int64_t n, int d, int64_t n_queries, int k, int dtype) {
  return query_neighbors_workspace(n, d, n_queries, k, dtype);
}

int ggp_query_preceding_neighbors(const void* points, const int32_t* split_dims, int64_t n, int d, int dtype,
                                  int64_t n0, int k, int32_t* neighbors, void* workspace, size_t workspace_bytes,
                                  void* stream) {
  GGP_TRY(check_dtype(dtype, "query_preceding_neighbors"));
  if (n0 < k) {
    set_error("n0 must be at least k. Got n0=" + std::to_string(n0) + ", k=" + std::to_string(k) + ".");
    return kInvalidArgument;
  }
  GGP_CHECK_ARG(n0 <= n, "query_preceding_neighbors: n0 > n");
  return query_neighbors(points, split_dims, n, d, dtype, nullptr, nullptr, n0, n - n0, k, neighbors, workspace,
                         workspace_bytes, S(stream));
}

int ggp_query_neighbors(const void* points, const int32_t* split_dims, int64_t n, int d, int dtype,
                        const int32_t* query_indices, const int32_t* max_indices, int64_t n_queries, int k,
                        int32_t* neighbors, void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(dtype, "query_neighbors"));
  GGP_CHECK_ARG(query_indices && max_indices, "query_neighbors: query and max indices are required");
  return query_neighbors(points, split_dims, n, d, dtype, query_indices, max_indices, 0, n_queries, k, neighbors,
                         workspace, workspace_bytes, S(stream));
}

size_t ggp_compute_depths_workspace(int64_t n_ref, int k) { return compute_depths_workspace(n_ref, k); }

int ggp_compute_depths(const int32_t* neighbors, int64_t n_ref, int k, int64_t n0, int32_t* depths, void* workspace,
                       size_t workspace_bytes, void* stream) {
  return compute_depths(neighbors, n_ref, k, n0, depths, kDepthsFlow, workspace, workspace_bytes, S(stream));
}

int ggp_compute_depths_method(const int32_t* neighbors, int64_t n_ref, int k, int64_t n0, int32_t* depths,
                              int method, void* workspace, size_t workspace_bytes, void* stream) {
  return compute_depths(neighbors, n_ref, k, n0, depths, method, workspace, workspace_bytes, S(stream));
}

int ggp_compute_depths_status(void* workspace, uint32_t* host_status, void* stream) {
  uint32_t* ctr = static_cast<uint32_t*>(workspace);
  return read_to_host(host_status, ctr + 1, 4, S(stream));
}

size_t ggp_order_by_depth_workspace(int64_t n) { return order_by_depth_workspace(n); }

int ggp_order_by_depth(const void* points, int64_t n, int d, int dtype, const int32_t* indices,
                       const int32_t* neighbors, int64_t n_ref, int k, const int32_t* depths, int32_t max_depth,
                       void* points_out, int32_t* indices_out, int32_t* neighbors_out, int32_t* depths_out,
                       void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(dtype, "order_by_depth"));
  return order_by_depth(points, n, d, dtype, indices, neighbors, n_ref, k, depths, max_depth, points_out, indices_out,
                        neighbors_out, depths_out, workspace, workspace_bytes, S(stream));
}

int ggp_depth_offsets(const int32_t* depths_sorted, int64_t n, int64_t n_offsets, int64_t* offsets, void* stream) {
  return depth_offsets(depths_sorted, n, n_offsets, offsets, S(stream));
}

size_t ggp_build_graph_workspace(int64_t n, int d, int dtype, int64_t n0, int k) {
  Workspace w(nullptr, 0);
  GraphBufs g;
  carve_graph(w, n, d, dtype, n0, k, g);
  return w.used + 256;
}

int ggp_build_graph(const void* points, int64_t n, int d, int dtype, int64_t n0, int k, void* points_out,
                    int32_t* indices_out, int32_t* neighbors_out, int32_t* depths_out, int32_t* max_depth_host,
                    void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(dtype, "build_graph"));
  if (n0 < k) {
    set_error("n0 must be at least k. Got n0=" + std::to_string(n0) + ", k=" + std::to_string(k) + ".");
    return kInvalidArgument;
  }
  GGP_CHECK_ARG(n0 <= n && n >= 1, "build_graph: need n0 <= n");
  Workspace w(workspace, workspace_bytes);
  GraphBufs g;
  if (!carve_graph(w, n, d, dtype, n0, k, g)) {
    set_error("build_graph: workspace too small");
    return kWorkspaceTooSmall;
  }
  hipStream_t s = S(stream);
  std::optional<StageScope> stage;
  stage.emplace("tree", s);
  GGP_TRY(build_tree(points, n, d, dtype, g.tpts, g.split, g.idx, g.sub, g.sub_bytes, s));
  stage.emplace("knn", s);
  GGP_TRY(query_neighbors(g.tpts, g.split, n, d, dtype, nullptr, nullptr, n0, n - n0, k, g.nbrs, g.sub, g.sub_bytes,
                          s));
  stage.emplace("depths", s);
  GGP_TRY(compute_depths(g.nbrs, n - n0, k, n0, g.depths, kDepthsFlow, g.sub, g.sub_bytes, s));
  uint32_t status = 0;
  GGP_TRY(ggp_compute_depths_status(g.sub, &status, stream));
  if (status != 0) {
    set_error("build_graph: neighbour graph is not in topological order (internal error)");
    return kHipError;
  }
  stage.emplace("order", s);
  GGP_TRY(fill_u32(g.dmax, 0u, 1, s));
  hipLaunchKernelGGL(max_kernel, dim3(std::min<int64_t>(grid_for(n, 256), 1024)), dim3(256), 0, s, g.depths, n,
                     g.dmax);
  GGP_LAUNCH_CHECK();
  int32_t mx = 0;
  GGP_TRY(read_to_host(&mx, g.dmax, 4, s));
  GGP_TRY(order_by_depth(g.tpts, n, d, dtype, g.idx, g.nbrs, n - n0, k, g.depths, mx, points_out, indices_out,
                         neighbors_out, depths_out, g.sub, g.sub_bytes, s));
  if (max_depth_host) *max_depth_host = mx;
  return kOk;
}

int ggp_check_graph(const int32_t* neighbors, int64_t n_ref, int k, int64_t n, const int64_t* offsets,
                    int64_t n_offsets, uint32_t* status_device, void* stream) {
  return check_graph(neighbors, n_ref, k, n, offsets, n_offsets, status_device, S(stream));
}

size_t ggp_refine_workspace(int64_t n, int d, int64_t n0, int k, int n_bins, int precision, int with_seed) {
  RefineProblem p = problem(nullptr, 0, n, d, nullptr, n0, k, nullptr, nullptr, nullptr, n_bins, precision,
                            with_seed);
  size_t a = refine_workspace(p, 0, t64_of(precision));
  size_t b = refine_workspace(p, 1, t64_of(precision));
  size_t c = refine_workspace(p, 2, t64_of(precision));
  return std::max(a, std::max(b, c));
}

int ggp_generate(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                 const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                 const void* xi, int xi_dtype, void* values, int values_dtype, int precision, void* workspace,
                 size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_TRY(check_dtype(xi_dtype, "generate xi"));
  GGP_CHECK_ARG(values_dtype == (precision == GGP_PREC_FP64 ? GGP_F64 : GGP_F32),
                "generate: values dtype must be the compute dtype (FP64 for GGP_PREC_FP64, FP32 otherwise)");
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, indices, cov_bins, cov_vals, n_bins, precision, 1);
  RefineIO io{};
  io.xi = xi;
  io.xi_f64 = xi_dtype == GGP_F64;
  io.out = values;
  io.out_f64 = values_dtype == GGP_F64;
  return refine_run(p, 0, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_generate_sched(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                       const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                       const void* xi, int xi_dtype, void* values, int values_dtype, int precision,
                       const int32_t* schedule, const int32_t* schedule_neighbors, const void* schedule_points,
                       void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_TRY(check_dtype(xi_dtype, "generate xi"));
  GGP_CHECK_ARG(values_dtype == (precision == GGP_PREC_FP64 ? GGP_F64 : GGP_F32),
                "generate: values dtype must be the compute dtype (FP64 for GGP_PREC_FP64, FP32 otherwise)");
  GGP_CHECK_ARG((schedule == nullptr) == (schedule_neighbors == nullptr) &&
                    (schedule == nullptr) == (schedule_points == nullptr),
                "generate: schedule, schedule_neighbors and schedule_points go together");
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, indices, cov_bins, cov_vals, n_bins, precision, 1);
  p.sched = schedule;
  p.sched_nbrs = schedule_neighbors;
  p.sched_pts = schedule_points;
  RefineIO io{};
  io.xi = xi;
  io.xi_f64 = xi_dtype == GGP_F64;
  io.out = values;
  io.out_f64 = values_dtype == GGP_F64;
  return refine_run(p, 0, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

size_t ggp_graph_schedule_workspace(int64_t n, int64_t n0) { return graph_schedule_workspace(n - n0); }

int ggp_graph_schedule(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                       const int64_t* offsets, int64_t n_offsets, int32_t* schedule, int32_t* schedule_neighbors,
                       void* schedule_points, void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(pdtype, "graph_schedule points"));
  StageScope stage("schedule", S(stream));
  return graph_schedule(points, pdtype == GGP_F64 ? 1 : 0, n, d, neighbors, n0, k, offsets, n_offsets, schedule,
                        schedule_neighbors, schedule_points, workspace, workspace_bytes, S(stream));
}

size_t ggp_generate_vjp_workspace(int64_t n, int d, int64_t n0, int k, int n_bins, int precision) {
  RefineProblem p = problem(nullptr, 0, n, d, nullptr, n0, k, nullptr, nullptr, nullptr, n_bins, precision, 1);
  return vjp_workspace(p, t64_of(precision));
}

int ggp_generate_vjp(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                     const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                     const void* xi, int xi_dtype, const void* values, int values_dtype, const void* values_bar,
                     int values_bar_dtype, void* xi_bar, int xi_bar_dtype, double* cov_vals_bar, int precision,
                     void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_TRY(check_dtype(xi_dtype, "generate_vjp xi"));
  GGP_TRY(check_dtype(values_dtype, "generate_vjp values"));
  GGP_TRY(check_dtype(values_bar_dtype, "generate_vjp values_bar"));
  GGP_TRY(check_dtype(xi_bar_dtype, "generate_vjp xi_bar"));
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, indices, cov_bins, cov_vals, n_bins, precision, 1);
  VjpIO io{};
  io.xi = xi;
  io.xi_f64 = xi_dtype == GGP_F64;
  io.values = values;
  io.values_f64 = values_dtype == GGP_F64;
  io.gvalues = values_bar;
  io.g_f64 = values_bar_dtype == GGP_F64;
  io.gxi = xi_bar;
  io.gxi_f64 = xi_bar_dtype == GGP_F64;
  io.gtab = cov_vals_bar;
  return vjp_run(p, 0, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_generate_inv_vjp(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0,
                         int k, const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                         const void* values, int values_dtype, const void* xi_bar, int xi_bar_dtype,
                         void* values_bar, int values_bar_dtype, double* cov_vals_bar, int precision,
                         void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_TRY(check_dtype(values_dtype, "generate_inv_vjp values"));
  GGP_TRY(check_dtype(xi_bar_dtype, "generate_inv_vjp xi_bar"));
  GGP_TRY(check_dtype(values_bar_dtype, "generate_inv_vjp values_bar"));
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, indices, cov_bins, cov_vals, n_bins, precision, 1);
  VjpIO io{};
  io.values = values;
  io.values_f64 = values_dtype == GGP_F64;
  io.gvalues = xi_bar;
  io.g_f64 = xi_bar_dtype == GGP_F64;
  io.gxi = values_bar;
  io.gxi_f64 = values_bar_dtype == GGP_F64;
  io.gtab = cov_vals_bar;
  return vjp_run(p, 1, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_generate_logdet_vjp(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0,
                            int k, const double* cov_bins, const double* cov_vals, int n_bins,
                            const double* logdet_bar, double* cov_vals_bar, int precision, void* workspace,
                            size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_CHECK_ARG(logdet_bar != nullptr, "generate_logdet_vjp: logdet_bar is required");
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, nullptr, cov_bins, cov_vals, n_bins, precision, 1);
  VjpIO io{};
  io.gscalar = logdet_bar;
  io.gtab = cov_vals_bar;
  return vjp_run(p, 2, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_generate_inv(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                     const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                     const void* values, int values_dtype, void* xi, int xi_dtype, int precision, void* workspace,
                     size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_TRY(check_dtype(values_dtype, "generate_inv values"));
  GGP_TRY(check_dtype(xi_dtype, "generate_inv xi"));
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, indices, cov_bins, cov_vals, n_bins, precision, 1);
  RefineIO io{};
  io.values = values;
  io.values_f64 = values_dtype == GGP_F64;
  io.out = xi;
  io.out_f64 = xi_dtype == GGP_F64;
  return refine_run(p, 1, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_generate_logdet(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                        const double* cov_bins, const double* cov_vals, int n_bins, double* logdet_device,
                        int precision, void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, nullptr, cov_bins, cov_vals, n_bins, precision, 1);
  RefineIO io{};
  io.out = logdet_device;
  return refine_run(p, 2, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_refine(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
               const double* cov_bins, const double* cov_vals, int n_bins, const void* initial_values,
               int init_dtype, const void* xi, int xi_dtype, void* values, int precision, void* workspace,
               size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_TRY(check_dtype(init_dtype, "refine initial_values"));
  GGP_TRY(check_dtype(xi_dtype, "refine xi"));
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, nullptr, cov_bins, cov_vals, n_bins, precision, 0);
  RefineIO io{};
  io.xi = xi;
  io.xi_f64 = xi_dtype == GGP_F64;
  io.init = const_cast<void*>(initial_values);
  io.init_f64 = init_dtype == GGP_F64;
  io.out = values;
  io.out_f64 = t64_of(precision);
  return refine_run(p, 0, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_refine_inv(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                   const double* cov_bins, const double* cov_vals, int n_bins, const void* values, int values_dtype,
                   double* initial_values, void* xi, int xi_dtype, int precision, void* workspace,
                   size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  GGP_TRY(check_dtype(values_dtype, "refine_inv values"));
  GGP_TRY(check_dtype(xi_dtype, "refine_inv xi"));
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, nullptr, cov_bins, cov_vals, n_bins, precision, 0);
  RefineIO io{};
  io.values = values;
  io.values_f64 = values_dtype == GGP_F64;
  io.init = initial_values;
  io.out = xi;
  io.out_f64 = xi_dtype == GGP_F64;
  return refine_run(p, 1, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

int ggp_refine_logdet(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                      const double* cov_bins, const double* cov_vals, int n_bins, double* logdet_device,
                      int precision, void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_refine_args(n, n0, k, precision, pdtype));
  RefineProblem p = problem(points, pdtype, n, d, neighbors, n0, k, nullptr, cov_bins, cov_vals, n_bins, precision, 0);
  RefineIO io{};
  io.out = logdet_device;
  return refine_run(p, 2, t64_of(precision), io, workspace, workspace_bytes, S(stream));
}

size_t ggp_dense_workspace(int64_t n) { return dense_flow_workspace(n); }

int ggp_generate_dense(const void* points, int pdtype, int64_t n, int d, const double* cov_bins,
                       const double* cov_vals, int n_bins, const void* xi, int xi_dtype, void* values,
                       int values_dtype, void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(pdtype, "generate_dense points"));
  return dense_flow(points, pdtype == GGP_F64, d, n, cov_bins, cov_vals, n_bins, 0, xi, xi_dtype == GGP_F64, values,
                    values_dtype == GGP_F64, workspace, workspace_bytes, S(stream));
}

int ggp_generate_dense_inv(const void* points, int pdtype, int64_t n, int d, const double* cov_bins,
                           const double* cov_vals, int n_bins, const void* values, int values_dtype, void* xi,
                           int xi_dtype, void* workspace, size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(pdtype, "generate_dense_inv points"));
  return dense_flow(points, pdtype == GGP_F64, d, n, cov_bins, cov_vals, n_bins, 1, values, values_dtype == GGP_F64,
                    xi, xi_dtype == GGP_F64, workspace, workspace_bytes, S(stream));
}

int ggp_generate_dense_logdet(const void* points, int pdtype, int64_t n, int d, const double* cov_bins,
                              const double* cov_vals, int n_bins, double* logdet_device, void* workspace,
                              size_t workspace_bytes, void* stream) {
  GGP_TRY(check_dtype(pdtype, "generate_dense_logdet points"));
  return dense_flow(points, pdtype == GGP_F64, d, n, cov_bins, cov_vals, n_bins, 2, nullptr, 0, logdet_device, 1,
                    workspace, workspace_bytes, S(stream));
}

int ggp_cov_matrix(const void* points_a, int64_t na, const void* points_b, int64_t nb, int d, int pdtype,
                   const double* cov_bins, const double* cov_vals, int n_bins, double* out, void* stream) {
  GGP_TRY(check_dtype(pdtype, "compute_cov_matrix"));
  GGP_CHECK_ARG(n_bins >= 2 && d >= 1, "compute_cov_matrix: bad sizes");
  return cov_matrix(points_a, na, points_b, nb, d, pdtype == GGP_F64, cov_bins, cov_vals, n_bins, out, S(stream));
}

int ggp_refine_counters(const void* workspace, uint32_t* host_counters, void* stream) {
  return read_to_host(host_counters, workspace, 16, S(stream));
}

int ggp_profile_events(void* start_event, void* stop_event) {
  ggp::g_prof_start = static_cast<hipEvent_t>(start_event);
  ggp::g_prof_stop = static_cast<hipEvent_t>(stop_event);
  return 0;
}

}  // extern "C"
