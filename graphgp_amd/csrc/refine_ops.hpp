// refine_ops.hpp -- host-side refinement flows (refine.hip).
#pragma once
#include "common.hpp"

namespace ggp {

struct RefineProblem {
  const void* pts;  // graph-order points (n, d)
  int pts_f64;
  int64_t n;
  int d;
  const int32_t* nbrs;  // (n_ref, k)
  int64_t n_ref;
  int k;
  const int32_t* perm;  // graph.indices or null
  const double* bins;
  const double* vals;
  int nb;
  int increment;  // FP32 increment formulation
  int with_seed;  // 1: generate/generate_inv/generate_logdet (dense seed block); 0: raw refine*
  // generate in the FP32 increment mode: the dataflow's schedule (ggp_graph_schedule) or null
  const int32_t* sched = nullptr;
  const int32_t* sched_nbrs = nullptr;
  const void* sched_pts = nullptr;  // points in slot order (dtype of pts)
};

struct RefineIO {
  const void* xi;  // generate: xi (caller order, N) or raw refine: xi (n_ref)
  int xi_f64;
  const void* values;  // inverse: values (caller order)
  int values_f64;
  void* init;  // raw refine: initial values in (n0); raw refine_inv: initial values out (n0, FP64)
  int init_f64;
  void* out;  // generate: values; inverse: xi; log-det: one FP64 scalar
  int out_f64;
};

struct VjpIO {
  const void* xi;  // generate: xi of the forward pass (caller order, N)
  int xi_f64;
  const void* values;  // generate: values of the forward pass; inverse: its input values (caller order, N)
  int values_f64;
  const void* gvalues;  // generate: cotangent of the values; inverse: cotangent of xi (caller order, N)
  int g_f64;
  const double* gscalar;  // log-det: cotangent of the log-det (device scalar)
  void* gxi;  // out -- generate: cotangent of xi; inverse: cotangent of the values (caller order, N)
  int gxi_f64;
  double* gtab;  // out: cotangent of the covariance table values (n_bins)
};

size_t vjp_workspace(const RefineProblem& p, int t64);
// mode: kGenerate (0), kInverse (1) or kLogdet (2) of refine_impl.hpp's RefineMode
int vjp_run(const RefineProblem& p, int mode, int t64, const VjpIO& io, void* ws, size_t ws_bytes, hipStream_t s);

int refine_points_per_tile(int t64, int kk, int d);
size_t refine_workspace(const RefineProblem& p, int mode, int t64);
int refine_run(const RefineProblem& p, int mode, int t64, const RefineIO& io, void* ws, size_t ws_bytes,
               hipStream_t s);
size_t dense_flow_workspace(int64_t n0);
int dense_flow(const void* pts, int pts_f64, int d, int64_t n0, const double* bins, const double* vals, int nb,
               int mode, const void* in, int in_f64, void* out, int out_f64, void* ws, size_t ws_bytes,
               hipStream_t s);

}  // namespace ggp
