/* graphgp_amd.h -- C ABI of the MI355X (gfx950) GraphGP hot path.
 *
 * One entry point per kernel family behind the reference's Python API
 * (stanford-ism/graphgp).  The reference crosses into its optional native extension
 * (`graphgp_cuda`) at nine call sites; each entry below names the call site it replaces.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers; sizes are int64_t; `stream` is a
 *    hipStream_t (NULL = default stream).  Work is enqueued on `stream` only; no entry
 *    point synchronises except where a host result is documented (marked "syncs").
 *  - No entry point allocates device memory.  Scratch comes from a caller-provided
 *    `workspace` of at least `*_workspace(...)` bytes (256-byte aligned is enough).
 *  - dtype codes: GGP_F32 = 0 (float32), GGP_F64 = 1 (float64).
 *  - Return value: 0 on success, negative on error; ggp_last_error() returns the
 *    message (per host thread).  Numerical failure (a non-positive-definite local
 *    covariance) is not an error: as in the reference (refine.py:55, :182) the whole
 *    output of generate / generate_inv becomes NaN.
 *  - Index arrays are int32 (N < 2^31).
 */
#ifndef GRAPHGP_AMD_H
#define GRAPHGP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GGP_F32 0
#define GGP_F64 1

/* Precision of the refinement arithmetic. */
#define GGP_PREC_FP64 0      /* FP64 local systems (the reference formulation)               */
#define GGP_PREC_FP32 1      /* FP32 local systems (the reference formulation)               */
#define GGP_PREC_FP32_INC 2  /* FP32 in increments relative to the nearest neighbour, with FP64
                                recomputation of the points whose FP32 pivots lose accuracy   */

const char* ggp_last_error(void);
int ggp_abi_version(void);

/* ---- k-d tree: replaces graphgp_cuda.build_tree(points), graphgp/tree.py:37 ------------
 * points (n, d) -> points_out (n, d) in tree order, split_dims (n), indices (n). */
size_t ggp_build_tree_workspace(int64_t n, int d, int dtype);
int ggp_build_tree(const void* points, int64_t n, int d, int dtype, void* points_out, int32_t* split_dims,
                   int32_t* indices, void* workspace, size_t workspace_bytes, void* stream);

/* ---- preceding kNN: replaces graphgp_cuda.query_preceding_neighbors(points, split_dims,
 * n0=, k=), graphgp/tree.py:61.  neighbors (n - n0, k). */
size_t ggp_query_neighbors_workspace(int64_t n, int d, int64_t n_queries, int k, int dtype);
int ggp_query_preceding_neighbors(const void* points, const int32_t* split_dims, int64_t n, int d, int dtype,
                                  int64_t n0, int k, int32_t* neighbors, void* workspace, size_t workspace_bytes,
                                  void* stream);

/* ---- kNN with explicit query / max indices: replaces graphgp_cuda.query_neighbors(points,
 * split_dims, query_indices, max_indices, k=), graphgp/tree.py:73.  neighbors (nq, k). */
int ggp_query_neighbors(const void* points, const int32_t* split_dims, int64_t n, int d, int dtype,
                        const int32_t* query_indices, const int32_t* max_indices, int64_t n_queries, int k,
                        int32_t* neighbors, void* workspace, size_t workspace_bytes, void* stream);

/* ---- depths: replaces graphgp_cuda.compute_depths_parallel(neighbors, n0=),
 * graphgp/graph.py:106.  depths (n0 + n_ref).  A malformed graph (a neighbour that does
 * not precede its point) is reported by ggp_compute_depths_status (syncs): status 1.
 * ggp_compute_depths computes them as a dataflow in index order (one launch, no host
 * synchronisation, a workspace of 64 bytes; for k > 64 it falls back to Kahn rounds when
 * the workspace holds ggp_compute_depths_workspace bytes, else to sweeps).
 * ggp_compute_depths_method (no reference counterpart) selects the method:
 * GGP_DEPTHS_FLOW, GGP_DEPTHS_ROUNDS (Kahn rounds over the transposed graph, workspace of
 * ggp_compute_depths_workspace) or GGP_DEPTHS_SWEEPS (in-place sweeps, 64 bytes; for
 * graphs whose transpose would not fit).  Every method gives the same depths. */
enum { GGP_DEPTHS_FLOW = 0, GGP_DEPTHS_ROUNDS = 1, GGP_DEPTHS_SWEEPS = 2 };
size_t ggp_compute_depths_workspace(int64_t n_ref, int k);
int ggp_compute_depths(const int32_t* neighbors, int64_t n_ref, int k, int64_t n0, int32_t* depths, void* workspace,
                       size_t workspace_bytes, void* stream);
int ggp_compute_depths_method(const int32_t* neighbors, int64_t n_ref, int k, int64_t n0, int32_t* depths,
                              int method, void* workspace, size_t workspace_bytes, void* stream);
int ggp_compute_depths_status(void* workspace, uint32_t* host_status, void* stream);

/* ---- depth order: replaces graphgp_cuda.order_by_depth(points, indices, neighbors,
 * depths), graphgp/graph.py:128.  Stable by depth; indices may be NULL (identity). */
size_t ggp_order_by_depth_workspace(int64_t n);
int ggp_order_by_depth(const void* points, int64_t n, int d, int dtype, const int32_t* indices,
                       const int32_t* neighbors, int64_t n_ref, int k, const int32_t* depths, int32_t max_depth,
                       void* points_out, int32_t* indices_out, int32_t* neighbors_out, int32_t* depths_out,
                       void* workspace, size_t workspace_bytes, void* stream);

/* ---- batch offsets of graphgp/graph.py:95-96: offsets[b] = #(depth <= b), b < n_offsets. */
int ggp_depth_offsets(const int32_t* depths_sorted, int64_t n, int64_t n_offsets, int64_t* offsets, void* stream);

/* ---- whole graph build: replaces graphgp_cuda.build_graph(points, n0=, k=),
 * graphgp/graph.py:89 (tree, preceding kNN, depths, depth order).  Outputs are in graph
 * (depth) order; *max_depth_host receives the largest depth (syncs once), from which
 * the caller sizes offsets (max_depth + 1 entries) for ggp_depth_offsets. */
size_t ggp_build_graph_workspace(int64_t n, int d, int dtype, int64_t n0, int k);
int ggp_build_graph(const void* points, int64_t n, int d, int dtype, int64_t n0, int k, void* points_out,
                    int32_t* indices_out, int32_t* neighbors_out, int32_t* depths_out, int32_t* max_depth_host,
                    void* workspace, size_t workspace_bytes, void* stream);

/* ---- graph validation of graphgp/graph.py:44-70 (check_graph).  *status_device receives
 * bit 0: malformed offsets, bit 1: a neighbour does not precede its point, bit 2: a
 * neighbour lies in its point's own batch. */
int ggp_check_graph(const int32_t* neighbors, int64_t n_ref, int k, int64_t n, const int64_t* offsets,
                    int64_t n_offsets, uint32_t* status_device, void* stream);

/* ---- refinement ------------------------------------------------------------------------
 * Shared arguments: points (n, d) in graph order (dtype pdtype), neighbors (n - n0, k),
 * cov_bins / cov_vals (n_bins) FP64, precision = GGP_PREC_*.  `indices` is the graph's
 * permutation (graph order -> caller order) or NULL.  Values / xi in caller order have
 * the dtype given by their own *_dtype argument.  The refined points' values come from a
 * single dataflow launch; `offsets` of the reference are therefore not an input. */
size_t ggp_refine_workspace(int64_t n, int d, int64_t n0, int k, int n_bins, int precision, int with_seed);

/* generate: replaces graphgp.generate (refine.py:21-56) including the dense seed block
 * (refine.py:59-76) and graphgp_cuda.refine (refine.py:119).  xi (n) -> values (n). */
int ggp_generate(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                 const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                 const void* xi, int xi_dtype, void* values, int values_dtype, int precision, void* workspace,
                 size_t workspace_bytes, void* stream);

/* Schedule of a graph's dataflow (no reference counterpart; an order of evaluation, not a
 * result): the refined points of every batch (offsets: the graph's batch ends, a device
 * int64 array of n_offsets = batches + 1 entries) ordered by the Morton code of their
 * points.  schedule (n - n0) holds refined indices; schedule_neighbors (n - n0, k) the
 * neighbours of schedule[q] as value slots (a seed keeps its index, refined point r moves
 * to n0 + its slot); schedule_points (n, d, dtype pdtype; may be NULL) the points in slot
 * order.  At most 65535 batches.  Computed once per graph. */
size_t ggp_graph_schedule_workspace(int64_t n, int64_t n0);
int ggp_graph_schedule(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                       const int64_t* offsets, int64_t n_offsets, int32_t* schedule, int32_t* schedule_neighbors,
                       void* schedule_points, void* workspace, size_t workspace_bytes, void* stream);

/* generate with a schedule (no reference counterpart; same results as ggp_generate): the
 * refinement runs in slot space (points and neighbours from schedule_points /
 * schedule_neighbors, values kept in schedule order, xi and output through the schedule),
 * so a tile's neighbours share cache lines.  In GGP_PREC_FP32_INC all three phases do; the
 * configurations beyond k + 1 = 40 (or d > 8) ignore the schedule. */
int ggp_generate_sched(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                       const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                       const void* xi, int xi_dtype, void* values, int values_dtype, int precision,
                       const int32_t* schedule, const int32_t* schedule_neighbors, const void* schedule_points,
                       void* workspace, size_t workspace_bytes, void* stream);

/* generate_inv: graphgp.generate_inv (refine.py:162-183), incl. graphgp_cuda.refine_inv
 * (refine.py:216) and generate_dense_inv (refine.py:186-195).  values (n) -> xi (n). */
int ggp_generate_inv(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                     const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                     const void* values, int values_dtype, void* xi, int xi_dtype, int precision, void* workspace,
                     size_t workspace_bytes, void* stream);

/* generate_logdet: graphgp.generate_logdet (refine.py:234-240), incl.
 * graphgp_cuda.refine_logdet (refine.py:265).  One FP64 scalar at logdet_device. */
int ggp_generate_logdet(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                        const double* cov_bins, const double* cov_vals, int n_bins, double* logdet_device,
                        int precision, void* workspace, size_t workspace_bytes, void* stream);

/* generate_vjp: reverse mode of generate -- what JAX autodiff of graphgp.generate
 * (refine.py:21-56, 79-159) returns for the cotangent values_bar (n, caller order):
 * xi_bar (n, caller order, dtype xi_bar_dtype) and cov_vals_bar (n_bins, FP64, the cotangent
 * of cov_vals).  xi and values are the forward pass's input and output.  precision
 * GGP_PREC_FP64 works in FP64; the FP32 modes in FP32 (FP64 seed block).  Any k and d (a
 * register kernel for d <= 3, k + 1 <= 32, else one lane per point); n0 <= 20480.  A
 * malformed graph (a wait that gives up) makes xi_bar and cov_vals_bar all NaN. */
size_t ggp_generate_vjp_workspace(int64_t n, int d, int64_t n0, int k, int n_bins, int precision);
int ggp_generate_vjp(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                     const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                     const void* xi, int xi_dtype, const void* values, int values_dtype, const void* values_bar,
                     int values_bar_dtype, void* xi_bar, int xi_bar_dtype, double* cov_vals_bar, int precision,
                     void* workspace, size_t workspace_bytes, void* stream);

/* generate_inv_vjp: reverse mode of generate_inv -- what JAX autodiff of graphgp.generate_inv
 * (refine.py:162-183, 186-195, 198-231) returns for the cotangent xi_bar (n, caller order):
 * values_bar (n, caller order) and cov_vals_bar (n_bins, FP64).  values is the inverse's
 * input.  Workspace: ggp_generate_vjp_workspace.  Precision as for ggp_generate_vjp. */
int ggp_generate_inv_vjp(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0,
                         int k, const int32_t* indices, const double* cov_bins, const double* cov_vals, int n_bins,
                         const void* values, int values_dtype, const void* xi_bar, int xi_bar_dtype,
                         void* values_bar, int values_bar_dtype, double* cov_vals_bar, int precision,
                         void* workspace, size_t workspace_bytes, void* stream);

/* generate_logdet_vjp: reverse mode of generate_logdet -- what JAX autodiff of
 * graphgp.generate_logdet (refine.py:234-240, 243-248, 251-275) returns for the cotangent
 * *logdet_bar (a device FP64 scalar): cov_vals_bar (n_bins, FP64).  The seed block's part
 * is (logdet_bar / 2) K^{-1} through its Cholesky factor (NaN when that factor fails).
 * Workspace: ggp_generate_vjp_workspace. */
int ggp_generate_logdet_vjp(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0,
                            int k, const double* cov_bins, const double* cov_vals, int n_bins,
                            const double* logdet_bar, double* cov_vals_bar, int precision, void* workspace,
                            size_t workspace_bytes, void* stream);

/* refine: graphgp_cuda.refine(points, neighbors, offsets, cov_bins, cov_vals, initial_values,
 * xi), refine.py:119.  initial_values (n0), xi (n - n0) -> values (n, graph order) with the
 * compute dtype (FP64 for GGP_PREC_FP64, FP32 otherwise). */
int ggp_refine(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
               const double* cov_bins, const double* cov_vals, int n_bins, const void* initial_values,
               int init_dtype, const void* xi, int xi_dtype, void* values, int precision, void* workspace,
               size_t workspace_bytes, void* stream);

/* refine_inv: graphgp_cuda.refine_inv(...), refine.py:216.  values (n, graph order) ->
 * initial_values (n0, FP64) and xi (n - n0, dtype xi_dtype). */
int ggp_refine_inv(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                   const double* cov_bins, const double* cov_vals, int n_bins, const void* values, int values_dtype,
                   double* initial_values, void* xi, int xi_dtype, int precision, void* workspace,
                   size_t workspace_bytes, void* stream);

/* refine_logdet: graphgp_cuda.refine_logdet(...), refine.py:265.  One FP64 scalar. */
int ggp_refine_logdet(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors, int64_t n0, int k,
                      const double* cov_bins, const double* cov_vals, int n_bins, double* logdet_device,
                      int precision, void* workspace, size_t workspace_bytes, void* stream);

/* ---- dense seed block (always FP64): generate_dense / generate_dense_inv /
 * generate_dense_logdet, refine.py:59-76, 186-195, 243-248. */
size_t ggp_dense_workspace(int64_t n);
int ggp_generate_dense(const void* points, int pdtype, int64_t n, int d, const double* cov_bins,
                       const double* cov_vals, int n_bins, const void* xi, int xi_dtype, void* values,
                       int values_dtype, void* workspace, size_t workspace_bytes, void* stream);
int ggp_generate_dense_inv(const void* points, int pdtype, int64_t n, int d, const double* cov_bins,
                           const double* cov_vals, int n_bins, const void* values, int values_dtype, void* xi,
                           int xi_dtype, void* workspace, size_t workspace_bytes, void* stream);
int ggp_generate_dense_logdet(const void* points, int pdtype, int64_t n, int d, const double* cov_bins,
                              const double* cov_vals, int n_bins, double* logdet_device, void* workspace,
                              size_t workspace_bytes, void* stream);

/* ---- covariance matrix of two point sets: graphgp.compute_cov_matrix(covariance, a, b),
 * refine.py:288-305 (r = ||a - b||, jnp.interp on the table).  out (na, nb) FP64. */
int ggp_cov_matrix(const void* points_a, int64_t na, const void* points_b, int64_t nb, int d, int pdtype,
                   const double* cov_bins, const double* cov_vals, int n_bins, double* out, void* stream);

/* ---- generate_logdet without a workspace (no counterpart in the reference; for callers that
 * cannot allocate device memory: code under HIP graph capture, serving and fitting loops,
 * harnesses that forbid hipMalloc).  Same arguments and result as ggp_generate_logdet, minus
 * the workspace: its scratch is a static device array of GGP_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends.  A problem whose log-det scratch exceeds that size
 * (ggp_refine_workspace's log-det share: about 8 MB for n0 = 1000 plus N/4 bytes) is
 * rejected with status -1.  Allocates nothing.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several streams or host threads) race. */
#define GGP_NOALLOC_SCRATCH_BYTES (64u << 20)
int ggp_generate_logdet_noalloc(const void* points, int pdtype, int64_t n, int d, const int32_t* neighbors,
                                int64_t n0, int k, const double* cov_bins, const double* cov_vals, int n_bins,
                                double* logdet_device, int precision, void* stream);

/* ---- counters of the last refinement flow that used `workspace` (syncs): [0] tiles issued,
 * [1] NaN produced, [2] malformed graph, [3] points recomputed in FP64 by the
 * GGP_PREC_FP32_INC mode (no reference counterpart). */
int ggp_refine_counters(const void* workspace, uint32_t* host_counters, void* stream);

/* ---- profiling hook (no reference counterpart): when both are non-NULL hipEvent_t,
 * every refinement flow on this host thread records start_event just before and
 * stop_event just after its refinement kernel, on the call's stream (for generate in
 * GGP_PREC_FP32_INC: just after the first, dominant kernel of its three).  Pass NULLs to
 * switch off. */
int ggp_profile_events(void* start_event, void* stop_event);

/* ---- stage timing (no reference counterpart; for benchmarks): ggp_stage_profile(1) switches
 * it on for the calling host thread and clears its sums; from then on the entry points record
 * a pair of events around each of their stages on the call's stream ("tree", "knn",
 * "depths", "order" in ggp_build_graph; "schedule"; "tables", "seed_block", "refine" and, in
 * GGP_PREC_FP32_INC generate, "weights", "fp64_list", "dataflow"; "nan_fill",
 * "logdet_sum").  ggp_stage_times waits for the recorded events and returns the number of
 * distinct stages seen, writing up to `capacity` of them: name (a static string), total
 * milliseconds and number of occurrences, in order of first appearance.
 * ggp_stage_profile(0) switches it off. */
int ggp_stage_profile(int enable);
int ggp_stage_times(int capacity, const char** names, double* total_ms, int* counts);

/* CU budget (no reference counterpart; for a serving loop that runs small and large
 * requests concurrently): the persistent refinement and adjoint launches, and the kNN
 * query (then a ticketed persistent launch), issued by the calling host thread use at most
 * `cus` CUs' worth of resident workgroups (0 = all CUs), leaving room for other streams'
 * kernels.  Results do not depend on it.  Returns the previous value. */
int ggp_set_cu_budget(int cus);

/* Yield (no reference counterpart; for a serving loop that reserves CUs for small requests
 * only while there are some): the persistent launches issued by the calling host thread
 * (refinement phases, and the kNN query, then ticketed) may fill the device, but while the
 * 32-bit word at `flag` is nonzero their workgroups beyond the first `keep_cus` CUs' worth
 * stop taking work and exit, leaving those slots to other streams; grid-stride passes
 * launched while it is nonzero stay within `keep_cus` CUs.  `flag` is pinned host memory
 * (hipHostMalloc) that the host writes; NULL switches the setting off.  Results do not
 * depend on it. */
int ggp_set_yield(const uint32_t* flag, int keep_cus);

#ifdef __cplusplus
}
#endif

#endif /* GRAPHGP_AMD_H */
